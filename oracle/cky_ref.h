/*
 * Oracle (test infrastructure only): grammar tables shared by the per-sentence restatement
 * (cky_ref.c) and its batch drivers (cky_batch.c). Never part of the product.
 */
#ifndef CKY_REF_H_
#define CKY_REF_H_

#include <stdint.h>

typedef struct {
  int32_t n_b;
  const int32_t *p, *l, *r;
  const float *q, *w;
} RefBin;

typedef struct {
  int32_t n_u;
  const int32_t *p, *c;
  const float *q, *w;
} RefUn;

typedef struct {
  int32_t nN, nT, cell_size, num_syms, root, num_coarse, mask_words, semiring;
  RefBin bin[4];   /* NN, NT, TN, TT */
  RefUn un[2];     /* unary N child, unary T child */
  const int32_t *proj_n, *proj_t;
  /* Block gating: [which 0 inside, 1 out-left, 2 out-right][family]: block_of_coarse and masks. */
  int32_t num_blocks[3][4];
  const int32_t *block_of_coarse[3][4];
  const uint32_t *block_masks[3][4];
  int32_t final_layer;
} RefGram;

void ref_tags(const RefGram* g, int n, const int32_t* words, const int32_t* lex_off, const int32_t* lex_tag,
              const double* lex_logp, const float* in_scale, float* chart, double* partition_scale);
void ref_inside(const RefGram* g, int n, float* chart, const uint32_t* masks, const float* in_scale);
void ref_outside(const RefGram* g, int n, const float* in, float* out, const uint32_t* masks,
                 const float* in_scale, const float* out_scale);
void ref_masks(const RefGram* g, int n, const float* in, const float* out, float thresh, uint32_t* masks_out);
void ref_scaling(const RefGram* g, int cells, const float* chart, float* scale);
void ref_viterbi(const RefGram* g, int n, const float* in, const int32_t* offsets, const float* scores,
                 const int32_t* lefts, const int32_t* rights, int32_t* tree);
void ref_mbr_cells(const RefGram* g, int n, const float* in, const float* out, int32_t* dec);
int ref_mbr_tree(int n, const int32_t* dec, const float* in_scale, const float* out_scale, int has_scales,
                 int32_t* nodes);

#endif
