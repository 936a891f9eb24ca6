/*
 * Oracle (test infrastructure only): batch drivers over the per-sentence restatement in cky_ref.c,
 * parallel over sentences with OpenMP. Tests use them to check the HIP path on whole workloads;
 * bench.py times them as the CPU baseline. They are never the product path.
 *
 *   ref_parse_viterbi_batch  exhaustive Viterbi parse of one layer: tags, inside, viterbi decode
 *                            (CLParser.parse on a single-layer parser, CLParser.scala:145-158)
 *   ref_parse_batch          the coarse-to-fine fold of CLParser.parse / insideOutside over any
 *                            number of layers (CLParser.scala:50-67,145-169,208-250): every coarse
 *                            layer runs inside + outside under the masks so far and computes new
 *                            masks at -9 (plus scaling constants when the next layer is Real); the
 *                            final layer runs inside, then the Viterbi decode, or outside + MBR
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "cky_ref.h"

static int set_threads(int threads) {
  int used = 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
  }
#else
  (void)threads;
#endif
  return used;
}

static float zero_of_semiring(int semiring) {
  return semiring == 0 ? -INFINITY : (semiring == 2 ? -1000000.0f : 0.0f);
}

static float* new_chart(const RefGram* g, int n) {
  const size_t k = (size_t)n * (n + 1) * g->cell_size;
  float* c = (float*)malloc(k * sizeof(float));
  const float z = zero_of_semiring(g->semiring);
  for (size_t i = 0; i < k; ++i) c[i] = z;
  return c;
}

/* Returns the number of threads used. trees: 4 ints per node, sentence i at tree_off[i]. */
int ref_parse_viterbi_batch(const RefGram* g, int cell_size, int num_sent, const int32_t* words,
                            const int32_t* offsets, const int32_t* lex_off, const int32_t* lex_tag,
                            const double* lex_logp, const int32_t* dec_off, const float* dec_scores,
                            const int32_t* dec_left, const int32_t* dec_right, int32_t* trees,
                            const int64_t* tree_off, int threads) {
  (void)cell_size;
  const int used = set_threads(threads);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int i = 0; i < num_sent; ++i) {
    const int n = offsets[i + 1] - offsets[i];
    float* chart = new_chart(g, n);
    ref_tags(g, n, words + offsets[i], lex_off, lex_tag, lex_logp, NULL, chart, NULL);
    ref_inside(g, n, chart, NULL, NULL);
    ref_viterbi(g, n, chart, dec_off, dec_scores, dec_left, dec_right, trees + 4 * tree_off[i]);
    free(chart);
  }
  return used;
}

/* One layer of ref_parse_batch. */
typedef struct {
  const RefGram* g;
  const int32_t* lex_off;
  const int32_t* lex_tag;
  const double* lex_logp;
} RefLayer;

/*
 * Coarse-to-fine parse of num_sent sentences through nl layers (the last one decodes).
 *   masks_out   (nullable) masks the last coarse layer computes: mask_words per cell, sentence i at
 *               cell_off[i] cells (n(n+1) cells per sentence)
 *   trees       Viterbi final layer: 4 ints per node, sentence i at tree_off[i]
 *   mbr         Real / LogSum final layer: 5 ints per node (top, bot, begin, end, score bits)
 *   partitions  sentence log-likelihood (CLParser.scala:189-206: Real = log(root) + top scale +
 *               partition scale; Viterbi / LogSum = the root value)
 *   with_outside  a Viterbi final layer also runs its outside pass (insideOutside,
 *               CLParser.scala:171-182); the trees do not depend on it
 * Returns the number of threads used.
 */
int ref_parse_batch(const RefLayer* layers, int nl, int num_sent, const int32_t* words, const int32_t* offsets,
                    const int32_t* dec_off, const float* dec_scores, const int32_t* dec_left,
                    const int32_t* dec_right, uint32_t* masks_out, const int64_t* cell_off, int32_t* trees,
                    int32_t* mbr, const int64_t* tree_off, float* partitions, int with_outside, int threads) {
  const int used = set_threads(threads);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int i = 0; i < num_sent; ++i) {
    const int n = offsets[i + 1] - offsets[i];
    const int cells = n * (n + 1);
    const int32_t* w = words + offsets[i];
    uint32_t* masks = NULL;
    float* isc = NULL;
    float* osc = NULL;
    int has_scales = 0;
    for (int li = 0; li < nl; ++li) {
      const RefGram* g = layers[li].g;
      const int last = li == nl - 1;
      float* in = new_chart(g, n);
      double ps = 0.0;
      ref_tags(g, n, w, layers[li].lex_off, layers[li].lex_tag, layers[li].lex_logp, has_scales ? isc : NULL, in,
               &ps);
      ref_inside(g, n, in, masks, has_scales ? isc : NULL);
      if (!last) {
        float* out = new_chart(g, n);
        ref_outside(g, n, in, out, masks, NULL, NULL);
        uint32_t* nm = (uint32_t*)calloc((size_t)cells * g->mask_words, sizeof(uint32_t));
        ref_masks(g, n, in, out, -9.0f, nm);
        free(masks);
        masks = nm;
        free(isc);
        free(osc);
        isc = osc = NULL;
        has_scales = layers[li + 1].g->semiring == 1;
        if (has_scales) {
          isc = (float*)malloc(sizeof(float) * cells);
          osc = (float*)malloc(sizeof(float) * cells);
          ref_scaling(g, cells, in, isc);
          ref_scaling(g, cells, out, osc);
        }
        if (li == nl - 2 && masks_out)
          memcpy(masks_out + (size_t)cell_off[i] * g->mask_words, masks, sizeof(uint32_t) * cells * g->mask_words);
        free(out);
        free(in);
        continue;
      }
      const int top_root = n * (n + 1) / 2 + (n - 1) * n - (n - 1) * (n - 2) / 2;
      const float root = in[(size_t)top_root * g->cell_size + g->root];
      if (g->semiring == 1) {
        const float top = has_scales ? isc[top_root] : 0.0f;
        if (partitions) partitions[i] = ((float)log((double)root) + top) + (float)ps;
      } else if (partitions) {
        partitions[i] = root;
      }
      if (g->semiring == 0) {
        if (with_outside) {
          float* out = new_chart(g, n);
          ref_outside(g, n, in, out, masks, NULL, NULL);
          free(out);
        }
        if (trees) ref_viterbi(g, n, in, dec_off, dec_scores, dec_left, dec_right, trees + 4 * tree_off[i]);
      } else {
        float* zs = (float*)calloc(cells, sizeof(float));
        const float* si = has_scales ? isc : zs;
        const float* so = has_scales ? osc : zs;
        float* out = new_chart(g, n);
        ref_outside(g, n, in, out, masks, si, so);
        int32_t* dec = (int32_t*)malloc(sizeof(int32_t) * 2 * cells);
        ref_mbr_cells(g, n, in, out, dec);
        if (mbr) ref_mbr_tree(n, dec, si, so, g->semiring == 1 && has_scales, mbr + 5 * tree_off[i]);
        free(dec);
        free(out);
        free(zs);
      }
      free(in);
    }
    free(masks);
    free(isc);
    free(osc);
  }
  return used;
}
