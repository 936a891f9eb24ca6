/*
 * Oracle (test infrastructure only): batch driver over the per-sentence restatement in cky_ref.c.
 * Runs the reference's exhaustive Viterbi parse (tags, inside, viterbi decode) for many sentences,
 * parallel over sentences with OpenMP. bench.py times it as the CPU baseline; it is never the
 * product path.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct RefGramOpaque RefGramOpaque;
/* Declarations from cky_ref.c (the RefGram struct is passed through as an opaque pointer). */
void ref_tags(const void* g, int n, const int32_t* words, const int32_t* lex_off, const int32_t* lex_tag,
              const double* lex_logp, const float* in_scale, float* chart, double* partition_scale);
void ref_inside(const void* g, int n, float* chart, const uint32_t* masks, const float* in_scale);
void ref_viterbi(const void* g, int n, const float* in, const int32_t* offsets, const float* scores,
                 const int32_t* lefts, const int32_t* rights, int32_t* tree);

/* Returns the number of threads used. trees: 4 ints per node, sentence i at tree_off[i]. */
int ref_parse_viterbi_batch(const void* g, int cell_size, int num_sent, const int32_t* words,
                            const int32_t* offsets, const int32_t* lex_off, const int32_t* lex_tag,
                            const double* lex_logp, const int32_t* dec_off, const float* dec_scores,
                            const int32_t* dec_left, const int32_t* dec_right, int32_t* trees,
                            const int64_t* tree_off, int threads) {
  int used = 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
  }
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int i = 0; i < num_sent; ++i) {
    const int n = offsets[i + 1] - offsets[i];
    const size_t cells = (size_t)n * (n + 1);
    float* chart = (float*)malloc(cells * cell_size * sizeof(float));
    for (size_t k = 0; k < cells * cell_size; ++k) chart[k] = -INFINITY;
    ref_tags(g, n, words + offsets[i], lex_off, lex_tag, lex_logp, NULL, chart, NULL);
    ref_inside(g, n, chart, NULL, NULL);
    ref_viterbi(g, n, chart, dec_off, dec_scores, dec_left, dec_right, trees + 4 * tree_off[i]);
    free(chart);
  }
  return used;
}
