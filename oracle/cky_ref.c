/*
 * Oracle (test infrastructure only, never linked into the product): plain-C restatement of the
 * reference's parse-path algorithms, one sentence at a time, evaluated in the order the reference
 * defines them. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Chart layout per sentence: [bottom half: n(n+1)/2 cells][top half], cell (b,e) at
 * chartIndex(b,e,n) = min(b,e) + s*n - s*(s-1)/2 with s = |e-b|-1 (ParseChart.scala:58-65),
 * cell_size floats per cell (Batch.scala:23,50-58).
 *
 * Restated functions:
 *   ref_tags       computeTagScores + writeTagScores        CLParser.scala:411-467
 *   ref_inside     ActualParser.inside + insideBinaryPass   CLParser.scala:256-314,489-504
 *                  rule kernel math                         SimpleGenRuleMultiply.java:151,213-216,304-314
 *                  block / span gating                      CLParser.scala:783-785,845-855,867,892-893
 *   ref_outside    ActualParser.outside + outsideBinaryPass CLParser.scala:316-369,506-522
 *   ref_masks      computeMasks kernel                      CLMaskKernels.scala:172-236
 *   ref_scaling    getScalingConstants kernel               CLScalingKernels.scala:51-69
 *   ref_viterbi    viterbi kernel (WG size 1)               CLViterbi.scala:364-605
 *   ref_mbr_cells  computeMBR kernel                        CLMBRKernels.scala:118-171
 *   ref_mbr_tree   extractMBRParses                         CLParser.scala:589-696
 * Semiring 0 = Viterbi (max, +), 1 = Real (+, *), 2 = LogSum (log-add, +) (RuleSemiring.scala:22-98);
 * LogSum's semiring_mad / semiring_add are restated from RuleSemiring.scala:58-73 (zero -1e6).
 * REAL is float by default; cky_ref64.c includes this file with REAL = double for the FP64 oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#define SUFFIX(x) x
#define REF_FMA fmaf
#define REF_EXP expf
#define REF_LOG logf
#endif

#include "cky_ref.h"

static inline int tri_(int n) { return n * (n + 1) / 2; }
static inline int cidx(int b, int e, int n) {
  int lo = b < e ? b : e;
  int s = (b < e ? e - b : b - e) - 1;
  return lo + s * n - s * (s - 1) / 2;
}
static inline int bot_cell(int b, int e, int n) { return cidx(b, e, n); }
static inline int top_cell(int b, int e, int n) { return tri_(n) + cidx(b, e, n); }

static int mask_any(const uint32_t* masks, int cell, int words) {
  if (!masks) return 1;
  for (int i = 0; i < words; ++i)
    if (masks[cell * words + i]) return 1;
  return 0;
}

/* Active blocks of a parent cell: determineBlocksToDo (CLParser.scala:845-855). With no masks every
 * rule applies. Non-final layers merge all kernels into one block keyed by the first kernel's mask. */
static int block_active(const RefGram* g, int which, int fam, const uint32_t* masks, int cell, int block) {
  if (!masks) return 1;
  const uint32_t* m = masks + (size_t)cell * g->mask_words;
  const uint32_t* bm = g->block_masks[which][fam] + (size_t)(g->final_layer ? block : 0) * g->mask_words;
  for (int i = 0; i < g->mask_words; ++i)
    if (m[i] & bm[i]) return 1;
  return 0;
}

static inline REAL zero_of(const RefGram* g) {
  return g->semiring == 0 ? (REAL)-INFINITY : (g->semiring == 2 ? (REAL)-1000000.0 : (REAL)0);
}

/* LogSum semiring_add (RuleSemiring.scala:66-72): larger + log(1 + exp(smaller - larger)). */
static inline REAL log_add(REAL x, REAL y) {
  const REAL lo = x < y ? x : y, hi = x < y ? y : x;
  return hi + (REAL)REF_LOG((REAL)1 + (REAL)REF_EXP(lo - hi));
}

/* Per-cell log scale; a missing scale array means every scale is 0 (PruningMask.scala:51-58). */
#define SC(arr, i) ((arr) ? (REAL)(arr)[i] : (REAL)0)

/* Tag scores of span-1 bottom cells. pos_scale receives log(sum/2) per position (Real). */
void SUFFIX(ref_tags)(const RefGram* g, int n, const int32_t* words, const int32_t* lex_off,
                      const int32_t* lex_tag, const double* lex_logp, const float* in_scale,
                      REAL* chart, double* partition_scale) {
  double scale_sum = 0.0;
  for (int pos = 0; pos < n; ++pos) {
    REAL* dst = chart + (size_t)bot_cell(pos, pos + 1, n) * g->cell_size;
    const int w = words[pos];
    if (g->semiring != 1) {
      for (int e = lex_off[w]; e < lex_off[w + 1]; ++e) dst[lex_tag[e]] = (REAL)(float)lex_logp[e];
    } else {
      const float my_scale = in_scale ? (float)exp(-(double)in_scale[bot_cell(pos, pos + 1, n)]) : 1.0f;
      REAL sum = 0;
      for (int e = lex_off[w]; e < lex_off[w + 1]; ++e) {
        REAL v = (REAL)(float)exp((double)(float)lex_logp[e]) * (REAL)my_scale;
        dst[lex_tag[e]] = v;
        sum = sum + v;
      }
      REAL half = sum / 2;
      for (int e = lex_off[w]; e < lex_off[w + 1]; ++e) dst[lex_tag[e]] = dst[lex_tag[e]] / half;
      scale_sum += log((double)half);
    }
  }
  if (partition_scale) *partition_scale = scale_sum;
}

/* One binary rule family applied to one parent cell over the given splits.
 * kind: 0 IN (A=(b,s), B=(s,e)), 1 OUT_L (A=(b,s), B=(e,s)), 2 OUT_R (A=(s,b), B=(s,e)). */
static void apply_binary(const RefGram* g, int which, int fam, int kind, REAL* tgt_chart, int tgt_cell,
                         const REAL* a_chart, int a_top, const REAL* b_chart, int b_top,
                         const float* t_scale, const float* a_scale, const float* b_scale,
                         int b, int e, int n, int s_lo, int s_hi, const uint32_t* masks,
                         int t_is_term, int a_mask_top, int b_mask_top, int t_mask_top) {
  const RefBin* R = &g->bin[fam];
  const int cs = g->cell_size;
  /* rotated roles: target, A, B per rule */
  const int32_t *tt, *aa, *bb;
  if (which == 0) { tt = R->p; aa = R->l; bb = R->r; }
  else if (which == 1) { tt = R->l; aa = R->p; bb = R->r; }
  else { tt = R->r; aa = R->l; bb = R->p; }
  const int tcell_mask = t_mask_top ? top_cell(b, e, n) : bot_cell(b, e, n);
  REAL* tgt = tgt_chart + (size_t)tgt_cell * cs;
  REAL* tmp = (REAL*)malloc(sizeof(REAL) * cs);
  for (int s = s_lo; s <= s_hi; ++s) {
    if (s < 0 || s > n) continue;
    int ax, ay, bx, by;
    if (kind == 0) { ax = b; ay = s; bx = s; by = e; }
    else if (kind == 1) { ax = b; ay = s; bx = e; by = s; }
    else { ax = s; ay = b; bx = s; by = e; }
    const int acell = (a_top ? tri_(n) : 0) + cidx(ax, ay, n);
    const int bcell = (b_top ? tri_(n) : 0) + cidx(bx, by, n);
    if (masks) {
      const int am = (a_mask_top ? tri_(n) : 0) + cidx(ax, ay, n);
      const int bm = (b_mask_top ? tri_(n) : 0) + cidx(bx, by, n);
      if (!mask_any(masks, am, g->mask_words) || !mask_any(masks, bm, g->mask_words)) continue;
    }
    const REAL* A = a_chart + (size_t)acell * cs;
    const REAL* B = b_chart + (size_t)bcell * cs;
    for (int i = 0; i < cs; ++i) tmp[i] = zero_of(g);
    for (int k = 0; k < R->n_b; ++k) {
      const int t = tt[k];
      const int coarse = t_is_term ? g->proj_t[t] : g->proj_n[t];
      const int blk = g->block_of_coarse[which][fam][coarse];
      if (!block_active(g, which, fam, masks, tcell_mask, blk)) continue;
      const REAL q = (REAL)R->q[k];
      if (g->semiring == 0) {
        REAL v = (A[aa[k]] + B[bb[k]]) + q;
        if (v > tmp[t]) tmp[t] = v;
      } else if (g->semiring == 2) {
        tmp[t] = log_add(tmp[t], (A[aa[k]] + B[bb[k]]) + q);
      } else {
        tmp[t] = REF_FMA(A[aa[k]] * B[bb[k]], q, tmp[t]);
      }
    }
    if (g->semiring == 0) {
      for (int i = 0; i < cs; ++i)
        if (tmp[i] > tgt[i]) tgt[i] = tmp[i];
    } else if (g->semiring == 2) {
      for (int i = 0; i < cs; ++i) tgt[i] = log_add(tgt[i], tmp[i]);
    } else {
      const REAL sc = (REAL)REF_EXP(-SC(t_scale, tgt_cell) + SC(b_scale, bcell) + SC(a_scale, acell));
      for (int i = 0; i < cs; ++i) tgt[i] += tmp[i] * sc;
    }
  }
  free(tmp);
}

/* Unary family on one cell: target cell overwritten (UnaryUpdateManager.flushQueue writes the whole
 * work row back, CLParser.scala:707-737), gated by the cell's top mask (:747). */
static void apply_unary(const RefGram* g, int which_rot, int fam, REAL* tgt_chart, int tgt_cell,
                        const REAL* src_chart, int src_cell, const float* t_scale, const float* s_scale,
                        const uint32_t* masks, int b, int e, int n, int overwrite) {
  const RefUn* U = &g->un[fam];
  const int cs = g->cell_size;
  if (masks && !mask_any(masks, top_cell(b, e, n), g->mask_words)) return;
  REAL* tgt = tgt_chart + (size_t)tgt_cell * cs;
  const REAL* src = src_chart + (size_t)src_cell * cs;
  REAL* tmp = (REAL*)malloc(sizeof(REAL) * cs);
  for (int i = 0; i < cs; ++i) tmp[i] = zero_of(g);
  for (int k = 0; k < U->n_u; ++k) {
    const int t = which_rot ? U->c[k] : U->p[k];
    const int a = which_rot ? U->p[k] : U->c[k];
    const REAL q = (REAL)U->q[k];
    if (g->semiring == 0) {
      REAL v = src[a] + q;
      if (v > tmp[t]) tmp[t] = v;
    } else if (g->semiring == 2) {
      tmp[t] = log_add(tmp[t], src[a] + q);
    } else {
      tmp[t] = REF_FMA(src[a], q, tmp[t]);
    }
  }
  const REAL sc = g->semiring != 1 ? (REAL)1 : (REAL)REF_EXP(-SC(t_scale, tgt_cell) + SC(s_scale, src_cell));
  for (int i = 0; i < cs; ++i) {
    REAL v = g->semiring != 1 ? tmp[i] : tmp[i] * sc;
    if (overwrite) tgt[i] = v;
    else if (g->semiring == 0) tgt[i] = v > tgt[i] ? v : tgt[i];
    else if (g->semiring == 2) tgt[i] = log_add(tgt[i], v);
    else tgt[i] = tgt[i] + v;
  }
  free(tmp);
}

/* inside: chart must already hold the tag scores in span-1 bottom cells and zero elsewhere. */
void SUFFIX(ref_inside)(const RefGram* g, int n, REAL* chart, const uint32_t* masks, const float* in_scale) {
  const int T = tri_(n);
  /* TU at span 1 */
  for (int b = 0; b < n; ++b)
    apply_unary(g, 0, 1, chart, T + cidx(b, b + 1, n), chart, cidx(b, b + 1, n), in_scale, in_scale, masks,
                b, b + 1, n, 1);
  for (int w = 2; w <= n; ++w) {
    for (int b = 0; b + w <= n; ++b) {
      const int e = b + w;
      const int tc = cidx(b, e, n);
      /* family order: TT (span 2), NT, TN, NN (CLParser.scala:495-502) */
      if (w == 2)
        apply_binary(g, 0, 3, 0, chart, tc, chart, 0, chart, 0, in_scale, in_scale, in_scale, b, e, n, b + 1, b + 1,
                     masks, 0, 0, 0, 0);
      apply_binary(g, 0, 1, 0, chart, tc, chart, 1, chart, 0, in_scale, in_scale, in_scale, b, e, n, e - 1, e - 1,
                   masks, 0, 1, 0, 0);
      apply_binary(g, 0, 2, 0, chart, tc, chart, 0, chart, 1, in_scale, in_scale, in_scale, b, e, n, b + 1, b + 1,
                   masks, 0, 0, 1, 0);
      apply_binary(g, 0, 0, 0, chart, tc, chart, 1, chart, 1, in_scale, in_scale, in_scale, b, e, n, b + 1, e - 1,
                   masks, 0, 1, 1, 0);
    }
    for (int b = 0; b + w <= n; ++b)
      apply_unary(g, 0, 0, chart, T + cidx(b, b + w, n), chart, cidx(b, b + w, n), in_scale, in_scale, masks, b,
                  b + w, n, 1);
  }
}

/* outside: `out` zero-filled by the caller; `in` is the finished inside chart. */
void SUFFIX(ref_outside)(const RefGram* g, int n, const REAL* in, REAL* out, const uint32_t* masks,
                         const float* in_scale, const float* out_scale) {
  const int T = tri_(n);
  const int cs = g->cell_size;
  /* setRootScores */
  out[(size_t)(T + cidx(0, n, n)) * cs + g->root] = g->semiring == 1 ? (REAL)1 : (REAL)0;
  /* NU at the root span. A one-word sentence gets TU there: the reference runs NU at the batch's
   * longest span and TU in its loop at span 1 (CLParser.scala:328-341), which is what a one-word
   * sentence receives in any batch that also holds a longer sentence. */
  apply_unary(g, 1, n == 1 ? 1 : 0, out, cidx(0, n, n), out, T + cidx(0, n, n), out_scale, out_scale, masks, 0, n,
              n, 1);
  for (int w = n - 1; w >= 1; --w) {
    for (int b = 0; b + w <= n; ++b) {
      const int e = b + w;
      const int tc = T + cidx(b, e, n);
      /* TN_R, NT_L, NN_L, NN_R (CLParser.scala:511-519), target = outside top */
      apply_binary(g, 2, 2, 2, out, tc, in, 0, out, 0, out_scale, in_scale, out_scale, b, e, n, b - 1, b - 1, masks,
                   0, 0, 0, 1);
      apply_binary(g, 1, 1, 1, out, tc, out, 0, in, 0, out_scale, out_scale, in_scale, b, e, n, e + 1, e + 1, masks,
                   0, 0, 0, 1);
      apply_binary(g, 1, 0, 1, out, tc, out, 0, in, 1, out_scale, out_scale, in_scale, b, e, n, e + 1, n, masks,
                   0, 0, 1, 1);
      apply_binary(g, 2, 0, 2, out, tc, in, 1, out, 0, out_scale, in_scale, out_scale, b, e, n, 0, b - 1, masks,
                   0, 1, 0, 1);
    }
    for (int b = 0; b + w <= n; ++b) {
      const int e = b + w;
      apply_unary(g, 1, w == 1 ? 1 : 0, out, cidx(b, e, n), out, T + cidx(b, e, n), out_scale, out_scale, masks, b, e,
                  n, 1);
    }
  }
  /* terminal families after the loop: TT_L, NT_R, TN_L, TT_R (CLParser.scala:344-355), target = outside bot */
  for (int b = 0; b < n; ++b) {
    const int e = b + 1;
    const int tc = cidx(b, e, n);
    apply_binary(g, 1, 3, 1, out, tc, out, 0, in, 0, out_scale, out_scale, in_scale, b, e, n, e + 1, e + 1, masks, 1,
                 0, 0, 0);
    apply_binary(g, 2, 1, 2, out, tc, in, 1, out, 0, out_scale, in_scale, out_scale, b, e, n, 0, b - 1, masks, 1, 1,
                 0, 0);
    apply_binary(g, 1, 2, 1, out, tc, out, 0, in, 1, out_scale, out_scale, in_scale, b, e, n, e + 1, n, masks, 1, 0,
                 1, 0);
    apply_binary(g, 2, 3, 2, out, tc, in, 0, out, 0, out_scale, in_scale, out_scale, b, e, n, b - 1, b - 1, masks, 1,
                 0, 0, 0);
  }
}

#ifndef REF_NO_DECODERS
/* computeMasks over every cell of one sentence; masks_out gets mask_words per cell. */
void ref_masks(const RefGram* g, int n, const float* in, const float* out, float thresh, uint32_t* masks_out) {
  const int cells = n * (n + 1);
  const int cs = g->cell_size;
  const float root_score = in[(size_t)(cells - 1) * cs + g->root];
  const float cutoff = root_score + thresh;
  for (int cell = 0; cell < cells; ++cell) {
    const int32_t* proj = cell < n ? g->proj_t : g->proj_n;
    const int nproj = cell < n ? g->nT : g->nN;
    for (int i = 0; i < g->mask_words; ++i) masks_out[cell * g->mask_words + i] = 0;
    for (int sym = 0; sym < g->num_syms; ++sym) {
      const float score = in[(size_t)cell * cs + sym] + out[(size_t)cell * cs + sym];
      const int field = sym < nproj ? proj[sym] : 0; /* projections padded with 0 */
      if (score >= cutoff) masks_out[cell * g->mask_words + field / 32] |= 1u << (field % 32);
    }
  }
}

void ref_scaling(const RefGram* g, int cells, const float* chart, float* scale) {
  for (int c = 0; c < cells; ++c) {
    float m = -100000.0f;
    for (int sym = 0; sym < g->cell_size; ++sym) {
      const float v = chart[(size_t)c * g->cell_size + sym];
      m = v > m ? v : (m != m ? v : m);
    }
    if (m < -9000.0f) m = 0;
    scale[c] = m;
  }
}

/* viterbi kernel with WG_SIZE 1: sequential scans, '>=' keeps the last best candidate.
 * dec_offsets: 6*nN+1, groups NN, NT, TN, TT, U, U_T. tree: 4 ints per node (top, bot, width, score bits). */
void ref_viterbi(const RefGram* g, int n, const float* in, const int32_t* offsets, const float* scores,
                 const int32_t* lefts, const int32_t* rights, int32_t* tree) {
  const int cs = g->cell_size, nN = g->nN;
  const float* ibot = in;
  const float* itop = in + (size_t)tri_(n) * cs;
#define T_TOP(p) tree[4 * (p)]
#define T_BOT(p) tree[4 * (p) + 1]
#define T_W(p) tree[4 * (p) + 2]
#define T_SC(p) (*(float*)&tree[4 * (p) + 3])
  for (int p = 0; p < 2 * n - 1; ++p) { T_TOP(p) = 0; T_BOT(p) = 0; T_W(p) = 0; T_SC(p) = 0.0f; }
  T_TOP(0) = g->root;
  T_W(0) = n;
  int begin = 0;
  for (int p = 0; p < 2 * n - 1; ++p) {
    const int width = T_W(p);
    const int end = begin + width;
    int botSym = T_BOT(p);
    if (T_TOP(p) != -1) {
      const int off = (width == 1) ? 5 * nN : 4 * nN;
      const int first = offsets[off + T_TOP(p)], last = offsets[off + T_TOP(p) + 1];
      const float* cell = ibot + (size_t)cidx(begin, end, n) * cs;
      float best = -30000.0f;
      int bestSym = 0;
      for (int r = first; r < last; ++r) {
        const int child = lefts[r];
        const float rScore = scores[r];
        const float cScore = cell[child];
        if (rScore + cScore >= best) {
          best = rScore + cScore;
          bestSym = child;
        }
      }
      T_SC(p) = best;
      botSym = bestSym;
      T_BOT(p) = botSym;
    }
    if (width == 1) {
      T_SC(p) = ibot[(size_t)cs * begin + botSym];
      begin += 1;
    } else {
      int il = 10, ir = 10, isplit = begin + 1;
      float iscore = -300000.0f;
#define BEST_BINARY(LC, RC, GOFF, SPLITV)                                      \
      do {                                                                      \
        const int f0 = offsets[(GOFF) + botSym], f1 = offsets[(GOFF) + botSym + 1]; \
        float bs = iscore;                                                      \
        for (int r = f0; r < f1; ++r) {                                         \
          const int lc = lefts[r], rc = rights[r];                              \
          const float rS = scores[r], lS = (LC)[lc], cS = (RC)[rc];             \
          if (rS + lS + cS >= bs) {                                             \
            bs = rS + lS + cS;                                                  \
            il = lc; ir = rc; isplit = (SPLITV); iscore = bs;                   \
          }                                                                     \
        }                                                                       \
      } while (0)
      BEST_BINARY(itop + (size_t)cidx(begin, end - 1, n) * cs, ibot + (size_t)cidx(end - 1, end, n) * cs, 1 * nN, -1);
      BEST_BINARY(ibot + (size_t)cidx(begin, begin + 1, n) * cs, itop + (size_t)cidx(begin + 1, end, n) * cs, 2 * nN, -2);
      if (width == 2)
        BEST_BINARY(ibot + (size_t)cidx(begin, begin + 1, n) * cs, ibot + (size_t)cidx(begin + 1, end, n) * cs, 3 * nN, -3);
      for (int split = begin + 1; split < end; ++split)
        BEST_BINARY(itop + (size_t)cidx(begin, split, n) * cs, itop + (size_t)cidx(split, end, n) * cs, 0, split);
#undef BEST_BINARY
      T_SC(p) = iscore;
      int bestSplit = isplit;
      const int bestConfig = bestSplit < 0 ? -bestSplit : 0;
      bestSplit = (bestConfig == 0) ? bestSplit : bestConfig == 2 ? begin + 1 : end - 1;
      const int lcp = p + 1, rcp = p + 2 * (bestSplit - begin);
      if ((bestConfig >> 1) == 0) {
        T_TOP(lcp) = il;
      } else {
        T_TOP(lcp) = -1;
        T_BOT(lcp) = il;
      }
      if ((bestConfig & 1) == 0) {
        T_TOP(rcp) = ir;
      } else {
        T_TOP(rcp) = -1;
        T_BOT(rcp) = ir;
      }
      T_W(lcp) = bestSplit - begin;
      T_W(rcp) = end - bestSplit;
    }
  }
#undef T_TOP
#undef T_BOT
#undef T_W
#undef T_SC
}

/* computeMBR per cell: out gets (score bits, symbol) pairs. */
void ref_mbr_cells(const RefGram* g, int n, const float* in, const float* out, int32_t* dec) {
  const int cells = n * (n + 1);
  const int cs = g->cell_size;
  const float root_score = in[(size_t)(cells - 1) * cs + g->root];
  float* margs = (float*)malloc(sizeof(float) * g->num_coarse);
  for (int cell = 0; cell < cells; ++cell) {
    const int32_t* proj = cell < n ? g->proj_t : g->proj_n;
    const int nproj = cell < n ? g->nT : g->nN;
    for (int c = 0; c < g->num_coarse; ++c) margs[c] = 0.0f;
    for (int sym = 0; sym < g->num_syms; ++sym) {
      const int c = sym < nproj ? proj[sym] : 0;
      if (g->semiring == 2) /* CLMBRKernels.scala:146-150 */
        margs[c] += expf(in[(size_t)cell * cs + sym] - root_score + out[(size_t)cell * cs + sym]);
      else
        margs[c] += in[(size_t)cell * cs + sym] / root_score * out[(size_t)cell * cs + sym];
    }
    float best = -INFINITY;
    int sym = -1;
    for (int c = 0; c < g->num_coarse; ++c)
      if (margs[c] > best) {
        best = margs[c];
        sym = c;
      }
    memcpy(&dec[2 * cell], &best, 4);
    dec[2 * cell + 1] = sym;
  }
  free(margs);
}

/* extractMBRParses: pre-order nodes (top coarse or -1, bot coarse, begin, end, score bits). */
int ref_mbr_tree(int n, const int32_t* dec, const float* in_scale, const float* out_scale, int has_scales,
                 int32_t* nodes) {
  const int W = n + 1, T = tri_(n);
  float* best = (float*)calloc((size_t)W * W, sizeof(float));
  int* split = (int*)calloc((size_t)W * W, sizeof(int));
  char* uu = (char*)calloc((size_t)W * W, 1);
  const float root_scale = in_scale[T + cidx(0, n, n)];
  for (int span = 1; span <= n; ++span)
    for (int b = 0; b + span <= n; ++b) {
      const int e = b + span;
      const int tc = T + cidx(b, e, n), bc = cidx(b, e, n);
      float top, bot;
      memcpy(&top, &dec[2 * tc], 4);
      memcpy(&bot, &dec[2 * bc], 4);
      if (has_scales) {
        top *= (float)exp((double)(in_scale[tc] + out_scale[tc] - root_scale));
        bot *= (float)exp((double)(in_scale[bc] + out_scale[bc] - root_scale));
      }
      if (top != top) top = 0.0f;
      if (bot != bot) bot = 0.0f;
      if (top > 1.0f) top = 1.0f;
      if (bot > 1.0f) bot = 1.0f;
      uu[b * W + e] = (b + 1 < e) || top > 0.4;
      float sc = uu[b * W + e] ? top + bot : bot;
      if (span > 1) {
        float bs = 0.0f;
        int bsp = b + 1;
        for (int s = b + 1; s < e; ++s) {
          const float v = best[b * W + s] + best[s * W + e];
          if (v > bs) {
            bsp = s;
            bs = v;
          }
        }
        split[b * W + e] = bsp;
        sc += bs;
      }
      best[b * W + e] = sc;
    }
  /* pre-order walk */
  int* stack = (int*)malloc(sizeof(int) * 2 * (2 * n + 2));
  int sp = 0, k = 0;
  stack[sp++] = 0;
  stack[sp++] = n;
  while (sp) {
    const int e = stack[--sp];
    const int b = stack[--sp];
    const int tc = T + cidx(b, e, n), bc = cidx(b, e, n);
    nodes[5 * k] = uu[b * W + e] ? dec[2 * tc + 1] : -1;
    nodes[5 * k + 1] = dec[2 * bc + 1];
    nodes[5 * k + 2] = b;
    nodes[5 * k + 3] = e;
    memcpy(&nodes[5 * k + 4], &best[b * W + e], 4);
    ++k;
    if (e - b > 1) {
      const int s = split[b * W + e];
      stack[sp++] = s;
      stack[sp++] = e;
      stack[sp++] = b;
      stack[sp++] = s;
    }
  }
  free(stack);
  free(best);
  free(split);
  free(uu);
  return k;
}
#endif
