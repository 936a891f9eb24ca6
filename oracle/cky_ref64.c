/*
 * Oracle (test infrastructure only): the inside/outside restatement of cky_ref.c evaluated in
 * double precision. It takes the same fp32 rule literals and tag scores and measures how far the
 * fp32 pipeline drifts from the same arithmetic carried out in FP64.
 */
#define REAL double
#define SUFFIX(x) x##_f64
#define REF_FMA fma
#define REF_EXP exp
#define REF_LOG log
#define REF_NO_DECODERS
#include "cky_ref.c"
