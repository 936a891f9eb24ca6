// Seeded inputs for the CKY engine: a synthetic grammar with the published wsj_6 shape and PCFG
// sentence sampling.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "grammar.h"

namespace puck {

// Portable PRNG (splitmix64 stream) so that seeded inputs are identical on every machine.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal();
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

// Writes out_prefix.{numstates,binary,unary,lexicon}: symbols and refined rule counts of the wsj_6
// shape files in data_dir (wsj_6.gr.numstates / .num.binary), coarse rules and probabilities
// derived from wsj_0 / wsj_1, seeded perturbation. Returns a one-line summary.
std::string synth_wsj6(const std::string& data_dir, const std::string& out_prefix, uint64_t seed);

// Length modes for sample_sentences.
enum LengthMode : int32_t {
  kLenRange = 0,      // PCFG samples rejection-filtered to [min_len, max_len]
  kLenLogNormal = 1,  // target length ~ round(exp(N(3.049, 0.416))) truncated to [min_len, max_len]
  kLenMixed = 2,      // 70% kLenLogNormal over [1, 40], 30% uniform over [41, 80]
};

// Top-down PCFG samples from ROOT_0 (identity unaries excluded), words drawn from P(w|t).
// Returns false if the sampler could not produce n sentences within its attempt budget.
bool sample_sentences(const Grammar& g, uint64_t seed, int32_t n, int32_t mode, int32_t min_len,
                      int32_t max_len, std::vector<int32_t>* words, std::vector<int32_t>* offsets);

}  // namespace puck
