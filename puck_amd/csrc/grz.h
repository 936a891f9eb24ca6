// grammar.grz: compiled grammar layers in one zip container.
//
// The reference writes its compiled parser data with CLParserData.writeSequence / readSequence
// (CLParser.scala:1249-1317) into a zip through ZipUtil (ZipUtil.scala:13-98): an entry
// `parserNames` listing the layer directories `parser-i`, then per layer the entries `grammar`,
// `structure`, `semiring` and those of its kernels. Its payloads are Java-serialized objects and
// OpenCL binaries. This container keeps the zip layout and every entry name, with payloads of this
// library (little-endian binaries; int arrays are raw int32):
//   parserNames                        text, "parser-0" ... one per line, coarse layer first
//   parser-i/grammar                   symbols, coarse symbols, rule families with scores and
//                                      kernel literals, lexicon (field order in grz.cpp)
//   parser-i/structure                 CoarseParent block partitions
//   parser-i/semiring                  text: "Viterbi", "Real" or "LogSum"
//   parser-i/inside{NT,NN,TN,TT,NU,TU}/..., outside_{L,R}_{NN,NT,TN,TT}/..., outside{NU,TU}/...
//                                      rule-kernel bundles: numKernels (= blocks), per block j
//                                      `j/bits` (coarse mask), `j/rules` (rule indices), `j/wgSize`,
//                                      `j/entries` and the kernel `j/0`; binary bundles also
//                                      `enqueuer/...`, `extra`, `directWrite`
//   parser-i/setRootScoresKernel, getRootScoresKernel, computeMasksKernel + MasksInts,
//   computeMBRKernel + MBRInts, computeViterbiKernel + ViterbiWGSize, ViterbiParents, ViterbiLeft,
//   ViterbiRight, ViterbiScores, ViterbiOffsets, getScalingConstantsKernel
//                                      kernels as `<name>.name` (the HIP kernel that replaces the
//                                      reference's) and `<name>.program`; the Viterbi decoder tables
//                                      are read back from their entries
// Device code is not stored: it lives in libpuck.so, and the rule streams are rebuilt from the
// tables when a parser is made. Entries are stored uncompressed (zip method 0) with CRC-32 checks.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "grammar.h"

namespace puck {

void save_grz(const std::string& path, const std::vector<const Grammar*>& layers);
std::vector<std::shared_ptr<Grammar>> load_grz(const std::string& path);

}  // namespace puck
