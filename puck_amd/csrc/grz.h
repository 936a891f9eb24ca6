// grammar.grz: compiled grammar layers in one zip container.
//
// The reference writes its compiled parser data with CLParserData.writeSequence / readSequence
// (CLParser.scala:1249-1317) into a zip through ZipUtil (ZipUtil.scala:13-98): an entry
// `parserNames`, then per layer i the entries `parser-i/grammar`, `parser-i/structure`,
// `parser-i/semiring` and the generated kernels. Its payloads are Java-serialized objects and
// OpenCL binaries. This container keeps the zip layout and the entry names of the grammar data and
// replaces the payloads with documented little-endian binaries of the compiled tables:
//   parserNames           text, one layer name (its text grammar prefix) per line
//   parser-i/grammar      symbols, coarse symbols, rule families with scores and kernel literals,
//                         lexicon (see grz.cpp for the field order)
//   parser-i/structure    CoarseParent block partitions and the Viterbi decoder tables
//   parser-i/semiring     text: "Viterbi", "Real" or "LogSum"
// Device code is not stored: the rule streams are rebuilt from the tables when a parser is made.
// Entries are stored uncompressed (zip method 0) with CRC-32 checks.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "grammar.h"

namespace puck {

void save_grz(const std::string& path, const std::vector<const Grammar*>& layers);
std::vector<std::shared_ptr<Grammar>> load_grz(const std::string& path);

}  // namespace puck
