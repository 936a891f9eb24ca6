// Batched parse engine: thousands of sentences of uneven length, bucketed by length and run as
// concurrent batches on several HIP streams.
//
// The reference parses one batch at a time: WorkSpace.getBatches closes a batch when its chart
// cells exceed the workspace (WorkSpace.scala:20-41), in input order, and CLParser.parse runs the
// batches one after the other on one OpenCL queue (CLParser.scala:50-53,145-206). Here sentences
// are sorted by length and cut into buckets of a tunable chart-cell budget, so a bucket's span
// loop stops at its own longest sentence; buckets are dealt longest-processing-time-first to
// `streams` workers, each owning a Parser (stream + workspace) with persistent rule launches and
// per-shape HIP graphs, so that one bucket's short late-span launches overlap another's work.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "engine.h"

namespace puck {

struct BatchConfig {
  int32_t streams = 3;            // concurrent buckets (one Parser and HIP stream each)
  int64_t bucket_cells = 1 << 20; // chart cells per bucket (sum of n(n+1) over its sentences)
  bool persistent = true;         // persistent rule launches (else one launch per family and span)
  bool graphs = true;             // capture each bucket shape's device work into a HIP graph
  int32_t piece_target = 512;
  int32_t items_per_lane = 2;
  bool outside_final = false;     // Viterbi final layer: outside pass before the decode (insideOutside)
  bool device_schedule = false;   // one persistent launch per pass with device dependency counters
  int32_t reduce_targets = 4;     // chart targets (waves) per reduce workgroup: 1, 2, 4, 8 or 16
  bool fold = true;               // Viterbi layers fold partial rows per target (Parser::set_fold)
};

struct BatchStats {
  double device_ms = 0;  // from the first bucket's first kernel to the last bucket's last one
  int64_t buckets = 0;
  int64_t graph_captures = 0, graph_replays = 0;
};

class BatchParser {
 public:
  BatchParser(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
              const BatchConfig& cfg);
  ~BatchParser();
  BatchParser(const BatchParser&) = delete;
  BatchParser& operator=(const BatchParser&) = delete;

  // Same results as Parser::parse on the same input, in input order.
  void parse(const int32_t* words, const int32_t* offsets, int32_t n,
             std::vector<std::vector<TreeNode>>* viterbi, std::vector<std::vector<MbrNode>>* mbr);
  BatchStats stats;
  const BatchConfig& config() const { return cfg_; }

 private:
  BatchConfig cfg_;
  int32_t device_;
  std::vector<std::unique_ptr<Parser>> parsers_;
  hipEvent_t ref_ = nullptr;
};

}  // namespace puck
