// Device kernel argument blocks and launchers shared by engine.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace puck {

constexpr int kMaxGroupFamilies = 8;
constexpr int kBlockThreads = 1024;   // rule kernels: 16 waves share one LDS item tile
constexpr int kLdsBudgetBytes = 150 * 1024;
constexpr int kRuleUnroll = 4;        // rules per pipelined group; slot rule lists are padded to it
constexpr uint32_t kSlotEnd = 0x80000000u;  // flag in the last rule record of a slot
constexpr int kReduceTargets = 16;    // reduce kernel: targets (waves) per workgroup

// Arguments of the rule-family kernel (one launch = one family, one span, one cell slice).
struct FamArgs {
  const float* chA;   // chart buffer holding operand A (inside or outside base pointer)
  const float* chB;
  const float* scT;   // per-cell log scales of the target / operand charts (Real semiring)
  const float* scA;
  const float* scB;
  const uint32_t* masks;  // coarse pruning masks (maskWords per cell), null = no pruning
  const int32_t* cell_offsets;
  const int32_t* lengths;
  const int32_t* cell_list;  // slice cells, packed (sentence << 8) | begin
  const int32_t* prefix;     // slice item prefix (num_cells + 1), null = uniform count
  const int32_t* chunk_range;
  const int4* wave_work;  // per (chunk, wave): {first slot, first rule, end rule, 0}
  const uint4* rules;  // {A row byte offset, B row byte offset, bits(q), kSlotEnd | block or 0}
  const int32_t* block_of_target;
  const uint32_t* block_masks;
  float* partial;
  int32_t num_cells, item_base, num_items, uniform_count;
  int32_t span, kind, range_split, cell_size;
  int32_t a_half, b_half, t_half;
  int32_t mask_words, num_blocks, unary_gate;
  int32_t pitch;
  int32_t row_stride;  // floats per LDS row
};

struct RedArgs {
  float* chT;
  const int32_t* cell_offsets;
  const int32_t* lengths;
  const int32_t* cell_list;
  const int32_t* off;
  const int32_t* fam;
  const int32_t* slot;
  const float* partial[kMaxGroupFamilies];
  const int32_t* prefix[kMaxGroupFamilies];
  int32_t uniform_count[kMaxGroupFamilies];
  int32_t base[kMaxGroupFamilies];
  int32_t pitch[kMaxGroupFamilies];
  int32_t num_cells, cell_size, n_t, t_half, span;
  int32_t root_init, root;
  float one;
};

}  // namespace puck
