// Device kernel argument blocks and launchers shared by engine.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace puck {

constexpr int kMaxGroupFamilies = 8;
constexpr int kBlockThreads = 1024;   // rule kernels: 16 waves share one LDS item tile
constexpr int kLdsBudgetBytes = 150 * 1024;
// Operand rows of an 8-wave workgroup, two of which share a CU's 160 KB with their static LDS.
constexpr int kLdsHalfBudgetBytes = 77 * 1024;
constexpr int kRuleUnroll = 4;        // slot rule lists are padded to a multiple of this
constexpr int kGroupRules = 8;        // rules per pipelined group of the rule stream
// A group is stored as kGroupWords dwords: A-row words [0, 8), B-row byte offsets [8, 16), weight
// bits [16, 24). An A-row word holds the byte offset in bits [0, 18); the words of rules 3 and 7
// (the last rule of each half group) carry the half group's tag in bits [18, 32): kSlotEnd when
// a slot ends there, with the slot's target symbol in bits [18, 31). Streams that tiles may skip
// through (block-major layouts) use kSkipGroupWords: the same words plus, at [24] and [25], the
// id of the slot ending with each half group, and padding.
constexpr int kGroupWords = 24;
constexpr int kSkipGroupWords = 32;
constexpr uint32_t kOffsetMask = (1u << 18) - 1;
constexpr uint32_t kSlotEnd = 0x80000000u;
constexpr int kTagTargetShift = 18;
constexpr uint32_t kTagTargetMask = (1u << 13) - 1;  // targets of a family: at most 8,191
// A launch may cut every wave's rule stream into P = 1, 2, 4, ..., kMaxPieces pieces (grid z),
// at precomputed places no slot spans, so that few tiles still spread over the whole chip.
constexpr int kMaxPieces = 16;
constexpr int kLog2MaxPieces = 4;
constexpr int kPieceTableInts = 2 * (2 * kMaxPieces + kLog2MaxPieces);  // sum over P of 2 (P + 1)
constexpr int kReduceTargets = 16;    // reduce kernel: targets (waves) per workgroup
constexpr int kMinFoldChunks = 3;     // binary families with this many operand chunks fold partials
constexpr int kMbrMaxCell = 1024;     // largest cell (floats) the MBR kernel stages in LDS
constexpr int kMbrMaxSyms = 2048;     // largest grammar (symbols) whose MBR symbol lists it stages
constexpr int kMbrCellsPerBlock = 16; // cells per MBR workgroup (four per wave)

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
  // Slice item prefix (num_cells + 1; items of the slice are [prefix[0], prefix[num_cells])), null
  // = uniform count. With plan_items, the prefix is the mask plan's and items are listed there.
  const int32_t* prefix;
  const int32_t* plan_items;  // planned items of the span: (cell << 8) | split, null = all items
  const uint32_t* plan_act;   // planned binary families: active-block bits, 4 rows of num_cells words
  const int32_t* chunk_range;
  // Per (chunk, wave): num_block_ranges + 1 group starts, the wave's first slot id, then the
  // split tables: for P = 1, 2, 4, ..., P + 1 (group, first slot) pairs.
  const int32_t* wave_blocks;
  const uint32_t* rules;  // rule groups, kGroupWords dwords each
  const int32_t* block_of_target;
  // Partial rows: one per slot (row = slot id, plain stores), or with `fold` one per target symbol
  // (row = the slot's target, from its tag), into which every slot of that target folds its values
  // by an integer atomic max on order-preserving encodings (see enc_ordered); Viterbi only.
  float* partial;
  int32_t fold;
  int32_t num_cells, num_items, uniform_count, slice_cell0;
  int32_t num_block_ranges, sentinel_group;
  int32_t pieces_log2;  // the launch cuts streams into 2^pieces_log2 pieces (grid z)
  int32_t span, kind, range_split, cell_size;
  int32_t a_half, b_half, t_half;
  int32_t pitch;
  int32_t row_stride;  // floats per LDS row
  int32_t num_chunks;  // operand chunks of the family (grid y of the plain launch)
  int32_t pair_splits; // uniform NN: items 2m, 2m+1 are splits 2p, 2p+1 of one cell (pair m)
  int32_t skip_layout; // block-major rule stream (masked layers with several blocks)
};

// Arguments of the persistent rule kernel: the families of one span that share a layout.
constexpr int kMaxQueueFamilies = 4;
struct QueueArgs {
  FamArgs fam[kMaxQueueFamilies];
  int32_t* counter;  // queue head, zero before the launch
  int32_t nfam;
};

struct RedArgs {
  float* chT;
  const int32_t* cell_offsets;
  const int32_t* lengths;
  const int32_t* cell_list;
  const int32_t* off;
  const int32_t* fam;
  const int32_t* slot;  // partial row of each entry: a slot id, or for a folding family the target
  const float* partial[kMaxGroupFamilies];
  // Per family: 0x7fffffff when its partial rows fold (values are order-preserving encodings, see
  // FamArgs::fold), else 0.
  int32_t dmask[kMaxGroupFamilies];
  const int32_t* prefix[kMaxGroupFamilies];  // slice item prefix, null = uniform count
  int32_t uniform_count[kMaxGroupFamilies];
  int32_t pitch[kMaxGroupFamilies];
  // Masked layers: a (cell, target) pair takes a family's partials only when the target's block
  // is active in the cell (act: the plan's active-block bits, 4 rows of act_stride words, slice-
  // relative); inactive pairs were never computed. Zero results are not written (the chart half
  // was filled with the semiring zero).
  const uint32_t* act[kMaxGroupFamilies];  // null = unmasked
  const int32_t* block_of_target[kMaxGroupFamilies];  // null = block 0
  int32_t act_stride;
  // Planned groups visit cells in the order of the group's largest family (perm: position -> cell,
  // null = identity); every family finds a cell's items at its own position of the cell (posof).
  const int32_t* perm;
  const int32_t* posof[kMaxGroupFamilies];
  int32_t num_cells, cell_size, n_t, t_half, span;
  int32_t root_init, root;
  float one;
  int32_t nfam;  // families of the group (entries name families 0 .. nfam - 1)
};

// Start value of folded partial rows: rows [0, rows) x values [0, len) of each family's region get
// the encoded semiring zero, len = pre[num_cells] for planned slices (their item count is on the
// device) and `len` otherwise.
struct FoldFillArgs {
  float* base[kMaxGroupFamilies];
  const int32_t* pre[kMaxGroupFamilies];
  int32_t num_cells[kMaxGroupFamilies];
  int32_t len[kMaxGroupFamilies];
  int32_t rows[kMaxGroupFamilies];
  int32_t pitch[kMaxGroupFamilies];
  int32_t nfam;
};

// Device-scheduled pass (k_pass): every (reduce group, span) of one pass becomes two stages, its
// rule families then its reduce, in schedule order; one persistent launch pulls units (a rule tile x
// chunk x piece, or a reduce block of 64 cells x 16 targets) of all stages from one device queue in
// that order. Dependencies are tracked per 64-cell block of a stage's cells with device counters:
//   rc[b]  rule units of the group done that touched block b
//   dc[b]  reduce units of the group done on block b (complete at `gy`, one per 16 targets)
//   rtot, dtot  totals of the group's rule and reduce units
// A reduce unit of block b waits until rc[b] reaches the rule units that touch b (host-counted); a
// rule tile waits until the previous group's reduce blocks that hold its sentences' cells are
// complete (the previous group finished its chain of earlier spans for those sentences), and until
// the group two back has finished reducing (its half of the partial buffer is reused). Units only
// wait on units earlier in the queue, and every unit in the queue belongs to a running workgroup, so
// the waits always end; each is also bounded, and a wait that times out sets an error flag that
// makes every later wait pass at once (the host reports it).
constexpr int kMaxPassFamilies = 5;
constexpr int kPassSpinLimit = 1 << 20;  // polls of a counter before a wait gives up (about a second)
struct PassStage {
  int32_t kind;          // 0: rule families, 1: reduce
  int32_t first_unit, num_units;
  int32_t cnt, nb;       // this group's counters: rc [cnt, cnt + nb), dc [cnt + nb, cnt + 2 nb), rtot, dtot
  int32_t gy;            // reduce units per block (target groups)
  int32_t rule_units;    // units of the group's rule stage
  int32_t dep_cnt, dep_nb, dep_gy, dep_units;  // previous group's counters (dep_cnt < 0: none)
  int32_t war_cnt, war_nb, war_units;          // the group two back (war_cnt < 0: none)
  int32_t exp_off;       // reduce: expected rule units per block at expected[exp_off + b]; < 0: wait on rtot
  int32_t nfam;
  int32_t fam_units[kMaxPassFamilies];  // rule units before family f
  int32_t fam_tiles[kMaxPassFamilies];
  int32_t tile_base[kMaxPassFamilies];  // family f's first entry in the tile table
  FamArgs fam[kMaxPassFamilies];
  RedArgs red;
};
// Per rule tile: output blocks [x, y] (x < 0: none, the group's reduce waits on rtot) and the
// previous group's blocks it waits on [z, w] (z == -2: all of them, by dtot; z > w: none).
struct PassTile {
  int32_t out_lo, out_hi, dep_lo, dep_hi;
};
struct PassArgs {
  const PassStage* stages;
  const PassTile* tiles;
  const int32_t* expected;
  int32_t* counters;  // zero before the launch
  int32_t* queue;     // queue head, zero before the launch
  int32_t* error;     // set when a wait timed out
  int32_t num_stages, total_units;
};

// Mask plans. A plan job is one family of one partial-buffer slice of one (reduce group, span) of
// a masked pass: the items the family must compute under the pruning masks, in the family's own
// cell order. Each family orders its cells by a stable radix sort on which of its kPlanKeyBlocks
// largest blocks the cell allows (the largest block is the most significant bit), so that a tile's
// items share their active blocks and the rule kernel's walk over the union of its items' blocks
// wastes little. Every job of a layer's masked passes is planned by the same four launches (count,
// sort, scan, fill) on the parser's stream, the jobs spread over the grids.
constexpr int kPlanKeyBlocks = 16;
constexpr int kPlanThreads = 256;  // threads per block of the count and fill launches
// Per-family data of a layer that every job of the family shares.
struct PlanFamStatic {
  int32_t kind, unary, a_half, b_half, t_half, num_blocks;
  const uint32_t* block_masks;
  const double* block_rules;  // rules per block (binary) or the rule count (unary, one entry)
  int32_t key_blocks[kPlanKeyBlocks];  // the family's largest blocks, largest first (-1: none)
};
// One job. Its arrays lie back to back at `mp` ints into the plan buffer (kc = num_cells):
//   keys[kc]   sort key of each cell
//   perm[kc]   sorted position -> cell
//   posof[kc]  cell -> sorted position
//   counts[kc] items per cell
//   pre[kc+1]  item prefix in the family's cell order
//   act[4 kc]  active-block bits per cell, 4 rows of kc words
//   items[..]  planned items in the family's cell order, (cell << 8) | split
struct PlanJob {
  int32_t fid;          // family id within the layer
  int32_t num_cells;
  int32_t span;
  int32_t cells_start;  // first cell of the slice in the batch's cell list
  int32_t block0;       // first block of this job in the flattened count and fill grids
  int32_t apps_slot;    // plan counter: the rule applications the plan implies
  int64_t mp;
};
struct PlanBatchArgs {
  const PlanJob* jobs;
  const PlanFamStatic* fams;
  int32_t* mplan;
  double* apps;
  const uint32_t* masks;
  const int32_t* cell_offsets;
  const int32_t* lengths;
  const int32_t* cells;
  int32_t num_jobs, mask_words;
};

}  // namespace puck
