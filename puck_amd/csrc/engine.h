// CKY engine: device tables for compiled grammar layers and the parse pipeline over batches.
//
// Host-side counterpart of CLParser / ActualParser (CLParser.scala:46-952) and WorkSpace
// (WorkSpace.scala:15-123). The schedule (inside span loop, outside span loop, families per span,
// coarse-to-fine fold over layers) follows CLParser.scala:50-67,256-369,489-522.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <memory>
#include <string>
#include <vector>

#include "grammar.h"
#include "kernels.h"

namespace puck {
constexpr const char* kDeviceScheduleDisabled =
    "device-scheduled passes are disabled: their one launch per pass did not finish on the GPU (cause not "
    "found); parse with per-span launches";
}

namespace puck {

// One node of a Viterbi tree, pre-order, 2n-1 nodes per sentence (CLViterbi.scala:378).
struct TreeNode {
  int32_t top_sym;  // nonterminal id of the unary top, -1 if the node has no unary layer
  int32_t bot_sym;  // nonterminal id (width > 1) or terminal id (width == 1)
  int32_t width;
  float score;
};

// Per-cell MBR decode record (CLMBRKernels.scala:102: {score, symbol, _, _}).
struct MbrCell {
  float score;
  int32_t symbol;  // coarse id
};

// One node of an MBR tree, pre-order: coarse labels, -1 top when the unary layer is skipped.
struct MbrNode {
  int32_t top_coarse;
  int32_t bot_coarse;
  int32_t begin, end;
  float score;
};

enum ChartHalf : int32_t { kInBot = 0, kInTop = 1, kOutBot = 2, kOutTop = 3 };
enum SplitKind : int32_t { kSkIn = 0, kSkOutL = 1, kSkOutR = 2, kSkUnary = 3 };

// Family-directions of the rule passes (inside families, rotated outside families, unaries).
enum FamilyId : int32_t {
  kFInNN = 0, kFInNT, kFInTN, kFInTT,
  kFOutLNN, kFOutLNT, kFOutLTN, kFOutLTT,
  kFOutRNN, kFOutRNT, kFOutRTN, kFOutRTT,
  kFInUN, kFInUT, kFOutUN, kFOutUT,
  kNumFamilies
};

struct DevFamily {
  int32_t id = 0;
  int32_t kind = 0;          // SplitKind
  int32_t range_split = 0;   // 1: split ranges over an interval, 0: one split
  int32_t t_half = 0, a_half = 0, b_half = 0;  // ChartHalf of target and operands
  int32_t n_t = 0, n_a = 0, n_b = 0;           // symbol space sizes
  bool unary = false;
  // Binary families with at least kMinFoldChunks operand chunks fold their partial rows per target
  // when the layer folds; unary families and binary ones with few chunks keep one row per slot.
  bool foldable = false;
  int32_t num_chunks = 0, num_slots = 0, num_rules = 0, max_rows = 0;
  std::vector<int32_t> h_chunk_range;       // 4 per chunk: a_lo, a_hi, b_lo, b_hi
  std::vector<int32_t> h_chunk_slot_start;  // num_chunks + 1
  std::vector<int32_t> h_slot_target;
  int32_t* d_chunk_range = nullptr;
  int32_t* d_slot_target = nullptr;
  std::vector<int32_t> h_slot_block;
  int32_t num_block_ranges = 1;  // block ranges per wave stream (blocks in block-major layouts)
  int32_t sentinel_group = 0;    // index of the null group after the last wave's range
  bool skip_layout = false;      // block-major stream of kSkipGroupWords groups
  int32_t* d_wave_blocks = nullptr;  // per (chunk, wave): num_block_ranges + 1 group starts, first slot
  int32_t* d_slot_block = nullptr;   // block bit of each slot
  uint32_t* d_rules = nullptr;  // rule groups (kGroupWords dwords each)
  // Block gating (CLParser.scala:783-785,845-855): block of each target symbol and block masks.
  int32_t num_blocks = 0;
  int32_t* d_block_of_target = nullptr;
  uint32_t* d_block_masks = nullptr;
  double rule_apps_per_item = 0;  // rules in the family (for roofline accounting)
  double* d_block_rules = nullptr;  // rules per block (per the whole family for unaries)
  std::vector<int32_t> h_key_blocks;  // the blocks with most rules, most first (mask plan sort key)
};

// Reduction group: families whose partials are combined into one target chart half.
struct DevReduce {
  std::vector<int32_t> families;  // FamilyId list
  int32_t n_t = 0;
  int32_t* d_off = nullptr;   // n_t + 1
  int32_t* d_fam = nullptr;   // index into families
  int32_t* d_slot = nullptr;  // slot within that family
  // The same when the layer folds: a foldable family has one entry per target it feeds (row = the
  // target), the others one per slot as above.
  int32_t* d_off_f = nullptr;
  int32_t* d_fam_f = nullptr;
  int32_t* d_slot_f = nullptr;
};

enum ReduceId : int32_t { kRInBin = 0, kRInBin2, kRInUN, kRInUT, kROutBin, kROutUN, kROutW1, kNumReduce };

class Layer {
 public:
  // masked: the layer runs under the pruning masks of the layer before it (every layer but the
  // first); its binary rule streams are then laid out block-major.
  Layer(Grammar g, bool final_layer, bool masked, int32_t items_per_lane);
  ~Layer();
  Layer(const Layer&) = delete;
  Layer& operator=(const Layer&) = delete;

  Grammar g;
  bool final_layer;
  bool masked;
  int32_t ipl;  // items per lane of the rule kernels (1 or 2)
  // Waves per rule workgroup: 16, or 8 for unmasked layers whose every family's operand rows fit
  // one chunk in half the LDS (two workgroups per CU then overlap one's unit prologue with the
  // other's rules: the wsj_1 shape).
  int32_t waves = 16;
  int32_t lds_budget = 0;  // bytes of operand rows per workgroup
  DevFamily fam[kNumFamilies];
  DevReduce red[kNumReduce];
  PlanFamStatic* d_plan_static = nullptr;  // mask plan data of every family (kNumFamilies)
  // Decoder tables.
  int32_t* d_dec_offsets = nullptr;
  float* d_dec_scores = nullptr;
  int32_t* d_dec_left = nullptr;
  int32_t* d_dec_right = nullptr;
  // Lexicon.
  int32_t* d_lex_offsets = nullptr;
  int32_t* d_lex_tags = nullptr;
  double* d_lex_logp = nullptr;
  int32_t* d_proj_n = nullptr;
  int32_t* d_proj_t = nullptr;
  // MBR: per coarse symbol the symbols projecting to it (offsets num_coarse + 1, symbol ids), for
  // the terminal (span-1 bottom) and nonterminal projections.
  int32_t *d_mbr_t_off = nullptr, *d_mbr_t_syms = nullptr, *d_mbr_n_off = nullptr, *d_mbr_n_syms = nullptr;

 private:
  void build_family(int32_t id);
  void build_reduce(int32_t rid, std::vector<int32_t> fams, int32_t n_t);
  std::vector<void*> owned_;
  template <class T> T* upload(const std::vector<T>& v);
};

// Timing and work counters of the last call (HIP events on the engine's stream).
struct PassStats {
  double rule_kernel_ms = 0;   // sum of rule-kernel durations
  double reduce_ms = 0;
  double other_ms = 0;
  double rule_apps = 0;        // algorithmic rule applications executed by rule kernels
  int64_t rule_launches = 0;
  double total_ms = 0;          // device time of the passes (inside, outside, decode)
  double device_ms = 0;         // device time from input upload to the first result download
  std::vector<float> launch_ms;      // per rule-kernel launch duration
  std::vector<double> launch_apps;   // per rule-kernel launch rule applications
  std::vector<int32_t> launch_family;
  int64_t graph_captures = 0;  // batches whose device work was captured into a graph
  int64_t graph_replays = 0;   // batches replayed from a captured graph
  double span_lo_ms = 1e300, span_hi_ms = -1e300;  // device interval of the batches from ref_event
};

class Parser {
 public:
  Parser(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
         int32_t items_per_lane);
  ~Parser();

  // CLParser.parse: Viterbi trees (final layer Viterbi) or MBR trees (final layer Real).
  void parse(const int32_t* words, const int32_t* offsets, int32_t n,
             std::vector<std::vector<TreeNode>>* viterbi, std::vector<std::vector<MbrNode>>* mbr);
  // parse() without any device allocation: every workspace buffer is carved from the library's
  // file-scope scratch array on this device (kNoAllocScratchBytes; 1/4 of it holds the rule kernels'
  // partial rows). Inputs whose workspace would not fit are rejected before any device work. No
  // graph capture, no timing events beyond a pool made when the parser was created. Calls that
  // overlap on one GPU share that GPU's array and race; the kernels clear what they need at the start
  // of every batch (fill kernels), so a call never depends on what the previous one left behind.
  static constexpr int64_t kNoAllocScratchBytes = (int64_t)256 << 20;
  void parse_noalloc(const int32_t* words, const int32_t* offsets, int32_t n,
                     std::vector<std::vector<TreeNode>>* viterbi, std::vector<std::vector<MbrNode>>* mbr);
  // CLParser.partitions: sentence inside log-likelihood under the final layer.
  void partitions(const int32_t* words, const int32_t* offsets, int32_t n, float* out);
  // CLParser.insideOutside: inside, outside and the decode kernels, results kept on device.
  void inside_outside(const int32_t* words, const int32_t* offsets, int32_t n);
  // Pruning masks of the last coarse layer for every cell of every sentence (maskWords per cell),
  // plus its inside/outside scaling constants when the next layer is Real.
  void masks(const int32_t* words, const int32_t* offsets, int32_t n, std::vector<uint32_t>* out,
             std::vector<int32_t>* cell_offsets);

  // Debug access for parity tests: run the named pipeline on one batch and copy back charts of
  // layer `layer`: inside, outside (cells x cellSize), scales, masks of the coarse layer.
  struct Charts {
    std::vector<float> inside, outside, in_scale, out_scale;
    std::vector<uint32_t> masks;
    std::vector<int32_t> cell_offsets;
    std::vector<double> partition_scales;
    int32_t cell_size = 0;
  };
  void debug_charts(const int32_t* words, const int32_t* offsets, int32_t n, int32_t layer,
                    bool outside, Charts* out);

  PassStats stats;
  // Device bytes the parser's workspace holds now (every buffer it has grown).
  int64_t workspace_bytes() const;
  // Device bytes one call on sentences of these lengths makes the workspace hold (an upper bound
  // that counts every masked item as planned): charts, scales, masks, mask plans, the partial
  // buffer and the batch tables. Serving admits requests by it.
  int64_t bytes_needed(const std::vector<int32_t>& lengths) const { return bytes_needed(lengths, 0); }
  int64_t bytes_needed(const std::vector<int32_t>& lengths, int64_t partial_bytes) const;
  // When set, each batch's device interval is also measured from this event (recorded on any
  // stream of the same device before the call): stats.span_lo_ms / span_hi_ms.
  hipEvent_t ref_event = nullptr;
  const Layer& layer(int i) const { return *layers_[i]; }
  int32_t num_layers() const { return (int32_t)layers_.size(); }
  // Per-launch timing (HIP events around every rule launch); off for serving.
  bool timing = true;
  // Persistent form: the rule families of a span share one k_rules_q launch over a device queue
  // of work units (otherwise one k_rules launch per family and span).
  bool persistent = false;
  // Launch shape knobs (autotuned): a plain launch cuts rule streams into pieces while
  // tiles x chunks x 2 x pieces stays within piece_target_; a persistent launch runs
  // queue_wgs_per_cu_ workgroups per CU.
  void set_launch_shape(int32_t piece_target, int32_t queue_wgs_per_cu) {
    piece_target_ = piece_target;
    queue_wgs_per_cu_ = queue_wgs_per_cu;
    drop_graphs();
  }
  void set_persistent(bool on) {
    persistent = on;
    drop_graphs();
  }
  // parse() with a Viterbi final layer also runs that layer's outside pass before the decode, as
  // CLParser.insideOutside does (CLParser.scala:171-182); the trees do not depend on it.
  void set_outside_final(bool on) {
    if (outside_final == on) return;
    outside_final = on;
    drop_graphs();
  }
  bool outside_final = false;
  // Device-scheduled passes: every pass whose (group, span) stages each fit in one slice of half the
  // partial buffer runs as one persistent k_pass launch that pulls the rule and reduce units of all
  // its spans from one device queue, with per-block dependency counters (see PassStage); other
  // passes keep one launch per span.
  // Disabled: on the GPU the first device-scheduled parse did not return within 150 s (twice, the
  // second time with every cross-unit access agent-coherent and no L2-wide fences), and no cause has
  // been shown; enabling it is rejected until one is.
  void set_device_schedule(bool on) {
    if (on) throw std::invalid_argument(kDeviceScheduleDisabled);
    if (device_schedule == on) return;
    device_schedule = on;
    drop_graphs();
  }
  bool device_schedule = false;
  // Size of the rule kernels' partial buffer (default: 30% of mem_bytes, at least 64 MiB); a
  // smaller buffer cuts a span's cells into more slices, never changing results.
  void set_partial_bytes(int64_t bytes) {
    partial_bytes_ = std::max<int64_t>(bytes, (int64_t)64 << 20);
    drop_graphs();
  }
  // Drops every captured graph (their launches no longer match the configuration).
  void drop_graphs();
  // Workgroup size of the reduce launches: targets (waves) per workgroup, one of 1, 2, 4, 8, 16 (64
  // to 1,024 threads); 4 measured fastest on the exhaustive and coarse-to-fine workloads. The
  // device-scheduled pass always reduces 16 targets per unit.
  void set_reduce_targets(int32_t targets) {
    if (targets != 1 && targets != 2 && targets != 4 && targets != 8 && targets != kReduceTargets)
      throw std::invalid_argument("reduce targets per workgroup must be 1, 2, 4, 8 or 16");
    reduce_targets_ = targets;
    drop_graphs();
  }
  int32_t piece_target_ = 512;
  int32_t queue_wgs_per_cu_ = 1;
  int32_t reduce_targets_ = 4;
  // Folded partial rows (Viterbi semiring, on by default): every slot of the rule kernels folds its
  // values into one partial row per target symbol with an integer atomic max on order-preserving
  // encodings (exact and independent of order), so the reduce reads one row per (family, target)
  // instead of one per (operand chunk, target) slot. Off: one row per slot, written with plain
  // stores. Results are identical either way.
  void set_fold(bool on) {
    if (fold_ == on) return;
    fold_ = on;
    drop_graphs();
  }
  bool fold_ = true;
  // Whether the rule launches of layer L fold (Viterbi layers with fold_ set; never in the
  // device-scheduled pass, whose units exchange partial rows with coherent plain accesses).
  bool folds(const Layer& L) const { return fold_ && !device_schedule && L.g.semiring == kViterbi; }
  // Graph mode of parse(): the device work of a batch is captured into a HIP graph the second
  // time a batch of the same sentence lengths (in order) is seen, and replayed after that.
  bool use_graphs = false;

 private:
  struct Batch;
  struct CaptureAbort {};
  struct GraphEntry {
    int32_t seen = 0;       // eager runs of this shape
    bool failed = false;    // capture abandoned: stay eager
    hipGraphExec_t exec = nullptr;
    int64_t generation = 0; // workspace generation the graph's addresses belong to
  };
  static constexpr size_t kMaxGraphs = 64;
  std::map<std::vector<int32_t>, GraphEntry> graphs_;
  bool capturing_ = false;
  int64_t ws_generation_ = 0;  // bumped whenever a workspace buffer is reallocated
  GraphEntry* graph_entry(const std::vector<int32_t>& lengths);
  std::vector<std::unique_ptr<Layer>> layers_;
  int32_t device_;
  int64_t mem_bytes_;
  int64_t partial_bytes_ = 0;  // 0: 30% of mem_bytes_
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  // apps_slot >= 0: the launch's rule applications add the plan counts in
  // apps_[apps_slot, apps_slot + apps_count).
  struct TimedLaunch { hipEvent_t a, b; double apps; int32_t family; int32_t apps_slot; int32_t apps_count; };
  int32_t apps_next_ = 0;
  std::vector<hipEvent_t> ev_pool_;
  size_t ev_next_ = 0;
  std::vector<TimedLaunch> pending_;
  struct Phase { hipEvent_t a, b; double* acc; };
  std::vector<Phase> phases_;
  hipEvent_t phase_start_ = nullptr;
  hipEvent_t dev_start_ = nullptr, dev_end_ = nullptr;
  bool dev_marked_ = true;
  void mark_device_end();
  hipEvent_t next_event();
  void flush_timing();

  // Workspace (grown on demand).
  struct Buf { void* p = nullptr; size_t bytes = 0; };
  Buf inside_, outside_, in_scale_, out_scale_, masks_, partial_, plan_, tags_, trees_, mbr_,
      pos_scale_, aux_, mplan_, apps_, queue_, jobs_, pass_, pcnt_;
  int32_t queue_next_ = 0;
  // No-alloc mode: ensure() carves buffers from [base, base + size) instead of allocating.
  struct Arena { char* base = nullptr; size_t size = 0, used = 0; };
  Arena arena_;
  char* scratch_ = nullptr;  // this device's scratch array (device address), set at construction
  std::string scratch_err_;  // why scratch_ is null, if it is
  void reset_arena();
  int32_t num_cus_ = 256;
  void ensure(Buf& b, size_t bytes);

  void run_batch(const int32_t* words, const int32_t* offsets, int32_t first, int32_t count,
                 int32_t what, std::vector<std::vector<TreeNode>>* viterbi,
                 std::vector<std::vector<MbrNode>>* mbr, float* partitions, Charts* charts,
                 int32_t chart_layer, std::vector<uint32_t>* masks_out);
  std::vector<std::pair<int32_t, int32_t>> make_batches(const int32_t* offsets, int32_t n);
  void parse_batch(Batch& b, int32_t first, std::vector<std::vector<TreeNode>>* viterbi,
                   std::vector<std::vector<MbrNode>>* mbr);
  void parse_device(Batch& b);
  void finish_parse(Batch& b, int32_t first, std::vector<std::vector<TreeNode>>* viterbi,
                    std::vector<std::vector<MbrNode>>* mbr);
  void finish_batch(Batch& b, int32_t what, float* partitions, Charts* charts, int32_t chart_layer,
                    std::vector<uint32_t>* masks_out);
  void reset_counters(Batch& b);
  void launch_masks_and_scales(Batch& b, int32_t li, int32_t mask_words);
  void launch_viterbi(Batch& b, int32_t li);
  void launch_mbr(Batch& b, int32_t li);
  void inside_pass(Batch& b, int32_t li, bool use_masks);
  void outside_pass(Batch& b, int32_t li, bool use_masks);
  bool family_args(Batch& b, int32_t li, int32_t fid, int32_t span, bool use_masks, int32_t k0,
                   int32_t kc, float* partial, const int32_t* plan_counts, const int32_t* plan_items,
                   const uint32_t* plan_act, FamArgs* out, dim3* grid, size_t* lds);
  void launch_family(Batch& b, int32_t li, int32_t fid, int32_t span, bool use_masks,
                     int32_t slice_cell0, int32_t slice_cells, float* partial_base,
                     const int32_t* plan_counts, const int32_t* plan_items, const uint32_t* plan_act,
                     int32_t apps_slot);
  void launch_queue(Batch& b, int32_t li, const std::vector<int32_t>& fids, int32_t span, bool use_masks,
                    int32_t k0, int32_t kc, const std::vector<float*>& partials,
                    const std::vector<const int32_t*>& plan_counts,
                    const std::vector<const int32_t*>& plan_items,
                    const std::vector<const uint32_t*>& plan_acts, const std::vector<int32_t>& apps_slots);
  // Host view of the mask plan of one slice of a group's cells at one span: per family of the
  // group, device pointers into the plan buffer (see PlanJob).
  struct PlanPtrs {
    int32_t* pre = nullptr;
    int32_t* items = nullptr;
    uint32_t* act = nullptr;
    int32_t* posof = nullptr;
    int32_t* perm = nullptr;
  };
  struct PlannedSlice {
    int32_t k0 = 0, kc = 0;
    std::vector<PlanPtrs> pfs;
    std::vector<int32_t> apps_slot;
    int32_t* d_perm = nullptr;  // visiting order of the reduce: the group's largest family's order
  };
  // One slice of a (reduce group, span): its families' rule-kernel arguments and its reduce.
  struct GroupSlice {
    int32_t nfam = 0;
    int32_t fid[kMaxGroupFamilies] = {};
    bool valid[kMaxGroupFamilies] = {};  // the family has items in the slice
    FamArgs fa[kMaxGroupFamilies] = {};
    dim3 grid[kMaxGroupFamilies];        // plain launch grid: tiles, chunks, pieces
    size_t lds[kMaxGroupFamilies] = {};
    float* partial[kMaxGroupFamilies] = {};
    int32_t apps_slot[kMaxGroupFamilies] = {};
    int64_t row_items[kMaxGroupFamilies] = {};  // values per partial row (unpruned upper bound)
    RedArgs ra = {};
    dim3 rgrid;
  };
  size_t group_slice_args(Batch& b, int32_t li, int32_t rid, int32_t span, bool use_masks, const PlannedSlice* ps,
                          int32_t k0, int32_t kc, float* base, GroupSlice* out);
  std::vector<std::pair<int32_t, int32_t>> group_slices(Batch& b, int32_t li, int32_t rid, int32_t span,
                                                        int64_t budget_floats = -1);
  int64_t family_items(Batch& b, int32_t li, int32_t fid, int32_t span, int32_t k0, int32_t kc);
  // Passes of a call in execution order: (layer, outside pass). `what` as in run_batch.
  std::vector<std::pair<int32_t, bool>> pass_schedule(int32_t what, int32_t chart_layer) const;
  static std::vector<std::pair<int32_t, int32_t>> inside_groups(const Batch& b);
  static std::vector<std::pair<int32_t, int32_t>> outside_groups(const Batch& b);
  // Lays out the mask plans of every masked pass of the call (before any device work, so that the
  // plan job table is uploaded with the batch and graphs replay it unchanged).
  void prepare_plans(Batch& b, const std::vector<std::pair<int32_t, bool>>& schedule);
  // The four plan launches over every job of layer li's masked passes, on the parser's stream.
  void launch_plans(Batch& b, int32_t li);
  void run_groups(Batch& b, int32_t li, bool outside, bool use_masks);
  // Device-scheduled passes of a call: laid out with the batch (prepare_dev_passes), launched in
  // schedule order in place of run_groups (run_dev_pass returns false for a pass that keeps its
  // per-span launches).
  void prepare_dev_passes(Batch& b, const std::vector<std::pair<int32_t, bool>>& schedule);
  bool build_dev_pass(Batch& b, int32_t li, bool outside, bool use_masks, const void* pass_plan,
                      std::vector<PassStage>* stages, std::vector<PassTile>* tiles, std::vector<int32_t>* expected,
                      int32_t* counters, void* dev_pass);
  bool run_dev_pass(Batch& b, int32_t li, bool outside, bool use_masks);
  bool dev_pass_ran_ = false;  // a k_pass ran since the last error check
  void run_group(Batch& b, int32_t li, int32_t rid, int32_t span, bool use_masks,
                 const std::vector<PlannedSlice>* planned);
  void timed_begin();
  void timed_end(double* acc);
};

}  // namespace puck
