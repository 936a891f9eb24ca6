// CKY engine for gfx950: rule-family kernels, reductions, tag scores, masks, scaling, Viterbi and
// MBR decode kernels, and the host schedule that drives them.
//
// Reference kernels replaced here (all OpenCL C generated as strings in the reference):
//   binary/unary rule kernels  SimpleGenRuleMultiply.java:60-319
//   gather/scatter             CLMatrixTransposeCopy.scala:123-220 (not needed: operands are staged
//                              straight from the chart into LDS)
//   computeMasks               CLMaskKernels.scala:154-238
//   getScalingConstants        CLScalingKernels.scala:49-71
//   viterbi                    CLViterbi.scala:364-608
//   computeMBR                 CLMBRKernels.scala:101-173
//   set/getRootScores          CLParserUtils.scala:66-77
//   mem_zero                   ZeroMemoryKernel.scala:18-30
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>

#include "engine.h"
#include "kernels.h"

namespace puck {

#define HIP_CHECK(x)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                               #x);                                                          \
  } while (0)

// Device scratch of Parser::parse_noalloc (see engine.h): one array per GPU, held until the process
// ends. Host code never passes `scratch` itself to a kernel: the array's device address on the
// current device comes from hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) unsigned char scratch[Parser::kNoAllocScratchBytes];

namespace {

constexpr int kMaxLength = 255;  // cell lists pack the begin position in 8 bits

__host__ __device__ inline int tri(int n) { return n * (n + 1) / 2; }

// ChartHalf.chartIndex (ParseChart.scala:58-65): span-diagonal-major cell index.
__host__ __device__ inline int chart_index(int b, int e, int n) {
  int lo = b < e ? b : e;
  int s = (b < e ? e - b : b - e) - 1;
  return lo + s * n - s * (s - 1) / 2;
}

__device__ inline int cell_abs(const int32_t* cell_offsets, int sent, int n, int half, int b, int e) {
  return cell_offsets[sent] + ((half & 1) ? tri(n) : 0) + chart_index(b, e, n);
}

// Operand cells of item j of cell (b, e) for a family kind: A = (ax0, ay0), B = (bx0, by0).
// Split ranges: CLParser.scala:345-350,492-518 (CLWorkQueueKernels.scala:138-168).
__device__ inline void item_operands(int kind, int b, int e, int j, int& ax0, int& ay0, int& bx0, int& by0) {
  ax0 = b; ay0 = e; bx0 = 0; by0 = 0;
  switch (kind) {
    case 0: { int s = b + 1 + j; ax0 = b; ay0 = s; bx0 = s; by0 = e; } break;      // IN range
    case 1: { int s = e - 1; ax0 = b; ay0 = s; bx0 = s; by0 = e; } break;          // IN e-1
    case 2: { int s = b + 1; ax0 = b; ay0 = s; bx0 = s; by0 = e; } break;          // IN b+1
    case 3: { int s = e + 1 + j; ax0 = b; ay0 = s; bx0 = e; by0 = s; } break;      // OUT_L range
    case 4: { int s = e + 1; ax0 = b; ay0 = s; bx0 = e; by0 = s; } break;          // OUT_L e+1
    case 5: { int s = j; ax0 = s; ay0 = b; bx0 = s; by0 = e; } break;              // OUT_R range
    case 6: { int s = b - 1; ax0 = s; ay0 = b; bx0 = s; by0 = e; } break;          // OUT_R b-1
    default: break;                                                                 // unary
  }
}

// Items a family generates at cell (b, e) of a sentence of length n.
__host__ __device__ inline int items_for(int kind, int b, int e, int n) {
  switch (kind) {
    case 0: return e - b - 1;
    case 1: case 2: return 1;
    case 3: return n - e;
    case 4: return e + 1 <= n ? 1 : 0;
    case 5: return b;
    case 6: return b >= 1 ? 1 : 0;
    default: return 1;
  }
}

__device__ inline bool mask_any(const uint32_t* m, int words) {
  uint32_t x = 0;
  for (int i = 0; i < words; ++i) x |= m[i];
  return x != 0;
}

template <int SEMI>
__host__ __device__ inline float semi_zero() {
  return SEMI == kViterbi ? -INFINITY : (SEMI == kLogSum ? -1000000.0f : 0.0f);
}

inline float zero_of(int semiring) {
  return semiring == kViterbi ? -INFINITY : (semiring == kLogSum ? -1000000.0f : 0.0f);
}

// LogSum semiring_add (RuleSemiring.scala:66-72): larger + log(1 + exp(smaller - larger)), with
// the fast hardware exp/log as the reference's native_exp/native_log.
__device__ inline float log_add(float x, float y) {
  const float lo = fminf(x, y), hi = fmaxf(x, y);
  return hi + __logf(1.0f + __expf(lo - hi));
}

// Semiring add of a lane's two paired splits in the rule kernel's flush.
template <int SEMI>
__device__ inline float semi_pair(float a, float b) {
  // Viterbi: one v_max_f32 (fmaxf first quiets each operand, two more instructions; max-plus
  // values are never NaN).
  if constexpr (SEMI == kViterbi) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  } else if constexpr (SEMI == kLogSum) {
    return log_add(a, b);
  } else {
    return a + b;
  }
}

// Semiring add of the reduce: max, +, or log-add.
template <int SEMI>
__device__ inline float semi_add(float acc, float v) {
  if constexpr (SEMI == kViterbi) return fmaxf(acc, v);
  else if constexpr (SEMI == kLogSum) return log_add(acc, v);
  else return acc + v;
}

// Order-preserving 32-bit encoding of a float: for any two non-NaN floats x < y (with -0 < +0),
// enc_ordered(x) < enc_ordered(y) as signed integers. Negative floats have their magnitude bits
// flipped, so the encoding is an involution (decode = encode). Folding partial values with an
// integer atomic max on encodings gives exactly the IEEE maximum, in any order.
__device__ __forceinline__ int32_t enc_ordered(float v) {
  const int32_t b = __float_as_int(v);
  return b ^ ((b >> 31) & 0x7fffffff);
}
// Decodes when `mask` is 0x7fffffff and passes the bits through when it is 0.
__device__ __forceinline__ float dec_ordered(float x, int32_t mask) {
  const int32_t b = __float_as_int(x);
  return __int_as_float(b ^ ((b >> 31) & mask));
}
typedef __attribute__((address_space(1))) int32_t GlobalI32;
__device__ __forceinline__ void fold_max(float* p, float v) {
  // A global (not flat) atomic: a flat one would also count on the LDS wait counter that the rule
  // loop drains every step.
  __hip_atomic_fetch_max((GlobalI32*)p, enc_ordered(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Encoding of the Viterbi zero (-inf): the fill value of folded partial rows.
constexpr uint32_t kEncNegInf = 0xff800000u ^ 0x7fffffffu;

// Loads and stores of data that another workgroup of the same launch produces or consumes (the
// device-scheduled pass, COH = true): relaxed agent-scope atomics, which the compiler emits as
// global loads and stores with the sc1 bit, coherent across the XCDs' L2 caches without a
// whole-cache writeback or invalidate. Other launches use plain accesses (COH = false). Both go
// through the global address space: a flat access would also count on the LDS wait counter.
typedef __attribute__((address_space(1))) float GlobalF32;
template <bool COH>
__device__ __forceinline__ float ld_f(const float* p) {
  if constexpr (COH) return __hip_atomic_load((GlobalF32*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *(const GlobalF32*)p;
}
template <bool COH>
__device__ __forceinline__ void st_f(float* p, float v) {
  if constexpr (COH) __hip_atomic_store((GlobalF32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *(GlobalF32*)p = v;
}

// ------------------------------------------------------------------------------------------------
// s_waitcnt immediate (gfx9 encoding) that waits for lgkmcnt 0 only: vmcnt 63, expcnt 7.
constexpr int kWaitLgkm0 = 0xC07F;

// Rule-family kernel. One workgroup = one tile of ITEMS items (target cell, split) x one chunk
// (row range of operand A x row range of operand B). Operand rows are staged item-minor in LDS so
// that a wave reads one row for IPL items per lane with one ds_read. Each wave streams its own
// share of the chunk's target symbols (slots, assigned longest-first when the grammar is loaded):
// the slots' rules, uniform across the wave, arrive in groups of eight by scalar loads. Each slot
// writes one partial value per item; the reduce kernel combines partials over splits, chunks and
// families. Rule semantics per SimpleGenRuleMultiply.java:151 (parent = semiring_mad(parent,
// left (x) right, q)) and :213-216 (Real: *= scale).
// Per-tile bookkeeping of the rule kernels in static LDS (the operand rows use dynamic LDS).
template <int ITEMS>
struct RulesLds {
  int32_t cell_a[ITEMS], cell_b[ITEMS];
  float scale[ITEMS];
  uint32_t uni[4];  // blocks active in any item of the tile (block-major streams)
  // Prefix-counted slices: item prefix of the cells from the tile's first cell on (slice-relative)
  // and that cell.
  int32_t pre[2 * ITEMS + 2];
  int32_t pre_c0;
};

// One work unit of a rule family: tile `tile` of ITEMS items x operand chunk `chunk` x piece
// `piece` of every wave's rule stream. Every thread of the workgroup calls it with the same
// arguments; it contains workgroup barriers before the rule loop only. COH: operand rows are read
// and partial rows written with agent-coherent accesses (see ld_f).
template <int IPL, int SEMI, bool UNARY, bool SKIP, bool COH = false, bool FOLD = false, int WAVES = 16>
__device__ __forceinline__ void rules_unit(const FamArgs& a, int tile, int chunk, int piece, float* smem,
                                           RulesLds<64 * IPL>& sh) {
  static_assert(!FOLD || SEMI == kViterbi, "only max-plus partials fold by atomic max");
  static_assert(WAVES == 16 || WAVES == 8, "workgroups of 16 or 8 waves");
  constexpr int ITEMS = 64 * IPL;
  int32_t* s_cell_a = sh.cell_a;
  int32_t* s_cell_b = sh.cell_b;
  float* s_scale = sh.scale;
  uint32_t* s_union = sh.uni;
  const float zero = semi_zero<SEMI>();
  const int tid = threadIdx.x;
  const int tile0 = tile * ITEMS;
  // Items of the slice: [item_begin, item_begin + num_items). Prefix-counted and planned slices
  // read their bounds on the device (the grid covers the largest possible count).
  int item_begin = 0, num_items = a.num_items;
  if (a.prefix != nullptr) {
    item_begin = a.prefix[0];
    num_items = a.prefix[a.num_cells] - item_begin;
  }
  if (tile0 >= num_items) return;
  if constexpr (SKIP) {
    if (tid < 4) s_union[tid] = 0u;
    __syncthreads();
  }
  const int a_lo = a.chunk_range[4 * chunk], a_hi = a.chunk_range[4 * chunk + 1];
  const int b_lo = a.chunk_range[4 * chunk + 2], b_hi = a.chunk_range[4 * chunk + 3];
  const int rows_a = a_hi - a_lo;
  const int rows_b = UNARY ? 0 : b_hi - b_lo;
  float* lds_a = smem;
  float* lds_b = smem + rows_a * a.row_stride;

  // Prefix-counted (unplanned) slices: wave 0 finds the tile's first cell by a 64-way search (a few
  // dependent loads where a binary search per item takes one per halving) and copies the prefix
  // of the 2 ITEMS + 2 cells from there on into LDS, where each item finds its cell. Cells without
  // items (at most one per sentence) are why the window holds two cells per item; an item past
  // the window falls back to searching the slice.
  constexpr int kWin = 2 * ITEMS + 2;
  const bool pre_win = a.prefix != nullptr && a.plan_items == nullptr;
  if (pre_win) {
    if (tid < 64) {
      int lo = 0, hi = a.num_cells;  // largest c in [lo, hi) with prefix[c] - item_begin <= tile0
      while (hi - lo > 64) {
        const int step = (hi - lo + 63) >> 6;
        const int c = lo + tid * step;
        const uint64_t m = __ballot(c < hi && a.prefix[c] - item_begin <= tile0);
        lo += (63 - __clzll(m)) * step;  // lane 0 (c = lo) always qualifies
        hi = min(lo + step, hi);
      }
      const int c = lo + tid;
      const uint64_t m = __ballot(c < hi && a.prefix[c] - item_begin <= tile0);
      const int c0 = lo + (63 - __clzll(m));
      for (int i = tid; i < kWin; i += 64) sh.pre[i] = a.prefix[min(c0 + i, a.num_cells)] - item_begin;
      if (tid == 0) sh.pre_c0 = c0;
    }
    __syncthreads();
  }

  if (tid < ITEMS) {
    const int item = tile0 + tid;
    // An invalid item (past the slice, or the missing partner of a lone last split pair) keeps
    // cell -1: its operand rows are staged as the semiring zero, so its value stays zero.
    int ca = -1, cb = -1;
    uint32_t gate[4] = {0u, 0u, 0u, 0u};
    float scale = 1.0f;
    int k = 0, j = 0;
    bool valid = item < num_items;
    if (valid) {
      if (a.plan_items != nullptr) {
        const int code = a.plan_items[item_begin + item];
        k = (code >> 8) - a.slice_cell0;
        j = code & 255;
      } else if (a.prefix == nullptr && a.pair_splits) {
        // Paired block-major items: pair m = (block of 64 cells, split pair p, cell of the block);
        // items 2m and 2m+1 (one lane) are splits 2p and 2p+1 of that cell, and the lane folds its
        // pair before writing the partial.
        const int m = item >> 1;
        const int per_block = 64 * a.uniform_count;
        const int bx = m / per_block;
        const int r = m - bx * per_block;
        k = bx * 64 + (r & 63);
        j = 2 * (r >> 6) + (item & 1);
        valid = j < a.span - 1 && k < a.num_cells;
      } else if (a.prefix == nullptr) {
        // Uniform families are block-major: item = (block of 64 cells, split, cell of the block),
        // so that a tile lies in one or two blocks and the reduce reads a block's split rows of
        // consecutive cells coalesced. The last block's rows are padded to 64 cells.
        const int per_block = 64 * a.uniform_count;
        const int bx = item / per_block;
        const int r = item - bx * per_block;
        j = r >> 6;
        k = bx * 64 + (r & 63);
        valid = k < a.num_cells;
      } else {
        // Largest i with pre[i] <= item in the LDS window (pre[0] <= tile0 <= item).
        int lo = 0, hi = kWin;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (sh.pre[mid] <= item) lo = mid; else hi = mid;
        }
        k = sh.pre_c0 + lo;
        j = item - sh.pre[lo];
        if (lo == kWin - 1) {
          // Possibly past the window: search the slice's prefix from there.
          const int abs_item = item + item_begin;
          int l2 = k, h2 = a.num_cells;  // largest k with prefix[k] <= abs_item
          while (h2 - l2 > 1) {
            const int mid = (l2 + h2) >> 1;
            if (a.prefix[mid] <= abs_item) l2 = mid; else h2 = mid;
          }
          k = l2;
          j = abs_item - a.prefix[k];
        }
      }
    }
    if (valid) {
      const int packed = a.cell_list[k];
      const int sent = packed >> 8, b = packed & 255, e = b + a.span;
      const int n = a.lengths[sent];
      int ax0, ay0, bx0, by0;
      item_operands(a.kind, b, e, j, ax0, ay0, bx0, by0);
      ca = cell_abs(a.cell_offsets, sent, n, a.a_half, ax0, ay0);
      cb = UNARY ? 0 : cell_abs(a.cell_offsets, sent, n, a.b_half, bx0, by0);
      const int ct = cell_abs(a.cell_offsets, sent, n, a.t_half, b, e);
      // Block-major streams walk the union of their items' active blocks: every block without
      // masks; under masks (masked launches are always planned) the blocks the plan kept for the
      // target cell, whose items have allowed children and an allowed target span.
      if constexpr (SKIP) {
        if (a.plan_items == nullptr) {
          gate[0] = gate[1] = gate[2] = gate[3] = 0xffffffffu;
        } else if (UNARY) {
          gate[0] = 1u;
        } else {
#pragma unroll
          for (int w = 0; w < 4; ++w) gate[w] = a.plan_act[(size_t)w * a.num_cells + k];
        }
      }
      if (SEMI == kReal) {
        // native_exp(-parentScale + rightScale + leftScale) (SimpleGenRuleMultiply.java:214,280)
        scale = UNARY ? expf(-a.scT[ct] + a.scA[ca]) : expf(-a.scT[ct] + a.scB[cb] + a.scA[ca]);
      }
    }
    s_cell_a[tid] = ca;
    s_cell_b[tid] = cb;
    if constexpr (SKIP) {
      // Union of the items' block bits: an OR across the wave (DPP row shifts, then row broadcasts
      // that leave the total in lane 63), then one LDS atomic per wave and word instead of one per
      // item on the same address. Every lane of the wave is active here (ITEMS is a multiple of 64).
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        int x = (int)gate[w];
        x |= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);   // row_shr:1
        x |= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);   // row_shr:2
        x |= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);   // row_shr:4
        x |= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);   // row_shr:8
        x |= __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
        x |= __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
        const uint32_t all = (uint32_t)__builtin_amdgcn_readlane(x, 63);
        if ((tid & 63) == 0 && all != 0u) atomicOr(&s_union[w], all);
      }
    }
    s_scale[tid] = scale;
  }
  __syncthreads();

  // Stage operand rows: wave w copies items w, w + WAVES, ... of the tile, eight at a time; its
  // lanes take 64 consecutive rows of an item's cell at a time, so each load is one contiguous run
  // of the cell's rows from a wave-uniform base, and one row block's loads for the batch's items
  // are in flight together. An invalid item (cell -1) stages the semiring zero.
  constexpr int kStageWaves = WAVES;
  constexpr int IPW = ITEMS / kStageWaves;  // items per wave
  constexpr int IPB = IPW < 16 ? IPW : 16;  // items per batch
  constexpr int kStageBlocks = 2;
  static_assert(ITEMS % kStageWaves == 0 && IPW % IPB == 0, "items split evenly over the waves");
  auto stage = [&](const float* ch, const int32_t* cells, int lo, int rows, float* dst) {
    const int sw = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t sl = tid & 63;
#pragma unroll
    for (int h = 0; h < IPW; h += IPB) {
      const char* base[IPB];
      bool ok[IPB];
#pragma unroll
      for (int k = 0; k < IPB; ++k) {
        const int c = __builtin_amdgcn_readfirstlane(cells[sw + kStageWaves * (h + k)]);
        ok[k] = c >= 0;
        base[k] = reinterpret_cast<const char*>(ch + (size_t)(ok[k] ? c : 0) * a.cell_size + lo);
      }
      // Up to kStageBlocks row blocks of 64 per step: kStageBlocks x IPB loads in flight before the
      // first LDS store (an operand's rows of a 150 KB tile take one or two steps).
      for (int r0 = 0; r0 < rows; r0 += 64 * kStageBlocks) {
        float v[kStageBlocks][IPB];
#pragma unroll
        for (int q = 0; q < kStageBlocks; ++q) {
          // 32-bit byte offsets from the uniform bases: one load with a scalar base per item. Lanes
          // past the last row load that row (no branch between the loads) and store nothing.
          const uint32_t row = (uint32_t)r0 + 64u * q + sl;
          const uint32_t b = min(row, (uint32_t)rows - 1u) * 4u;
#pragma unroll
          for (int k = 0; k < IPB; ++k) v[q][k] = ld_f<COH>(reinterpret_cast<const float*>(base[k] + b));
        }
#pragma unroll
        for (int q = 0; q < kStageBlocks; ++q) {
          const uint32_t row = (uint32_t)r0 + 64u * q + sl;
          if (row < (uint32_t)rows) {
            float* d = dst + row * a.row_stride + sw + kStageWaves * h;
#pragma unroll
            for (int k = 0; k < IPB; ++k) d[kStageWaves * k] = ok[k] ? v[q][k] : zero;
          }
        }
      }
    }
  };
  // Block ranges of this wave's stream (skip layouts): lane l loads the bounds of blocks l and
  // 64 + l here, so the loads are in flight with the staging.
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nbr = a.num_block_ranges;
  const int32_t* bt = a.wave_blocks + (size_t)(chunk * WAVES + wave) * (nbr + 2 + kPieceTableInts);
  int bt0[2] = {0, 0}, bt1[2] = {0, 0};
  if constexpr (SKIP) {
    if (lane < nbr) {
      bt0[0] = bt[lane];
      bt1[0] = bt[lane + 1];
    }
    if (64 + lane < nbr) {
      bt0[1] = bt[64 + lane];
      bt1[1] = bt[65 + lane];
    }
  }
  stage(a.chA, s_cell_a, a_lo, rows_a, lds_a);
  if constexpr (!UNARY) stage(a.chB, s_cell_b, b_lo, rows_b, lds_b);
  __syncthreads();

  const char* lane_base = reinterpret_cast<const char*>(smem) + lane * IPL * 4;
  constexpr int G = kGroupRules;
  constexpr int H = kRuleUnroll;
  constexpr int kWaves = WAVES;
  // Rule groups are wave-uniform and never written while a kernel runs: reading them through the
  // constant address space lets the compiler use scalar loads into SGPRs (two s_load_dwordx16 per
  // group of eight rules for the whole wave).
  typedef uint32_t W16 __attribute__((ext_vector_type(16)));
  typedef uint32_t W8 __attribute__((ext_vector_type(8)));
  typedef uint32_t W2 __attribute__((ext_vector_type(2)));
  typedef const __attribute__((address_space(4))) W16 ConstW16;
  typedef const __attribute__((address_space(4))) W8 ConstW8;
  typedef const __attribute__((address_space(4))) W2 ConstW2;
  typedef const __attribute__((address_space(4))) char ConstByte;
  typedef const __attribute__((address_space(4))) int32_t ConstI32;
  // Groups are addressed by 32-bit byte offsets from the start of the family's stream.
  constexpr uint32_t GB = (SKIP ? kSkipGroupWords : kGroupWords) * 4;
  ConstByte* gbase = (ConstByte*)a.rules;
  auto group_off = [&](uint32_t o) { return *(ConstW16*)(gbase + o); };
  auto group_q = [&](uint32_t o) { return *(ConstW8*)(gbase + o + 64); };
  auto group_slots = [&](uint32_t o) { return *(ConstW2*)(gbase + o + 96); };

  // Block ranges of this wave's stream, visited in order. A skip layout lets a masked tile pass
  // over the blocks that none of its items allows; the cursor's range end c_end belongs to the
  // furthest group fetched so far. Otherwise the wave streams one range and counts its slots.
  // This workgroup's piece of the stream: [p_lo, p_hi) in groups, first slot p_slot.
  const int npieces = 1 << a.pieces_log2;
  ConstI32* sp = (ConstI32*)bt + nbr + 2 + 2 * ((npieces - 1) + a.pieces_log2) + 2 * piece;
  const int p_lo = sp[0], p_slot = sp[1], p_hi = sp[2];
  uint64_t un_lo = ~0ull, un_hi = ~0ull;  // block bits 0-63 and 64-127 of the tile's union
  if (SKIP && a.masks != nullptr) {
    un_lo = (uint64_t)__builtin_amdgcn_readfirstlane(s_union[0]) |
            (uint64_t)__builtin_amdgcn_readfirstlane(s_union[1]) << 32;
    un_hi = (uint64_t)__builtin_amdgcn_readfirstlane(s_union[2]) |
            (uint64_t)__builtin_amdgcn_readfirstlane(s_union[3]) << 32;
  }
  // The wave's block ranges, clipped to its piece, in vector registers: lane l holds blocks l and
  // 64 + l as byte offsets, and a 64-bit mask per half marks the ranges the tile visits (in the
  // union, non-empty). A range change then takes the next set bit and two lane reads, with no
  // memory access: a scalar load of the bounds there would wait on the LDS reads in flight too.
  uint32_t c_end = 0;  // byte offset of the end of the cursor's range
  const uint32_t sentinel = (uint32_t)a.sentinel_group * GB;
  // (Named registers, not arrays: an array indexed in the loop would live in scratch memory.)
  uint32_t lo_b0 = 0u, lo_b1 = 0u, hi_b0 = 0u, hi_b1 = 0u;
  uint64_t visit_lo = 0ull, visit_hi = 0ull;
  if constexpr (SKIP) {
    const int b0 = max(bt0[0], p_lo), b1 = min(bt1[0], p_hi);
    lo_b0 = (uint32_t)b0 * GB;
    lo_b1 = (uint32_t)b1 * GB;
    visit_lo = __ballot(lane < nbr && b0 < b1) & un_lo;
    if (nbr > 64) {
      const int c0 = max(bt0[1], p_lo), c1 = min(bt1[1], p_hi);
      hi_b0 = (uint32_t)c0 * GB;
      hi_b1 = (uint32_t)c1 * GB;
      visit_hi = __ballot(64 + lane < nbr && c0 < c1) & un_hi;
    }
  }
  auto next_range = [&]() -> uint32_t {
    if (visit_lo != 0ull) {
      const int l = __builtin_ctzll(visit_lo);
      visit_lo &= visit_lo - 1ull;
      c_end = (uint32_t)__builtin_amdgcn_readlane((int)lo_b1, l);
      return (uint32_t)__builtin_amdgcn_readlane((int)lo_b0, l);
    }
    if (visit_hi != 0ull) {
      const int l = __builtin_ctzll(visit_hi);
      visit_hi &= visit_hi - 1ull;
      c_end = (uint32_t)__builtin_amdgcn_readlane((int)hi_b1, l);
      return (uint32_t)__builtin_amdgcn_readlane((int)hi_b0, l);
    }
    return sentinel;
  };
  auto advance = [&](uint32_t o) { return o + GB < c_end ? o + GB : next_range(); };
  uint32_t gi, g1, g_end = 0;
  uint32_t slot_ctr = 0;
  if constexpr (SKIP) {
    gi = next_range();
    if (gi == sentinel) return;
    g1 = advance(gi);
  } else {
    gi = (uint32_t)p_lo * GB;
    g_end = (uint32_t)p_hi * GB;
    slot_ctr = (uint32_t)p_slot;
    if (gi >= g_end) return;
    g1 = gi + GB;
  }

  // Per-item scale, kept in registers for the slot flushes.
  float scl[IPL];
#pragma unroll
  for (int i = 0; i < IPL; ++i) scl[i] = s_scale[lane * IPL + i];
  // Operand reads of one group: word u = A-row byte offset, word G + u = B-row byte offset.
  auto read_group = [&](const W16& off, float (*va)[IPL], float (*vb)[IPL]) {
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const uint32_t xo = (u % H == H - 1) ? (off[u] & kOffsetMask) : off[u];
      if constexpr (IPL == 2) {
        const float2 x = *reinterpret_cast<const float2*>(lane_base + xo);
        va[u][0] = x.x;
        va[u][IPL - 1] = x.y;
        if constexpr (!UNARY) {
          const float2 y = *reinterpret_cast<const float2*>(lane_base + off[G + u]);
          vb[u][0] = y.x;
          vb[u][IPL - 1] = y.y;
        }
      } else {
        va[u][0] = *reinterpret_cast<const float*>(lane_base + xo);
        if constexpr (!UNARY) vb[u][0] = *reinterpret_cast<const float*>(lane_base + off[G + u]);
      }
    }
  };
  float acc[IPL];
#pragma unroll
  for (int i = 0; i < IPL; ++i) acc[i] = zero;
  // The flush's loop-invariant fields, moved into vector registers once per unit. Reloaded from
  // the argument block inside the group loop (the compiler's choice under scalar-register
  // pressure) they cost a scalar load whose wait also drains the LDS reads already in flight for
  // the next group.
  // Block-major (masked) streams never pair splits (their items are planned).
  uint32_t out_lo, out_hi, pitch_v, pair_v = 0;
  {
    const uint64_t base = reinterpret_cast<uint64_t>(a.partial);
    asm volatile("v_mov_b32 %0, %1" : "=v"(out_lo) : "s"((uint32_t)base));
    asm volatile("v_mov_b32 %0, %1" : "=v"(out_hi) : "s"((uint32_t)(base >> 32)));
    asm volatile("v_mov_b32 %0, %1" : "=v"(pitch_v) : "s"((uint32_t)a.pitch));
    if constexpr (!SKIP) asm volatile("v_mov_b32 %0, %1" : "=v"(pair_v) : "s"((uint32_t)a.pair_splits));
  }
  // Plain streams flush their slots in order from p_slot on, so the partial row advances by one
  // pitch per flush; block-major streams name each flushed slot.
  float* row = reinterpret_cast<float*>(((uint64_t)out_hi << 32) | out_lo) + (uint64_t)slot_ctr * pitch_v;
  auto flush = [&](uint32_t slot, uint32_t tgt) {
    // Slot done: scale (Real) and write the partial. Under masks, an item whose cell does not
    // allow the slot's block also gets a value; the reduce never reads it, since it takes a
    // (cell, target) partial only where the plan's active-block bits allow the target's block.
    float v[IPL];
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
      v[i] = SEMI == kReal ? acc[i] * scl[i] : acc[i];
      acc[i] = zero;
    }
    if constexpr (FOLD) {
      {
        // Folded rows: the slot's target (from its tag) names the row.
        float* out = reinterpret_cast<float*>(((uint64_t)out_hi << 32) | out_lo) + (uint64_t)tgt * pitch_v;
        if constexpr (IPL == 2 && !SKIP) {
          if (__builtin_amdgcn_readfirstlane(pair_v)) {
            const int m = (tile0 >> 1) + lane;
            if (2 * m < num_items) fold_max(out + m, semi_pair<SEMI>(v[0], v[1]));
            return;
          }
        }
#pragma unroll
        for (int i = 0; i < IPL; ++i) {
          const int item = tile0 + lane * IPL + i;
          if (item < num_items) fold_max(out + item, v[i]);
        }
        return;
      }
    }
    float* out;
    if constexpr (SKIP) {
      out = reinterpret_cast<float*>(((uint64_t)out_hi << 32) | out_lo) + (uint64_t)slot * pitch_v;
    } else {
      out = row;
      row += pitch_v;
    }
    if constexpr (IPL == 2 && !SKIP) {
      if (__builtin_amdgcn_readfirstlane(pair_v)) {
        const int m = (tile0 >> 1) + lane;
        if (2 * m < num_items) st_f<COH>(out + m, semi_pair<SEMI>(v[0], v[1]));
        return;
      }
    }
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
      const int item = tile0 + lane * IPL + i;
      if (item < num_items) st_f<COH>(out + item, v[i]);
    }
  };
  auto half = [&](const W8& qs, uint32_t tag, uint32_t slot, float (*da)[IPL], float (*db)[IPL], int h) {
#pragma unroll
    for (int u = h * H; u < h * H + H; ++u) {
      const float q = __uint_as_float(qs[u]);
#pragma unroll
      for (int i = 0; i < IPL; ++i) {
        if constexpr (SEMI == kLogSum) {
          // semiring_mad(x, y, z) = log_add(x, y + z) (RuleSemiring.scala:58-64)
          acc[i] = UNARY ? log_add(acc[i], da[u][i] + q) : log_add(acc[i], (da[u][i] + db[u][i]) + q);
        } else if constexpr (UNARY) {
          acc[i] = SEMI == kViterbi ? fmaxf(acc[i], da[u][i] + q) : fmaf(da[u][i], q, acc[i]);
        } else {
          acc[i] = SEMI == kViterbi ? fmaxf(acc[i], (da[u][i] + db[u][i]) + q)
                                    : fmaf(da[u][i] * db[u][i], q, acc[i]);
        }
      }
    }
    if (tag & kSlotEnd) {
      flush(slot, (tag >> kTagTargetShift) & kTagTargetMask);
    }
  };
  // Software pipeline over groups: while group gi is computed from (da, db) with weights cur_q
  // and tags (cur_t0, cur_t1), the LDS reads of group g1 (offsets nxt_off) go to (la, lb), and the
  // offsets of the group after g1 and the weights (and, in skip layouts, slot ids) of g1 are
  // loaded into SGPRs. Scalar loads share the LDS wait counter, so the one wait the compiler
  // places before the next step's LDS reads covers both. Register sets alternate between two
  // unrolled steps, so nothing is copied.
  auto step = [&](const W8& cur_q, const W2& cur_s, uint32_t cur_t0, uint32_t cur_t1, const W16& nxt_off,
                  uint32_t& nxt_t0, uint32_t& nxt_t1, W16& nn_off, W8& nxt_q, W2& nxt_s,
                  float (*da)[IPL], float (*db)[IPL], float (*la)[IPL], float (*lb)[IPL]) {
    // Wait for the previous step's LDS reads and scalar loads before this step's reads issue; a
    // wait the compiler places later (before the first use of da/db) would drain them as well.
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    read_group(nxt_off, la, lb);
    nxt_t0 = nxt_off[H - 1];
    nxt_t1 = nxt_off[G - 1];
    const uint32_t g2 = SKIP ? advance(g1) : g1 + GB;
    nn_off = group_off(g2);
    nxt_q = group_q(g1);
    if constexpr (SKIP) nxt_s = group_slots(g1);
    __builtin_amdgcn_sched_barrier(0);
    half(cur_q, cur_t0, cur_s.x, da, db, 0);
    half(cur_q, cur_t1, cur_s.y, da, db, 1);
    __builtin_amdgcn_sched_barrier(0);
    gi = g1;
    g1 = g2;
  };
  // Skip layouts end at the null sentinel group; plain layouts at the wave's range end (the
  // groups after it, read but never used, belong to the next wave or to the null tail).
  auto done = [&]() { return SKIP ? gi == sentinel : gi >= g_end; };
  W16 oA = group_off(gi), oB = group_off(g1);
  W8 qA = group_q(gi), qB;
  W2 sA = {0u, 0u}, sB = {0u, 0u};
  if constexpr (SKIP) sA = group_slots(gi);
  uint32_t tA0 = oA[H - 1], tA1 = oA[G - 1], tB0 = 0, tB1 = 0;
  float a0[G][IPL], b0[G][IPL], a1[G][IPL], b1[G][IPL];
  read_group(oA, a0, b0);
  for (;;) {
    step(qA, sA, tA0, tA1, oB, tB0, tB1, oA, qB, sB, a0, b0, a1, b1);
    if (done()) break;
    step(qB, sB, tB0, tB1, oA, tA0, tA1, oB, qA, sA, a1, b1, a0, b0);
    if (done()) break;
  }
}

// Plain form: one launch per rule family and span, grid (tiles, chunks, pieces).
template <int IPL, int SEMI, bool UNARY, bool SKIP, bool FOLD, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_rules(FamArgs a) {
  extern __shared__ __align__(16) float smem[];
  __shared__ RulesLds<64 * IPL> sh;
  rules_unit<IPL, SEMI, UNARY, SKIP, false, FOLD, WAVES>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem, sh);
}

// Persistent form: the rule families of one span share one launch of about one workgroup per CU.
// Work units (family, tile, chunk, piece) form a queue, family-major; each workgroup pops the next
// unit from a device counter (zeroed before the launch) until the queue is empty. Plain streams
// run a tile's chunks and pieces back to back (chunk-fastest), so the tile's operand cells and
// folded partial rows are reused from cache while its units run together (exhaustive
// insideOutside step 711 -> 668 ms); block-major (masked) streams run a chunk's tiles back to back
// (tile-fastest), so consecutive units share the chunk's rule stream (the coarse-to-fine parse
// was 0.6% slower chunk-fastest). Planned families take their item counts, hence their tile
// counts, from the device, so the launch needs no host knowledge of the pruned work.
template <int IPL, int SEMI, bool UNARY, bool SKIP, bool FOLD, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_rules_q(QueueArgs q) {
  constexpr int ITEMS = 64 * IPL;
  extern __shared__ __align__(16) float smem[];
  __shared__ RulesLds<ITEMS> sh;
  __shared__ int32_t s_unit;
  // Tiles of family f (planned families read their item count on the device).
  auto tiles_of = [&](int f) {
    const FamArgs& a = q.fam[f];
    const int items = a.prefix != nullptr ? a.prefix[a.num_cells] - a.prefix[0] : a.num_items;
    return (items + ITEMS - 1) / ITEMS;
  };
  int total = 0;
  for (int f = 0; f < q.nfam; ++f) total += tiles_of(f) * q.fam[f].num_chunks * (1 << q.fam[f].pieces_log2);
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(q.counter, 1);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    __syncthreads();  // every thread has read s_unit before thread 0 writes the next one
    if (u >= total) break;
    int f = 0, r = u, tiles = tiles_of(0);
    for (;;) {
      const int n = tiles * q.fam[f].num_chunks * (1 << q.fam[f].pieces_log2);
      if (r < n || f + 1 == q.nfam) break;
      r -= n;
      tiles = tiles_of(++f);
    }
    const int nc = q.fam[f].num_chunks;
    int tile, chunk, piece;
    if constexpr (SKIP) {
      tile = r % tiles;
      const int cp = r / tiles;
      chunk = cp % nc;
      piece = cp / nc;
    } else {
      chunk = r % nc;
      const int rest = r / nc;
      piece = rest & ((1 << q.fam[f].pieces_log2) - 1);
      tile = rest >> q.fam[f].pieces_log2;
    }
    rules_unit<IPL, SEMI, UNARY, SKIP, false, FOLD, WAVES>(q.fam[f], tile, chunk, piece, smem, sh);
  }
}

// Mask plans (see PlanJob): the items each family must compute under the pruning masks, as the
// reference's update managers enumerate them (CLParser.scala:838-925): a binary item needs a
// target cell whose mask meets some block of the family (:845-855) and allowed child spans
// (:892-893); a unary item needs a non-empty top mask (:747). Counts, a stable sort of each job's
// cells by key, an exclusive scan per job in that order, then the item lists ((cell << 8) | split).
// The count and fill launches flatten every job's cells into one grid of kPlanThreads-thread
// blocks (job j owns blocks [block0_j, block0_{j+1})); sort and scan take one workgroup per job.
struct PlanView {
  int32_t *keys, *perm, *posof, *counts, *pre, *items;
  uint32_t* act;
};

__device__ inline PlanView plan_view(int32_t* mplan, const PlanJob& J) {
  const int64_t kc = J.num_cells;
  int32_t* b = mplan + J.mp;
  return PlanView{b, b + kc, b + 2 * kc, b + 3 * kc, b + 4 * kc, b + 9 * kc + 1,
                  reinterpret_cast<uint32_t*>(b + 5 * kc + 1)};
}

// Job owning flattened block `blk`: the last job whose block0 <= blk.
__device__ inline int plan_job_of(const PlanJob* jobs, int num_jobs, int blk) {
  int lo = 0, hi = num_jobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (jobs[mid].block0 <= blk) lo = mid; else hi = mid;
  }
  return lo;
}

template <bool FILL>
__global__ __launch_bounds__(kPlanThreads) void k_plan(PlanBatchArgs p) {
  const PlanJob& J = p.jobs[plan_job_of(p.jobs, p.num_jobs, blockIdx.x)];
  const PlanFamStatic& f = p.fams[J.fid];
  const PlanView v = plan_view(p.mplan, J);
  const int k = (blockIdx.x - J.block0) * kPlanThreads + threadIdx.x;
  if (k >= J.num_cells) return;
  const int packed = p.cells[J.cells_start + k];
  const int sent = packed >> 8, b = packed & 255, e = b + J.span;
  const int n = p.lengths[sent];
  int count = 0;
  int32_t* out = FILL ? v.items + v.pre[v.posof[k]] : nullptr;
  uint32_t act[4] = {0u, 0u, 0u, 0u};
  if (f.unary) {
    if (mask_any(p.masks + (size_t)cell_abs(p.cell_offsets, sent, n, 1, b, e) * p.mask_words, p.mask_words)) {
      if (FILL) out[0] = k << 8;
      count = 1;
      act[0] = 1u;
    }
  } else {
    const uint32_t* tm = p.masks + (size_t)cell_abs(p.cell_offsets, sent, n, f.t_half, b, e) * p.mask_words;
    bool active = false;
    for (int bb = 0; bb < f.num_blocks; ++bb) {
      uint32_t hit = 0u;
      for (int w = 0; w < p.mask_words; ++w) hit |= tm[w] & f.block_masks[bb * p.mask_words + w];
      if (hit) {
        act[bb >> 5] |= 1u << (bb & 31);
        active = true;
      }
    }
    if (active) {
      const int m = items_for(f.kind, b, e, n);
      for (int j = 0; j < m; ++j) {
        int ax0, ay0, bx0, by0;
        item_operands(f.kind, b, e, j, ax0, ay0, bx0, by0);
        const int ca = cell_abs(p.cell_offsets, sent, n, f.a_half, ax0, ay0);
        const int cb = cell_abs(p.cell_offsets, sent, n, f.b_half, bx0, by0);
        if (mask_any(p.masks + (size_t)ca * p.mask_words, p.mask_words) &&
            mask_any(p.masks + (size_t)cb * p.mask_words, p.mask_words)) {
          if (FILL) out[count] = (k << 8) | j;
          ++count;
        }
      }
    }
  }
  if (!FILL) {
    // Rule applications the reference semantics perform here: every rule of every active block
    // for every item (CLParser.scala:899-901).
    double rules = 0.0;
    if (count > 0) {
      if (f.unary) {
        rules = f.block_rules[0];
      } else {
        for (int bb = 0; bb < f.num_blocks; ++bb)
          if ((act[bb >> 5] >> (bb & 31)) & 1u) rules += f.block_rules[bb];
      }
    }
    double apps = rules * count;
    // One atomic per full wave instead of one per cell on the same address: the lanes' counts are
    // summed first (integers held exactly in doubles, so the order does not matter). A wave cut
    // short by the slice's end keeps one atomic per cell.
    double* counter = p.apps + J.apps_slot;
    if (__ballot(1) == ~0ull) {
      for (int m = 32; m >= 1; m >>= 1) apps += __shfl_xor(apps, m);
      if ((threadIdx.x & 63) == 0 && apps != 0.0) atomicAdd(counter, apps);
    } else if (apps != 0.0) {
      atomicAdd(counter, apps);
    }
    v.counts[k] = count;
    // A cell whose children all fail keeps no item: no block is active there.
#pragma unroll
    for (int w = 0; w < 4; ++w) v.act[(size_t)w * J.num_cells + k] = count ? act[w] : 0u;
    // Sort key: the family's largest blocks active here, the largest as the most significant bit.
    int key = 0;
    if (count > 0) {
#pragma unroll
      for (int i = 0; i < kPlanKeyBlocks; ++i) {
        const int bb = f.key_blocks[i];
        if (bb >= 0 && ((act[bb >> 5] >> (bb & 31)) & 1u)) key |= 1 << (kPlanKeyBlocks - 1 - i);
      }
    }
    v.keys[k] = key;
    // A family without key blocks keeps the slice order (its sort returns at once).
    if (f.key_blocks[0] < 0) {
      v.perm[k] = k;
      v.posof[k] = k;
    }
  }
}

// Stable sort of a job's cells by key (one workgroup per job): counting-sort passes over the key's
// 4-bit digits, low digit first (an LSD radix sort). Passes carry cell indices, alternating between
// posof (even passes write it) and perm (odd passes), and read each cell's digit from keys; the
// last pass writes perm, then posof is rebuilt as perm's inverse.
__global__ __launch_bounds__(512) void k_plan_sort(PlanBatchArgs p) {
  constexpr int T = 512, D = 16;
  constexpr int kPasses = kPlanKeyBlocks / 4;
  static_assert(kPlanKeyBlocks % 8 == 0, "an even number of 4-bit digits: the last pass writes perm");
  __shared__ int32_t s_a[D * T];
  __shared__ int32_t s_sum[T];
  const PlanJob& J = p.jobs[blockIdx.x];
  if (p.fams[J.fid].key_blocks[0] < 0) return;  // order already written by the count pass
  const PlanView v = plan_view(p.mplan, J);
  const int nc = J.num_cells;
  const int t = threadIdx.x;
  const int per = (nc + T - 1) / T;
  const int lo = min(t * per, nc), hi = min(lo + per, nc);
  int32_t* tmp = v.posof;
  for (int pass = 0; pass < kPasses; ++pass) {
    const int32_t* src = pass == 0 ? nullptr : ((pass & 1) ? tmp : v.perm);
    int32_t* dst = (pass & 1) ? v.perm : tmp;
    auto cell_at = [&](int i) -> int { return src != nullptr ? src[i] : i; };
    auto digit = [&](int c) -> int { return (v.keys[c] >> (4 * pass)) & (D - 1); };
    for (int b = 0; b < D; ++b) s_a[b * T + t] = 0;
    for (int i = lo; i < hi; ++i) ++s_a[digit(cell_at(i)) * T + t];
    __syncthreads();
    // Exclusive scan of the digit-major table: entry (d, t) becomes the sorted position of thread
    // t's first cell with digit d.
    int32_t local = 0;
    for (int i = 0; i < D; ++i) local += s_a[t * D + i];
    s_sum[t] = local;
    __syncthreads();
    for (int d = 1; d < T; d <<= 1) {
      const int32_t x = t >= d ? s_sum[t - d] : 0;
      __syncthreads();
      s_sum[t] += x;
      __syncthreads();
    }
    int32_t run = s_sum[t] - local;
    for (int i = 0; i < D; ++i) {
      const int32_t x = s_a[t * D + i];
      s_a[t * D + i] = run;
      run += x;
    }
    __syncthreads();
    for (int i = lo; i < hi; ++i) {
      const int c = cell_at(i);
      dst[s_a[digit(c) * T + t]++] = c;
    }
    __syncthreads();
  }
  for (int i = lo; i < hi; ++i) v.posof[v.perm[i]] = i;
}

// Exclusive scan of a job's item counts in its sorted cell order into pre[0, num_cells]; one
// workgroup per job.
__global__ __launch_bounds__(1024) void k_plan_scan(PlanBatchArgs p) {
  __shared__ int32_t s_sum[1024];
  const PlanJob& J = p.jobs[blockIdx.x];
  const PlanView v = plan_view(p.mplan, J);
  const int nc = J.num_cells;
  const int t = threadIdx.x;
  const int per = (nc + 1023) / 1024;
  const int lo = min(t * per, nc), hi = min(lo + per, nc);
  int32_t sum = 0;
  for (int i = lo; i < hi; ++i) sum += v.counts[v.perm[i]];
  s_sum[t] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int32_t x = t >= d ? s_sum[t - d] : 0;
    __syncthreads();
    s_sum[t] += x;
    __syncthreads();
  }
  int32_t run = s_sum[t] - sum;  // exclusive prefix of this thread's range
  for (int i = lo; i < hi; ++i) {
    v.pre[i] = run;
    run += v.counts[v.perm[i]];
  }
  if (t == 1023) v.pre[nc] = s_sum[1023];
}

// Combine partials: target[cell][t] = (+) over (family, slot) of t over the cell's items.
// One workgroup = 64 consecutive cells x TGT targets, one wave per target (TGT = 1..16). For each
// partial row feeding its target, a wave streams the items of its 64 cells (one contiguous range)
// through a per-wave LDS buffer with coalesced loads; each lane then folds its own cell's items.
// The block's results are transposed through LDS so that each cell's targets are written as one
// contiguous run of the chart row.
// Partial values a reduce wave keeps in flight within one family's run of a target's entries.
constexpr int kReduceBatch = 8;

constexpr int kReduceCap = 512;  // floats per wave buffer of the reduce

// One reduce unit: cells [64 bx, 64 bx + 64) of the slice x targets [TGT by, TGT by + TGT), by a
// workgroup of TGT waves (k_reduce's block, or with TGT = kReduceTargets a unit of the
// device-scheduled pass). LDS: s_buf (TGT x kReduceCap floats) and s_out (TGT x 65). COH: partial
// rows are read and the chart written with agent-coherent accesses (see ld_f).
template <int SEMI, int TGT, bool COH = false>
__device__ __forceinline__ void reduce_unit(const RedArgs& a, int bx, int by, float (*s_buf)[kReduceCap],
                                            float (*s_out)[65]) {
  constexpr int CAP = kReduceCap;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int k0 = bx * 64;
  const int kend = min(k0 + 64, a.num_cells);
  const int t = by * TGT + w;
  const int my_k = k0 + lane;  // position in the visiting order
  const int my_c = (a.perm != nullptr && my_k < kend) ? a.perm[my_k] : my_k;  // slice cell
  float acc = semi_zero<SEMI>();
  float* buf = s_buf[w];
  // Folded partial rows hold order-preserving encodings: every loaded value of family f is
  // decoded with that family's mask.
  auto ldp = [&](const float* p, int32_t dm) { return dec_ordered(ld_f<COH>(p), dm); };
  if (t < a.n_t && a.act[0] != nullptr) {
    // Masked (planned) groups: a cell's items are contiguous in the plan's order and few, so each
    // lane reads its own items straight from the partial rows, four slots of a family at a time.
    // The target's entries (family, slot) are read 64 at a time with one vector load each and taken
    // lane by lane.
    // Whether the target's block is active in the lane's cell, for every family at once (their
    // loads in flight together); a wave none of whose cells has the target active in any family
    // (most of them under pruning) is done after this one round of loads.
    uint32_t use_bits = 0u;
#pragma unroll
    for (int g = 0; g < kMaxGroupFamilies; ++g) {
      if (g < a.nfam && my_k < kend) {
        const int blk = a.block_of_target[g] != nullptr ? a.block_of_target[g][t] : 0;
        use_bits |= ((a.act[g][(size_t)(blk >> 5) * a.act_stride + my_c] >> (blk & 31)) & 1u) << g;
      }
    }
    const int e0 = a.off[t], ne = __ballot(use_bits != 0u) != 0ull ? a.off[t + 1] - e0 : 0;
    for (int b0 = 0; b0 < ne; b0 += 64) {
      const int cnt = min(64, ne - b0);
      const int li = e0 + b0 + min(lane, cnt - 1);
      const int lf = a.fam[li], ls = a.slot[li];
      for (int k = 0; k < cnt;) {
        const int f = __builtin_amdgcn_readlane(lf, k);
        int kr = k + 1;
        while (kr < cnt && __builtin_amdgcn_readlane(lf, kr) == f) ++kr;
        const bool use = (use_bits >> f) & 1u;
        if (__ballot(use) != 0ull && use) {
          const int32_t* pre = a.prefix[f];
          const int pos = a.posof[f] != nullptr ? a.posof[f][my_c] : my_k;
          const int lo = pre[pos], hi = pre[pos + 1];
          const size_t pitch = (size_t)a.pitch[f];
          const float* base = a.partial[f];
          const int32_t dm = a.dmask[f];
          int q = k;
          for (; q + 4 <= kr; q += 4) {
            const float* p0 = base + (size_t)__builtin_amdgcn_readlane(ls, q) * pitch;
            const float* p1 = base + (size_t)__builtin_amdgcn_readlane(ls, q + 1) * pitch;
            const float* p2 = base + (size_t)__builtin_amdgcn_readlane(ls, q + 2) * pitch;
            const float* p3 = base + (size_t)__builtin_amdgcn_readlane(ls, q + 3) * pitch;
            // Two items at a time: eight loads in flight, folded in (item, slot) order.
            int j = lo;
            for (; j + 2 <= hi; j += 2) {
              const float v0 = ldp(p0 + j, dm), v1 = ldp(p1 + j, dm), v2 = ldp(p2 + j, dm),
                          v3 = ldp(p3 + j, dm);
              const float w0 = ldp(p0 + j + 1, dm), w1 = ldp(p1 + j + 1, dm), w2 = ldp(p2 + j + 1, dm),
                          w3 = ldp(p3 + j + 1, dm);
              if (SEMI == kViterbi) {
                acc = fmaxf(fmaxf(acc, fmaxf(fmaxf(v0, v1), fmaxf(v2, v3))), fmaxf(fmaxf(w0, w1), fmaxf(w2, w3)));
              } else {
                acc = semi_add<SEMI>(acc, v0);
                acc = semi_add<SEMI>(acc, v1);
                acc = semi_add<SEMI>(acc, v2);
                acc = semi_add<SEMI>(acc, v3);
                acc = semi_add<SEMI>(acc, w0);
                acc = semi_add<SEMI>(acc, w1);
                acc = semi_add<SEMI>(acc, w2);
                acc = semi_add<SEMI>(acc, w3);
              }
            }
            if (j < hi) {
              const float v0 = ldp(p0 + j, dm), v1 = ldp(p1 + j, dm), v2 = ldp(p2 + j, dm),
                          v3 = ldp(p3 + j, dm);
              if (SEMI == kViterbi) {
                acc = fmaxf(acc, fmaxf(fmaxf(v0, v1), fmaxf(v2, v3)));
              } else {
                acc = semi_add<SEMI>(acc, v0);
                acc = semi_add<SEMI>(acc, v1);
                acc = semi_add<SEMI>(acc, v2);
                acc = semi_add<SEMI>(acc, v3);
              }
            }
          }
          for (; q < kr; ++q) {
            const float* p0 = base + (size_t)__builtin_amdgcn_readlane(ls, q) * pitch;
            for (int j = lo; j < hi; ++j) acc = semi_add<SEMI>(acc, ldp(p0 + j, dm));
          }
        }
        k = kr;
      }
    }
  } else if (t < a.n_t) {
    // The target's entries (family, slot) are read 64 at a time with one vector load each and
    // taken lane by lane; entries of one family are consecutive, so a family's fields and block
    // test are read once per run. Within a uniform family's run the (entry, split row) values are
    // loaded kReduceBatch at a time across entries and folded in entry order.
    const int e0 = a.off[t], ne = a.off[t + 1] - e0;
    for (int b0 = 0; b0 < ne; b0 += 64) {
      const int cnt = min(64, ne - b0);
      const int li = e0 + b0 + min(lane, cnt - 1);
      const int lf = a.fam[li], ls = a.slot[li];
      for (int k = 0; k < cnt;) {
        const int f = __builtin_amdgcn_readlane(lf, k);
        int kr = k + 1;
        while (kr < cnt && __builtin_amdgcn_readlane(lf, kr) == f) ++kr;
        bool use = true;
        if (a.act[f] != nullptr) {
          // All slots of the target in one family share its block: test it once per run, and skip
          // the run if no lane needs it.
          const int blk = a.block_of_target[f] != nullptr ? a.block_of_target[f][t] : 0;
          use = my_k < kend && ((a.act[f][(size_t)(blk >> 5) * a.act_stride + my_c] >> (blk & 31)) & 1u);
          if (__ballot(use) == 0ull) {
            k = kr;
            continue;
          }
        }
        const float* part = a.partial[f];
        const int32_t dm = a.dmask[f];
        const size_t pitch = (size_t)a.pitch[f];
        if (a.prefix[f] == nullptr) {
          // Block-major items: split row j of this block's 64 cells is item (block * uc + j) * 64 +
          // cell of the block; lanes read consecutive cells of one split row.
          const int uc = a.uniform_count[f];
          const int total = (kr - k) * uc;
          const bool live = my_k < kend && use;
          int ent = k, jj = 0;  // (entry, split row) of the next value to load
          // Batches of kReduceBatch values, loaded one batch ahead of the fold.
          auto load_batch = [&](int m, float* dst) {
#pragma unroll
            for (int u = 0; u < kReduceBatch; ++u) {
              dst[u] = semi_zero<SEMI>();
              if (m + u < total) {
                const int sl = __builtin_amdgcn_readlane(ls, ent);
                if (live) dst[u] = ldp(part + (size_t)sl * pitch + ((size_t)bx * uc + jj) * 64 + lane, dm);
                if (++jj == uc) {
                  jj = 0;
                  ++ent;
                }
              }
            }
          };
          float nxt[kReduceBatch];
          if (total > 0) load_batch(0, nxt);
          for (int m = 0; m < total; m += kReduceBatch) {
            float v[kReduceBatch];
#pragma unroll
            for (int u = 0; u < kReduceBatch; ++u) v[u] = nxt[u];
            if (m + kReduceBatch < total) load_batch(m + kReduceBatch, nxt);
            if (SEMI == kViterbi) {
#pragma unroll
              for (int u = 0; u < kReduceBatch; u += 2) acc = fmaxf(acc, fmaxf(v[u], v[u + 1]));
            } else {
#pragma unroll
              for (int u = 0; u < kReduceBatch; ++u)
                if (live && m + u < total) acc = semi_add<SEMI>(acc, v[u]);
            }
          }
          k = kr;
          continue;
        }
        int lo0, hi0, my_lo, my_hi;
        {
          const int32_t* pre = a.prefix[f];
          const int base = pre[0];
          lo0 = pre[k0] - base;
          hi0 = pre[kend] - base;
          my_lo = my_k < kend ? pre[my_k] - base : hi0;
          my_hi = my_k < kend ? pre[my_k + 1] - base : hi0;
        }
        // The run's (entry, piece) steps in order: piece q of entry e covers [lo0 + q CAP, ...) of
        // that entry's partial row. A piece's values are loaded into registers one step ahead, so
        // its global loads are in flight while the previous piece is folded from LDS.
        const int np = (hi0 - lo0 + CAP - 1) / CAP;
        const int steps = (kr - k) * np;
        float nxt[CAP / 64];
        auto load_step = [&](int t) {
          const int e = k + t / np;
          const int pb = lo0 + (t % np) * CAP;
          const int len = min(CAP, hi0 - pb);
          const float* p = part + (size_t)__builtin_amdgcn_readlane(ls, e) * pitch;
#pragma unroll
          for (int u = 0; u < CAP / 64; ++u) {
            const int i = lane + 64 * u;
            nxt[u] = i < len ? ldp(p + pb + i, dm) : 0.0f;
          }
        };
        if (steps > 0) load_step(0);
        for (int t = 0; t < steps; ++t) {
          const int pb = lo0 + (t % np) * CAP;
          const int len = min(CAP, hi0 - pb);
#pragma unroll
          for (int u = 0; u < CAP / 64; ++u) {
            const int i = lane + 64 * u;
            if (i < len) buf[i] = nxt[u];
          }
          // LDS operations of one wave complete in order; the fences keep the compiler from
          // moving the reads above the writes (or the next piece's writes above these reads).
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (t + 1 < steps) load_step(t + 1);
          const int jz = use ? min(my_hi, pb + len) : 0;
          for (int j = max(my_lo, pb); j < jz; ++j) acc = semi_add<SEMI>(acc, buf[j - pb]);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        k = kr;
      }
    }
  }
  s_out[w][lane] = acc;
  __syncthreads();
  // Write: thread i -> cell i / TGT, target i % TGT.
  const int tt = threadIdx.x % TGT;
  const int kk = threadIdx.x / TGT;
  const int t_out = by * TGT + tt;
  const int k = k0 + kk;
  if (k >= a.num_cells || t_out >= a.n_t) return;
  float v = s_out[tt][kk];
  const int packed = a.cell_list[a.perm != nullptr ? a.perm[k] : k];
  const int sent = packed >> 8, b = packed & 255, e = b + a.span;
  const int n = a.lengths[sent];
  if (a.root_init && b == 0 && e == n) {
    // setRootScores (CLParserUtils.scala:66-70) on the outside top root cell.
    v = (t_out == a.root) ? a.one : semi_zero<SEMI>();
  }
  if (a.act[0] != nullptr && v == semi_zero<SEMI>() && !(a.root_init && b == 0 && e == n)) return;
  const int c = cell_abs(a.cell_offsets, sent, n, a.t_half, b, e);
  st_f<COH>(a.chT + (size_t)c * a.cell_size + t_out, v);
}

template <int SEMI, int TGT>
__global__ __launch_bounds__(64 * TGT) void k_reduce(RedArgs a) {
  __shared__ float s_buf[TGT][kReduceCap];
  __shared__ float s_out[TGT][65];
  reduce_unit<SEMI, TGT>(a, blockIdx.x, blockIdx.y, s_buf, s_out);
}

// Device-scheduled pass (see PassStage). Waits are taken by thread 0 on device-scope counters. The
// data that crosses units (chart cells, partial rows) is written and read with agent-coherent
// accesses (rules_unit and reduce_unit with COH), so no unit needs an agent-scope fence, which on
// this GPU writes back or invalidates a whole L2 cache. A unit publishes by waiting in every wave
// until its stores are acknowledged (unit_done), then raising its counters from thread 0; a waiting
// unit's loads come after its wait and the workgroup barrier.
__device__ inline void pass_wait(int32_t* c, int32_t target, int32_t* err) {
  for (int it = 0;; ++it) {
    if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    if (it >= kPassSpinLimit) {
      __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(8);
  }
}

// Every wave: wait until this unit's memory operations have completed (stores acknowledged at the
// coherence point), then the workgroup barrier; only then does thread 0 raise the counters.
__device__ __forceinline__ void unit_done() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

template <int IPL, int SEMI, bool SKIP>
__global__ __launch_bounds__(kBlockThreads) void k_pass(PassArgs p) {
  constexpr int ITEMS = 64 * IPL;
  extern __shared__ __align__(16) float smem[];
  __shared__ RulesLds<ITEMS> sh;
  __shared__ int32_t s_unit;
  // The reduce's buffers live in the dynamic LDS the rule units stage operands in.
  float (*s_buf)[kReduceCap] = reinterpret_cast<float (*)[kReduceCap]>(smem);
  float (*s_out)[65] = reinterpret_cast<float (*)[65]>(smem + kReduceTargets * kReduceCap);
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(p.queue, 1);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    __syncthreads();  // every thread has read s_unit before thread 0 writes the next one
    if (u >= p.total_units) break;
    int lo = 0, hi = p.num_stages;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (p.stages[mid].first_unit <= u) lo = mid; else hi = mid;
    }
    const PassStage& st = p.stages[lo];
    int32_t* cnt = p.counters + st.cnt;
    int r = u - st.first_unit;
    if (st.kind == 0) {
      int f = 0;
      while (f + 1 < st.nfam && st.fam_units[f + 1] <= r) ++f;
      r -= st.fam_units[f];
      const FamArgs& fa = st.fam[f];
      const int tiles = st.fam_tiles[f];
      const int tile = r % tiles;
      const int cp = r / tiles;
      const int chunk = cp % fa.num_chunks;
      const int piece = cp / fa.num_chunks;
      const PassTile tt = p.tiles[st.tile_base[f] + tile];
      if (threadIdx.x == 0) {
        if (st.war_cnt >= 0) pass_wait(p.counters + st.war_cnt + 2 * st.war_nb + 1, st.war_units, p.error);
        if (st.dep_cnt >= 0) {
          int32_t* dc = p.counters + st.dep_cnt + st.dep_nb;
          if (tt.dep_lo == -2) {
            pass_wait(p.counters + st.dep_cnt + 2 * st.dep_nb + 1, st.dep_units, p.error);
          } else {
            for (int b = tt.dep_lo; b <= tt.dep_hi; ++b) pass_wait(dc + b, st.dep_gy, p.error);
          }
        }
      }
      __syncthreads();
      if (fa.kind == 7) rules_unit<IPL, SEMI, true, false, true>(fa, tile, chunk, piece, smem, sh);
      else if (SKIP && fa.skip_layout) rules_unit<IPL, SEMI, false, SKIP, true>(fa, tile, chunk, piece, smem, sh);
      else rules_unit<IPL, SEMI, false, false, true>(fa, tile, chunk, piece, smem, sh);
      unit_done();
      if (threadIdx.x == 0) {
        if (tt.out_lo >= 0)
          for (int b = tt.out_lo; b <= tt.out_hi; ++b) atomicAdd(cnt + b, 1);
        atomicAdd(cnt + 2 * st.nb, 1);
      }
    } else {
      const int bx = r / st.gy, by = r - (r / st.gy) * st.gy;
      if (threadIdx.x == 0) {
        if (st.exp_off >= 0) pass_wait(cnt + bx, p.expected[st.exp_off + bx], p.error);
        else pass_wait(cnt + 2 * st.nb, st.rule_units, p.error);
      }
      __syncthreads();
      reduce_unit<SEMI, kReduceTargets, true>(st.red, bx, by, s_buf, s_out);
      unit_done();
      if (threadIdx.x == 0) {
        atomicAdd(cnt + st.nb + bx, 1);
        atomicAdd(cnt + 2 * st.nb + 1, 1);
      }
    }
  }
}

__global__ void k_fill(float* p, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Folded partial rows of a slice start at the encoded semiring zero (FoldFillArgs): family
// blockIdx.y, rows strided over blockIdx.x, values over the threads.
__global__ __launch_bounds__(256) void k_fill_fold(FoldFillArgs f) {
  const int fi = blockIdx.y;
  const int len = f.pre[fi] != nullptr ? f.pre[fi][f.num_cells[fi]] : f.len[fi];
  uint32_t* base = reinterpret_cast<uint32_t*>(f.base[fi]);
  for (int r = blockIdx.x; r < f.rows[fi]; r += gridDim.x) {
    uint32_t* row = base + (size_t)r * f.pitch[fi];
    for (int i = threadIdx.x; i < len; i += blockDim.x) row[i] = kEncNegInf;
  }
}

// Span-1 tag scores (CLParser.scala:411-467): one thread per word position.
template <int SEMI>
__global__ void k_tags(float* inside, const int32_t* cell_offsets, const int32_t* lengths,
                       const int32_t* len_offsets, const int32_t* words, int32_t num_sent,
                       int32_t total_pos, const int32_t* lex_off, const int32_t* lex_tag,
                       const double* lex_logp, const float* in_scale, double* pos_scale,
                       int32_t cell_size) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total_pos) return;
  int lo = 0, hi = num_sent;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (len_offsets[mid] <= p) lo = mid; else hi = mid;
  }
  const int sent = lo, pos = p - len_offsets[sent], n = lengths[sent];
  const int cell = cell_offsets[sent] + chart_index(pos, pos + 1, n);
  float* dst = inside + (size_t)cell * cell_size;
  const int w = words[p];
  if (SEMI != kReal) {
    for (int e = lex_off[w]; e < lex_off[w + 1]; ++e) dst[lex_tag[e]] = (float)lex_logp[e];
  } else {
    const float my_scale = (float)exp(-(double)in_scale[cell]);
    float sum = 0.0f;
    for (int e = lex_off[w]; e < lex_off[w + 1]; ++e) {
      const float v = (float)exp((double)(float)lex_logp[e]) * my_scale;
      dst[lex_tag[e]] = v;
      sum = sum + v;
    }
    const float half = sum / 2;
    for (int e = lex_off[w]; e < lex_off[w + 1]; ++e) dst[lex_tag[e]] = dst[lex_tag[e]] / half;
    pos_scale[p] = log((double)half);
  }
}

// computeMasks (CLMaskKernels.scala:172-236): one wave per cell, bits OR-reduced across the wave.
// Sentence of a chart cell, for a whole wave (every lane active, `cell` wave-uniform): the largest s
// with cell_offsets[s] <= cell, by a 64-way ballot search, a few dependent loads where a binary
// search takes one per halving. Lane 0's probe (s = lo) always qualifies.
__device__ inline int sentence_of(const int32_t* cell_offsets, int num_sent, int cell, int lane) {
  int lo = 0, hi = num_sent;
  while (hi - lo > 1) {
    const int step = (hi - lo + 63) >> 6;
    const int sidx = lo + lane * step;
    const uint64_t m = __ballot(sidx < hi && cell_offsets[sidx] <= cell);
    lo += (63 - __clzll(m)) * step;
    hi = min(lo + step, hi);
  }
  return lo;
}

__global__ void k_masks(uint32_t* masks_out, const float* inside, const float* outside,
                        const int32_t* cell_offsets, const int32_t* lengths, int32_t num_sent,
                        int32_t num_cells, const int32_t* proj_t, const int32_t* proj_n,
                        int32_t num_syms, int32_t cell_size, int32_t root, float thresh,
                        int32_t mask_words) {
  const int cell = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (cell >= num_cells) return;
  const int sent = sentence_of(cell_offsets, num_sent, cell, lane);
  const int first = cell_offsets[sent], last = cell_offsets[sent + 1];
  const int n = lengths[sent];
  const float root_score = inside[(size_t)(last - 1) * cell_size + root];
  const float cutoff = root_score + thresh;
  const bool term = (cell - first) < n;
  const int32_t* proj = term ? proj_t : proj_n;
  const float* in = inside + (size_t)cell * cell_size;
  const float* out = outside + (size_t)cell * cell_size;
  uint32_t m[kMaxCoarseSymbols / 32] = {};
  for (int sym = lane; sym < num_syms; sym += 64) {
    const float score = in[sym] + out[sym];
    if (score >= cutoff) {
      const int field = proj[sym];
      m[field >> 5] |= 1u << (field & 31);
    }
  }
  for (int w = 0; w < mask_words; ++w) {
    uint32_t v = m[w];
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    if (lane == 0) masks_out[(size_t)cell * mask_words + w] = v;
  }
}

// getScalingConstants (CLScalingKernels.scala:51-69): scale = max over the cell's symbols (start
// -100000), 0 below -9000. One wave per cell: lanes read consecutive symbols (coalesced) and the
// maximum is combined across the wave; max is exact, so the order does not matter.
__global__ __launch_bounds__(256) void k_scaling(float* scale, const float* chart, int32_t num_syms, int32_t num_cells) {
  const int cell = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (cell >= num_cells) return;
  float m = -100000.0f;
  const float* c = chart + (size_t)cell * num_syms;
  for (int sym = lane; sym < num_syms; sym += 64) m = fmaxf(c[sym], m);
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (lane == 0) scale[cell] = m < -9000.0f ? 0.0f : m;
}

// computeMBR (CLMBRKernels.scala:118-171): per cell, the coarse marginals
// margs[c] = sum over the cell's symbols of coarse c of in / root * out (LogSum: exp(in - root + out),
// CLMBRKernels.scala:145-150), then the coarse symbol with the largest marginal, the first one on
// ties. One wave per cell: lane c sums coarse c's symbols in increasing symbol order (the order in
// which the reference's loop over symbols adds into margs[c]; symbols past the half's own space count
// as coarse 0, as their projection is padded with 0), and the wave picks the best (score, coarse).
__global__ __launch_bounds__(256) void k_mbr(MbrCell* out, const float* inside, const float* outside,
                                             const int32_t* cell_offsets, const int32_t* lengths, int32_t num_sent,
                                             int32_t num_cells, const int32_t* t_off, const int32_t* t_syms,
                                             const int32_t* n_off, const int32_t* n_syms, int32_t cell_size,
                                             int32_t root, int32_t num_coarse, int32_t num_syms, int32_t logsum) {
  // The symbol lists of both halves, once per block, and each wave's per-symbol terms, in LDS.
  __shared__ int32_t s_off[2][kMaxCoarseSymbols + 1];
  __shared__ int32_t s_syms[2][kMbrMaxSyms];
  __shared__ float s_rows[4][kMbrMaxCell];
  for (int i = threadIdx.x; i <= num_coarse; i += 256) {
    s_off[0][i] = t_off[i];
    s_off[1][i] = n_off[i];
  }
  for (int i = threadIdx.x; i < num_syms; i += 256) {
    s_syms[0][i] = t_syms[i];
    s_syms[1][i] = n_syms[i];
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  float* in = s_rows[wv];  // the cell's marginal term per symbol
  for (int j = 0; j < kMbrCellsPerBlock / 4; ++j) {
    const int cell = blockIdx.x * kMbrCellsPerBlock + 4 * j + wv;
    if (cell >= num_cells) return;
    const int sent = sentence_of(cell_offsets, num_sent, cell, lane);
    const int first = cell_offsets[sent], last = cell_offsets[sent + 1];
    const float root_score = inside[(size_t)(last - 1) * cell_size + root];
    const int half = (cell - first) < lengths[sent] ? 0 : 1;  // span-1 bottom cells hold tags
    const int32_t* off = s_off[half];
    const int32_t* syms = s_syms[half];
    // The previous cell's reads of the rows are done (LDS operations of a wave complete in order;
    // the fences keep the compiler from moving accesses across).
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Each symbol's marginal term, computed once across the lanes: in / root * out (Real) or
    // exp(in - root + out) (LogSum); the coarse sums below add them in symbol order.
    for (int i = lane; i < cell_size; i += 64) {
      const float x = inside[(size_t)cell * cell_size + i], y = outside[(size_t)cell * cell_size + i];
      in[i] = logsum ? expf(x - root_score + y) : x / root_score * y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float best = -INFINITY;
    int best_c = -1;
    for (int c = lane; c < num_coarse; c += 64) {
      float m = 0.0f;
      for (int k = off[c]; k < off[c + 1]; ++k) m += in[syms[k]];
      if (m > best) {  // lanes visit their coarse symbols in increasing order: first max wins
        best = m;
        best_c = c;
      }
    }
    for (int w = 32; w > 0; w >>= 1) {
      const float ob = __shfl_xor(best, w);
      const int oc = __shfl_xor(best_c, w);
      if (ob > best || (ob == best && oc >= 0 && (best_c < 0 || oc < best_c))) {
        best = ob;
        best_c = oc;
      }
    }
    if (lane == 0) out[cell] = MbrCell{best, best_c};
  }
}

// viterbi (CLViterbi.scala:495-605): top-down best tree, one wave per sentence. The reference scans
// candidates sequentially with '>=' (the last best candidate wins); here lanes scan strided subsets
// keeping their last best, and the wave picks the best score with ties going to the larger
// candidate index, which selects the same candidate.
struct Best {
  float s;
  int i;
};

__device__ inline Best better(Best a, Best b) {
  return (b.s > a.s || (b.s == a.s && b.i > a.i)) ? b : a;
}

__device__ inline Best wave_best(Best b) {
  for (int o = 32; o > 0; o >>= 1) {
    Best x{__shfl_xor(b.s, o), __shfl_xor(b.i, o)};
    b = better(b, x);
  }
  return b;
}

constexpr int kViterbiThreads = 256;

// Best of the workgroup's candidates: highest score, ties to the larger candidate index (the
// candidate the reference's sequential scan with >= keeps). Every thread returns the result.
__device__ inline Best block_best(Best b, Best* s_best) {
  b = wave_best(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_best[w] = b;
  __syncthreads();
  Best r = s_best[0];
  for (int k = 1; k < kViterbiThreads / 64; ++k) r = better(r, s_best[k]);
  __syncthreads();
  return r;
}

// Viterbi decode (CLViterbi.scala:364-608): one workgroup per sentence walks the tree top-down in
// pre-order; at each node its threads scan strided subsets of the candidates (unary rules of the
// node's top symbol, then NT, TN, TT and NN rules of its bottom symbol over the splits) and the
// workgroup keeps the best, ties to the larger candidate index -- the candidate the reference's
// sequential scan with >= selects.
__global__ __launch_bounds__(kViterbiThreads) void k_viterbi(
    TreeNode* tree_out, const int32_t* tree_offsets, const float* inside, const int32_t* cell_offsets,
    const int32_t* lengths, int32_t num_sent, int32_t cell_size, int32_t root, int32_t nn,
    const int32_t* offsets, const float* scores, const int32_t* lefts, const int32_t* rights) {
  constexpr int VT = kViterbiThreads;
  __shared__ int32_t s_top[2 * kMaxLength], s_bot[2 * kMaxLength], s_width[2 * kMaxLength];
  __shared__ float s_score[2 * kMaxLength];
  __shared__ Best s_best[VT / 64];
  const int sent = blockIdx.x;
  if (sent >= num_sent) return;
  const int tid = threadIdx.x;
  const int n = lengths[sent];
  const float* ibot = inside + (size_t)cell_offsets[sent] * cell_size;
  const float* itop = ibot + (size_t)tri(n) * cell_size;
  const int nodes = 2 * n - 1;
  for (int p = tid; p < nodes; p += VT) {
    s_top[p] = 0;
    s_bot[p] = 0;
    s_width[p] = 0;
    s_score[p] = 0.0f;
  }
  __syncthreads();
  if (tid == 0) {
    s_top[0] = root;
    s_width[0] = n;
  }
  __syncthreads();
  const int g_nn = 0, g_nt = nn, g_tn = 2 * nn, g_tt = 3 * nn, g_u = 4 * nn, g_ut = 5 * nn;
  int begin = 0;
  for (int p = 0; p < nodes; ++p) {
    const int width = s_width[p];
    const int end = begin + width;
    int bot = s_bot[p];
    const int top = s_top[p];
    if (top != -1) {
      // BestUnary (CLViterbi.scala:383-425): init -30000, sym 0.
      const int g = (width == 1) ? g_ut : g_u;
      const int r0 = offsets[g + top], r1 = offsets[g + top + 1];
      const float* cell = ibot + (size_t)chart_index(begin, end, n) * cell_size;
      Best b{-30000.0f, -1};
      for (int r = r0 + tid; r < r1; r += VT) {
        const float sc = scores[r] + cell[lefts[r]];
        if (sc >= b.s) {
          b.s = sc;
          b.i = r;
        }
      }
      b = block_best(b, s_best);
      bot = b.i >= 0 ? lefts[b.i] : 0;
      if (tid == 0) {
        s_bot[p] = bot;
        s_score[p] = b.s;
      }
    }
    if (width == 1) {
      if (tid == 0) s_score[p] = ibot[(size_t)cell_size * begin + bot];
      begin += 1;
    } else {
      const int nt0 = offsets[g_nt + bot], nt1 = offsets[g_nt + bot + 1];
      const int tn0 = offsets[g_tn + bot], tn1 = offsets[g_tn + bot + 1];
      const int tt0 = offsets[g_tt + bot], tt1 = offsets[g_tt + bot + 1];
      const int nn0 = offsets[g_nn + bot], nn1 = offsets[g_nn + bot + 1];
      const int c0 = nt1 - nt0;
      const int c1 = c0 + (tn1 - tn0);
      const int c2 = c1 + (width == 2 ? tt1 - tt0 : 0);
      const int nnn = nn1 - nn0;
      const float* nt_l = itop + (size_t)chart_index(begin, end - 1, n) * cell_size;
      const float* nt_r = ibot + (size_t)chart_index(end - 1, end, n) * cell_size;
      const float* tn_l = ibot + (size_t)chart_index(begin, begin + 1, n) * cell_size;
      const float* tn_r = itop + (size_t)chart_index(begin + 1, end, n) * cell_size;
      const float* tt_r = ibot + (size_t)chart_index(begin + 1, end, n) * cell_size;
      // Candidates c: [0, c0) NT, [c0, c1) TN, [c1, c2) TT, then c2 + split * nnn + rule for NN.
      Best b{-300000.0f, -1};
      for (int c = tid; c < c2; c += VT) {
        float sc;
        if (c < c0) {
          const int r = nt0 + c;
          sc = (scores[r] + nt_l[lefts[r]]) + nt_r[rights[r]];
        } else if (c < c1) {
          const int r = tn0 + (c - c0);
          sc = (scores[r] + tn_l[lefts[r]]) + tn_r[rights[r]];
        } else {
          const int r = tt0 + (c - c1);
          sc = (scores[r] + tn_l[lefts[r]]) + tt_r[rights[r]];
        }
        if (sc >= b.s) {
          b.s = sc;
          b.i = c;
        }
      }
      for (int si = 0; si < width - 1; ++si) {
        const int split = begin + 1 + si;
        const float* l = itop + (size_t)chart_index(begin, split, n) * cell_size;
        const float* rr = itop + (size_t)chart_index(split, end, n) * cell_size;
        for (int j = tid; j < nnn; j += VT) {
          const int r = nn0 + j;
          const float sc = (scores[r] + l[lefts[r]]) + rr[rights[r]];
          if (sc >= b.s) {
            b.s = sc;
            b.i = c2 + si * nnn + j;
          }
        }
      }
      b = block_best(b, s_best);
      if (tid == 0) {
        int config, best_split, left = 10, right = 10;
        if (b.i < 0) {
          config = 0;
          best_split = begin + 1;
        } else if (b.i < c0) {
          config = 1;
          best_split = end - 1;
          left = lefts[nt0 + b.i];
          right = rights[nt0 + b.i];
        } else if (b.i < c1) {
          config = 2;
          best_split = begin + 1;
          left = lefts[tn0 + b.i - c0];
          right = rights[tn0 + b.i - c0];
        } else if (b.i < c2) {
          config = 3;
          best_split = end - 1;
          left = lefts[tt0 + b.i - c1];
          right = rights[tt0 + b.i - c1];
        } else {
          const int cc = b.i - c2;
          const int si = cc / nnn;
          config = 0;
          best_split = begin + 1 + si;
          left = lefts[nn0 + cc - si * nnn];
          right = rights[nn0 + cc - si * nnn];
        }
        s_score[p] = b.s;
        const int lc = p + 1, rc = p + 2 * (best_split - begin);
        if ((config >> 1) == 0) {
          s_top[lc] = left;
        } else {
          s_top[lc] = -1;
          s_bot[lc] = left;
        }
        if ((config & 1) == 0) {
          s_top[rc] = right;
        } else {
          s_top[rc] = -1;
          s_bot[rc] = right;
        }
        s_width[lc] = best_split - begin;
        s_width[rc] = end - best_split;
      }
    }
    __syncthreads();
  }
  TreeNode* out = tree_out + tree_offsets[sent];
  for (int p = tid; p < nodes; p += VT) out[p] = TreeNode{s_top[p], s_bot[p], s_width[p], s_score[p]};
}

__global__ void k_get_root(float* out, const float* inside, const int32_t* cell_offsets,
                           int32_t num_sent, int32_t cell_size, int32_t root) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_sent) return;
  out[i] = inside[(size_t)(cell_offsets[i + 1] - 1) * cell_size + root];
}

// Every rule kernel instance of one (items per lane, semiring, fold) (plain and persistent forms, the
// unary, block-major and plain layouts), for setting their LDS limits.
template <int IPL, int S, bool FD>
void rule_kernels(std::vector<const void*>& v) {
  v.insert(v.end(), {(const void*)k_rules<IPL, S, false, false, FD, 16>, (const void*)k_rules<IPL, S, true, false, FD, 16>,
                     (const void*)k_rules<IPL, S, false, true, FD, 16>, (const void*)k_rules_q<IPL, S, false, false, FD, 16>,
                     (const void*)k_rules_q<IPL, S, true, false, FD, 16>, (const void*)k_rules_q<IPL, S, false, true, FD, 16>});
}
// The 8-wave instances (unmasked layers: plain and unary streams, two items per lane, no folding).
template <int S>
void rule_kernels_8(std::vector<const void*>& v) {
  v.insert(v.end(), {(const void*)k_rules<2, S, false, false, false, 8>, (const void*)k_rules<2, S, true, false, false, 8>,
                     (const void*)k_rules_q<2, S, false, false, false, 8>, (const void*)k_rules_q<2, S, true, false, false, 8>});
}

// Rule kernels: the semiring as above and, for the Viterbi semiring only, whether partial rows fold.
template <class F>
void launch_rules(int semi, bool fold, F&& f) {
  if (semi == kViterbi) {
    if (fold) f(std::integral_constant<int, kViterbi>{}, std::true_type{});
    else f(std::integral_constant<int, kViterbi>{}, std::false_type{});
  } else if (semi == kLogSum) {
    f(std::integral_constant<int, kLogSum>{}, std::false_type{});
  } else {
    f(std::integral_constant<int, kReal>{}, std::false_type{});
  }
}

template <class F>
void launch_semi(int semi, F&& f) {
  if (semi == kViterbi) f(std::integral_constant<int, kViterbi>{});
  else if (semi == kReal) f(std::integral_constant<int, kReal>{});
  else if (semi == kLogSum) f(std::integral_constant<int, kLogSum>{});
  else throw std::runtime_error("semiring not supported by the device engine");
}

}  // namespace

// ================================================================================================
// Layer: device tables for one compiled grammar.

template <class T>
T* Layer::upload(const std::vector<T>& v) {
  T* p = nullptr;
  const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
  HIP_CHECK(hipMalloc(&p, bytes));
  if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  owned_.push_back(p);
  return p;
}

Layer::Layer(Grammar gr, bool final_l, bool masked_l, int32_t items_per_lane)
    : g(std::move(gr)), final_layer(final_l), masked(masked_l), ipl(items_per_lane) {
  if (ipl != 1 && ipl != 2) throw std::runtime_error("items per lane must be 1 or 2");
  if ((int32_t)g.coarse_names.size() > kMaxCoarseSymbols)
    throw std::runtime_error("at most " + std::to_string(kMaxCoarseSymbols) + " coarse symbols are supported");
  {
    // A family's operands are two symbol spaces (one for unaries), so 2 max(#N, #T) rows bound
    // every family's single-chunk tile.
    const int stride = 64 * ipl + (ipl == 1 ? 1 : 2);
    const int rows = 2 * std::max(g.num_nonterms(), g.num_terms());
    if (ipl == 2 && !masked && rows * stride * 4 <= kLdsHalfBudgetBytes) {
      waves = 8;
      lds_budget = kLdsHalfBudgetBytes;
    } else {
      lds_budget = ipl == 1 ? kLdsBudgetBytes / 2 : kLdsBudgetBytes;
    }
  }
  for (int f = 0; f < kNumFamilies; ++f) build_family(f);
  build_reduce(kRInBin, {kFInNN, kFInNT, kFInTN}, g.num_nonterms());
  build_reduce(kRInBin2, {kFInNN, kFInNT, kFInTN, kFInTT}, g.num_nonterms());
  build_reduce(kRInUN, {kFInUN}, g.num_nonterms());
  build_reduce(kRInUT, {kFInUT}, g.num_nonterms());
  build_reduce(kROutBin, {kFOutRTN, kFOutLNT, kFOutLNN, kFOutRNN}, g.num_nonterms());
  build_reduce(kROutUN, {kFOutUN}, g.num_nonterms());
  build_reduce(kROutW1, {kFOutUT, kFOutLTT, kFOutRNT, kFOutLTN, kFOutRTT}, g.num_terms());
  d_dec_offsets = upload(g.dec_offsets);
  d_dec_scores = upload(g.dec_scores);
  d_dec_left = upload(g.dec_left);
  d_dec_right = upload(g.dec_right);
  d_lex_offsets = upload(g.lex_offsets);
  d_lex_tags = upload(g.lex_tags);
  d_lex_logp = upload(g.lex_logp);
  std::vector<PlanFamStatic> pst(kNumFamilies);
  for (int f = 0; f < kNumFamilies; ++f) {
    const DevFamily& F = fam[f];
    PlanFamStatic& ps = pst[f];
    ps.kind = F.kind;
    ps.unary = F.unary;
    ps.a_half = F.a_half;
    ps.b_half = F.b_half;
    ps.t_half = F.t_half;
    ps.num_blocks = F.unary ? 0 : F.num_blocks;
    ps.block_masks = F.unary ? nullptr : F.d_block_masks;
    ps.block_rules = F.d_block_rules;
    for (int i = 0; i < kPlanKeyBlocks; ++i) ps.key_blocks[i] = i < (int)F.h_key_blocks.size() ? F.h_key_blocks[i] : -1;
  }
  d_plan_static = upload(pst);
  std::vector<int32_t> pn = g.proj_n, pt = g.proj_t;
  pn.resize(g.num_syms, 0);
  pt.resize(g.num_syms, 0);
  d_proj_n = upload(pn);
  d_proj_t = upload(pt);
  // MBR symbol lists: per coarse symbol, the symbols projecting to it in increasing order, over the
  // padded projection (symbols past a half's own space project to 0).
  const int32_t nc = (int32_t)g.coarse_names.size();
  for (int half = 0; half < 2; ++half) {
    const std::vector<int32_t>& pr = half ? pn : pt;
    std::vector<int32_t> off(nc + 1, 0), syms(g.num_syms);
    for (int32_t sym = 0; sym < g.num_syms; ++sym) ++off[pr[sym] + 1];
    for (int32_t c = 0; c < nc; ++c) off[c + 1] += off[c];
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int32_t sym = 0; sym < g.num_syms; ++sym) syms[fill[pr[sym]]++] = sym;
    if (half) {
      d_mbr_n_off = upload(off);
      d_mbr_n_syms = upload(syms);
    } else {
      d_mbr_t_off = upload(off);
      d_mbr_t_syms = upload(syms);
    }
  }
}

Layer::~Layer() {
  for (void* p : owned_) (void)hipFree(p);
}

void Layer::build_family(int32_t id) {
  DevFamily& F = fam[id];
  F.id = id;
  const int nN = g.num_nonterms(), nT = g.num_terms();
  struct R4 { int32_t t, a, b; float q; };
  std::vector<R4> rules;
  const BlockPartition* bp = nullptr;
  bool t_term = false;
  auto binfam = [&](int f) { return g.binaries[f]; };
  if (id <= kFInTT) {
    const int f = id - kFInNN;
    for (auto& r : binfam(f)) rules.push_back({r.p, r.l, r.r, r.q});
    F.kind = (f == kNN) ? 0 : (f == kNT ? 1 : 2);
    F.range_split = f == kNN;
    F.t_half = kInBot;
    F.a_half = (f == kNN || f == kNT) ? kInTop : kInBot;
    F.b_half = (f == kNN || f == kTN) ? kInTop : kInBot;
    F.n_t = nN;
    F.n_a = (f == kTN || f == kTT) ? nT : nN;
    F.n_b = (f == kNT || f == kTT) ? nT : nN;
    bp = &g.inside_blocks[f];
  } else if (id <= kFOutLTT) {
    const int f = id - kFOutLNN;  // rotateLeftToParent: target = left child, A = parent, B = right
    for (auto& r : binfam(f)) rules.push_back({r.l, r.p, r.r, r.q});
    const bool range = (f == kNN || f == kTN);
    F.kind = range ? 3 : 4;
    F.range_split = range;
    t_term = (f == kTN || f == kTT);
    F.t_half = t_term ? kOutBot : kOutTop;
    F.a_half = kOutBot;
    F.b_half = (f == kNN || f == kTN) ? kInTop : kInBot;
    F.n_t = t_term ? nT : nN;
    F.n_a = nN;
    F.n_b = (f == kNT || f == kTT) ? nT : nN;
    bp = &g.outside_l_blocks[f];
  } else if (id <= kFOutRTT) {
    const int f = id - kFOutRNN;  // rotateRightToParent: target = right child, A = left, B = parent
    for (auto& r : binfam(f)) rules.push_back({r.r, r.l, r.p, r.q});
    const bool range = (f == kNN || f == kNT);
    F.kind = range ? 5 : 6;
    F.range_split = range;
    t_term = (f == kNT || f == kTT);
    F.t_half = t_term ? kOutBot : kOutTop;
    F.a_half = (f == kNN || f == kNT) ? kInTop : kInBot;
    F.b_half = kOutBot;
    F.n_t = t_term ? nT : nN;
    F.n_a = (f == kTN || f == kTT) ? nT : nN;
    F.n_b = nN;
    bp = &g.outside_r_blocks[f];
  } else {
    F.unary = true;
    F.kind = 7;
    F.range_split = 0;
    if (id == kFInUN || id == kFInUT) {
      const auto& us = id == kFInUN ? g.unary_n : g.unary_t;
      for (auto& u : us) rules.push_back({u.p, u.c, 0, u.q});
      F.t_half = kInTop;
      F.a_half = kInBot;
      F.n_t = nN;
      F.n_a = id == kFInUN ? nN : nT;
    } else {
      const auto& us = id == kFOutUN ? g.unary_n : g.unary_t;
      for (auto& u : us) rules.push_back({u.c, u.p, 0, u.q});  // rotateChildToParent
      t_term = id == kFOutUT;
      F.t_half = kOutBot;
      F.a_half = kOutTop;
      F.n_t = t_term ? nT : nN;
      F.n_a = nN;
    }
    F.n_b = 0;
    F.b_half = 0;
  }
  F.num_rules = (int32_t)rules.size();
  F.rule_apps_per_item = (double)rules.size();

  // Chunking of the operand symbol spaces so that one tile's rows fit the layer's LDS budget (one
  // 16-wave workgroup per CU with two items per lane; two workgroups per CU with one item per lane
  // or with 8 waves).
  const int stride = 64 * ipl + (ipl == 1 ? 1 : 2);
  const int max_rows = lds_budget / (stride * 4);
  int ca = 1, cb = 1;
  if (F.unary) {
    ca = (F.n_a + max_rows - 1) / max_rows;
  } else {
    int best = 1 << 30;
    for (int x = 1; x <= 64; ++x)
      for (int y = 1; y <= 64; ++y) {
        int ra = (F.n_a + x - 1) / x, rb = (F.n_b + y - 1) / y;
        if (ra + rb <= max_rows && x * y < best) {
          best = x * y;
          ca = x;
          cb = y;
        }
      }
  }
  ca = std::max(1, ca);
  const int ra = (std::max(1, F.n_a) + ca - 1) / ca;
  const int rb = F.unary ? 0 : (std::max(1, F.n_b) + cb - 1) / cb;
  F.max_rows = ra + rb;
  // Block gating tables (binary families): the block bit of each target symbol. Non-final layers
  // merge all kernels into one block whose mask is the first kernel's (CLParser.scala:783-785).
  std::vector<int32_t> bot(F.n_t, 0);
  if (!F.unary) {
    std::vector<uint32_t> masks;
    if (final_layer) {
      F.num_blocks = bp->num_blocks;
      masks = bp->masks;
      for (int t = 0; t < F.n_t; ++t) {
        const int c = t_term ? g.proj_t[t] : g.proj_n[t];
        bot[t] = std::max(0, bp->block_of_coarse[c]);
      }
    } else {
      F.num_blocks = bp->num_blocks > 0 ? 1 : 0;
      masks.assign(bp->masks.begin(), bp->masks.begin() + std::min<size_t>(bp->masks.size(), g.mask_words));
    }
    if (F.num_blocks > 128) throw std::runtime_error("more than 128 rule blocks in one family");
    masks.resize(std::max<size_t>(masks.size(), 1), 0u);
    F.d_block_of_target = upload(bot);
    F.d_block_masks = upload(masks);
  }

  if ((uint32_t)F.n_t > kTagTargetMask + 1u) throw std::runtime_error("more than 8,192 target symbols in one family");
  // Rule stream. Per chunk, the target symbols (slots) are assigned to the workgroup's waves by
  // longest-processing-time-first over their padded rule counts, and each wave streams its slots
  // back to back. A slot's rules are padded with null rules (semiring-zero weight) to a multiple of
  // kRuleUnroll; its last rule carries the slot-end tag and the slot id travels with the group
  // (see kGroupWords). In masked layers with several blocks the stream of each wave is block-major
  // (blocks padded to whole groups) and a per-wave table of block ranges lets a tile skip the
  // blocks that none of its cells allows.
  const int kWaves = waves;
  const bool block_major = masked && !F.unary && F.num_blocks > 1;
  F.skip_layout = block_major;
  F.num_block_ranges = block_major ? F.num_blocks : 1;
  const int gw = block_major ? kSkipGroupWords : kGroupWords;
  std::vector<uint32_t> packed;  // kGroupWords per group
  std::vector<int32_t> wave_blocks;
  const float null_q = zero_of(g.semiring);
  struct SRule { uint32_t x, y, q, tag; int32_t slot; };
  auto null_rule = [&]() {
    SRule u{0u, 0u, 0u, 0u, 0};
    std::memcpy(&u.q, &null_q, 4);
    return u;
  };
  F.h_chunk_slot_start.push_back(0);
  for (int ia = 0; ia < ca; ++ia)
    for (int ib = 0; ib < (F.unary ? 1 : cb); ++ib) {
      const int a_lo = ia * ra, a_hi = std::min(F.n_a, a_lo + ra);
      const int b_lo = F.unary ? 0 : ib * rb, b_hi = F.unary ? 0 : std::min(F.n_b, b_lo + rb);
      std::vector<R4> cr;
      for (auto& r : rules)
        if (r.a >= a_lo && r.a < a_hi && (F.unary || (r.b >= b_lo && r.b < b_hi))) cr.push_back(r);
      std::stable_sort(cr.begin(), cr.end(), [](const R4& x, const R4& y) {
        if (x.t != y.t) return x.t < y.t;
        if (x.a != y.a) return x.a < y.a;
        return x.b < y.b;
      });
      F.h_chunk_range.insert(F.h_chunk_range.end(), {a_lo, a_hi, b_lo, b_hi});
      const int rows_a = a_hi - a_lo;
      // Slots of this chunk: [first, last) ranges of cr with one target each.
      std::vector<std::pair<size_t, size_t>> slots;
      for (size_t i = 0; i < cr.size();) {
        size_t j = i;
        while (j < cr.size() && cr[j].t == cr[i].t) ++j;
        slots.push_back({i, j});
        i = j;
      }
      auto padded = [](size_t n) { return (n + kRuleUnroll - 1) / kRuleUnroll * kRuleUnroll; };
      const int nbr = F.num_block_ranges;
      // of_wave[w][b]: slots of block range b streamed by wave w.
      std::vector<std::vector<std::vector<size_t>>> of_wave(kWaves, std::vector<std::vector<size_t>>(nbr));
      // Block-major streams are balanced per block (a tile may run any subset of the blocks), ties
      // going to the wave with the least load over all blocks; otherwise over the whole stream.
      std::vector<int64_t> load(kWaves, 0);
      for (int br = 0; br < nbr; ++br) {
        std::vector<size_t> order;
        for (size_t k = 0; k < slots.size(); ++k)
          if (!block_major || bot[cr[slots[k].first].t] == br) order.push_back(k);
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
          return slots[x].second - slots[x].first > slots[y].second - slots[y].first;
        });
        std::vector<int64_t> bload(kWaves, 0);
        // Per slot the flush costs about two rule groups; count it in the balance.
        for (size_t k : order) {
          int w = 0;
          for (int v = 1; v < kWaves; ++v)
            if (std::make_pair(bload[v], load[v]) < std::make_pair(bload[w], load[w])) w = v;
          const int64_t c = (int64_t)padded(slots[k].second - slots[k].first) + 2 * kRuleUnroll;
          load[w] += c;
          if (block_major) bload[w] += c;
          of_wave[w][br].push_back(k);
        }
      }
      for (int w = 0; w < kWaves; ++w) {
        std::vector<SRule> stream;
        const int32_t first_slot = (int32_t)F.h_slot_target.size();
        for (int br = 0; br < nbr; ++br) {
          wave_blocks.push_back((int32_t)(packed.size() / gw + stream.size() / kGroupRules));
          for (size_t k : of_wave[w][br]) {
            const size_t i = slots[k].first, j = slots[k].second;
            const int t = cr[i].t;
            const int32_t slot_id = (int32_t)F.h_slot_target.size();
            F.h_slot_target.push_back(t);
            F.h_slot_block.push_back(bot[t]);
            for (size_t m = i; m < j; ++m) {
              SRule u{(uint32_t)((cr[m].a - a_lo) * stride * 4),
                      F.unary ? 0u : (uint32_t)((rows_a + cr[m].b - b_lo) * stride * 4), 0u, 0u, 0};
              std::memcpy(&u.q, &cr[m].q, 4);
              stream.push_back(u);
            }
            for (size_t m = j - i; m < padded(j - i); ++m) stream.push_back(null_rule());
            stream.back().tag = kSlotEnd | (uint32_t)t << kTagTargetShift;
            stream.back().slot = slot_id;
          }
          if (block_major || br == nbr - 1)
            while (stream.size() % kGroupRules) stream.push_back(null_rule());
        }
        const int32_t wave_g0 = (int32_t)(packed.size() / gw);
        const int32_t wave_len = (int32_t)(stream.size() / kGroupRules);
        wave_blocks.push_back(wave_g0 + wave_len);
        wave_blocks.push_back(first_slot);
        // Split tables: for P = 1, 2, 4, ..., kMaxPieces, P + 1 (group, first slot) points cutting
        // the wave's stream into P pieces at places no slot spans (after a group whose second half
        // ends a slot, at block starts, at the ends), each as close as possible to an even split.
        {
          std::vector<int32_t> cand{0};
          std::vector<int32_t> slots_before(wave_len + 1, 0);
          std::vector<char> is_block_start(wave_len + 1, 0);
          for (int br = 0; br < nbr; ++br) {
            const int32_t bs = wave_blocks[wave_blocks.size() - 2 - (nbr - br)] - wave_g0;
            if (bs >= 0 && bs <= wave_len) is_block_start[bs] = 1;
          }
          for (int32_t gg = 0; gg < wave_len; ++gg) {
            int32_t ends = 0;
            for (int u = 0; u < kGroupRules; ++u) ends += stream[(size_t)gg * kGroupRules + u].tag != 0;
            slots_before[gg + 1] = slots_before[gg] + ends;
            const bool last_ends = stream[(size_t)gg * kGroupRules + kGroupRules - 1].tag != 0;
            if (gg + 1 < wave_len && (last_ends || is_block_start[gg + 1])) cand.push_back(gg + 1);
          }
          if (wave_len > 0) cand.push_back(wave_len);
          for (int pp = 1; pp <= kMaxPieces; pp *= 2) {
            size_t ci = 0;
            for (int k = 0; k <= pp; ++k) {
              int32_t gpos = wave_len;
              if (k == 0) {
                gpos = 0;
              } else if (k < pp) {
                const int64_t target = ((int64_t)wave_len * k + pp / 2) / pp;
                while (ci + 1 < cand.size() && cand[ci + 1] <= target) ++ci;
                gpos = cand[ci];
              }
              wave_blocks.push_back(wave_g0 + gpos);
              wave_blocks.push_back(first_slot + slots_before[gpos]);
            }
          }
        }
        for (size_t g0 = 0; g0 < stream.size(); g0 += kGroupRules) {
          uint32_t words[kSkipGroupWords] = {};
          for (int u = 0; u < kGroupRules; ++u) {
            const SRule& r = stream[g0 + u];
            if (r.x > kOffsetMask || r.y > kOffsetMask) throw std::logic_error("LDS offset out of range");
            if (r.tag != 0 && u % kRuleUnroll != kRuleUnroll - 1)
              throw std::logic_error("slot end inside a half group");
            words[u] = r.x | r.tag;
            words[kGroupRules + u] = r.y;
            words[2 * kGroupRules + u] = r.q;
            if (r.tag && block_major) words[3 * kGroupRules + u / kRuleUnroll] = (uint32_t)r.slot;
          }
          packed.insert(packed.end(), words, words + gw);
        }
      }
      F.h_chunk_slot_start.push_back((int32_t)F.h_slot_target.size());
      ++F.num_chunks;
    }
  // The first group past the last wave's range is a null group: the rule loop parks its cursor
  // there when a wave runs out of work, and the two groups after it keep the prefetch in bounds.
  F.sentinel_group = (int32_t)(packed.size() / gw);
  packed.resize(packed.size() + 3 * gw, 0u);
  for (int k = 0; k < 3; ++k)
    for (int u = 0; u < kGroupRules; ++u)
      std::memcpy(&packed[(size_t)(F.sentinel_group + k) * gw + 2 * kGroupRules + u], &null_q, 4);
  F.num_slots = (int32_t)F.h_slot_target.size();
  // Folding pays where a target's rules spread over several operand chunks (the wsj_6 shape: 12-20
  // chunks); with one or two there is little to fold and atomics cost more than stores.
  F.foldable = !F.unary && F.num_chunks >= kMinFoldChunks;
  F.d_chunk_range = upload(F.h_chunk_range);
  F.d_slot_target = upload(F.h_slot_target);
  std::vector<double> block_rules(std::max(1, F.unary ? 1 : F.num_blocks), 0.0);
  for (auto& r : rules) block_rules[F.unary ? 0 : std::min(bot[r.t], (int)block_rules.size() - 1)] += 1.0;
  // The blocks with most rules key the mask plan's cell order.
  if (!F.unary && F.num_blocks > 1) {
    std::vector<int32_t> ord(block_rules.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int32_t)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return block_rules[x] > block_rules[y]; });
    ord.resize(std::min<size_t>(kPlanKeyBlocks, ord.size()));
    F.h_key_blocks = ord;
  }
  F.d_block_rules = upload(block_rules);
  F.d_rules = upload(packed);
  F.d_wave_blocks = upload(wave_blocks);
  F.d_slot_block = upload(F.h_slot_block);
}

void Layer::build_reduce(int32_t rid, std::vector<int32_t> fams, int32_t n_t) {
  DevReduce& R = red[rid];
  R.families = fams;
  R.n_t = n_t;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> per(n_t);
  for (size_t fi = 0; fi < fams.size(); ++fi) {
    const DevFamily& F = fam[fams[fi]];
    if (F.n_t != n_t) throw std::runtime_error("reduce group mixes target spaces");
    for (int s = 0; s < F.num_slots; ++s) per[F.h_slot_target[s]].push_back({(int32_t)fi, s});
  }
  std::vector<int32_t> off(1, 0), fv, sv;
  for (auto& v : per) {
    for (auto& pr : v) {
      fv.push_back(pr.first);
      sv.push_back(pr.second);
    }
    off.push_back((int32_t)fv.size());
  }
  R.d_off = upload(off);
  R.d_fam = upload(fv);
  R.d_slot = upload(sv);
  // Folding layers: a foldable family's slots of one target share one row, the target's.
  std::vector<int32_t> off_f(1, 0), fv_f, sv_f;
  for (int32_t t = 0; t < n_t; ++t) {
    int32_t last = -1;
    for (auto& pr : per[t]) {
      const bool f = fam[fams[pr.first]].foldable;
      if (f && pr.first == last) continue;
      fv_f.push_back(pr.first);
      sv_f.push_back(f ? t : pr.second);
      last = pr.first;
    }
    off_f.push_back((int32_t)fv_f.size());
  }
  R.d_off_f = upload(off_f);
  R.d_fam_f = upload(fv_f);
  R.d_slot_f = upload(sv_f);
}

// ================================================================================================
// Parser

struct Parser::Batch {
  int32_t n = 0, max_len = 0, num_cells = 0, total_len = 0;
  std::vector<int32_t> lengths, cell_offsets, len_offsets, words, tree_offsets;
  int32_t *d_lengths = nullptr, *d_cell_offsets = nullptr, *d_len_offsets = nullptr,
          *d_words = nullptr, *d_tree_offsets = nullptr, *d_cells = nullptr;
  std::vector<int32_t> span_start, span_count;  // cell list per span (index w)
  std::vector<int32_t> h_cells;
  // Item prefixes of variable-count families: (span, family) -> offset into d_prefix, -1 if none.
  std::vector<int64_t> prefix_off;
  int32_t* d_prefix = nullptr;
  std::vector<int32_t> h_prefix;
  std::vector<double> partition_scales;
  size_t apps_cap = 0;   // plan counters (doubles) of the batch; queue heads use 3x as many ints
  // Mask plans of the call's masked passes, in schedule order (prepare_plans); run_groups takes
  // them in turn from next_pass. Jobs of each masked layer are contiguous in the job table.
  struct PassPlan {
    int32_t li = 0;
    bool outside = false;
    std::vector<std::vector<PlannedSlice>> groups;  // per group of the pass (inside_groups order)
  };
  std::vector<PassPlan> pass_plans;
  size_t next_pass = 0;
  struct LayerJobs { int32_t li; int32_t first_job, num_jobs, total_blocks; bool any_key; };
  std::vector<LayerJobs> layer_jobs;
  std::vector<PlanJob> h_jobs;
  PlanJob* d_jobs = nullptr;
  // Device-scheduled passes, one entry per pass of the call in schedule order (on = false: the
  // pass keeps its per-span launches). Stages, tiles and expected counts of all passes travel in
  // one table (pass_), counters in pcnt_ ([0] is the error flag).
  struct DevPass {
    int32_t li = 0;
    bool outside = false;
    bool on = false;
    bool skip = false;          // some family streams block-major
    int32_t stage0 = 0, num_stages = 0, total_units = 0;
    int32_t cnt_off = 0, num_counters = 0;  // queue head at cnt_off, group counters after it
    double apps = 0;            // rule applications of unmasked families
    int32_t apps_first = -1, apps_count = 0;  // plan counters of masked families
  };
  std::vector<DevPass> dev_passes;
  size_t next_dev = 0;
  size_t tiles_byte_off = 0, exp_byte_off = 0;
  std::vector<char> h_pass_blob;
  int32_t num_pass_counters = 0;
  int32_t mask_in = 0;   // which half of masks_ holds the current input masks
  bool has_masks = false;
  bool has_scales = false;
};

namespace {
std::mutex g_scratch_mu;
std::map<int32_t, char*> g_scratch_ptr;

std::map<int32_t, std::string> g_scratch_err;

// Device address of the scratch array on `device`, looked up once per device with that device
// current; nullptr (and the reason in *err) if the runtime cannot give it.
char* scratch_address(int32_t device, std::string* err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *err = g_scratch_err[device];
    return it->second;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    g_scratch_err[device] = std::string("HIP error: ") + hipGetErrorString(e) + " at hipGetSymbolAddress(scratch)";
    p = nullptr;
  }
  g_scratch_ptr[device] = static_cast<char*>(p);
  *err = g_scratch_err[device];
  return static_cast<char*>(p);
}
}  // namespace

Parser::Parser(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
               int32_t items_per_lane)
    : device_(device), mem_bytes_(mem_bytes) {
  if (layers.empty()) throw std::runtime_error("parser needs at least one grammar layer");
  HIP_CHECK(hipSetDevice(device_));
  scratch_ = scratch_address(device_, &scratch_err_);
  for (size_t i = 0; i < layers.size(); ++i) {
    if (i + 1 < layers.size() && layers[i]->semiring != kViterbi)
      throw std::runtime_error("coarse layers must use the Viterbi semiring");
    if (layers[i]->mask_words != layers[0]->mask_words ||
        layers[i]->coarse_names != layers[0]->coarse_names)
      throw std::runtime_error("all layers must share the coarse symbol set");
    layers_.emplace_back(new Layer(*layers[i], i + 1 == layers.size(), i > 0, items_per_lane));
  }
  // Non-blocking: work of other parsers' streams (and of the null stream) never orders against
  // this one, so parsers on one device run concurrently.
  HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreate(&ev0_));
  HIP_CHECK(hipEventCreate(&ev1_));
  // Event pool: the no-alloc parse uses these and creates none.
  for (int i = 0; i < 64; ++i) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    ev_pool_.push_back(e);
  }
  // The rule kernels stage up to kLdsBudgetBytes of operand rows in dynamic LDS.
  const int lds = kLdsBudgetBytes;
  std::vector<const void*> fns;
  rule_kernels<2, kViterbi, false>(fns);
  rule_kernels<2, kViterbi, true>(fns);
  rule_kernels<2, kReal, false>(fns);
  rule_kernels<2, kLogSum, false>(fns);
  rule_kernels<1, kViterbi, false>(fns);
  rule_kernels<1, kViterbi, true>(fns);
  rule_kernels<1, kReal, false>(fns);
  rule_kernels<1, kLogSum, false>(fns);
  rule_kernels_8<kViterbi>(fns);
  rule_kernels_8<kReal>(fns);
  rule_kernels_8<kLogSum>(fns);
  for (const void* fn : fns) HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  for (const void* fn : {(const void*)k_pass<2, kViterbi, false>, (const void*)k_pass<2, kViterbi, true>,
                         (const void*)k_pass<2, kReal, false>, (const void*)k_pass<2, kReal, true>,
                         (const void*)k_pass<2, kLogSum, false>, (const void*)k_pass<2, kLogSum, true>,
                         (const void*)k_pass<1, kViterbi, false>, (const void*)k_pass<1, kViterbi, true>,
                         (const void*)k_pass<1, kReal, false>, (const void*)k_pass<1, kReal, true>,
                         (const void*)k_pass<1, kLogSum, false>, (const void*)k_pass<1, kLogSum, true>})
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  int cus = 0;
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_));
  num_cus_ = std::max(1, cus);
}

Parser::~Parser() {
  for (Buf* b : {&inside_, &outside_, &in_scale_, &out_scale_, &masks_, &partial_, &plan_, &tags_,
                 &trees_, &mbr_, &pos_scale_, &aux_, &mplan_, &apps_, &queue_, &jobs_, &pass_, &pcnt_})
    if (b->p) (void)hipFree(b->p);
  drop_graphs();
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  if (stream_) (void)hipStreamDestroy(stream_);
}

Parser::GraphEntry* Parser::graph_entry(const std::vector<int32_t>& lengths) {
  auto it = graphs_.find(lengths);
  if (it == graphs_.end()) {
    if (graphs_.size() >= kMaxGraphs) {
      for (auto& kv : graphs_)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      graphs_.clear();
    }
    it = graphs_.emplace(lengths, GraphEntry{}).first;
  }
  GraphEntry& e = it->second;
  if (e.exec != nullptr && e.generation != ws_generation_) {
    // A workspace buffer moved since the capture: capture again.
    HIP_CHECK(hipGraphExecDestroy(e.exec));
    e.exec = nullptr;
  }
  return &e;
}

void Parser::ensure(Buf& b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (b.bytes >= bytes) return;
  if (arena_.base != nullptr) {
    // No-alloc mode: carve from the scratch array (256-byte aligned, bump order).
    const size_t off = (arena_.used + 255) & ~(size_t)255;
    if (off + bytes > arena_.size)
      throw std::runtime_error("the no-alloc parse needs more than its " + std::to_string(arena_.size) +
                               " bytes of device scratch for this input");
    b.p = arena_.base + off;
    b.bytes = bytes;
    arena_.used = off + bytes;
    return;
  }
  // A graph being captured cannot take new memory; graphs captured before hold the old address.
  if (capturing_) throw CaptureAbort();
  ++ws_generation_;
  if (b.p) HIP_CHECK(hipFree(b.p));
  b.p = nullptr;
  b.bytes = 0;
  HIP_CHECK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
}

int64_t Parser::workspace_bytes() const {
  int64_t total = 0;
  for (const Buf* b : {&inside_, &outside_, &in_scale_, &out_scale_, &masks_, &partial_, &plan_, &tags_, &trees_,
                       &mbr_, &pos_scale_, &aux_, &mplan_, &apps_, &queue_, &jobs_, &pass_, &pcnt_})
    total += (int64_t)b->bytes;
  return total;
}

int64_t Parser::bytes_needed(const std::vector<int32_t>& lengths, int64_t partial_bytes) const {
  int32_t max_cell = 0;
  for (auto& l : layers_) max_cell = std::max(max_cell, l->g.cell_size);
  const int64_t mw = layers_[0]->g.mask_words;
  const bool masked = layers_.size() > 1;
  int64_t cells = 0, words = 0, plan_ints = 0, tables = 0;
  for (int32_t n64 : lengths) {
    const int64_t n = n64, c = n * (n + 1), tri_n = c / 2, c3 = (n + 1) * n * (n - 1) / 6;
    cells += c;
    words += n;
    tables += 4 * n + 2 + c / 2 + 8 * (tri_n + 1);
    if (masked) {
      // Per masked layer: six inside and ten outside plan jobs per span-cell at most, and every
      // item of their families (NN over all splits, the others one or a range per cell).
      const int64_t in_items = c3 + n * (n - 1) + (n - 1) + 2 * tri_n;
      const int64_t out_items = 2 * c3 + 2 * n * (n - 1) + 2 * n + c;
      plan_ints += (int64_t)(layers_.size() - 1) * (9 * 16 * tri_n + in_items + out_items);
    }
  }
  const int64_t partial = partial_bytes > 0    ? partial_bytes
                          : partial_bytes_ > 0 ? partial_bytes_
                                               : std::max<int64_t>((int64_t)(mem_bytes_ * 0.3), (int64_t)64 << 20);
  return cells * (2 * (int64_t)max_cell * 4 + 8 + 8 * mw + (int64_t)sizeof(MbrCell)) + words * 8 + 4 * plan_ints +
         4 * tables + partial + ((int64_t)16 << 20);
}

// WorkSpace.getBatches (WorkSpace.scala:20-41): close a batch when its chart cells exceed the
// chart budget.
std::vector<std::pair<int32_t, int32_t>> Parser::make_batches(const int32_t* offsets, int32_t n) {
  int32_t max_cell = 0;
  for (auto& l : layers_) max_cell = std::max(max_cell, l->g.cell_size);
  // Two charts per layer pass plus the rule partial buffer; keep 60% of the budget for charts.
  const int64_t chart_budget = (int64_t)(mem_bytes_ * 0.6) / 2;
  const int64_t max_cells = std::max<int64_t>(1, chart_budget / ((int64_t)max_cell * 4));
  std::vector<std::pair<int32_t, int32_t>> out;
  int32_t start = 0;
  int64_t cells = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t len = offsets[i + 1] - offsets[i];
    const int64_t c = (int64_t)len * (len + 1);
    if (i > start && cells + c > max_cells) {
      out.push_back({start, i - start});
      start = i;
      cells = 0;
    }
    cells += c;
  }
  if (n > start) out.push_back({start, n - start});
  return out;
}

// Phase timing with pooled events, resolved once per batch in flush_timing (no host syncs).
void Parser::timed_begin() {
  if (capturing_) return;
  phase_start_ = next_event();
  HIP_CHECK(hipEventRecord(phase_start_, stream_));
}
void Parser::timed_end(double* acc) {
  if (capturing_) return;
  hipEvent_t e = next_event();
  HIP_CHECK(hipEventRecord(e, stream_));
  phases_.push_back(Phase{phase_start_, e, acc});
}

namespace {
bool uniform_kind(int kind) { return kind <= 2 || kind == 7; }
// Items of a block-major uniform family over kc cells with `rows` rows (splits or split pairs) per
// cell: whole 64-cell blocks.
int64_t uniform_items(int64_t kc, int64_t rows) { return (kc + 63) / 64 * 64 * rows; }
// Unplanned inside NN launches with two items per lane fold split pairs in the rule kernel when
// the splits pair up without a lone last split or that lone split is a small share of the work.
bool pairs_splits(const Layer& L, const DevFamily& F, int span) {
  const int splits = span - 1;
  // Block-major streams (masked layers) never pair: their kernel variant writes one value per item.
  return L.ipl == 2 && F.kind == 0 && !F.skip_layout && splits >= 2 && (splits % 2 == 0 || splits >= 15);
}
}  // namespace

bool Parser::family_args(Batch& B, int32_t li, int32_t fid, int32_t span, bool use_masks, int32_t k0,
                         int32_t kc, float* partial, const int32_t* plan_counts, const int32_t* plan_items,
                         const uint32_t* plan_act, FamArgs* out, dim3* grid, size_t* lds) {
  Layer& L = *layers_[li];
  const DevFamily& F = L.fam[fid];
  const Grammar& g = L.g;
  FamArgs a{};
  float* inside = (float*)inside_.p;
  float* outside = (float*)outside_.p;
  a.chA = F.a_half >= 2 ? outside : inside;
  a.chB = F.b_half >= 2 ? outside : inside;
  float* isc = (float*)in_scale_.p;
  float* osc = (float*)out_scale_.p;
  a.scT = F.t_half >= 2 ? osc : isc;
  a.scA = F.a_half >= 2 ? osc : isc;
  a.scB = F.b_half >= 2 ? osc : isc;
  a.masks = use_masks ? (const uint32_t*)masks_.p + (size_t)B.mask_in * B.num_cells * g.mask_words : nullptr;
  a.cell_offsets = B.d_cell_offsets;
  a.lengths = B.d_lengths;
  a.cell_list = B.d_cells + B.span_start[span] + k0;
  a.num_cells = kc;
  int64_t nitems;
  if (use_masks && plan_items == nullptr) throw std::logic_error("masked rule launches need a mask plan");
  a.plan_items = plan_items;
  a.plan_act = plan_act;
  a.slice_cell0 = 0;  // planned item codes are slice-relative
  if (plan_items != nullptr) {
    // Planned slice: its item count is on the device; the grid covers the unpruned count.
    a.prefix = plan_counts;
    a.uniform_count = 1;
    if (uniform_kind(F.kind)) {
      nitems = (int64_t)kc * (F.kind == 0 ? span - 1 : 1);
    } else {
      const int64_t po = B.prefix_off[(size_t)span * kNumFamilies + fid];
      nitems = B.h_prefix[po + k0 + kc] - B.h_prefix[po + k0];
    }
  } else if (uniform_kind(F.kind)) {
    a.prefix = nullptr;
    a.uniform_count = F.kind == 0 ? span - 1 : 1;
    nitems = uniform_items(kc, a.uniform_count);
    if (pairs_splits(L, F, span)) {
      // Items 2m, 2m+1 are splits 2p, 2p+1 of one cell; a partial row holds one value per pair.
      a.pair_splits = 1;
      a.uniform_count = span / 2;  // pairs: ceil((span - 1) / 2)
      nitems = 2 * uniform_items(kc, a.uniform_count);
    }
  } else {
    const int64_t po = B.prefix_off[(size_t)span * kNumFamilies + fid];
    a.prefix = B.d_prefix + po + k0;
    nitems = B.h_prefix[po + k0 + kc] - B.h_prefix[po + k0];
    a.uniform_count = 1;
  }
  if (nitems <= 0) return false;
  a.num_items = (int32_t)nitems;
  a.chunk_range = F.d_chunk_range;
  a.wave_blocks = F.d_wave_blocks;
  a.num_block_ranges = F.num_block_ranges;
  a.sentinel_group = F.sentinel_group;
  a.rules = F.d_rules;
  a.block_of_target = F.unary ? nullptr : F.d_block_of_target;
  a.partial = partial;
  a.fold = folds(L) && F.foldable ? 1 : 0;
  const int items_tile = 64 * L.ipl;
  a.pitch = (int32_t)(((a.pair_splits ? nitems / 2 : nitems) + 255) / 256 * 256);
  a.span = span;
  a.kind = F.kind;
  a.range_split = F.range_split;
  a.cell_size = g.cell_size;
  a.a_half = F.a_half;
  a.b_half = F.b_half;
  a.t_half = F.t_half;
  a.row_stride = 64 * L.ipl + (L.ipl == 1 ? 1 : 2);
  a.num_chunks = F.num_chunks;
  a.skip_layout = F.skip_layout ? 1 : 0;
  // Few tiles: cut the rule streams into pieces so that the launch still fills the chip (about
  // two workgroups per CU); large launches run whole streams.
  const int64_t tiles = (nitems + items_tile - 1) / items_tile;
  int plog = 0;
  while (plog < kLog2MaxPieces && tiles * F.num_chunks * (2 << plog) <= piece_target_) ++plog;
  a.pieces_log2 = plog;
  *grid = dim3((unsigned)tiles, (unsigned)F.num_chunks, 1u << plog);
  *lds = (size_t)F.max_rows * a.row_stride * 4;
  *out = a;
  return true;
}

void Parser::launch_family(Batch& B, int32_t li, int32_t fid, int32_t span, bool use_masks,
                           int32_t k0, int32_t kc, float* partial, const int32_t* plan_counts,
                           const int32_t* plan_items, const uint32_t* plan_act, int32_t apps_slot) {
  Layer& L = *layers_[li];
  const DevFamily& F = L.fam[fid];
  FamArgs a;
  dim3 grid;
  size_t lds;
  if (!family_args(B, li, fid, span, use_masks, k0, kc, partial, plan_counts, plan_items, plan_act, &a, &grid, &lds))
    return;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (timing) {
    t0 = next_event();
    t1 = next_event();
    HIP_CHECK(hipEventRecord(t0, stream_));
  }
  launch_rules(L.g.semiring, a.fold != 0, [&](auto semi, auto fold) {
    constexpr int S = decltype(semi)::value;
    constexpr bool FD = decltype(fold)::value;
    if (L.waves == 8) {
      // 8-wave layers are unmasked with one chunk per family: plain streams, no folding.
      if constexpr (!FD) {
        const dim3 blk(64 * 8);
        if (F.unary) hipLaunchKernelGGL((k_rules<2, S, true, false, false, 8>), grid, blk, lds, stream_, a);
        else hipLaunchKernelGGL((k_rules<2, S, false, false, false, 8>), grid, blk, lds, stream_, a);
      } else {
        throw std::logic_error("8-wave rule layers do not fold");
      }
    } else if (L.ipl == 2) {
      if (F.unary) hipLaunchKernelGGL((k_rules<2, S, true, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, a);
      else if (F.skip_layout) hipLaunchKernelGGL((k_rules<2, S, false, true, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, a);
      else hipLaunchKernelGGL((k_rules<2, S, false, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, a);
    } else {
      if (F.unary) hipLaunchKernelGGL((k_rules<1, S, true, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, a);
      else if (F.skip_layout) hipLaunchKernelGGL((k_rules<1, S, false, true, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, a);
      else hipLaunchKernelGGL((k_rules<1, S, false, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, a);
    }
  });
  HIP_CHECK(hipGetLastError());
  // Planned launches take their count from the plan (the first slice of a span carries the whole
  // span's count); others count every rule for every item.
  const double apps =
      plan_items != nullptr ? 0.0 : (double)family_items(B, li, fid, span, k0, kc) * F.rule_apps_per_item;
  stats.rule_apps += apps;
  if (!timing) return;
  HIP_CHECK(hipEventRecord(t1, stream_));
  pending_.push_back(TimedLaunch{t0, t1, apps, fid + 16 * li, plan_items != nullptr ? apps_slot : -1, 1});
  stats.rule_launches += 1;
}

// Persistent form of a reduce group's rule families: families sharing a layout (unary or binary,
// block-major or not) run as one k_rules_q launch over a device queue of work units.
void Parser::launch_queue(Batch& B, int32_t li, const std::vector<int32_t>& fids, int32_t span,
                          bool use_masks, int32_t k0, int32_t kc, const std::vector<float*>& partials,
                          const std::vector<const int32_t*>& plan_counts,
                          const std::vector<const int32_t*>& plan_items,
                          const std::vector<const uint32_t*>& plan_acts, const std::vector<int32_t>& apps_slots) {
  Layer& L = *layers_[li];
  QueueArgs q{};
  size_t lds = 0;
  int64_t units = 0;
  double apps = 0;
  int32_t first_slot = -1, nslots = 0;
  for (size_t i = 0; i < fids.size(); ++i) {
    // Plan counters of the families (planned ones only), consecutive by construction.
    if (apps_slots[i] >= 0) {
      if (first_slot < 0) first_slot = apps_slots[i];
      if (apps_slots[i] != first_slot + nslots) throw std::logic_error("plan counters of a queue launch not contiguous");
      ++nslots;
    }
    dim3 grid;
    size_t l = 0;
    FamArgs a;
    if (!family_args(B, li, fids[i], span, use_masks, k0, kc, partials[i], plan_counts[i], plan_items[i],
                     plan_acts[i], &a, &grid, &l))
      continue;
    if (q.nfam == kMaxQueueFamilies) throw std::logic_error("too many families in one queue launch");
    q.fam[q.nfam++] = a;
    lds = std::max(lds, l);
    units += (int64_t)grid.x * grid.y * grid.z;
    if (plan_items[i] == nullptr)
      apps += (double)family_items(B, li, fids[i], span, k0, kc) * L.fam[fids[i]].rule_apps_per_item;
  }
  if (q.nfam == 0) return;
  if (units >= (int64_t)1 << 31) throw std::runtime_error("rule work queue exceeds 2^31 units");
  if ((size_t)(queue_next_ + 1) * 4 > queue_.bytes) throw std::runtime_error("rule queue counters exhausted");
  q.counter = (int32_t*)queue_.p + queue_next_++;
  const DevFamily& F0 = L.fam[fids[0]];
  // About one workgroup per CU (two for 8-wave layers).
  const dim3 grid((unsigned)std::min<int64_t>(units, (int64_t)num_cus_ * queue_wgs_per_cu_ * (16 / L.waves)));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (timing) {
    t0 = next_event();
    t1 = next_event();
    HIP_CHECK(hipEventRecord(t0, stream_));
  }
  launch_rules(L.g.semiring, q.fam[0].fold != 0, [&](auto semi, auto fold) {
    constexpr int S = decltype(semi)::value;
    constexpr bool FD = decltype(fold)::value;
    if (L.waves == 8) {
      if constexpr (!FD) {
        const dim3 blk(64 * 8);
        if (F0.unary) hipLaunchKernelGGL((k_rules_q<2, S, true, false, false, 8>), grid, blk, lds, stream_, q);
        else hipLaunchKernelGGL((k_rules_q<2, S, false, false, false, 8>), grid, blk, lds, stream_, q);
      } else {
        throw std::logic_error("8-wave rule layers do not fold");
      }
    } else if (L.ipl == 2) {
      if (F0.unary) hipLaunchKernelGGL((k_rules_q<2, S, true, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, q);
      else if (F0.skip_layout) hipLaunchKernelGGL((k_rules_q<2, S, false, true, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, q);
      else hipLaunchKernelGGL((k_rules_q<2, S, false, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, q);
    } else {
      if (F0.unary) hipLaunchKernelGGL((k_rules_q<1, S, true, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, q);
      else if (F0.skip_layout) hipLaunchKernelGGL((k_rules_q<1, S, false, true, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, q);
      else hipLaunchKernelGGL((k_rules_q<1, S, false, false, FD, 16>), grid, dim3(kBlockThreads), lds, stream_, q);
    }
  });
  HIP_CHECK(hipGetLastError());
  stats.rule_apps += apps;
  if (!timing) return;
  HIP_CHECK(hipEventRecord(t1, stream_));
  pending_.push_back(TimedLaunch{t0, t1, apps, fids[0] + 16 * li, first_slot, nslots});
  stats.rule_launches += 1;
}

hipEvent_t Parser::next_event() {
  if (ev_next_ == ev_pool_.size()) {
    if (arena_.base != nullptr) throw std::runtime_error("event pool exhausted in the no-alloc parse");
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    ev_pool_.push_back(e);
  }
  return ev_pool_[ev_next_++];
}

void Parser::mark_device_end() {
  if (dev_marked_) return;
  dev_marked_ = true;
  hipEvent_t e = next_event();
  HIP_CHECK(hipEventRecord(e, stream_));
  phases_.push_back(Phase{dev_start_, e, &stats.device_ms});
  dev_end_ = e;
}

void Parser::drop_graphs() {
  for (auto& kv : graphs_)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  graphs_.clear();
}

void Parser::flush_timing() {
  HIP_CHECK(hipStreamSynchronize(stream_));
  if (dev_pass_ran_) {
    dev_pass_ran_ = false;
    int32_t err = 0;
    HIP_CHECK(hipMemcpy(&err, pcnt_.p, 4, hipMemcpyDeviceToHost));
    if (err != 0) throw std::runtime_error("device-scheduled pass: a dependency wait timed out");
  }
  std::vector<double> plan_apps(apps_next_, 0.0);
  if (apps_next_ > 0 && !pending_.empty()) {
    HIP_CHECK(hipMemcpyAsync(plan_apps.data(), apps_.p, (size_t)apps_next_ * 8, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
  }
  for (const TimedLaunch& t : pending_) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
    double apps = t.apps;
    for (int32_t k = 0; t.apps_slot >= 0 && k < t.apps_count; ++k) {
      apps += plan_apps[t.apps_slot + k];
      stats.rule_apps += plan_apps[t.apps_slot + k];
    }
    stats.rule_kernel_ms += ms;
    stats.launch_ms.push_back(ms);
    stats.launch_apps.push_back(apps);
    stats.launch_family.push_back(t.family);
  }
  apps_next_ = 0;
  for (const Phase& ph : phases_) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ph.a, ph.b));
    *ph.acc += ms;
  }
  if (ref_event != nullptr && dev_start_ != nullptr && dev_end_ != nullptr) {
    float lo = 0, hi = 0;
    HIP_CHECK(hipEventElapsedTime(&lo, ref_event, dev_start_));
    HIP_CHECK(hipEventElapsedTime(&hi, ref_event, dev_end_));
    stats.span_lo_ms = std::min(stats.span_lo_ms, (double)lo);
    stats.span_hi_ms = std::max(stats.span_hi_ms, (double)hi);
  }
  dev_end_ = nullptr;
  pending_.clear();
  phases_.clear();
  ev_next_ = 0;
}

// Slices of a group's cells at one span whose partials fit the partial buffer (num_slots x pitch
// floats per family, pitch = items rounded up to 256).
std::vector<std::pair<int32_t, int32_t>> Parser::group_slices(Batch& B, int32_t li, int32_t rid, int32_t span,
                                                              int64_t budget_floats) {
  Layer& L = *layers_[li];
  const DevReduce& R = L.red[rid];
  const int32_t ncells = B.span_count[span];
  std::vector<std::pair<int32_t, int32_t>> out;
  if (budget_floats < 0) budget_floats = (int64_t)(partial_.bytes / 4);
  int64_t fixed = 0;
  // Partial rows per family: one per slot, or one per target symbol when the layer folds.
  const bool fold = folds(L);
  auto prow = [&](const DevFamily& F) { return (int64_t)(fold && F.foldable ? F.n_t : F.num_slots); };
  // Row padding: pitches round up to 256 values, block-major uniform rows to whole 64-cell blocks.
  for (int fid : R.families) {
    const DevFamily& F = L.fam[fid];
    const int64_t rows = uniform_kind(F.kind) ? (F.kind == 0 ? std::max(span - 1, 1) : 1) : 0;
    fixed += (256 + 64 * rows) * prow(F);
  }
  int32_t k0 = 0;
  while (k0 < ncells) {
    int32_t k1 = k0;
    int64_t used = 0;
    while (k1 < ncells) {
      int64_t add = 0;
      for (int fid : R.families) add += family_items(B, li, fid, span, k1, 1) * prow(L.fam[fid]);
      if (k1 > k0 && used + add + fixed > budget_floats) break;
      used += add;
      ++k1;
    }
    if (used + fixed > budget_floats) throw std::runtime_error("partial buffer too small for one cell");
    out.push_back({k0, k1 - k0});
    k0 = k1;
  }
  return out;
}

// Items of cells [k0, k0 + kc) of a span for one family, unpruned (an upper bound under masks).
int64_t Parser::family_items(Batch& B, int32_t li, int32_t fid, int32_t span, int32_t k0, int32_t kc) {
  const DevFamily& F = layers_[li]->fam[fid];
  if (uniform_kind(F.kind)) return (int64_t)kc * (F.kind == 0 ? span - 1 : 1);
  const int64_t po = B.prefix_off[(size_t)span * kNumFamilies + fid];
  return B.h_prefix[po + k0 + kc] - B.h_prefix[po + k0];
}

// Passes of a call in execution order (run_batch with `what`, parse_device for what = 0).
std::vector<std::pair<int32_t, bool>> Parser::pass_schedule(int32_t what, int32_t chart_layer) const {
  std::vector<std::pair<int32_t, bool>> s;
  const int32_t nl = (int32_t)layers_.size();
  const bool viterbi_final = layers_.back()->g.semiring == kViterbi;
  if (what == 0) {
    for (int32_t li = 0; li < nl; ++li) {
      s.push_back({li, false});
      if (li + 1 < nl || !viterbi_final || outside_final) s.push_back({li, true});
    }
    return s;
  }
  const int32_t last_layer = (what == 3) ? nl - 2 : nl - 1;
  for (int32_t li = 0; li <= last_layer; ++li) {
    s.push_back({li, false});
    if (li != nl - 1) {
      s.push_back({li, true});
      if (what == 4 && li == chart_layer) break;
      continue;
    }
    if (what == 1) break;
    if (!viterbi_final || what == 2 || what == 4) s.push_back({li, true});
  }
  return s;
}

// The (reduce group, span) list of a pass in schedule order (CLParser.scala:256-369): inside, TU at
// span 1 then per span the binary families and NU; outside, per span from the root down the binary
// families then NU (at span 1 TU with the terminal binary families).
std::vector<std::pair<int32_t, int32_t>> Parser::inside_groups(const Batch& B) {
  std::vector<std::pair<int32_t, int32_t>> groups{{kRInUT, 1}};
  for (int w = 2; w <= B.max_len; ++w) {
    groups.push_back({w == 2 ? kRInBin2 : kRInBin, w});
    groups.push_back({kRInUN, w});
  }
  return groups;
}

std::vector<std::pair<int32_t, int32_t>> Parser::outside_groups(const Batch& B) {
  std::vector<std::pair<int32_t, int32_t>> groups;
  for (int w = B.max_len; w >= 1; --w) {
    groups.push_back({kROutBin, w});
    groups.push_back({w >= 2 ? kROutUN : kROutW1, w});
  }
  return groups;
}

// Lays out one plan job per (masked pass, group, slice, family): its arrays in the plan buffer,
// its blocks in its layer's flattened count and fill grids, and its plan counter. Every layer but
// the first runs under masks; a layer's jobs are planned together at the start of its inside pass.
void Parser::prepare_plans(Batch& B, const std::vector<std::pair<int32_t, bool>>& schedule) {
  B.pass_plans.clear();
  B.layer_jobs.clear();
  B.h_jobs.clear();
  B.next_pass = 0;
  struct Pending { size_t pass, group, slice, fi; int64_t mp; };
  std::vector<Pending> where;
  int64_t mp = 0;
  for (auto [li, outside] : schedule) {
    if (li == 0) continue;  // the first layer has no masks
    Layer& L = *layers_[li];
    if (B.layer_jobs.empty() || B.layer_jobs.back().li != li)
      B.layer_jobs.push_back({li, (int32_t)B.h_jobs.size(), 0, 0, false});
    auto& lj = B.layer_jobs.back();
    Batch::PassPlan pp;
    pp.li = li;
    pp.outside = outside;
    const auto groups = outside ? outside_groups(B) : inside_groups(B);
    pp.groups.resize(groups.size());
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const auto [rid, span] = groups[gi];
      if (B.span_count[span] == 0) continue;
      const DevReduce& R = L.red[rid];
      for (auto [k0, kc] : group_slices(B, li, rid, span)) {
        if (kc >= (1 << 23)) throw std::runtime_error("mask plan slice exceeds 2^23 cells");
        PlannedSlice ps;
        ps.k0 = k0;
        ps.kc = kc;
        ps.pfs.resize(R.families.size());
        ps.apps_slot.resize(R.families.size());
        for (size_t fi = 0; fi < R.families.size(); ++fi) {
          const int fid = R.families[fi];
          const DevFamily& F = L.fam[fid];
          PlanJob j{};
          j.fid = fid;
          j.num_cells = kc;
          j.span = span;
          j.cells_start = B.span_start[span] + k0;
          j.block0 = lj.total_blocks;
          j.apps_slot = (int32_t)B.h_jobs.size();
          j.mp = mp;
          lj.total_blocks += (kc + kPlanThreads - 1) / kPlanThreads;
          lj.any_key |= !F.h_key_blocks.empty();
          ps.apps_slot[fi] = j.apps_slot;
          where.push_back({B.pass_plans.size(), gi, pp.groups[gi].size(), fi, mp});
          mp += 9 * (int64_t)kc + 1 + std::max<int64_t>(family_items(B, li, fid, span, k0, kc), 1);
          B.h_jobs.push_back(j);
          ++lj.num_jobs;
        }
        pp.groups[gi].push_back(std::move(ps));
      }
    }
    B.pass_plans.push_back(std::move(pp));
  }
  if (B.h_jobs.empty()) return;
  ensure(mplan_, (size_t)mp * 4);
  ensure(jobs_, B.h_jobs.size() * sizeof(PlanJob));
  int32_t* base = (int32_t*)mplan_.p;
  for (const Pending& w : where) {
    PlannedSlice& ps = B.pass_plans[w.pass].groups[w.group][w.slice];
    const int64_t kc = ps.kc;
    int32_t* b = base + w.mp;
    PlanPtrs& pf = ps.pfs[w.fi];
    pf.perm = b + kc;
    pf.posof = b + 2 * kc;
    pf.pre = b + 4 * kc;
    pf.act = (uint32_t*)(b + 5 * kc + 1);
    pf.items = b + 9 * kc + 1;
  }
  // The reduce visits a slice's cells in the order of its group's largest family.
  for (Batch::PassPlan& pp : B.pass_plans) {
    const auto groups = pp.outside ? outside_groups(B) : inside_groups(B);
    const Layer& L = *layers_[pp.li];
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const DevReduce& R = L.red[groups[gi].first];
      int key_fam = 0;
      for (size_t fi = 1; fi < R.families.size(); ++fi)
        if (L.fam[R.families[fi]].num_rules > L.fam[R.families[key_fam]].num_rules) key_fam = (int)fi;
      for (PlannedSlice& ps : pp.groups[gi]) ps.d_perm = ps.pfs[key_fam].perm;
    }
  }
}

// Lays out the device-scheduled passes of a call (see PassStage) and their table.
void Parser::prepare_dev_passes(Batch& B, const std::vector<std::pair<int32_t, bool>>& schedule) {
  B.dev_passes.clear();
  B.next_dev = 0;
  B.h_pass_blob.clear();
  std::vector<PassStage> stages;
  std::vector<PassTile> tiles;
  std::vector<int32_t> expected;
  int32_t counters = 64;  // [0, 64): the error flag, on a cache line of its own
  size_t masked_pass = 0;
  for (auto [li, outside] : schedule) {
    Batch::DevPass dp;
    dp.li = li;
    dp.outside = outside;
    const bool use_masks = li > 0;
    const Batch::PassPlan* pp = use_masks ? &B.pass_plans[masked_pass++] : nullptr;
    if (device_schedule) {
      const size_t s0 = stages.size(), t0 = tiles.size(), e0 = expected.size();
      const int32_t c0 = counters;
      if (!build_dev_pass(B, li, outside, use_masks, pp, &stages, &tiles, &expected, &counters, &dp)) {
        stages.resize(s0);
        tiles.resize(t0);
        expected.resize(e0);
        counters = c0;
        dp = Batch::DevPass{};
        dp.li = li;
        dp.outside = outside;
      }
    }
    B.dev_passes.push_back(dp);
  }
  B.num_pass_counters = counters;
  if (stages.empty()) return;
  auto align = [](size_t x) { return (x + 255) / 256 * 256; };
  B.tiles_byte_off = align(stages.size() * sizeof(PassStage));
  B.exp_byte_off = align(B.tiles_byte_off + tiles.size() * sizeof(PassTile));
  B.h_pass_blob.assign(B.exp_byte_off + std::max<size_t>(expected.size(), 1) * 4, 0);
  std::memcpy(B.h_pass_blob.data(), stages.data(), stages.size() * sizeof(PassStage));
  if (!tiles.empty()) std::memcpy(B.h_pass_blob.data() + B.tiles_byte_off, tiles.data(), tiles.size() * sizeof(PassTile));
  if (!expected.empty()) std::memcpy(B.h_pass_blob.data() + B.exp_byte_off, expected.data(), expected.size() * 4);
}

// One pass as stages; false (nothing appended) when some (group, span) needs more than one slice of
// half the partial buffer or has more families than a stage holds.
bool Parser::build_dev_pass(Batch& B, int32_t li, bool outside, bool use_masks, const void* pass_plan,
                            std::vector<PassStage>* stages, std::vector<PassTile>* tiles,
                            std::vector<int32_t>* expected, int32_t* counters, void* dev_pass) {
  const Batch::PassPlan* pp = static_cast<const Batch::PassPlan*>(pass_plan);
  Batch::DevPass& dp = *static_cast<Batch::DevPass*>(dev_pass);
  Layer& L = *layers_[li];
  if (L.waves != 16) return false;  // k_pass runs 16-wave rule units
  const auto groups = outside ? outside_groups(B) : inside_groups(B);
  const int64_t half = (int64_t)(partial_.bytes / 4 / 2);
  for (auto [rid, span] : groups) {
    if (B.span_count[span] == 0) continue;
    if (L.red[rid].families.size() > (size_t)kMaxPassFamilies) return false;
    try {
      if (group_slices(B, li, rid, span, half).size() != 1) return false;
    } catch (const std::runtime_error&) {
      return false;
    }
  }
  // first_cell[w][i]: cells of span w that belong to sentences before i (cells are sentence-major).
  std::map<int32_t, std::vector<int32_t>> first_cell;
  auto firsts = [&](int32_t w) -> const std::vector<int32_t>& {
    auto it = first_cell.find(w);
    if (it != first_cell.end()) return it->second;
    std::vector<int32_t> f(B.n + 1, 0);
    for (int32_t i = 0; i < B.n; ++i) f[i + 1] = f[i] + std::max(0, B.lengths[i] - w + 1);
    return first_cell.emplace(w, std::move(f)).first->second;
  };
  struct Prev { int32_t cnt, nb, gy, dunits, span; };
  std::vector<Prev> prev;
  dp.on = true;
  dp.stage0 = (int32_t)stages->size();
  dp.cnt_off = *counters;
  int32_t cnt = *counters + 1;  // after the queue head
  int32_t unit = 0;
  int32_t ei = 0;  // index among the pass's groups with cells
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const auto [rid, span] = groups[gi];
    const int32_t kc = B.span_count[span];
    if (kc == 0) continue;
    const DevReduce& R = L.red[rid];
    const PlannedSlice* ps = use_masks ? &pp->groups[gi][0] : nullptr;
    GroupSlice gs;
    float* base = (float*)partial_.p + (size_t)(ei % 2) * (size_t)half;
    group_slice_args(B, li, rid, span, use_masks, ps, 0, kc, base, &gs);
    const int32_t nb = (kc + 63) / 64;
    const int32_t gy = (R.n_t + kReduceTargets - 1) / kReduceTargets;
    PassStage rs{};
    rs.kind = 0;
    rs.first_unit = unit;
    rs.cnt = cnt;
    rs.nb = nb;
    rs.gy = gy;
    rs.dep_cnt = rs.war_cnt = -1;
    if (ei >= 1) {
      const Prev& p = prev[ei - 1];
      rs.dep_cnt = p.cnt;
      rs.dep_nb = p.nb;
      rs.dep_gy = p.gy;
      rs.dep_units = p.dunits;
    }
    if (ei >= 2) {
      const Prev& p = prev[ei - 2];
      rs.war_cnt = p.cnt;
      rs.war_nb = p.nb;
      rs.war_units = p.dunits;
    }
    std::vector<int32_t> exp_blocks(nb, 0);
    int32_t ru = 0;
    for (int32_t fi = 0; fi < gs.nfam; ++fi) {
      if (!gs.valid[fi]) continue;
      const FamArgs& a = gs.fa[fi];
      const int32_t k = rs.nfam++;
      const int32_t T = (int32_t)gs.grid[fi].x, C = (int32_t)gs.grid[fi].y, P = (int32_t)gs.grid[fi].z;
      rs.fam[k] = a;
      rs.fam_units[k] = ru;
      rs.fam_tiles[k] = T;
      rs.tile_base[k] = (int32_t)tiles->size();
      ru += T * C * P;
      dp.skip |= a.skip_layout != 0;
      if (use_masks) {
        // Planned items follow the family's own cell order: tiles touch any block.
        for (int32_t t = 0; t < T; ++t) tiles->push_back(PassTile{-1, -1, ei >= 1 ? -2 : 1, ei >= 1 ? -2 : 0});
        const int32_t slot = gs.apps_slot[fi];
        if (slot >= 0) {
          if (dp.apps_first < 0) dp.apps_first = slot;
          dp.apps_count = std::max(dp.apps_count, slot - dp.apps_first + 1);
        }
        continue;
      }
      dp.apps += (double)family_items(B, li, gs.fid[fi], span, 0, kc) * L.fam[gs.fid[fi]].rule_apps_per_item;
      const int32_t ti = 64 * L.ipl;
      for (int32_t t = 0; t < T; ++t) {
        int32_t b_lo, b_hi, c_lo, c_hi;
        if (a.prefix == nullptr) {
          const int32_t uc = a.uniform_count;
          if (a.pair_splits) {
            b_lo = b_hi = t / uc;  // 64 pairs: one row of one block
          } else {
            b_lo = (t * L.ipl) / uc;
            b_hi = std::min((t * L.ipl + L.ipl - 1) / uc, nb - 1);
          }
          c_lo = b_lo * 64;
          c_hi = std::min(b_hi * 64 + 63, kc - 1);
        } else {
          const int64_t po = B.prefix_off[(size_t)span * kNumFamilies + gs.fid[fi]];
          const int32_t* pre = B.h_prefix.data() + po;
          const int32_t n_items = pre[kc] - pre[0];
          const int32_t first = t * ti, last = std::min(first + ti, n_items) - 1;
          c_lo = (int32_t)(std::upper_bound(pre, pre + kc + 1, pre[0] + first) - pre) - 1;
          c_hi = (int32_t)(std::upper_bound(pre, pre + kc + 1, pre[0] + last) - pre) - 1;
          b_lo = c_lo / 64;
          b_hi = c_hi / 64;
        }
        PassTile pt{b_lo, b_hi, 1, 0};
        if (ei >= 1) {
          const int32_t i_lo = B.h_cells[B.span_start[span] + c_lo] >> 8;
          const int32_t i_hi = B.h_cells[B.span_start[span] + c_hi] >> 8;
          const std::vector<int32_t>& f = firsts(prev[ei - 1].span);
          const int32_t c0 = f[i_lo], c1 = f[i_hi + 1];
          if (c0 < c1) {
            pt.dep_lo = c0 / 64;
            pt.dep_hi = (c1 - 1) / 64;
          }
        }
        tiles->push_back(pt);
        for (int32_t b = b_lo; b <= b_hi; ++b) exp_blocks[b] += C * P;
      }
    }
    rs.num_units = ru;
    rs.rule_units = ru;
    unit += ru;
    PassStage ds = rs;
    ds.kind = 1;
    ds.first_unit = unit;
    ds.num_units = nb * gy;
    ds.nfam = 0;
    ds.red = gs.ra;
    ds.exp_off = -1;
    if (!use_masks) {
      ds.exp_off = (int32_t)expected->size();
      expected->insert(expected->end(), exp_blocks.begin(), exp_blocks.end());
    }
    unit += nb * gy;
    if (ru > 0) stages->push_back(rs);
    stages->push_back(ds);
    prev.push_back(Prev{cnt, nb, gy, nb * gy, span});
    cnt += 2 * nb + 2;
    ++ei;
  }
  dp.num_stages = (int32_t)stages->size() - dp.stage0;
  dp.total_units = unit;
  dp.num_counters = cnt - dp.cnt_off;
  *counters = cnt;
  return true;
}

// Runs the next pass of the call as one k_pass launch if it was laid out for it.
bool Parser::run_dev_pass(Batch& B, int32_t li, bool outside, bool use_masks) {
  if (B.next_dev >= B.dev_passes.size()) return false;
  const Batch::DevPass& dp = B.dev_passes[B.next_dev++];
  if (dp.li != li || dp.outside != outside) throw std::logic_error("pass out of the laid-out schedule");
  if (!dp.on) return false;
  if (use_masks) {
    if (B.next_pass >= B.pass_plans.size()) throw std::logic_error("masked pass without its plan");
    ++B.next_pass;
  }
  const Layer& L = *layers_[li];
  PassArgs pa{};
  char* blob = (char*)pass_.p;
  pa.stages = (const PassStage*)blob + dp.stage0;
  pa.tiles = (const PassTile*)(blob + B.tiles_byte_off);
  pa.expected = (const int32_t*)(blob + B.exp_byte_off);
  pa.counters = (int32_t*)pcnt_.p;
  pa.queue = pa.counters + dp.cnt_off;
  pa.error = pa.counters;
  pa.num_stages = dp.num_stages;
  pa.total_units = dp.total_units;
  hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, stream_, (float*)(pa.counters + dp.cnt_off),
                     (size_t)dp.num_counters, 0.0f);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (timing) {
    t0 = next_event();
    t1 = next_event();
    HIP_CHECK(hipEventRecord(t0, stream_));
  }
  const dim3 grid((unsigned)std::max(1, num_cus_));
  launch_semi(L.g.semiring, [&](auto semi) {
    constexpr int S = decltype(semi)::value;
    if (L.ipl == 2) {
      if (dp.skip) hipLaunchKernelGGL((k_pass<2, S, true>), grid, dim3(kBlockThreads), kLdsBudgetBytes, stream_, pa);
      else hipLaunchKernelGGL((k_pass<2, S, false>), grid, dim3(kBlockThreads), kLdsBudgetBytes, stream_, pa);
    } else {
      if (dp.skip) hipLaunchKernelGGL((k_pass<1, S, true>), grid, dim3(kBlockThreads), kLdsBudgetBytes, stream_, pa);
      else hipLaunchKernelGGL((k_pass<1, S, false>), grid, dim3(kBlockThreads), kLdsBudgetBytes, stream_, pa);
    }
  });
  HIP_CHECK(hipGetLastError());
  dev_pass_ran_ = true;
  stats.rule_apps += dp.apps;
  if (timing) {
    HIP_CHECK(hipEventRecord(t1, stream_));
    pending_.push_back(TimedLaunch{t0, t1, dp.apps, (outside ? 4 : 0) + 16 * li, dp.apps_first, dp.apps_count});
    stats.rule_launches += 1;
  }
  return true;
}

void Parser::launch_plans(Batch& B, int32_t li) {
  const Batch::LayerJobs* lj = nullptr;
  for (const auto& x : B.layer_jobs)
    if (x.li == li) lj = &x;
  if (lj == nullptr || lj->num_jobs == 0) return;
  const Layer& L = *layers_[li];
  PlanBatchArgs pa{};
  pa.jobs = B.d_jobs + lj->first_job;
  pa.fams = L.d_plan_static;
  pa.mplan = (int32_t*)mplan_.p;
  pa.apps = (double*)apps_.p;
  pa.masks = (const uint32_t*)masks_.p + (size_t)B.mask_in * B.num_cells * L.g.mask_words;
  pa.cell_offsets = B.d_cell_offsets;
  pa.lengths = B.d_lengths;
  pa.cells = B.d_cells;
  pa.num_jobs = lj->num_jobs;
  pa.mask_words = L.g.mask_words;
  hipLaunchKernelGGL((k_plan<false>), dim3(lj->total_blocks), dim3(kPlanThreads), 0, stream_, pa);
  if (lj->any_key) hipLaunchKernelGGL(k_plan_sort, dim3(lj->num_jobs), dim3(512), 0, stream_, pa);
  hipLaunchKernelGGL(k_plan_scan, dim3(lj->num_jobs), dim3(1024), 0, stream_, pa);
  hipLaunchKernelGGL((k_plan<true>), dim3(lj->total_blocks), dim3(kPlanThreads), 0, stream_, pa);
  HIP_CHECK(hipGetLastError());
}

// Arguments of one slice of a (reduce group, span): each family's rule-kernel arguments and
// partial rows (laid out from `base`), and the reduce's arguments. Returns the partial floats used.
size_t Parser::group_slice_args(Batch& B, int32_t li, int32_t rid, int32_t span, bool use_masks,
                                const PlannedSlice* ps, int32_t k0, int32_t kc, float* base, GroupSlice* out) {
  Layer& L = *layers_[li];
  const DevReduce& R = L.red[rid];
  const Grammar& g = L.g;
  const size_t nfam = R.families.size();
  GroupSlice& gs = *out;
  gs = GroupSlice{};
  gs.nfam = (int32_t)nfam;
  RedArgs& ra = gs.ra;
  size_t off = 0;
  const bool fold = folds(L);
  for (size_t fi = 0; fi < nfam; ++fi) {
    const int fid = R.families[fi];
    const DevFamily& F = L.fam[fid];
    const int64_t nitems = family_items(B, li, fid, span, k0, kc);
    if (use_masks) {
      ra.prefix[fi] = ps->pfs[fi].pre;
      ra.posof[fi] = ps->pfs[fi].posof;
      ra.uniform_count[fi] = 1;
      ra.act[fi] = ps->pfs[fi].act;
      ra.block_of_target[fi] = F.unary ? nullptr : F.d_block_of_target;
      ra.act_stride = kc;
    } else if (uniform_kind(F.kind)) {
      ra.prefix[fi] = nullptr;
      ra.uniform_count[fi] = F.kind == 0 ? span - 1 : 1;
    } else {
      const int64_t po = B.prefix_off[(size_t)span * kNumFamilies + fid];
      ra.prefix[fi] = B.d_prefix + po + k0;
      ra.uniform_count[fi] = 1;
    }
    int64_t row_items = nitems;  // partial values per slot row
    if (!use_masks && uniform_kind(F.kind)) row_items = uniform_items(kc, ra.uniform_count[fi]);
    if (!use_masks && pairs_splits(L, F, span)) {
      ra.uniform_count[fi] = span / 2;
      row_items = uniform_items(kc, span / 2);
    }
    const int32_t pitch = (int32_t)((row_items + 255) / 256 * 256);
    const bool ffold = fold && F.foldable;
    ra.partial[fi] = base + off;
    ra.pitch[fi] = pitch;
    ra.dmask[fi] = ffold ? 0x7fffffff : 0;
    gs.row_items[fi] = row_items;
    gs.fid[fi] = fid;
    gs.partial[fi] = base + off;
    gs.apps_slot[fi] = ps ? ps->apps_slot[fi] : -1;
    gs.valid[fi] = family_args(B, li, fid, span, use_masks, k0, kc, base + off, use_masks ? ps->pfs[fi].pre : nullptr,
                               use_masks ? ps->pfs[fi].items : nullptr, use_masks ? ps->pfs[fi].act : nullptr,
                               &gs.fa[fi], &gs.grid[fi], &gs.lds[fi]);
    off += (size_t)pitch * (ffold ? F.n_t : F.num_slots);
  }
  float* inside = (float*)inside_.p;
  float* outside = (float*)outside_.p;
  const DevFamily& F0 = L.fam[R.families[0]];
  ra.chT = F0.t_half >= 2 ? outside : inside;
  ra.t_half = F0.t_half;
  ra.perm = ps ? ps->d_perm : nullptr;
  ra.cell_offsets = B.d_cell_offsets;
  ra.lengths = B.d_lengths;
  ra.cell_list = B.d_cells + B.span_start[span] + k0;
  ra.off = fold ? R.d_off_f : R.d_off;
  ra.fam = fold ? R.d_fam_f : R.d_fam;
  ra.slot = fold ? R.d_slot_f : R.d_slot;
  ra.num_cells = kc;
  ra.cell_size = g.cell_size;
  ra.n_t = R.n_t;
  ra.span = span;
  ra.root_init = (rid == kROutBin);
  ra.root = g.root;
  ra.one = g.semiring == kReal ? 1.0f : 0.0f;
  ra.nfam = (int32_t)nfam;
  gs.rgrid = dim3((kc + 63) / 64, (R.n_t + reduce_targets_ - 1) / reduce_targets_);
  return off;
}

void Parser::run_group(Batch& B, int32_t li, int32_t rid, int32_t span, bool use_masks,
                       const std::vector<PlannedSlice>* planned) {
  Layer& L = *layers_[li];
  const DevReduce& R = L.red[rid];
  const Grammar& g = L.g;
  const int32_t ncells = B.span_count[span];
  if (ncells == 0) return;
  const size_t nfam = R.families.size();
  const auto slices = group_slices(B, li, rid, span);
  for (size_t si = 0; si < slices.size(); ++si) {
    const int32_t k0 = slices[si].first, kc = slices[si].second;
    const PlannedSlice* ps = nullptr;
    if (use_masks) {
      if (planned == nullptr || planned->size() != slices.size()) throw std::logic_error("masked group without its plan");
      ps = &(*planned)[si];
    }
    GroupSlice gs;
    group_slice_args(B, li, rid, span, use_masks, ps, k0, kc, (float*)partial_.p, &gs);
    // Folded rows start at the encoded semiring zero; every slot of a target maxes into them. Only
    // the values the rule kernels can write are filled: a planned family's planned item count (on
    // the device), else the row's items.
    {
      FoldFillArgs ff{};
      int32_t max_rows = 0;
      for (size_t fi = 0; fi < nfam; ++fi) {
        if (gs.ra.dmask[fi] == 0 || !gs.valid[fi]) continue;
        const int k = ff.nfam++;
        ff.base[k] = gs.partial[fi];
        ff.pre[k] = use_masks ? ps->pfs[fi].pre : nullptr;
        ff.num_cells[k] = kc;
        ff.len[k] = (int32_t)gs.row_items[fi];
        ff.rows[k] = L.fam[gs.fid[fi]].n_t;
        ff.pitch[k] = gs.ra.pitch[fi];
        max_rows = std::max(max_rows, ff.rows[k]);
      }
      if (ff.nfam > 0)
        hipLaunchKernelGGL(k_fill_fold, dim3((unsigned)std::min(max_rows, 1024), (unsigned)ff.nfam), dim3(256), 0,
                           stream_, ff);
    }
    if (!persistent) {
      for (size_t fi = 0; fi < nfam; ++fi)
        launch_family(B, li, gs.fid[fi], span, use_masks, k0, kc, gs.partial[fi], use_masks ? ps->pfs[fi].pre : nullptr,
                      use_masks ? ps->pfs[fi].items : nullptr, use_masks ? ps->pfs[fi].act : nullptr, gs.apps_slot[fi]);
    } else {
      // One queue launch per layout among the group's families (unary, block-major, plain) and per
      // partial-row form (folded or one row per slot: the kernel variant differs).
      for (int layout = 0; layout < 6; ++layout) {
        std::vector<int32_t> fids, slots;
        std::vector<float*> parts;
        std::vector<const int32_t*> pcs, pis;
        std::vector<const uint32_t*> pas;
        for (size_t fi = 0; fi < nfam; ++fi) {
          const DevFamily& F = L.fam[R.families[fi]];
          const int lay = 2 * (F.unary ? 0 : (F.skip_layout ? 1 : 2)) + (gs.ra.dmask[fi] != 0 ? 1 : 0);
          if (lay != layout) continue;
          fids.push_back(R.families[fi]);
          parts.push_back(gs.partial[fi]);
          pcs.push_back(use_masks ? ps->pfs[fi].pre : nullptr);
          pis.push_back(use_masks ? ps->pfs[fi].items : nullptr);
          pas.push_back(use_masks ? ps->pfs[fi].act : nullptr);
          slots.push_back(gs.apps_slot[fi]);
        }
        if (!fids.empty()) launch_queue(B, li, fids, span, use_masks, k0, kc, parts, pcs, pis, pas, slots);
      }
    }
    launch_semi(g.semiring, [&](auto semi) {
      constexpr int S = decltype(semi)::value;
      const dim3 blk(64 * reduce_targets_);
      switch (reduce_targets_) {
        case 1: hipLaunchKernelGGL((k_reduce<S, 1>), gs.rgrid, blk, 0, stream_, gs.ra); break;
        case 2: hipLaunchKernelGGL((k_reduce<S, 2>), gs.rgrid, blk, 0, stream_, gs.ra); break;
        case 4: hipLaunchKernelGGL((k_reduce<S, 4>), gs.rgrid, blk, 0, stream_, gs.ra); break;
        case 8: hipLaunchKernelGGL((k_reduce<S, 8>), gs.rgrid, blk, 0, stream_, gs.ra); break;
        default: hipLaunchKernelGGL((k_reduce<S, kReduceTargets>), gs.rgrid, blk, 0, stream_, gs.ra); break;
      }
    });
    HIP_CHECK(hipGetLastError());
  }
}

void Parser::inside_pass(Batch& B, int32_t li, bool use_masks) {
  Layer& L = *layers_[li];
  const Grammar& g = L.g;
  float* inside = (float*)inside_.p;
  const float zero = zero_of(g.semiring);
  const size_t nfl = (size_t)B.num_cells * g.cell_size;
  hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, stream_, inside, nfl, zero);
  if (g.semiring == kReal && !B.has_scales) {
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, stream_, (float*)in_scale_.p, (size_t)B.num_cells, 0.0f);
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, stream_, (float*)out_scale_.p, (size_t)B.num_cells, 0.0f);
  }
  launch_semi(g.semiring, [&](auto semi) {
    constexpr int S = decltype(semi)::value;
    hipLaunchKernelGGL((k_tags<S>), dim3((B.total_len + 127) / 128), dim3(128), 0, stream_, inside,
                       B.d_cell_offsets, B.d_lengths, B.d_len_offsets, B.d_words, B.n, B.total_len,
                       L.d_lex_offsets, L.d_lex_tags, L.d_lex_logp, (const float*)in_scale_.p,
                       (double*)pos_scale_.p, g.cell_size);
  });
  HIP_CHECK(hipGetLastError());
  // A masked layer plans every masked pass of its own up front (the plans read only the masks).
  if (use_masks) launch_plans(B, li);
  if (!run_dev_pass(B, li, false, use_masks)) run_groups(B, li, false, use_masks);
}

void Parser::outside_pass(Batch& B, int32_t li, bool use_masks) {
  Layer& L = *layers_[li];
  const Grammar& g = L.g;
  const float zero = zero_of(g.semiring);
  const size_t nfl = (size_t)B.num_cells * g.cell_size;
  hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, stream_, (float*)outside_.p, nfl, zero);
  if (!run_dev_pass(B, li, true, use_masks)) run_groups(B, li, true, use_masks);
}

// Runs a pass's (group, span) list in order, each masked group with its slices' plans.
void Parser::run_groups(Batch& B, int32_t li, bool outside, bool use_masks) {
  const auto groups = outside ? outside_groups(B) : inside_groups(B);
  const Batch::PassPlan* pp = nullptr;
  if (use_masks) {
    if (B.next_pass >= B.pass_plans.size() || B.pass_plans[B.next_pass].li != li ||
        B.pass_plans[B.next_pass].outside != outside)
      throw std::logic_error("masked pass out of the planned schedule");
    pp = &B.pass_plans[B.next_pass++];
  }
  for (size_t k = 0; k < groups.size(); ++k)
    run_group(B, li, groups[k].first, groups[k].second, use_masks, pp ? &pp->groups[k] : nullptr);
}

// what: 0 = parse, 1 = partitions, 2 = insideOutside, 3 = masks of the last coarse layer,
// 4 = debug charts.
void Parser::run_batch(const int32_t* words, const int32_t* offsets, int32_t first, int32_t count,
                       int32_t what, std::vector<std::vector<TreeNode>>* viterbi,
                       std::vector<std::vector<MbrNode>>* mbr, float* partitions, Charts* charts,
                       int32_t chart_layer, std::vector<uint32_t>* masks_out) {
  if (arena_.base != nullptr) reset_arena();  // the previous batch has finished with its buffers
  Batch B;
  B.n = count;
  B.cell_offsets.push_back(0);
  B.len_offsets.push_back(0);
  B.tree_offsets.push_back(0);
  for (int32_t i = 0; i < count; ++i) {
    const int32_t len = offsets[first + i + 1] - offsets[first + i];
    if (len < 1 || len > kMaxLength) throw std::runtime_error("sentence length must be in [1, 255]");
    B.lengths.push_back(len);
    B.max_len = std::max(B.max_len, len);
    B.cell_offsets.push_back(B.cell_offsets.back() + len * (len + 1));
    B.len_offsets.push_back(B.len_offsets.back() + len);
    B.tree_offsets.push_back(B.tree_offsets.back() + 2 * len - 1);
    for (int32_t j = 0; j < len; ++j) {
      const int32_t w = words[offsets[first + i] + j];
      for (auto& l : layers_)
        if (w < 0 || w + 1 >= (int32_t)l->g.lex_offsets.size())
          throw std::runtime_error("word id out of range of the grammar lexicon");
      B.words.push_back(w);
    }
  }
  B.num_cells = B.cell_offsets.back();
  B.total_len = B.len_offsets.back();
  // Cell lists per span: sentence-major, begin-minor.
  B.span_start.assign(B.max_len + 2, 0);
  B.span_count.assign(B.max_len + 2, 0);
  for (int w = 1; w <= B.max_len; ++w) {
    B.span_start[w] = (int32_t)B.h_cells.size();
    for (int32_t i = 0; i < count; ++i)
      for (int b = 0; b + w <= B.lengths[i]; ++b) B.h_cells.push_back((i << 8) | b);
    B.span_count[w] = (int32_t)B.h_cells.size() - B.span_start[w];
  }
  const bool need_outside = (what != 1) && !(what == 0 && layers_.back()->g.semiring == kViterbi &&
                                             layers_.size() == 1);
  B.prefix_off.assign((size_t)(B.max_len + 2) * kNumFamilies, -1);
  if (what != 1 || layers_.size() > 1) {
    for (int w = 1; w <= B.max_len; ++w)
      for (int fid = kFOutLNN; fid <= kFOutRTT; ++fid) {
        const DevFamily& F = layers_[0]->fam[fid];
        B.prefix_off[(size_t)w * kNumFamilies + fid] = (int64_t)B.h_prefix.size();
        int32_t acc = 0;
        B.h_prefix.push_back(0);
        for (int k = 0; k < B.span_count[w]; ++k) {
          const int32_t packed = B.h_cells[B.span_start[w] + k];
          const int i = packed >> 8, b = packed & 255;
          acc += items_for(F.kind, b, b + w, B.lengths[i]);
          B.h_prefix.push_back(acc);
        }
      }
  }
  (void)need_outside;

  // Device buffers for the batch plan.
  std::vector<int32_t> plan;
  auto put = [&](const std::vector<int32_t>& v) {
    size_t o = plan.size();
    plan.insert(plan.end(), v.begin(), v.end());
    return o;
  };
  const size_t o_len = put(B.lengths), o_co = put(B.cell_offsets), o_lo = put(B.len_offsets),
               o_w = put(B.words), o_to = put(B.tree_offsets), o_c = put(B.h_cells),
               o_p = put(B.h_prefix);
  ensure(plan_, plan.size() * 4 + 256);
  HIP_CHECK(hipMemcpyAsync(plan_.p, plan.data(), plan.size() * 4, hipMemcpyHostToDevice, stream_));
  dev_start_ = next_event();
  HIP_CHECK(hipEventRecord(dev_start_, stream_));
  dev_marked_ = false;
  int32_t* dp = (int32_t*)plan_.p;
  B.d_lengths = dp + o_len;
  B.d_cell_offsets = dp + o_co;
  B.d_len_offsets = dp + o_lo;
  B.d_words = dp + o_w;
  B.d_tree_offsets = dp + o_to;
  B.d_cells = dp + o_c;
  B.d_prefix = dp + o_p;

  int32_t max_cell = 0, mask_words = layers_[0]->g.mask_words;
  for (auto& l : layers_) max_cell = std::max(max_cell, l->g.cell_size);
  ensure(inside_, (size_t)B.num_cells * max_cell * 4);
  ensure(outside_, (size_t)B.num_cells * max_cell * 4);
  ensure(in_scale_, (size_t)B.num_cells * 4);
  ensure(out_scale_, (size_t)B.num_cells * 4);
  ensure(masks_, (size_t)2 * B.num_cells * mask_words * 4);
  ensure(pos_scale_, (size_t)B.total_len * 8);
  const size_t partial_bytes = partial_bytes_ > 0 ? (size_t)partial_bytes_
                                                 : std::max<size_t>((size_t)(mem_bytes_ * 0.3), (size_t)64 << 20);
  ensure(partial_, partial_bytes);
  // Mask plans of every masked pass of the call, laid out now so that the job table travels with
  // the batch's other uploads (and a captured graph replays it unchanged).
  const auto schedule = pass_schedule(what, chart_layer);
  prepare_plans(B, schedule);
  if (!B.h_jobs.empty()) {
    HIP_CHECK(hipMemcpyAsync(jobs_.p, B.h_jobs.data(), B.h_jobs.size() * sizeof(PlanJob), hipMemcpyHostToDevice,
                             stream_));
    B.d_jobs = (PlanJob*)jobs_.p;
  }
  prepare_dev_passes(B, schedule);
  if (!B.h_pass_blob.empty()) {
    ensure(pass_, B.h_pass_blob.size());
    ensure(pcnt_, (size_t)B.num_pass_counters * 4);
    HIP_CHECK(hipMemcpyAsync(pass_.p, B.h_pass_blob.data(), B.h_pass_blob.size(), hipMemcpyHostToDevice, stream_));
  }
  // Plan counters: one double per plan job. Queue heads: one per persistent launch, at most three
  // layouts per (layer, pass, group, slice); sized with room for eight slices per group.
  B.apps_cap = std::max<size_t>((size_t)layers_.size() * 4 * (B.max_len + 2) * kMaxGroupFamilies * 8,
                                B.h_jobs.size());
  ensure(apps_, B.apps_cap * 8);
  // Heads of the persistent rule queues: one per queue launch of the batch (at most three layouts
  // per reduce group and slice).
  if (persistent) ensure(queue_, B.apps_cap * 3 * 4);

  if (what == 0) {
    parse_batch(B, first, viterbi, mbr);
    return;
  }
  reset_counters(B);

  const int32_t nl = (int32_t)layers_.size();
  const int32_t last_layer = (what == 3) ? nl - 2 : nl - 1;
  B.has_masks = false;
  B.has_scales = false;
  for (int32_t li = 0; li <= last_layer; ++li) {
    Layer& L = *layers_[li];
    const Grammar& g = L.g;
    const bool is_last = (li == nl - 1);
    const bool use_masks = B.has_masks;
    timed_begin();
    inside_pass(B, li, use_masks);
    timed_end(&stats.total_ms);
    if (!is_last) {
      timed_begin();
      outside_pass(B, li, use_masks);
      launch_masks_and_scales(B, li, mask_words);
      timed_end(&stats.total_ms);
      if (what == 4 && li == chart_layer) break;
      continue;
    }
    // Final layer.
    if (what == 1) {
      std::vector<float> roots(B.n);
      ensure(aux_, (size_t)B.n * 4);
      hipLaunchKernelGGL(k_get_root, dim3((B.n + 127) / 128), dim3(128), 0, stream_, (float*)aux_.p,
                         (const float*)inside_.p, B.d_cell_offsets, B.n, g.cell_size, g.root);
      mark_device_end();
      HIP_CHECK(hipMemcpyAsync(roots.data(), aux_.p, (size_t)B.n * 4, hipMemcpyDeviceToHost, stream_));
      std::vector<double> ps(B.total_len, 0.0);
      std::vector<float> isc(B.num_cells, 0.0f);
      if (g.semiring == kReal) {
        HIP_CHECK(hipMemcpyAsync(ps.data(), pos_scale_.p, (size_t)B.total_len * 8, hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipMemcpyAsync(isc.data(), in_scale_.p, (size_t)B.num_cells * 4, hipMemcpyDeviceToHost, stream_));
      }
      HIP_CHECK(hipStreamSynchronize(stream_));
      for (int32_t i = 0; i < B.n; ++i) {
        if (g.semiring == kReal) {
          double sc = 0;
          for (int j = B.len_offsets[i]; j < B.len_offsets[i + 1]; ++j) sc += ps[j];
          const int n = B.lengths[i];
          const float top_scale = isc[B.cell_offsets[i] + tri(n) + chart_index(0, n, n)];
          // RealSemiring.toLogSpace(x, scale) + partitionScales (CLParser.scala:197-199)
          const float ls = (float)std::log((double)roots[i]) + top_scale;
          partitions[first + i] = ls + (float)sc;
        } else {
          partitions[first + i] = roots[i];
        }
      }
      break;
    }
    if (g.semiring == kViterbi) {
      if (what == 2 || what == 4) {
        timed_begin();
        outside_pass(B, li, use_masks);
        timed_end(&stats.total_ms);
      }
      if (what == 2) {
        ensure(trees_, (size_t)B.tree_offsets.back() * sizeof(TreeNode));
        timed_begin();
        launch_viterbi(B, li);
        timed_end(&stats.other_ms);
      }
    } else {
      timed_begin();
      outside_pass(B, li, use_masks);
      timed_end(&stats.total_ms);
      if (what == 2) {
        ensure(mbr_, (size_t)B.num_cells * sizeof(MbrCell));
        launch_mbr(B, li);
      }
    }
  }
  finish_batch(B, what, partitions, charts, chart_layer, masks_out);
}

// computePruningMasks: masks at threshold -9 (CLParser.scala:223-229) into the half of masks_ not
// holding this layer's input masks, and the next layer's scaling constants when it is Real
// (CLParser.scala:236-245).
void Parser::launch_masks_and_scales(Batch& B, int32_t li, int32_t mask_words) {
  const Layer& L = *layers_[li];
  const Grammar& g = L.g;
  const int32_t mask_out = B.has_masks ? 1 - B.mask_in : 0;
  uint32_t* mo = (uint32_t*)masks_.p + (size_t)mask_out * B.num_cells * mask_words;
  hipLaunchKernelGGL(k_masks, dim3((B.num_cells + 3) / 4), dim3(256), 0, stream_, mo,
                     (const float*)inside_.p, (const float*)outside_.p, B.d_cell_offsets,
                     B.d_lengths, B.n, B.num_cells, L.d_proj_t, L.d_proj_n, g.num_syms,
                     g.cell_size, g.root, -9.0f, mask_words);
  HIP_CHECK(hipGetLastError());
  if (layers_[li + 1]->g.semiring == kReal) {
    hipLaunchKernelGGL(k_scaling, dim3((B.num_cells + 3) / 4), dim3(256), 0, stream_,
                       (float*)in_scale_.p, (const float*)inside_.p, g.cell_size, B.num_cells);
    hipLaunchKernelGGL(k_scaling, dim3((B.num_cells + 3) / 4), dim3(256), 0, stream_,
                       (float*)out_scale_.p, (const float*)outside_.p, g.cell_size, B.num_cells);
    HIP_CHECK(hipGetLastError());
    B.has_scales = true;
  } else {
    B.has_scales = false;
  }
  B.mask_in = mask_out;
  B.has_masks = true;
}

void Parser::launch_viterbi(Batch& B, int32_t li) {
  const Layer& L = *layers_[li];
  const Grammar& g = L.g;
  hipLaunchKernelGGL(k_viterbi, dim3(B.n), dim3(kViterbiThreads), 0, stream_, (TreeNode*)trees_.p,
                     B.d_tree_offsets, (const float*)inside_.p, B.d_cell_offsets, B.d_lengths,
                     B.n, g.cell_size, g.root, g.num_nonterms(), L.d_dec_offsets,
                     L.d_dec_scores, L.d_dec_left, L.d_dec_right);
  HIP_CHECK(hipGetLastError());
}

void Parser::launch_mbr(Batch& B, int32_t li) {
  const Layer& L = *layers_[li];
  const Grammar& g = L.g;
  if (g.cell_size > kMbrMaxCell || g.num_syms > kMbrMaxSyms)
    throw std::runtime_error("MBR decode supports cells of at most 1,024 and grammars of at most 2,048 symbols");
  hipLaunchKernelGGL(k_mbr, dim3((B.num_cells + kMbrCellsPerBlock - 1) / kMbrCellsPerBlock), dim3(256), 0, stream_,
                     (MbrCell*)mbr_.p, (const float*)inside_.p, (const float*)outside_.p,
                     B.d_cell_offsets, B.d_lengths, B.n, B.num_cells, L.d_mbr_t_off, L.d_mbr_t_syms,
                     L.d_mbr_n_off, L.d_mbr_n_syms, g.cell_size, g.root, (int32_t)g.coarse_names.size(), g.num_syms,
                     g.semiring == kLogSum);
  HIP_CHECK(hipGetLastError());
}

void Parser::reset_counters(Batch& B) {
  hipLaunchKernelGGL(k_fill, dim3(16), dim3(256), 0, stream_, (float*)apps_.p, B.apps_cap * 2, 0.0f);
  if (persistent)
    hipLaunchKernelGGL(k_fill, dim3(16), dim3(256), 0, stream_, (float*)queue_.p, B.apps_cap * 3, 0.0f);
  apps_next_ = (int32_t)B.h_jobs.size();  // every plan job owns one counter
  queue_next_ = 0;
  B.next_pass = 0;
  B.next_dev = 0;
  if (B.num_pass_counters > 64)  // the error flag of the device-scheduled passes
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, stream_, (float*)pcnt_.p, (size_t)64, 0.0f);
}

// Device work of CLParser.parse on one batch: every layer's passes, the masks between layers and
// the final decode kernel. Enqueues only (no host waits), so it can be captured into a graph.
void Parser::parse_device(Batch& B) {
  reset_counters(B);
  B.has_masks = false;
  B.has_scales = false;
  B.mask_in = 0;
  const int32_t nl = (int32_t)layers_.size();
  const int32_t mask_words = layers_[0]->g.mask_words;
  for (int32_t li = 0; li < nl; ++li) {
    const Grammar& g = layers_[li]->g;
    timed_begin();
    inside_pass(B, li, B.has_masks);
    timed_end(&stats.total_ms);
    if (li + 1 < nl) {
      timed_begin();
      outside_pass(B, li, B.has_masks);
      launch_masks_and_scales(B, li, mask_words);
      timed_end(&stats.total_ms);
      continue;
    }
    if (g.semiring == kViterbi) {
      if (outside_final) {
        // CLParser.insideOutside (CLParser.scala:171-182): the outside pass before the decode.
        timed_begin();
        outside_pass(B, li, B.has_masks);
        timed_end(&stats.total_ms);
      }
      timed_begin();
      launch_viterbi(B, li);
      timed_end(&stats.other_ms);
    } else {
      timed_begin();
      outside_pass(B, li, B.has_masks);
      timed_end(&stats.total_ms);
      launch_mbr(B, li);
    }
  }
}

// CLParser.parse on one batch: the device work (replayed from a graph when this batch's shape
// was captured before), then the result download and the host side of the decoders.
void Parser::parse_batch(Batch& B, int32_t first, std::vector<std::vector<TreeNode>>* viterbi,
                         std::vector<std::vector<MbrNode>>* mbr) {
  const bool viterbi_final = layers_.back()->g.semiring == kViterbi;
  if (viterbi_final) ensure(trees_, (size_t)B.tree_offsets.back() * sizeof(TreeNode));
  else ensure(mbr_, (size_t)B.num_cells * sizeof(MbrCell));
  GraphEntry* ge = use_graphs ? graph_entry(B.lengths) : nullptr;
  if (ge != nullptr && ge->exec != nullptr) {
    HIP_CHECK(hipGraphLaunch(ge->exec, stream_));
    ++stats.graph_replays;
  } else if (ge != nullptr && ge->seen > 0 && !ge->failed) {
    // Second sighting of this shape: capture the device work and launch the graph. Workspace
    // buffers were sized by the first (eager) run; a capture that would need to grow one is
    // abandoned and the batch runs eagerly.
    const PassStats saved = stats;
    const bool saved_timing = timing;
    timing = false;
    capturing_ = true;
    HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    bool aborted = false;
    try {
      parse_device(B);
    } catch (const CaptureAbort&) {
      aborted = true;
    } catch (...) {
      hipGraph_t gdrop = nullptr;
      (void)hipStreamEndCapture(stream_, &gdrop);
      if (gdrop) (void)hipGraphDestroy(gdrop);
      capturing_ = false;
      timing = saved_timing;
      throw;
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream_, &graph);
    capturing_ = false;
    timing = saved_timing;
    if (!aborted && ec == hipSuccess && graph != nullptr &&
        hipGraphInstantiate(&ge->exec, graph, nullptr, nullptr, 0) == hipSuccess) {
      ge->generation = ws_generation_;
      HIP_CHECK(hipGraphDestroy(graph));
      HIP_CHECK(hipGraphLaunch(ge->exec, stream_));
      ++stats.graph_captures;
    } else {
      if (graph != nullptr) (void)hipGraphDestroy(graph);
      ge->exec = nullptr;
      ge->failed = true;
      (void)hipGetLastError();
      stats = saved;
      parse_device(B);
    }
  } else {
    parse_device(B);
    if (ge != nullptr) ge->seen += 1;
  }
  // Host state the device work leaves behind (also when it was replayed).
  B.has_masks = layers_.size() > 1;
  B.has_scales = layers_.size() > 1 && layers_.back()->g.semiring == kReal;
  finish_parse(B, first, viterbi, mbr);
  flush_timing();
}

void Parser::finish_parse(Batch& B, int32_t first, std::vector<std::vector<TreeNode>>* viterbi,
                          std::vector<std::vector<MbrNode>>* mbr) {
  if (layers_.back()->g.semiring == kViterbi) {
    std::vector<TreeNode> all(B.tree_offsets.back());
    mark_device_end();
    HIP_CHECK(hipMemcpyAsync(all.data(), trees_.p, all.size() * sizeof(TreeNode), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (int32_t i = 0; i < B.n; ++i)
      (*viterbi)[first + i].assign(all.begin() + B.tree_offsets[i], all.begin() + B.tree_offsets[i + 1]);
    return;
  }
  {
    {
      {
        std::vector<MbrCell> dec(B.num_cells);
        std::vector<float> isc(B.num_cells), osc(B.num_cells);
        mark_device_end();
        HIP_CHECK(hipMemcpyAsync(dec.data(), mbr_.p, dec.size() * sizeof(MbrCell), hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipMemcpyAsync(isc.data(), in_scale_.p, (size_t)B.num_cells * 4, hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipMemcpyAsync(osc.data(), out_scale_.p, (size_t)B.num_cells * 4, hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        // extractMBRParses (CLParser.scala:589-696).
        for (int32_t i = 0; i < B.n; ++i) {
          const int n = B.lengths[i];
          const int co = B.cell_offsets[i];
          auto bot_cell = [&](int b, int e) { return co + chart_index(b, e, n); };
          auto top_cell = [&](int b, int e) { return co + tri(n) + chart_index(b, e, n); };
          const int W = n + 1;
          std::vector<float> best(W * W, 0.0f);
          std::vector<int> split(W * W, 0);
          std::vector<char> use_unary(W * W, 0);
          const float root_top_scale = isc[top_cell(0, n)];
          for (int span = 1; span <= n; ++span)
            for (int b = 0; b + span <= n; ++b) {
              const int e = b + span;
              float top = dec[top_cell(b, e)].score, bot = dec[bot_cell(b, e)].score;
              if (B.has_scales) {
                top *= (float)std::exp((double)(isc[top_cell(b, e)] + osc[top_cell(b, e)] - root_top_scale));
                bot *= (float)std::exp((double)(isc[bot_cell(b, e)] + osc[bot_cell(b, e)] - root_top_scale));
              }
              if (std::isnan(top)) top = 0.0f;
              if (std::isnan(bot)) bot = 0.0f;
              top = std::min(top, 1.0f);
              bot = std::min(bot, 1.0f);
              const bool uu = (b + 1 < e) || top > 0.4;
              use_unary[b * W + e] = uu;
              float sc = uu ? top + bot : bot;
              if (span > 1) {
                float bs = 0.0f;
                int bsplit = b + 1;
                for (int s = b + 1; s < e; ++s) {
                  const float v = best[b * W + s] + best[s * W + e];
                  if (v > bs) {
                    bsplit = s;
                    bs = v;
                  }
                }
                split[b * W + e] = bsplit;
                sc += bs;
              }
              best[b * W + e] = sc;
            }
          std::vector<MbrNode>& out = (*mbr)[first + i];
          out.clear();
          std::vector<std::pair<int, int>> stack{{0, n}};
          while (!stack.empty()) {
            auto [b, e] = stack.back();
            stack.pop_back();
            MbrNode nd;
            nd.begin = b;
            nd.end = e;
            nd.bot_coarse = dec[bot_cell(b, e)].symbol;
            nd.top_coarse = use_unary[b * (n + 1) + e] ? dec[top_cell(b, e)].symbol : -1;
            nd.score = best[b * (n + 1) + e];
            out.push_back(nd);
            if (e - b > 1) {
              const int s = split[b * (n + 1) + e];
              stack.push_back({s, e});
              stack.push_back({b, s});
            }
          }
        }
      }
    }
  }
}

// Downloads of the non-parse calls: debug charts and coarse masks; resolves the timing events.
void Parser::finish_batch(Batch& B, int32_t what, float* partitions, Charts* charts, int32_t chart_layer,
                          std::vector<uint32_t>* masks_out) {
  (void)what;
  (void)partitions;
  const int32_t mask_words = layers_[0]->g.mask_words;
  mark_device_end();
  if (charts != nullptr) {
    const Grammar& g = layers_[chart_layer]->g;
    HIP_CHECK(hipStreamSynchronize(stream_));
    charts->cell_size = g.cell_size;
    charts->inside.resize((size_t)B.num_cells * g.cell_size);
    charts->outside.resize((size_t)B.num_cells * g.cell_size);
    charts->in_scale.resize(B.num_cells);
    charts->out_scale.resize(B.num_cells);
    charts->cell_offsets = B.cell_offsets;
    HIP_CHECK(hipMemcpy(charts->inside.data(), inside_.p, charts->inside.size() * 4, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(charts->outside.data(), outside_.p, charts->outside.size() * 4, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(charts->in_scale.data(), in_scale_.p, (size_t)B.num_cells * 4, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(charts->out_scale.data(), out_scale_.p, (size_t)B.num_cells * 4, hipMemcpyDeviceToHost));
    if (B.has_masks) {
      charts->masks.resize((size_t)B.num_cells * mask_words);
      HIP_CHECK(hipMemcpy(charts->masks.data(), (uint32_t*)masks_.p + (size_t)B.mask_in * B.num_cells * mask_words,
                          charts->masks.size() * 4, hipMemcpyDeviceToHost));
    }
    if (g.semiring == kReal) {
      std::vector<double> ps(B.total_len);
      HIP_CHECK(hipMemcpy(ps.data(), pos_scale_.p, (size_t)B.total_len * 8, hipMemcpyDeviceToHost));
      charts->partition_scales.assign(B.n, 0.0);
      for (int32_t i = 0; i < B.n; ++i)
        for (int j = B.len_offsets[i]; j < B.len_offsets[i + 1]; ++j) charts->partition_scales[i] += ps[j];
    }
  }
  if (masks_out != nullptr && B.has_masks) {
    std::vector<uint32_t> m((size_t)B.num_cells * mask_words);
    HIP_CHECK(hipMemcpyAsync(m.data(), (uint32_t*)masks_.p + (size_t)B.mask_in * B.num_cells * mask_words,
                             m.size() * 4, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    masks_out->insert(masks_out->end(), m.begin(), m.end());
  }
  flush_timing();
}

void Parser::parse(const int32_t* words, const int32_t* offsets, int32_t n,
                   std::vector<std::vector<TreeNode>>* viterbi, std::vector<std::vector<MbrNode>>* mbr) {
  HIP_CHECK(hipSetDevice(device_));
  stats = PassStats{};
  if (viterbi) viterbi->assign(n, {});
  if (mbr) mbr->assign(n, {});
  if (layers_.back()->g.semiring == kViterbi && !viterbi) throw std::runtime_error("viterbi output required");
  if (layers_.back()->g.semiring != kViterbi && !mbr) throw std::runtime_error("mbr output required");
  for (auto [f, c] : make_batches(offsets, n)) run_batch(words, offsets, f, c, 0, viterbi, mbr, nullptr, nullptr, 0, nullptr);
}

void Parser::reset_arena() {
  for (Buf* b : {&inside_, &outside_, &in_scale_, &out_scale_, &masks_, &partial_, &plan_, &tags_, &trees_, &mbr_,
                 &pos_scale_, &aux_, &mplan_, &apps_, &queue_, &jobs_, &pass_, &pcnt_})
    *b = Buf{};
  arena_.used = 0;
}

void Parser::parse_noalloc(const int32_t* words, const int32_t* offsets, int32_t n,
                           std::vector<std::vector<TreeNode>>* viterbi, std::vector<std::vector<MbrNode>>* mbr) {
  HIP_CHECK(hipSetDevice(device_));
  if (scratch_ == nullptr) throw std::runtime_error("no scratch array on this device: " + scratch_err_);
  std::vector<int32_t> lens;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t len = offsets[i + 1] - offsets[i];
    if (len < 1 || len > kMaxLength) throw std::runtime_error("sentence length must be in [1, 255]");
    lens.push_back(len);
  }
  const int64_t partial = kNoAllocScratchBytes / 4;
  const int64_t need = bytes_needed(lens, partial);
  if (need > kNoAllocScratchBytes)
    throw std::runtime_error("input needs about " + std::to_string(need) + " bytes of device workspace; the no-alloc "
                             "parse holds at most " + std::to_string(kNoAllocScratchBytes));
  // Swap the parser's own workspace out for the scratch array, and back afterwards.
  Buf* bufs[] = {&inside_, &outside_, &in_scale_, &out_scale_, &masks_, &partial_, &plan_, &tags_, &trees_, &mbr_,
                 &pos_scale_, &aux_, &mplan_, &apps_, &queue_, &jobs_, &pass_, &pcnt_};
  std::vector<Buf> saved;
  for (Buf* b : bufs) saved.push_back(*b);
  const bool saved_graphs = use_graphs, saved_timing = timing;
  const int64_t saved_partial = partial_bytes_;
  const int64_t saved_mem = mem_bytes_;
  use_graphs = false;
  timing = false;
  partial_bytes_ = partial;
  mem_bytes_ = kNoAllocScratchBytes * 10;  // one batch: the size check above already bounds the input
  arena_ = Arena{scratch_, (size_t)kNoAllocScratchBytes, 0};
  reset_arena();
  ++ws_generation_;
  auto restore = [&]() {
    for (size_t k = 0; k < saved.size(); ++k) *bufs[k] = saved[k];
    use_graphs = saved_graphs;
    timing = saved_timing;
    partial_bytes_ = saved_partial;
    mem_bytes_ = saved_mem;
    arena_ = Arena{};
    ++ws_generation_;
  };
  try {
    parse(words, offsets, n, viterbi, mbr);
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    restore();
    throw;
  }
  restore();
}

void Parser::partitions(const int32_t* words, const int32_t* offsets, int32_t n, float* out) {
  HIP_CHECK(hipSetDevice(device_));
  stats = PassStats{};
  for (auto [f, c] : make_batches(offsets, n)) run_batch(words, offsets, f, c, 1, nullptr, nullptr, out, nullptr, 0, nullptr);
}

void Parser::inside_outside(const int32_t* words, const int32_t* offsets, int32_t n) {
  HIP_CHECK(hipSetDevice(device_));
  stats = PassStats{};
  for (auto [f, c] : make_batches(offsets, n)) run_batch(words, offsets, f, c, 2, nullptr, nullptr, nullptr, nullptr, 0, nullptr);
}

void Parser::masks(const int32_t* words, const int32_t* offsets, int32_t n, std::vector<uint32_t>* out,
                   std::vector<int32_t>* cell_offsets) {
  HIP_CHECK(hipSetDevice(device_));
  if (layers_.size() < 2) throw std::runtime_error("masks need a coarse layer");
  stats = PassStats{};
  out->clear();
  cell_offsets->assign(1, 0);
  for (int32_t i = 0; i < n; ++i) {
    const int32_t len = offsets[i + 1] - offsets[i];
    cell_offsets->push_back(cell_offsets->back() + len * (len + 1));
  }
  for (auto [f, c] : make_batches(offsets, n)) run_batch(words, offsets, f, c, 3, nullptr, nullptr, nullptr, nullptr, 0, out);
}

void Parser::debug_charts(const int32_t* words, const int32_t* offsets, int32_t n, int32_t layer,
                          bool outside, Charts* out) {
  HIP_CHECK(hipSetDevice(device_));
  stats = PassStats{};
  if (layer < 0 || layer >= (int32_t)layers_.size()) throw std::runtime_error("bad layer index");
  // Charts of one batch only: the caller keeps the sentences within one batch.
  auto batches = make_batches(offsets, n);
  if (batches.size() != 1) throw std::runtime_error("debug_charts input does not fit one batch");
  // Run layers up to `layer`; for the final layer run inside (and outside if asked).
  const int32_t what = (layer == (int32_t)layers_.size() - 1) ? (outside ? 4 : 1) : 4;
  std::vector<float> parts(n);
  run_batch(words, offsets, 0, n, what, nullptr, nullptr, parts.data(), out, layer, nullptr);
}

}  // namespace puck
