// Batched parse engine. See batch.h.
#include "batch.h"

#include <algorithm>
#include <exception>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>

namespace puck {

namespace {
void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at " + what);
}
}  // namespace

BatchParser::BatchParser(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
                         const BatchConfig& cfg)
    : cfg_(cfg), device_(device) {
  if (cfg_.streams < 1 || cfg_.streams > 8) throw std::runtime_error("streams must be in [1, 8]");
  if (cfg_.bucket_cells < 1) throw std::runtime_error("bucket_cells must be positive");
  const int64_t per = mem_bytes / cfg_.streams;
  for (int32_t s = 0; s < cfg_.streams; ++s) {
    parsers_.emplace_back(new Parser(layers, device, per, cfg_.items_per_lane));
    Parser& p = *parsers_.back();
    p.timing = false;
    p.set_persistent(cfg_.persistent);
    p.set_launch_shape(cfg_.piece_target, 1);
    p.set_outside_final(cfg_.outside_final);
    p.set_device_schedule(cfg_.device_schedule);
    p.set_reduce_targets(cfg_.reduce_targets);
    p.set_fold(cfg_.fold);
    p.use_graphs = cfg_.graphs;
  }
  hip_ok(hipSetDevice(device_), "hipSetDevice");
  hip_ok(hipEventCreate(&ref_), "hipEventCreate");
}

BatchParser::~BatchParser() {
  if (ref_) (void)hipEventDestroy(ref_);
}

void BatchParser::parse(const int32_t* words, const int32_t* offsets, int32_t n,
                        std::vector<std::vector<TreeNode>>* viterbi, std::vector<std::vector<MbrNode>>* mbr) {
  hip_ok(hipSetDevice(device_), "hipSetDevice");
  stats = BatchStats{};
  if (viterbi) viterbi->assign(n, {});
  if (mbr) mbr->assign(n, {});
  // Longest first (stable), then buckets of at most bucket_cells chart cells.
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  auto len = [&](int32_t i) { return offsets[i + 1] - offsets[i]; };
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return len(a) > len(b); });
  struct Bucket { int32_t begin, end; double cost; };
  std::vector<Bucket> buckets;
  int64_t cells = 0;
  double cost = 0;
  int32_t start = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int64_t l = len(order[k]);
    const int64_t c = l * (l + 1);
    if (k > start && cells + c > cfg_.bucket_cells) {
      buckets.push_back({start, k, cost});
      start = k;
      cells = 0;
      cost = 0;
    }
    cells += c;
    cost += (double)l * l * l + 64.0 * l;  // NN work grows as n^3, per-span launches as n
  }
  if (n > start) buckets.push_back({start, n, cost});
  stats.buckets = (int64_t)buckets.size();
  // Longest-processing-time-first assignment of buckets to the workers.
  const int32_t S = std::min<int32_t>(cfg_.streams, std::max<int32_t>(1, (int32_t)buckets.size()));
  std::vector<std::vector<int32_t>> plan(S);
  {
    std::vector<int32_t> idx(buckets.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return buckets[a].cost > buckets[b].cost; });
    std::vector<double> load(S, 0.0);
    for (int32_t b : idx) {
      const int32_t w = (int32_t)(std::min_element(load.begin(), load.end()) - load.begin());
      load[w] += buckets[b].cost;
      plan[w].push_back(b);
    }
  }
  for (auto& p : parsers_) p->ref_event = ref_;
  hip_ok(hipEventRecord(ref_, nullptr), "hipEventRecord");
  hip_ok(hipEventSynchronize(ref_), "hipEventSynchronize");
  std::mutex mu;
  std::exception_ptr error;
  std::vector<double> lo(S, 1e300), hi(S, -1e300);
  std::vector<int64_t> captures(S, 0), replays(S, 0);
  auto worker = [&](int32_t w) {
    try {
      hip_ok(hipSetDevice(device_), "hipSetDevice");
      Parser& p = *parsers_[w];
      for (int32_t b : plan[w]) {
        const Bucket& bk = buckets[b];
        std::vector<int32_t> bw, bo{0};
        for (int32_t k = bk.begin; k < bk.end; ++k) {
          const int32_t i = order[k];
          bw.insert(bw.end(), words + offsets[i], words + offsets[i + 1]);
          bo.push_back((int32_t)bw.size());
        }
        const int32_t m = bk.end - bk.begin;
        std::vector<std::vector<TreeNode>> vt;
        std::vector<std::vector<MbrNode>> mt;
        p.parse(bw.data(), bo.data(), m, viterbi ? &vt : nullptr, mbr ? &mt : nullptr);
        lo[w] = std::min(lo[w], p.stats.span_lo_ms);
        hi[w] = std::max(hi[w], p.stats.span_hi_ms);
        captures[w] += p.stats.graph_captures;
        replays[w] += p.stats.graph_replays;
        for (int32_t k = 0; k < m; ++k) {
          const int32_t i = order[bk.begin + k];
          if (viterbi) (*viterbi)[i] = std::move(vt[k]);
          if (mbr) (*mbr)[i] = std::move(mt[k]);
        }
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu);
      if (!error) error = std::current_exception();
    }
  };
  std::vector<std::thread> threads;
  for (int32_t w = 1; w < S; ++w) threads.emplace_back(worker, w);
  worker(0);
  for (auto& t : threads) t.join();
  for (auto& p : parsers_) p->ref_event = nullptr;
  if (error) std::rethrow_exception(error);
  const double l = *std::min_element(lo.begin(), lo.end());
  const double h = *std::max_element(hi.begin(), hi.end());
  stats.device_ms = h > l ? h - l : 0.0;
  for (int32_t w = 0; w < S; ++w) {
    stats.graph_captures += captures[w];
    stats.graph_replays += replays[w];
  }
}

}  // namespace puck
