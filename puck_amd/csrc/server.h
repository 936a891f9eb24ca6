// Continuous-batching parse server over the CKY engine.
//
// Replaces the reference's BatchFunctionAnnotatorService (AnnotatorService.scala:29-78), which
// accumulates requests and runs one CLParser.parse per flush (20,000 sentences or 1 s). Here
// requests are admitted into batches as soon as one of `max_inflight` batch slots is free; each slot
// owns a Parser with its own HIP stream and workspace, so batches overlap on the device. Admission
// is by free device memory: a slot takes queued requests in arrival order while the memory its batch
// needs (Parser::bytes_needed) fits into what the slot already holds plus the device's free memory
// (hipMemGetInfo, less what other slots have claimed but not yet allocated and a reserve), and the
// whole server stays within its memory budget (WorkSpace.allocate's rule, WorkSpace.scala:88-122,
// applied per admission instead of once per call). A slot parses its batch in two length classes,
// short first, publishing each class's results as soon as it is parsed.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "engine.h"

namespace puck {

struct ServerStats {
  int64_t completed = 0;
  int64_t batches = 0;
  double mean_batch = 0;
  double p50_ms = 0, p90_ms = 0, p99_ms = 0, max_ms = 0;
};

class Server {
 public:
  Server(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
         int32_t max_inflight, int32_t max_batch, int32_t items_per_lane, bool persistent);
  ~Server();

  int64_t submit(const int32_t* words, int32_t len);
  // Returns true and fills `nodes` when the request is done (status < 0: the parse failed, see
  // `error`). With wait = true blocks until done.
  bool result(int64_t ticket, bool wait, std::vector<TreeNode>* nodes, std::vector<MbrNode>* mbr,
              int32_t* status, std::string* error);
  void flush();
  ServerStats stats();
  void reset_stats();

 private:
  struct Request {
    int64_t ticket;
    std::vector<int32_t> words;
    std::chrono::steady_clock::time_point submitted;
  };
  struct Done {
    std::vector<TreeNode> nodes;
    std::vector<MbrNode> mbr;
    int32_t status = 0;
    std::string error;
  };
  void worker(int32_t slot);

  std::vector<std::unique_ptr<Parser>> parsers_;
  std::vector<std::thread> threads_;
  int32_t max_batch_;
  int64_t max_cells_;
  int32_t num_words_ = 0;  // word ids every layer's lexicon covers: [0, num_words_)
  bool mbr_final_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Request> queue_;
  std::unordered_map<int64_t, Done> done_;
  int64_t next_ticket_ = 1;
  int64_t outstanding_ = 0;
  bool stop_ = false;
  // Latencies of the last kLatencyWindow completed requests (a ring).
  static constexpr size_t kLatencyWindow = 1 << 20;
  std::vector<double> latencies_ms_;
  size_t latency_next_ = 0;
  int64_t completed_ = 0;
  int64_t batches_ = 0, batched_requests_ = 0;
  std::unordered_set<int64_t> issued_;  // tickets submitted and not yet collected
  int64_t budget_ = 0;                  // device bytes all slots together may hold
  int64_t reserve_ = (int64_t)256 << 20;
  int64_t claimed_ = 0;                 // growth admitted by slots but not yet allocated
  int32_t device_ = 0;
  // Length class boundary of a slot's batch (words): the short class is parsed and published first.
  static constexpr int32_t kSplitLength = 40;
  bool fits(int32_t slot, const std::vector<int32_t>& lengths, int64_t* growth);
  void record_latency(double ms);
};

}  // namespace puck
