// Continuous-batching parse server. See server.h.
#include "server.h"

#include <algorithm>
#include <climits>
#include <numeric>
#include <stdexcept>

namespace puck {

Server::Server(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
               int32_t max_inflight, int32_t max_batch, int32_t items_per_lane, bool persistent)
    : max_batch_(std::max(1, max_batch)), device_(device) {
  if (max_inflight < 1) throw std::runtime_error("max_inflight must be >= 1");
  mbr_final_ = layers.back()->semiring != kViterbi;
  int32_t max_cell = 0;
  num_words_ = INT32_MAX;
  for (auto& g : layers) {
    max_cell = std::max(max_cell, g->cell_size);
    num_words_ = std::min(num_words_, (int32_t)g->lex_offsets.size() - 1);
  }
  if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw std::runtime_error("hipMemGetInfo failed");
  budget_ = std::min<int64_t>(mem_bytes, (int64_t)free_b - reserve_);
  if (budget_ <= 0) throw std::runtime_error("no free device memory for the server");
  // Every slot's parser may take a batch up to the whole budget in one go (its make_batches limit);
  // admission keeps the slots together within it.
  for (int32_t i = 0; i < max_inflight; ++i) {
    parsers_.emplace_back(new Parser(layers, device, budget_, items_per_lane));
    parsers_.back()->timing = false;
    parsers_.back()->set_persistent(persistent);
    parsers_.back()->set_partial_bytes(std::max<int64_t>((int64_t)64 << 20, budget_ / 10 / max_inflight));
  }
  // A batch never exceeds what one parser keeps in one batch (its chart budget).
  max_cells_ = std::max<int64_t>(1, (int64_t)(budget_ * 0.6) / 2 / ((int64_t)max_cell * 4));
  for (int32_t i = 0; i < max_inflight; ++i) threads_.emplace_back(&Server::worker, this, i);
}

// Whether a batch of these lengths fits on `slot`: the bytes beyond what the slot holds must fit into
// the device's free memory not claimed by other slots, and all slots together into the budget.
bool Server::fits(int32_t slot, const std::vector<int32_t>& lengths, int64_t* growth) {
  const int64_t held = parsers_[slot]->workspace_bytes();
  const int64_t need = parsers_[slot]->bytes_needed(lengths);
  *growth = std::max<int64_t>(0, need - held);
  if (*growth == 0) return true;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  int64_t all_held = 0;
  for (auto& p : parsers_) all_held += p->workspace_bytes();
  return *growth <= (int64_t)free_b - claimed_ - reserve_ && all_held + claimed_ + *growth <= budget_;
}

void Server::record_latency(double ms) {
  if (latencies_ms_.size() < kLatencyWindow) latencies_ms_.push_back(ms);
  else latencies_ms_[latency_next_] = ms;
  latency_next_ = (latency_next_ + 1) % kLatencyWindow;
  ++completed_;
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& t : threads_) t.join();
}

int64_t Server::submit(const int32_t* words, int32_t len) {
  if (len < 1 || len > 255) throw std::runtime_error("sentence length must be in [1, 255]");
  // Rejected here, a bad request cannot fail the batch it would have joined.
  for (int32_t i = 0; i < len; ++i)
    if (words[i] < 0 || words[i] >= num_words_) throw std::runtime_error("word id out of range of the grammar lexicon");
  std::lock_guard<std::mutex> lk(mu_);
  Request r;
  r.ticket = next_ticket_++;
  r.words.assign(words, words + len);
  r.submitted = std::chrono::steady_clock::now();
  issued_.insert(r.ticket);
  queue_.push_back(std::move(r));
  ++outstanding_;
  cv_work_.notify_one();
  return queue_.back().ticket;
}

void Server::worker(int32_t slot) {
  Parser& p = *parsers_[slot];
  if (hipSetDevice(device_) != hipSuccess) return;
  while (true) {
    std::vector<Request> batch;
    int64_t growth = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
      if (stop_ && queue_.empty()) return;
      // Admit requests in arrival order while the batch fits this slot's memory and the device's
      // free memory; the first request is always taken (if it cannot be allocated, it fails alone).
      int64_t cells = 0;
      std::vector<int32_t> lengths;
      while (!queue_.empty() && (int32_t)batch.size() < max_batch_) {
        const int64_t n = (int64_t)queue_.front().words.size();
        if (!batch.empty() && cells + n * (n + 1) > max_cells_) break;
        lengths.push_back((int32_t)n);
        int64_t g = 0;
        const bool ok = fits(slot, lengths, &g);
        if (!ok && !batch.empty()) break;
        growth = g;
        cells += n * (n + 1);
        batch.push_back(std::move(queue_.front()));
        queue_.pop_front();
      }
      claimed_ += growth;
      if (!queue_.empty()) cv_work_.notify_one();
    }
    // Length classes, shortest first: sentences of at most kSplitLength words are parsed (and their
    // results published) before the longer ones, so a short request does not wait for the span
    // loop of a long one admitted with it. Within a class, arrival order.
    std::vector<size_t> order(batch.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_partition(order.begin(), order.end(),
                          [&](size_t i) { return (int32_t)batch[i].words.size() <= kSplitLength; });
    size_t first = 0;
    while (first < order.size()) {
      const bool short_class = (int32_t)batch[order[first]].words.size() <= kSplitLength;
      size_t last = first;
      while (last < order.size() && ((int32_t)batch[order[last]].words.size() <= kSplitLength) == short_class) ++last;
      std::vector<int32_t> words, offsets{0};
      for (size_t k = first; k < last; ++k) {
        const Request& r = batch[order[k]];
        words.insert(words.end(), r.words.begin(), r.words.end());
        offsets.push_back((int32_t)words.size());
      }
      const int32_t n = (int32_t)(last - first);
      std::vector<std::vector<TreeNode>> vit;
      std::vector<std::vector<MbrNode>> mbr;
      int32_t status = 0;
      std::string err;
      try {
        if (mbr_final_) p.parse(words.data(), offsets.data(), n, nullptr, &mbr);
        else p.parse(words.data(), offsets.data(), n, &vit, nullptr);
      } catch (const std::exception& e) {
        status = -5;
        err = e.what();
      }
      const auto now = std::chrono::steady_clock::now();
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (last == order.size()) claimed_ -= growth;  // allocated by now (or failed)
        for (size_t k = first; k < last; ++k) {
          const Request& r = batch[order[k]];
          Done d;
          d.status = status;
          d.error = err;
          if (status == 0) {
            if (mbr_final_) d.mbr = std::move(mbr[k - first]);
            else d.nodes = std::move(vit[k - first]);
          }
          done_[r.ticket] = std::move(d);
          record_latency(std::chrono::duration<double, std::milli>(now - r.submitted).count());
        }
        outstanding_ -= n;
        ++batches_;
        batched_requests_ += n;
      }
      cv_done_.notify_all();
      first = last;
    }
  }
}

bool Server::result(int64_t ticket, bool wait, std::vector<TreeNode>* nodes, std::vector<MbrNode>* mbr,
                    int32_t* status, std::string* error) {
  std::unique_lock<std::mutex> lk(mu_);
  // Only a ticket that was issued and not yet collected can complete: anything else would wait
  // forever.
  if (!issued_.count(ticket)) throw std::runtime_error("unknown or already collected ticket");
  if (wait) {
    cv_done_.wait(lk, [&] { return done_.count(ticket) > 0; });
  } else if (!done_.count(ticket)) {
    return false;
  }
  Done d = std::move(done_[ticket]);
  done_.erase(ticket);
  issued_.erase(ticket);
  if (nodes) *nodes = std::move(d.nodes);
  if (mbr) *mbr = std::move(d.mbr);
  if (status) *status = d.status;
  if (error) *error = d.error;
  return true;
}

void Server::flush() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return outstanding_ == 0; });
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  ServerStats s;
  s.completed = completed_;
  s.batches = batches_;
  s.mean_batch = batches_ ? (double)batched_requests_ / batches_ : 0.0;
  if (!latencies_ms_.empty()) {
    std::vector<double> v = latencies_ms_;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
    s.p50_ms = pct(0.50);
    s.p90_ms = pct(0.90);
    s.p99_ms = pct(0.99);
    s.max_ms = v.back();
  }
  return s;
}

void Server::reset_stats() {
  std::lock_guard<std::mutex> lk(mu_);
  latencies_ms_.clear();
  latency_next_ = 0;
  completed_ = 0;
  batches_ = 0;
  batched_requests_ = 0;
}

}  // namespace puck
