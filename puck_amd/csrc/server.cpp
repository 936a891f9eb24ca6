// Continuous-batching parse server. See server.h.
#include "server.h"

#include <algorithm>
#include <climits>
#include <stdexcept>

namespace puck {

Server::Server(std::vector<std::shared_ptr<Grammar>> layers, int32_t device, int64_t mem_bytes,
               int32_t max_inflight, int32_t max_batch, int32_t items_per_lane, bool persistent)
    : max_batch_(std::max(1, max_batch)) {
  if (max_inflight < 1) throw std::runtime_error("max_inflight must be >= 1");
  mbr_final_ = layers.back()->semiring != kViterbi;
  const int64_t per_slot = mem_bytes / max_inflight;
  int32_t max_cell = 0;
  num_words_ = INT32_MAX;
  for (auto& g : layers) {
    max_cell = std::max(max_cell, g->cell_size);
    num_words_ = std::min(num_words_, (int32_t)g->lex_offsets.size() - 1);
  }
  // Same split as Parser::make_batches: 30% of a slot's memory for two charts per layer.
  max_cells_ = std::max<int64_t>(1, (int64_t)(per_slot * 0.3) / ((int64_t)max_cell * 4));
  for (int32_t i = 0; i < max_inflight; ++i) {
    parsers_.emplace_back(new Parser(layers, device, per_slot, items_per_lane));
    parsers_.back()->timing = false;
    parsers_.back()->set_persistent(persistent);
  }
  for (int32_t i = 0; i < max_inflight; ++i) threads_.emplace_back(&Server::worker, this, i);
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
  queue_.push_back(std::move(r));
  ++outstanding_;
  cv_work_.notify_one();
  return queue_.back().ticket;
}

void Server::worker(int32_t slot) {
  Parser& p = *parsers_[slot];
  while (true) {
    std::vector<Request> batch;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
      if (stop_ && queue_.empty()) return;
      // Admit requests in arrival order while they fit this slot's chart budget.
      int64_t cells = 0;
      while (!queue_.empty() && (int32_t)batch.size() < max_batch_) {
        const int64_t n = (int64_t)queue_.front().words.size();
        const int64_t c = n * (n + 1);
        if (!batch.empty() && cells + c > max_cells_) break;
        cells += c;
        batch.push_back(std::move(queue_.front()));
        queue_.pop_front();
      }
      if (!queue_.empty()) cv_work_.notify_one();
    }
    std::vector<int32_t> words, offsets{0};
    for (auto& r : batch) {
      words.insert(words.end(), r.words.begin(), r.words.end());
      offsets.push_back((int32_t)words.size());
    }
    std::vector<std::vector<TreeNode>> vit;
    std::vector<std::vector<MbrNode>> mbr;
    int32_t status = 0;
    std::string err;
    try {
      if (mbr_final_) p.parse(words.data(), offsets.data(), (int32_t)batch.size(), nullptr, &mbr);
      else p.parse(words.data(), offsets.data(), (int32_t)batch.size(), &vit, nullptr);
    } catch (const std::exception& e) {
      status = -5;
      err = e.what();
    }
    const auto now = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (size_t i = 0; i < batch.size(); ++i) {
        Done d;
        d.status = status;
        d.error = err;
        if (status == 0) {
          if (mbr_final_) d.mbr = std::move(mbr[i]);
          else d.nodes = std::move(vit[i]);
        }
        done_[batch[i].ticket] = std::move(d);
        latencies_ms_.push_back(std::chrono::duration<double, std::milli>(now - batch[i].submitted).count());
      }
      outstanding_ -= (int64_t)batch.size();
      ++batches_;
      batched_requests_ += (int64_t)batch.size();
    }
    cv_done_.notify_all();
  }
}

bool Server::result(int64_t ticket, bool wait, std::vector<TreeNode>* nodes, std::vector<MbrNode>* mbr,
                    int32_t* status, std::string* error) {
  std::unique_lock<std::mutex> lk(mu_);
  if (wait) {
    cv_done_.wait(lk, [&] { return done_.count(ticket) > 0; });
  } else if (!done_.count(ticket)) {
    return false;
  }
  Done d = std::move(done_[ticket]);
  done_.erase(ticket);
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
  s.completed = (int64_t)latencies_ms_.size();
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
  batches_ = 0;
  batched_requests_ = 0;
}

}  // namespace puck
