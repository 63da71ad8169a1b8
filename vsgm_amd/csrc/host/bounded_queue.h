// Bounded blocking queue of holder ids between the pm_pipeline producers
// and consumers (the role of rigtorp::MPMCQueue in
// src/app/pattern_matching_pipeline.cu:106-110, without its busy waiting).
#pragma once

#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <queue>

namespace vsgm {

class BoundedQueue {
 public:
  explicit BoundedQueue(size_t cap) : cap_(cap) {}
  // Blocks while the queue is full; returns false (and drops v) once closed.
  bool push(int v) {
    std::unique_lock<std::mutex> lk(mu_);
    not_full_.wait(lk, [&] { return q_.size() < cap_ || closed_; });
    if (closed_) return false;
    q_.push(v);
    not_empty_.notify_one();
    return true;
  }
  // Blocks until an item arrives, or returns false once the queue is
  // closed and drained.
  bool pop(int& v) {
    std::unique_lock<std::mutex> lk(mu_);
    not_empty_.wait(lk, [&] { return !q_.empty() || closed_; });
    if (q_.empty()) return false;
    v = q_.front();
    q_.pop();
    not_full_.notify_one();
    return true;
  }
  void close() {
    std::lock_guard<std::mutex> lk(mu_);
    closed_ = true;
    not_empty_.notify_all();
    not_full_.notify_all();
  }

 private:
  size_t cap_;
  bool closed_ = false;
  std::queue<int> q_;
  std::mutex mu_;
  std::condition_variable not_full_, not_empty_;
};

}  // namespace vsgm
