// Host copies split over a few threads (no HIP dependency).
#pragma once

#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

namespace tc {

// Host copies split over a few threads (the caller works too): one thread's
// memcpy bandwidth, not PCIe, would otherwise bound host-pointer prediction
// and the server's staging.
class CopyPool {
 public:
  using Piece = std::tuple<void*, const void*, size_t>;
  static constexpr size_t kInline = size_t(256) << 10;
  explicit CopyPool(int helpers) {
    for (int i = 0; i < helpers; ++i) th_.emplace_back([this] { run(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  // Copies every piece; returns when all are done.  One caller at a time.
  void copy(const std::vector<Piece>& pieces) {
    if (pieces.empty()) return;
    size_t total = 0;
    for (const Piece& p : pieces) total += std::get<2>(p);
    if (total <= kInline) {  // small: waking the helpers costs more than the copy
      for (const Piece& p : pieces) std::memcpy(std::get<0>(p), std::get<1>(p), std::get<2>(p));
      return;
    }
    std::lock_guard<std::mutex> one(call_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_ = &pieces;
      next_ = 0;
      left_ = pieces.size();
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return left_ == 0; });
    jobs_ = nullptr;
  }

 private:
  void work() {
    while (true) {
      const Piece* pc;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!jobs_ || next_ >= jobs_->size()) return;
        pc = &(*jobs_)[next_++];
      }
      std::memcpy(std::get<0>(*pc), std::get<1>(*pc), std::get<2>(*pc));
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  void run() {
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
      cv_.wait(lk, [this] { return stop_ || (jobs_ && next_ < jobs_->size()); });
      if (stop_) return;
      lk.unlock();
      work();
      lk.lock();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_, call_mu_;
  std::condition_variable cv_, done_;
  const std::vector<Piece>* jobs_ = nullptr;
  size_t next_ = 0, left_ = 0;
  bool stop_ = false;
};

// The process's pool (seven helpers).
CopyPool& copy_pool();
// dst <- src in 1 MiB pieces, for CopyPool::copy.
std::vector<CopyPool::Piece> split_copy(void* dst, const void* src, size_t bytes);

}  // namespace tc
