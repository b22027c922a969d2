// Host side of the continuous-batching prediction server, independent of the
// device: admission by a memory budget, packing of admitted requests (whole
// or in pieces) into batches, completion, cancellation, timeouts and failure
// propagation.  The device side is a ServeBackend: a fixed set of batch
// slots, each with host buffers the device reads and writes (the batch's
// rows, one flag per row, the outputs), a publish call that hands a filled
// slot to the device and a poll for its completion.  The GPU backend
// (tc_predict.hip) is one persistent kernel that reads the published slots;
// the core has no HIP dependency.
//
// A row's flag says whether the reference sums that row tree by tree (the
// rows of a request past its last full 8-row block), so a row's value depends
// only on the row, the model and its place in its own request, never on the
// batch the request was packed into.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tc {

struct ServeConfig {
  uint32_t F = 0, K = 1;
  uint32_t max_batch_rows = 0;
  uint64_t memory_budget = 0;    // bytes of admitted request rows (4F + 4K per row)
  uint32_t n_slots = 1;          // batches in flight
  uint32_t seq_group = 8;        // rows of a request past the last full group: tree by tree
  bool seq_all = false;          // every row tree by tree (an ensemble without splits)
  uint64_t fail_batch = 0;       // fault injection: batch number fail_batch - 1 fails (0: none)
};

class ServeBackend {
 public:
  virtual ~ServeBackend() = default;
  virtual float* stage_x(uint32_t slot) = 0;      // max_batch_rows * F floats
  virtual uint8_t* stage_flag(uint32_t slot) = 0;  // max_batch_rows: 1 = tree-by-tree row
  virtual float* stage_out(uint32_t slot) = 0;    // max_batch_rows * K floats
  // Whether batch number `seq` may be published now (a device-side ring
  // position may still hold an earlier batch in flight).
  virtual bool ready(uint64_t seq) { (void)seq; return true; }
  virtual void publish(uint32_t slot, uint64_t seq, uint32_t rows) = 0;
  virtual bool done(uint32_t slot, uint64_t seq) = 0;
  // Empty string while healthy; a message once the device side failed.
  virtual std::string health() = 0;
  virtual void heartbeat() = 0;
  // Copies of request rows into staging; a backend may spread them over
  // threads.  pieces: (dst, src, bytes).
  virtual void copy(const std::vector<std::tuple<void*, const void*, size_t>>& pieces) {
    for (const auto& p : pieces) std::memcpy(std::get<0>(p), std::get<1>(p), std::get<2>(p));
  }
  virtual void shutdown() = 0;  // the device side stops; no slot is read afterwards
  virtual uint64_t launches() const { return 0; }  // device launches so far
};

enum class ReqState { kPending, kAdmitted, kDone, kFailed, kCancelled };

class ServeCore {
 public:
  // Status codes returned to the C-ABI (same values as TC_*).
  static constexpr int32_t kOK = 0, kInval = -1, kDevice = -3, kState = -4, kBusy = -5,
                           kTimeout = -6;

  ServeCore(const ServeConfig& c, std::unique_ptr<ServeBackend> be)
      : cfg_(c), be_(std::move(be)), slots_(c.n_slots) {
    worker_ = std::thread([this] { loop(); });
  }
  ~ServeCore() { stop(); }

  int32_t submit(const float* X, uint32_t n, float* out, uint64_t* ticket, std::string* err) {
    if (!X || !out || n == 0) { *err = "bad arguments"; return kInval; }
    auto rq = std::make_shared<Req>();
    rq->X = X;
    rq->out = out;
    rq->n = n;
    rq->bytes = row_bytes() * n;
    rq->t_submit = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stop_) { *err = "server stopped" + (error_.empty() ? "" : ": " + error_); return kState; }
      rq->id = next_id_++;
      pending_.push_back(rq);
      live_[rq->id] = rq;
    }
    cv_work_.notify_one();
    *ticket = rq->id;
    return kOK;
  }

  // Waits for the request; timeout_us < 0: no limit.
  int32_t wait(uint64_t ticket, int64_t timeout_us, std::string* err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = live_.find(ticket);
    if (it == live_.end()) { *err = "unknown ticket"; return kInval; }
    auto rq = it->second;
    auto ready = [&] { return rq->state == ReqState::kDone || rq->state == ReqState::kFailed ||
                              rq->state == ReqState::kCancelled; };
    if (timeout_us < 0) {
      cv_done_.wait(lk, ready);
    } else if (!cv_done_.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
      *err = "timed out";
      return kTimeout;
    }
    live_.erase(ticket);
    if (rq->state == ReqState::kDone) return kOK;
    if (rq->state == ReqState::kCancelled) { *err = "request cancelled"; return kState; }
    *err = "server failed: " + rq->error;
    return kDevice;
  }

  // Cancels a request none of whose rows has been handed to the device yet;
  // a request already (partly) on the device cannot be cancelled (kBusy).
  int32_t cancel(uint64_t ticket, std::string* err) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = live_.find(ticket);
      if (it == live_.end()) { *err = "unknown ticket"; return kInval; }
      auto rq = it->second;
      if (rq->state != ReqState::kPending && rq->state != ReqState::kAdmitted) {
        *err = "request already finished";
        return kState;
      }
      if (rq->next > 0) { *err = "request already on the device"; return kBusy; }
      if (rq->state == ReqState::kAdmitted) {
        admitted_.erase(std::find(admitted_.begin(), admitted_.end(), rq));
        admitted_bytes_ -= rq->bytes;
      } else {
        pending_.erase(std::find(pending_.begin(), pending_.end(), rq));
      }
      rq->state = ReqState::kCancelled;
      ++n_cancelled_;
    }
    cv_done_.notify_all();
    cv_work_.notify_one();
    return kOK;
  }

  struct Stats {
    uint64_t batches, rows, peak_admitted_bytes, max_inflight, cancelled, failed, launches;
  };
  Stats stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return Stats{n_batches_, n_rows_,     peak_admitted_,  max_inflight_,
                 n_cancelled_, n_failed_, be_->launches()};
  }

  // Drains what was submitted (unless the device failed), then stops.
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stopped_) return;
      stop_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    be_->shutdown();
    std::lock_guard<std::mutex> lk(mu_);
    stopped_ = true;
  }

  std::string error() {
    std::lock_guard<std::mutex> lk(mu_);
    return error_;
  }

  // Latencies (microseconds) of the requests finished so far, from submit
  // to the result in the caller's buffer, in completion order (the first
  // kMaxLatencies).  Copies up to cap of them; returns how many there are.
  uint64_t latencies(float* out, uint64_t cap) {
    std::lock_guard<std::mutex> lk(mu_);
    const uint64_t n = std::min<uint64_t>(cap, latencies_.size());
    if (out && n) std::memcpy(out, latencies_.data(), 4 * n);
    return latencies_.size();
  }

 private:
  struct Req {
    uint64_t id = 0;
    const float* X = nullptr;
    float* out = nullptr;
    uint32_t n = 0;
    uint64_t bytes = 0;
    std::chrono::steady_clock::time_point t_submit;
    uint32_t next = 0;  // next row to place in a batch
    uint32_t done = 0;  // rows completed
    ReqState state = ReqState::kPending;
    std::string error;
  };
  struct Seg {
    std::shared_ptr<Req> req;
    uint32_t row0, count, slot_row;
  };
  struct Slot {
    bool busy = false;
    uint64_t seq = 0;
    uint32_t rows = 0;
    std::vector<Seg> segs;
  };

  uint64_t row_bytes() const { return 4ull * cfg_.F + 4ull * cfg_.K; }

  void finish_request(const std::shared_ptr<Req>& rq, ReqState s, const std::string& why) {
    if (rq->state == ReqState::kDone || rq->state == ReqState::kFailed ||
        rq->state == ReqState::kCancelled)
      return;
    if (rq->state == ReqState::kAdmitted) admitted_bytes_ -= rq->bytes;
    rq->state = s;
    rq->error = why;
    if (s == ReqState::kFailed) ++n_failed_;
    if (s == ReqState::kDone && latencies_.size() < kMaxLatencies)
      latencies_.push_back(static_cast<float>(
          std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - rq->t_submit)
              .count()));
  }

  // Fails every request not yet finished.  Called once nothing the device
  // could still write into is in flight (or the device side is gone).
  void fail_all(const std::string& why) {
    if (error_.empty()) error_ = why;
    stop_ = true;
    for (auto& sl : slots_) {
      for (auto& sg : sl.segs) finish_request(sg.req, ReqState::kFailed, why);
      sl.segs.clear();
      sl.busy = false;
    }
    inflight_ = 0;
    for (auto& rq : failed_segs_) finish_request(rq, ReqState::kFailed, why);
    failed_segs_.clear();
    for (auto& rq : admitted_) finish_request(rq, ReqState::kFailed, why);
    for (auto& rq : pending_) finish_request(rq, ReqState::kFailed, why);
    admitted_.clear();
    pending_.clear();
    cv_done_.notify_all();
  }

  void admit() {
    // a request larger than the whole budget is admitted alone when nothing
    // else is admitted
    while (!pending_.empty()) {
      const auto& rq = pending_.front();
      if (admitted_bytes_ + rq->bytes > cfg_.memory_budget && admitted_bytes_ > 0) break;
      admitted_bytes_ += rq->bytes;
      peak_admitted_ = std::max(peak_admitted_, admitted_bytes_);
      rq->state = ReqState::kAdmitted;
      admitted_.push_back(rq);
      pending_.pop_front();
    }
  }

  bool retire() {
    bool any = false;
    for (uint32_t s = 0; s < slots_.size(); ++s) {
      Slot& sl = slots_[s];
      if (!sl.busy || !be_->done(s, sl.seq)) continue;
      const float* so = be_->stage_out(s);
      for (auto& sg : sl.segs) {
        Req& rq = *sg.req;
        std::memcpy(rq.out + static_cast<size_t>(sg.row0) * cfg_.K,
                    so + static_cast<size_t>(sg.slot_row) * cfg_.K, 4ull * sg.count * cfg_.K);
        rq.done += sg.count;
        if (rq.done == rq.n) finish_request(sg.req, ReqState::kDone, "");
      }
      sl.segs.clear();
      sl.busy = false;
      --inflight_;
      any = true;
    }
    if (any) cv_done_.notify_all();
    return any;
  }

  // Packs the head of the admitted queue into a free slot and publishes it.
  // Called with the lock held; releases it around the staging copies.
  bool fill(std::unique_lock<std::mutex>& lk) {
    bool any = false;
    for (uint32_t s = 0; s < slots_.size(); ++s) {
      Slot& sl = slots_[s];
      if (sl.busy || admitted_.empty() || stop_failed_ || !be_->ready(next_batch_)) continue;
      uint32_t rows = 0;
      while (!admitted_.empty() && rows < cfg_.max_batch_rows) {
        auto rq = admitted_.front();
        const uint32_t take = std::min(rq->n - rq->next, cfg_.max_batch_rows - rows);
        sl.segs.push_back(Seg{rq, rq->next, take, rows});
        rq->next += take;
        rows += take;
        if (rq->next == rq->n) admitted_.pop_front();
      }
      if (rows == 0) continue;
      const uint64_t seq = next_batch_++;
      sl.busy = true;
      sl.seq = seq;
      sl.rows = rows;
      ++inflight_;
      max_inflight_ = std::max(max_inflight_, inflight_);
      ++n_batches_;
      n_rows_ += rows;
      // the slot is ours: build its rows with the queue unlocked
      std::vector<Seg> segs = sl.segs;
      lk.unlock();
      float* sx = be_->stage_x(s);
      uint8_t* sf = be_->stage_flag(s);
      std::vector<std::tuple<void*, const void*, size_t>> pieces;
      const uint32_t F = cfg_.F;
      for (const Seg& sg : segs) {
        const Req& rq = *sg.req;
        const uint32_t from = cfg_.seq_all ? 0u : rq.n / cfg_.seq_group * cfg_.seq_group;
        for (uint32_t k = 0; k < sg.count; ++k) sf[sg.slot_row + k] = sg.row0 + k >= from ? 1u : 0u;
        constexpr size_t kPiece = size_t(1) << 20;
        const char* src = reinterpret_cast<const char*>(rq.X + static_cast<size_t>(sg.row0) * F);
        char* dst = reinterpret_cast<char*>(sx + static_cast<size_t>(sg.slot_row) * F);
        const size_t bytes = 4ull * sg.count * F;
        for (size_t o = 0; o < bytes; o += kPiece)
          pieces.emplace_back(dst + o, src + o, std::min(kPiece, bytes - o));
      }
      be_->copy(pieces);
      const bool inject = cfg_.fail_batch != 0 && seq + 1 == cfg_.fail_batch;
      if (!inject) be_->publish(s, seq, rows);
      lk.lock();
      if (inject) {
        // as if the device had reported an error on this batch: its requests
        // fail once the other batches in flight (which may write into the
        // same requests' memory) have completed
        failing_ = "injected failure of batch " + std::to_string(seq);
        for (auto& sg : sl.segs) failed_segs_.push_back(sg.req);
        sl.segs.clear();
        sl.busy = false;
        --inflight_;
        stop_ = true;
        return true;
      }
      any = true;
    }
    return any;
  }

  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    auto last_health = std::chrono::steady_clock::now();
    while (true) {
      if (failing_.empty() && !stop_failed_) admit();
      bool progressed = retire();
      if (failing_.empty() && !stop_failed_) progressed |= fill(lk);
      const auto now = std::chrono::steady_clock::now();
      be_->heartbeat();
      if (!stop_failed_ && now - last_health > std::chrono::milliseconds(1)) {
        last_health = now;
        const std::string h = be_->health();
        if (!h.empty()) {
          // batches the device finished before it failed keep their results:
          // their done flags may have been set after this iteration's retire
          retire();
          fail_all(h);
          stop_failed_ = true;
        }
      }
      if (!failing_.empty() && inflight_ == 0) {
        fail_all(failing_);
        failing_.clear();
        stop_failed_ = true;
      }
      const bool drained = inflight_ == 0 && admitted_.empty() && pending_.empty();
      if (stop_ && inflight_ == 0 && failing_.empty() && (drained || stop_failed_)) break;
      if (!progressed) {
        if (inflight_ > 0) {
          // batches on the device: poll their completion flags (host memory)
          lk.unlock();
          std::this_thread::yield();
          lk.lock();
        } else {
          cv_work_.wait_for(lk, std::chrono::milliseconds(2));
        }
      }
    }
    cv_done_.notify_all();
  }

  ServeConfig cfg_;
  std::unique_ptr<ServeBackend> be_;
  std::vector<Slot> slots_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<std::shared_ptr<Req>> pending_, admitted_;
  std::unordered_map<uint64_t, std::shared_ptr<Req>> live_;
  uint64_t admitted_bytes_ = 0, peak_admitted_ = 0;
  uint64_t next_id_ = 1, next_batch_ = 0;
  uint64_t n_batches_ = 0, n_rows_ = 0, n_cancelled_ = 0, n_failed_ = 0;
  uint64_t inflight_ = 0, max_inflight_ = 0;
  bool stop_ = false, stop_failed_ = false, stopped_ = false;
  std::string error_;
  static constexpr size_t kMaxLatencies = size_t(1) << 20;
  std::vector<float> latencies_;
  std::string failing_;                          // a batch failed: fail once in-flight drains
  std::vector<std::shared_ptr<Req>> failed_segs_;  // requests of the failed batch
  std::thread worker_;
};

}  // namespace tc
