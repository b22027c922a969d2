// Test infrastructure only: the serving core (turbo-cat_amd/csrc/tc_serve_core.h)
// over a fake device backend, so that admission, batching, completion,
// cancellation, timeouts and failure handling are tested on the CPU.  Built
// by tests/test_serve_core.py with g++; nothing in the product uses it.
//
// The fake "device" is one thread that takes published batches in order and
// writes, for every batch row, out[c] = sum_f x[f] * (f + 1) + c + (tree-by-
// tree flag ? 0.5 : 0): a deterministic function of the row, its channel and
// the row's order flag, so a test can check each output alone.
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tc_serve_core.h"

using namespace tc;

namespace {

class FakeBackend : public ServeBackend {
 public:
  FakeBackend(const ServeConfig& c, uint64_t fail_after)
      : cfg_(c), fail_after_(fail_after), slots_(c.n_slots) {
    for (auto& s : slots_) {
      s.flag.resize(c.max_batch_rows);
      s.sx.resize(static_cast<size_t>(c.max_batch_rows) * c.F);
      s.so.resize(static_cast<size_t>(c.max_batch_rows) * c.K);
      s.done = UINT64_MAX;
    }
    dev_ = std::thread([this] { run(); });
  }
  ~FakeBackend() override { shutdown(); }

  float* stage_x(uint32_t s) override { return slots_[s].sx.data(); }
  uint8_t* stage_flag(uint32_t s) override { return slots_[s].flag.data(); }
  float* stage_out(uint32_t s) override { return slots_[s].so.data(); }
  void publish(uint32_t s, uint64_t seq, uint32_t rows) override {
    std::lock_guard<std::mutex> lk(mu_);
    queue_.push_back({s, seq, rows});
    ++published_;
  }
  bool done(uint32_t s, uint64_t seq) override {
    return slots_[s].done_seq.load(std::memory_order_acquire) == seq + 1;
  }
  std::string health() override {
    if (failed_.load()) return "fake device failed";
    return "";
  }
  void heartbeat() override {}
  void shutdown() override {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    if (dev_.joinable()) dev_.join();
  }

  void hold(bool h) { hold_.store(h); }
  uint64_t published() {
    std::lock_guard<std::mutex> lk(mu_);
    return published_;
  }

 private:
  struct Slot {
    std::vector<uint8_t> flag;
    std::vector<float> sx, so;
    uint64_t done;
    std::atomic<uint64_t> done_seq{0};
  };
  struct Job {
    uint32_t slot;
    uint64_t seq;
    uint32_t rows;
  };

  void run() {
    uint64_t processed = 0;
    while (true) {
      Job j{};
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) return;
        if (queue_.empty() || hold_.load() || failed_.load()) {
          j.rows = 0;
        } else {
          j = queue_.front();
          queue_.pop_front();
        }
      }
      if (j.rows == 0) {
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        continue;
      }
      if (fail_after_ && processed + 1 >= fail_after_) {  // the device dies on this batch
        failed_.store(true);
        continue;
      }
      Slot& s = slots_[j.slot];
      for (uint32_t i = 0; i < j.rows; ++i) {
        const float* x = s.sx.data() + static_cast<size_t>(i) * cfg_.F;
        float v = 0.f;
        for (uint32_t f = 0; f < cfg_.F; ++f) v += x[f] * static_cast<float>(f + 1);
        float* o = s.so.data() + static_cast<size_t>(i) * cfg_.K;
        for (uint32_t c = 0; c < cfg_.K; ++c)
          o[c] = v + static_cast<float>(c) + (s.flag[i] ? 0.5f : 0.f);
      }
      ++processed;
      s.done_seq.store(j.seq + 1, std::memory_order_release);
    }
  }

  ServeConfig cfg_;
  uint64_t fail_after_;
  std::vector<Slot> slots_;
  std::mutex mu_;
  std::deque<Job> queue_;
  std::atomic<bool> hold_{false}, failed_{false};
  bool stop_ = false;
  uint64_t published_ = 0;
  std::thread dev_;
};

struct Handle {
  FakeBackend* be = nullptr;
  std::unique_ptr<ServeCore> core;
  std::string err;
};

}  // namespace

extern "C" {

void* fake_start(uint32_t F, uint32_t K, uint32_t max_batch_rows, uint64_t budget,
                 uint32_t n_slots, uint64_t fail_batch, uint64_t device_fail_after, int hold) {
  ServeConfig c;
  c.F = F;
  c.K = K;
  c.max_batch_rows = max_batch_rows;
  c.memory_budget = budget;
  c.n_slots = n_slots;
  c.fail_batch = fail_batch;
  auto* h = new Handle();
  auto be = std::make_unique<FakeBackend>(c, device_fail_after);
  be->hold(hold != 0);
  h->be = be.get();
  h->core = std::make_unique<ServeCore>(c, std::move(be));
  return h;
}

void fake_hold(void* p, int hold) { static_cast<Handle*>(p)->be->hold(hold != 0); }
uint64_t fake_published(void* p) { return static_cast<Handle*>(p)->be->published(); }

int32_t fake_submit(void* p, const float* X, uint32_t n, float* out, uint64_t* ticket) {
  auto* h = static_cast<Handle*>(p);
  return h->core->submit(X, n, out, ticket, &h->err);
}
int32_t fake_wait(void* p, uint64_t ticket, int64_t timeout_us) {
  auto* h = static_cast<Handle*>(p);
  return h->core->wait(ticket, timeout_us, &h->err);
}
int32_t fake_cancel(void* p, uint64_t ticket) {
  auto* h = static_cast<Handle*>(p);
  return h->core->cancel(ticket, &h->err);
}
void fake_counters(void* p, uint64_t* out6) {
  const ServeCore::Stats s = static_cast<Handle*>(p)->core->stats();
  out6[0] = s.batches; out6[1] = s.rows; out6[2] = s.peak_admitted_bytes;
  out6[3] = s.max_inflight; out6[4] = s.cancelled; out6[5] = s.failed;
}
uint64_t fake_latencies(void* p, float* us, uint64_t cap) {
  return static_cast<Handle*>(p)->core->latencies(us, cap);
}
const char* fake_error(void* p) { return static_cast<Handle*>(p)->err.c_str(); }
void fake_stop(void* p) {
  auto* h = static_cast<Handle*>(p);
  h->core->stop();
  h->core.reset();
  delete h;
}

}  // extern "C"
