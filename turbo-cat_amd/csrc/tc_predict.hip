// A15-A18: symmetric-ensemble inference on gfx950 (kernels in
// tc_predict_kernels.h), the plain entry points, and the continuous-batching
// server: the host logic of tc_serve_core.h over a device backend whose
// serving kernel (k_serve) reads batches published in uncached pinned host
// memory.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "tc_copy_pool.h"
#include "tc_internal.h"
#include "tc_predict_kernels.h"
#include "tc_serve_core.h"

namespace tc {

CopyPool& copy_pool() {
  static CopyPool pool(7);
  return pool;
}

std::vector<CopyPool::Piece> split_copy(void* dst, const void* src, size_t bytes) {
  constexpr size_t kPiece = size_t(1) << 20;
  std::vector<CopyPool::Piece> v;
  for (size_t o = 0; o < bytes; o += kPiece)
    v.emplace_back(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o,
                   std::min(kPiece, bytes - o));
  return v;
}

// ---------------------------------------------------------------------------
// Pinned host memory handed out by tc_host_alloc: rows in it are transferred
// to the device without a staging copy.
static std::mutex g_host_mu;
static std::map<uintptr_t, size_t> g_host_ranges;  // start -> bytes

static bool host_reachable(const void* p, size_t bytes) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lk(g_host_mu);
  auto it = g_host_ranges.upper_bound(a);
  if (it == g_host_ranges.begin()) return false;
  --it;
  return a >= it->first && a + bytes <= it->first + it->second;
}

// The kernels address pinned host memory by its host address; refuse memory
// whose device address differs (rather than let a kernel read elsewhere).
static void check_same_address(void* p) {
  void* d = nullptr;
  TC_HIP(hipHostGetDevicePointer(&d, p, 0));
  if (d != p) {
    hfree(p);
    throw Error(TC_EDEVICE, "pinned host memory has a device address other than its host address");
  }
}

// Every running serving kernel's host stop word: at process exit they are all
// raised, so that no persistent kernel outlives a process that did not stop
// its servers (the runtime's teardown would wait for it).
static std::mutex g_serve_mu;
static std::vector<uint32_t*> g_serve_stops;
static void stop_all_servers() {
  std::lock_guard<std::mutex> lk(g_serve_mu);
  for (uint32_t* s : g_serve_stops) __atomic_store_n(s, 1u, __ATOMIC_RELEASE);
}
static void register_server(uint32_t* stop_word) {
  static bool hooked = false;
  std::lock_guard<std::mutex> lk(g_serve_mu);
  if (!hooked) {
    std::atexit(stop_all_servers);
    hooked = true;
  }
  g_serve_stops.push_back(stop_word);
}
static void unregister_server(uint32_t* stop_word) {
  std::lock_guard<std::mutex> lk(g_serve_mu);
  g_serve_stops.erase(std::remove(g_serve_stops.begin(), g_serve_stops.end(), stop_word),
                      g_serve_stops.end());
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free prediction entry point: a file-scope device
// array (one copy per GPU, held until the process ends).  Host code never
// passes the array's name to a kernel or a copy: it uses the address
// hipGetSymbolAddress returns on each device, taken once per device when the
// device's first context is created (scratch_setup, called by tc_ctx_create
// with the device current).
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_dev;  // per device: the address of `scratch`, or null
}  // namespace

void scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev.size() <= static_cast<size_t>(device)) g_scratch_dev.resize(device + 1, nullptr);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw Error(TC_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_dev[device] = p;
}

static unsigned char* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev.size() <= static_cast<size_t>(device) || !g_scratch_dev[device])
    throw Error(TC_ESTATE, "scratch of this device was not set up (no context created on it)");
  return static_cast<unsigned char*>(g_scratch_dev[device]);
}

// ---------------------------------------------------------------------------
// Ensemble model and launch shapes.

static int dmax_of(uint32_t D) { return D <= 6 ? 6 : (D <= 8 ? 8 : 16); }

template <int DMAX>
static void pack_records(uint32_t T, uint32_t D, const uint16_t* feature, const float* fthr,
                         const uint8_t* bin, const uint16_t* depth, std::vector<uint32_t>& rec) {
  using L = TreeRec<DMAX>;
  rec.assign(static_cast<size_t>(std::max(1u, T)) * L::kDwords, 0u);
  for (uint32_t t = 0; t < T; ++t) {
    uint32_t* r = rec.data() + static_cast<size_t>(t) * L::kDwords;
    const uint32_t dep = depth[t];
    for (uint32_t d = 0; d < dep; ++d) {
      const size_t k = static_cast<size_t>(t) * D + d;
      uint32_t fb;
      std::memcpy(&fb, fthr + k, 4);
      r[L::kThr + d] = fb;
      r[L::kFeat + d / 2] |= static_cast<uint32_t>(feature[k]) << (16 * (d & 1));
      r[L::kBin + d / 4] |= static_cast<uint32_t>(bin ? bin[k] : 0u) << (8 * (d & 3));
    }
    r[L::kDepth] = dep;
  }
}

EnsDev ens_dev(const tc_ensemble* e) {
  EnsDev E;
  E.rec = e->rec; E.leaves = e->leaves; E.weight = e->weight;
  E.T = e->T; E.D = e->D; E.K = e->K; E.base = e->base;
  return E;
}

// Rows per tile and LDS bytes: as many rows (<= 64) as keep three
// workgroups per CU, else two, else one.
void predict_shape(uint32_t F, uint32_t K, bool binned, uint32_t* R, size_t* lds) {
  const bool multi = K > 1;
  const size_t head = pred_head_bytes(multi, K);
  const size_t row = static_cast<size_t>(pred_stride(F, binned)) * (binned ? 1 : 4);
  constexpr size_t kLds = 160 * 1024, kStatic = 2048;
  for (uint32_t per : {3u, 2u, 1u}) {
    const size_t room = kLds / per - kStatic;
    if (room <= head + row) continue;
    const size_t r = std::min<size_t>(kPredRowsMax, (room - head) / row);
    if (r >= 16 || per == 1) {
      *R = static_cast<uint32_t>(r);
      *lds = head + r * row;
      return;
    }
  }
  throw Error(TC_EINVAL, "too many features for one LDS row tile");
}

// First row of a call of n rows that the reference sums tree by tree: past
// the last full 8-row block; every row of a binned call under 16 rows, or of
// an ensemble without splits.
uint32_t seq_rows_from(const tc_ensemble* e, bool binned, uint32_t n) {
  if (!e->any_split || (binned && n < 16)) return 0u;
  return n / 8u * 8u;
}

template <int DMAX>
static void launch_dmax(const tc_ensemble* e, const void* X, bool binned, uint32_t n, uint32_t F,
                        uint32_t ldc, int transform, float* out, hipStream_t s) {
  uint32_t R;
  size_t lds;
  predict_shape(F, e->K, binned, &R, &lds);
  const dim3 grid((n + R - 1) / R);
  const uint32_t from = seq_rows_from(e, binned, n);
  auto pick = [&](auto kern) {
    if (lds > 64 * 1024)
      TC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(kern, grid, dim3(kPredThreads), lds, s, ens_dev(e), X, n, F, ldc, R,
                       transform, out, from);
  };
  if (e->K > 1) {
    if (binned) pick(&k_predict<true, DMAX, true>);
    else pick(&k_predict<false, DMAX, true>);
  } else {
    if (binned) pick(&k_predict<true, DMAX, false>);
    else pick(&k_predict<false, DMAX, false>);
  }
  TC_HIP(hipGetLastError());
}

void launch_predict(const tc_ensemble* e, const void* X, bool binned, uint32_t n, uint32_t F,
                    uint32_t ldc, int transform, float* out, hipStream_t s) {
  if (n == 0) return;
  switch (e->dmax) {
    case 6: launch_dmax<6>(e, X, binned, n, F, ldc, transform, out, s); break;
    case 8: launch_dmax<8>(e, X, binned, n, F, ldc, transform, out, s); break;
    default: launch_dmax<16>(e, X, binned, n, F, ldc, transform, out, s); break;
  }
}

static void launch_leaf_index(const tc_ensemble* e, const float* X, uint32_t n, uint32_t F,
                              uint32_t* li, hipStream_t s) {
  const dim3 grid((n + 255) / 256);
  switch (e->dmax) {
    case 6: hipLaunchKernelGGL(k_leaf_index<6>, grid, dim3(256), 0, s, ens_dev(e), X, n, F, li); break;
    case 8: hipLaunchKernelGGL(k_leaf_index<8>, grid, dim3(256), 0, s, ens_dev(e), X, n, F, li); break;
    default: hipLaunchKernelGGL(k_leaf_index<16>, grid, dim3(256), 0, s, ens_dev(e), X, n, F, li); break;
  }
  TC_HIP(hipGetLastError());
}

// Device buffer and pinned staging of the host-pointer entry points, kept
// with the ensemble and grown on demand (one host thread per context).
static float* ens_device_buf(tc_ensemble* e, size_t floats) {
  if (floats > e->d_cap) {
    dfree(e->d_buf);
    e->d_buf = nullptr;
    e->d_cap = 0;
    e->d_buf = dmalloc<float>(floats);
    e->d_cap = floats;
  }
  return e->d_buf;
}
static float* ens_host_buf(tc_ensemble* e, size_t floats) {
  if (floats > e->h_cap) {
    if (e->h_buf) hfree(e->h_buf);
    e->h_buf = nullptr;
    e->h_cap = 0;
    TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&e->h_buf), floats * 4, hipHostMallocDefault));
    e->h_cap = floats;
  }
  return e->h_buf;
}

// ---------------------------------------------------------------------------
// The server's device backend.  Everything the host writes and the device
// reads (the descriptor ring, the staged rows and flags) and the outputs the
// device writes live in uncached pinned host memory (never held in a GPU
// cache, so the device cannot read a line older than the host's last write);
// the persistent kernel indexes it by (slot, row) and checks every
// host-written index before use.
template <typename T>
static T* uncached_alloc(size_t count) {
  void* p = nullptr;
  TC_HIP(hipHostMalloc(&p, std::max<size_t>(1, count) * sizeof(T),
                       hipHostMallocUncached | hipHostMallocMapped));
  check_same_address(p);
  return static_cast<T*>(p);
}

class GpuServeBackend : public ServeBackend {
 public:
  GpuServeBackend(tc_ensemble* e, const ServeConfig& c, int transform) : ens_(ref(e)), cfg_(c) {
    try {
      setup(transform);
    } catch (...) {
      release();
      unref(ens_);
      throw;
    }
  }
  ~GpuServeBackend() override {
    release();
    unref(ens_);
  }

  float* stage_x(uint32_t s) override {
    return sx_ + static_cast<size_t>(s) * cfg_.max_batch_rows * cfg_.F;
  }
  uint8_t* stage_flag(uint32_t s) override {
    return sf_ + static_cast<size_t>(s) * cfg_.max_batch_rows;
  }
  float* stage_out(uint32_t s) override {
    return so_ + static_cast<size_t>(s) * cfg_.max_batch_rows * cfg_.K;
  }

  // A descriptor position is free once the batch kServeRing before it is done.
  bool ready(uint64_t seq) override {
    if (seq < kServeRing) return true;
    return __atomic_load_n(&host_->done[seq % kServeRing], __ATOMIC_ACQUIRE) >= seq - kServeRing + 1;
  }

  void publish(uint32_t s, uint64_t seq, uint32_t rows) override {
    const uint32_t pos = static_cast<uint32_t>(seq % kServeRing);
    ServeHostDesc& d = host_->ring[pos];
    tile0_[pos] = tiles_;
    d.tile0 = tiles_;
    d.slot = s;
    d.rows = rows;
    __atomic_store_n(&d.seq, seq, __ATOMIC_RELEASE);  // after the other fields
    tiles_ += (rows + R_ - 1) / R_;
    published_ = seq + 1;
    pump();
  }

  bool done(uint32_t, uint64_t seq) override {
    return __atomic_load_n(&host_->done[seq % kServeRing], __ATOMIC_ACQUIRE) >= seq + 1;
  }

  std::string health() override {
    if (running_) reap();
    return error_;
  }

  void heartbeat() override {
    __atomic_store_n(&host_->heartbeat, ++beats_, __ATOMIC_RELEASE);
    if (published_ > __atomic_load_n(&host_->taken, __ATOMIC_ACQUIRE)) pump();
  }

  void copy(const std::vector<std::tuple<void*, const void*, size_t>>& pieces) override {
    copy_pool().copy(pieces);
  }

  uint64_t launches() const override { return launches_; }

  void shutdown() override {
    if (!began_) return;
    __atomic_store_n(&host_->stop, 1u, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(stream_);  // a running launch ends on the stop word
    running_ = false;
    unregister_server(&host_->stop);
    began_ = false;
    serving_end();
  }

 private:
  // When a published batch is not taken and no launch is running, launches
  // the kernel from the first batch not taken.
  void pump() {
    if (running_ && !reap()) return;
    if (!error_.empty()) return;
    const uint64_t taken = __atomic_load_n(&host_->taken, __ATOMIC_ACQUIRE);
    if (published_ <= taken) return;
    try {
      launch(taken, tile0_[taken % kServeRing]);
    } catch (const std::exception& ex) {
      error_ = std::string("serving kernel launch failed: ") + ex.what();
    }
  }

  // True once the running launch has ended; a fault it raised, or a launch
  // failure, becomes the server's error.
  bool reap() {
    const hipError_t q = hipEventQuery(end_);
    if (q == hipErrorNotReady) return false;
    running_ = false;
    if (q != hipSuccess) {
      error_ = std::string("serving kernel failed: ") + hipGetErrorString(q);
      return true;
    }
    uint32_t fault = 0;
    if (hipMemcpyAsync(&fault, &ctl_->fault, 4, hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
      error_ = "serving kernel state could not be read";
    else if (fault & 2u)
      error_ = "serving kernel stopped: a batch descriptor failed its checks";
    else if (fault & 1u)
      error_ = "serving kernel stopped: no published batch held a claimed tile";
    return true;
  }

  void launch(uint64_t batch0, uint64_t tile0) {
    TC_HIP(hipMemsetAsync(ctl_, 0, sizeof(ServeCtl), stream_));
    ServeLaunch L;
    L.batch0 = batch0;
    L.tile0 = tile0;
    L.hb_ticks = hb_ticks_;
    L.idle_ticks = idle_ticks_;
    L.slice_ticks = slice_ticks_;
    hipLaunchKernelGGL(kern_, dim3(grid_), dim3(kPredThreads), lds_, stream_, ens_dev(ens_), ctl_,
                       host_, stage_, cfg_.F, R_, transform_, L);
    TC_HIP(hipGetLastError());
    TC_HIP(hipEventRecord(end_, stream_));
    running_ = true;
    ++launches_;
  }

  void setup(int transform) {
    TC_HIP(hipSetDevice(ens_->ctx->device));
    const uint32_t F = cfg_.F, K = cfg_.K;
    const size_t rows = static_cast<size_t>(cfg_.n_slots) * cfg_.max_batch_rows;
    predict_shape(F, K, false, &R_, &lds_);
    transform_ = transform;
    host_ = uncached_alloc<ServeHost>(1);
    std::memset(static_cast<void*>(host_), 0, sizeof(ServeHost));
    for (auto& d : host_->ring) d.seq = kServeNoBatch;
    sx_ = uncached_alloc<float>(rows * F);
    sf_ = uncached_alloc<uint8_t>(rows);
    so_ = uncached_alloc<float>(rows * K);
    ctl_ = dmalloc<ServeCtl>(1);
    TC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    TC_HIP(hipEventCreateWithFlags(&end_, hipEventDisableTiming));
    int cus = 0, rate_khz = 0;
    TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ens_->ctx->device));
    TC_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ens_->ctx->device));
    if (rate_khz <= 0) rate_khz = 100000;  // the gfx9 wall clock runs at 100 MHz
    // one worker per CU: the rest of each CU stays free for other work
    grid_ = 1u + static_cast<uint32_t>(std::max(1, cus));
    const unsigned long long per_ms = static_cast<unsigned long long>(rate_khz);
    hb_ticks_ = 2000ull * per_ms;    // 2 s without a heartbeat: the host is gone
    idle_ticks_ = per_ms / 2;        // 0.5 ms without a batch
    slice_ticks_ = 10ull * per_ms;   // 10 ms per launch at most
    stage_.x = sx_;
    stage_.flag = sf_;
    stage_.out = so_;
    stage_.n_slots = cfg_.n_slots;
    stage_.max_rows = cfg_.max_batch_rows;
    const bool multi = K > 1;
    switch (ens_->dmax) {
      case 6: kern_ = multi ? &k_serve<6, true> : &k_serve<6, false>; break;
      case 8: kern_ = multi ? &k_serve<8, true> : &k_serve<8, false>; break;
      default: kern_ = multi ? &k_serve<16, true> : &k_serve<16, false>; break;
    }
    if (lds_ > 64 * 1024)
      TC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern_),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_)));
    serving_begin();  // from here on, the library's frees wait for the server's stop
    began_ = true;
    register_server(&host_->stop);
  }

  void release() {
    shutdown();
    if (end_) (void)hipEventDestroy(end_);
    if (stream_) (void)hipStreamDestroy(stream_);
    dfree(ctl_);
    for (void* p : {static_cast<void*>(sx_), static_cast<void*>(sf_), static_cast<void*>(so_),
                    static_cast<void*>(host_)})
      if (p) hfree(p);
    end_ = nullptr;
    stream_ = nullptr;
    ctl_ = nullptr;
    host_ = nullptr;
    sx_ = nullptr;
    sf_ = nullptr;
    so_ = nullptr;
  }

  tc_ensemble* ens_;
  ServeConfig cfg_;
  uint32_t R_ = 64;
  ServeHost* host_ = nullptr;
  ServeCtl* ctl_ = nullptr;
  float* sx_ = nullptr;
  uint8_t* sf_ = nullptr;
  float* so_ = nullptr;
  hipStream_t stream_ = nullptr;
  hipEvent_t end_ = nullptr;
  using ServeKernel = void (*)(EnsDev, ServeCtl*, ServeHost*, ServeStage, uint32_t, uint32_t, int,
                               ServeLaunch);
  ServeKernel kern_ = nullptr;
  ServeStage stage_{};
  size_t lds_ = 0;
  uint32_t grid_ = 1;
  int transform_ = 0;
  unsigned long long hb_ticks_ = 0, idle_ticks_ = 0, slice_ticks_ = 0;
  bool began_ = false;    // serving_begin called, stop word registered
  bool running_ = false;  // a launch whose end has not been seen
  std::string error_;
  uint64_t published_ = 0, launches_ = 0;
  unsigned long long tile0_[kServeRing] = {};
  unsigned long long tiles_ = 0, beats_ = 0;
};

}  // namespace tc

struct tc_server {
  std::unique_ptr<tc::ServeCore> core;
};

namespace tc {
void unref(tc_ensemble* e) {
  if (!e || e->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  if (e->ctx) (void)hipStreamSynchronize(e->ctx->stream);
  dfree(e->rec); dfree(e->leaves); dfree(e->weight); dfree(e->d_buf);
  hfree(e->h_buf);
  unref(e->ctx);
  delete e;
}
}  // namespace tc

using namespace tc;

extern "C" {

int32_t tc_ensemble_create_multi(tc_ctx* ctx, uint32_t n_trees, uint32_t max_depth,
                                 uint32_t n_classes, const uint16_t* feature, const float* fthr,
                                 const uint8_t* bin, const uint16_t* depth, const float* leaves,
                                 const float* weights, float base, uint32_t n_features,
                                 tc_ensemble** out) {
  return guard([&] {
    TC_REQUIRE(ctx && out && feature && fthr && depth && leaves && weights, "tc_ensemble_create: NULL argument");
    TC_REQUIRE(max_depth >= 1 && max_depth <= 16, "max_depth must be in [1, 16]");
    TC_REQUIRE(n_classes >= 1 && n_classes <= static_cast<uint32_t>(kMaxClasses),
               "n_classes must be in [1, 16]");
    TC_REQUIRE(n_features >= 1 && n_features <= 65535, "n_features must be in [1, 65535]");
    const uint32_t K = n_classes;
    *out = nullptr;
    for (uint32_t t = 0; t < n_trees; ++t) {
      TC_REQUIRE(depth[t] <= max_depth, "tree depth exceeds max_depth");
      for (uint32_t d = 0; d < depth[t]; ++d)
        TC_REQUIRE(feature[t * max_depth + d] < n_features, "split feature out of range");
    }
    auto* e = new tc_ensemble();
    e->ctx = ref(ctx); e->T = n_trees; e->D = max_depth; e->F = n_features; e->K = K; e->base = base;
    e->dmax = dmax_of(max_depth);
    for (uint32_t t = 0; t < n_trees; ++t) e->any_split = e->any_split || depth[t] > 0;
    try {
      std::vector<uint32_t> rec;
      switch (e->dmax) {
        case 6: pack_records<6>(n_trees, max_depth, feature, fthr, bin, depth, rec); break;
        case 8: pack_records<8>(n_trees, max_depth, feature, fthr, bin, depth, rec); break;
        default: pack_records<16>(n_trees, max_depth, feature, fthr, bin, depth, rec); break;
      }
      const size_t TL = std::max<size_t>(1, (static_cast<size_t>(n_trees) << max_depth) * K);
      e->rec = dmalloc<uint32_t>(rec.size());
      e->leaves = dmalloc<float>(TL);
      e->weight = dmalloc<float>(std::max(1u, n_trees));
      hipStream_t s = ctx->stream;
      copy_sync(e->rec, rec.data(), 4 * rec.size(), hipMemcpyHostToDevice, s);
      if (n_trees) {
        copy_sync(e->leaves, leaves, (static_cast<size_t>(n_trees) << max_depth) * K * 4,
                  hipMemcpyHostToDevice, s);
        copy_sync(e->weight, weights, n_trees * 4ull, hipMemcpyHostToDevice, s);
      }
    } catch (...) {
      tc_ensemble_destroy(e);
      throw;
    }
    *out = e;
  });
}

int32_t tc_ensemble_create(tc_ctx* ctx, uint32_t n_trees, uint32_t max_depth,
                           const uint16_t* feature, const float* fthr, const uint8_t* bin,
                           const uint16_t* depth, const float* leaves, const float* weights,
                           float base, uint32_t n_features, tc_ensemble** out) {
  return tc_ensemble_create_multi(ctx, n_trees, max_depth, 1, feature, fthr, bin, depth, leaves,
                                  weights, base, n_features, out);
}

int32_t tc_ensemble_destroy(tc_ensemble* e) {
  return guard([&] { unref(e); });
}

int32_t tc_predict_dev(tc_ensemble* e, const float* X_dev, uint32_t n, uint32_t F, int32_t transform,
                       float* out_dev) {
  return guard([&] {
    TC_REQUIRE(e && (X_dev || n == 0) && (out_dev || n == 0), "tc_predict_dev: NULL argument");
    TC_REQUIRE(F == e->F, "feature count differs from the model's");
    launch_predict(e, X_dev, false, n, F, 0, transform, out_dev, e->ctx->stream);
  });
}

int32_t tc_predict(tc_ensemble* e, const float* X, uint32_t n, uint32_t F, int32_t transform,
                   float* out, uint32_t* leaf_idx) {
  return guard([&] {
    TC_REQUIRE(e && (X || n == 0) && (out || n == 0), "tc_predict: NULL argument");
    TC_REQUIRE(F == e->F, "feature count differs from the model's");
    if (n == 0) return;
    std::lock_guard<std::mutex> lk(e->buf_mu);
    hipStream_t s = e->ctx->stream;
    const size_t nx = static_cast<size_t>(n) * F, no = static_cast<size_t>(n) * e->K;
    const size_t nl = leaf_idx ? static_cast<size_t>(n) * std::max(1u, e->T) : 0;
    float* d = ens_device_buf(e, nx + no + nl);
    float* dout = d + nx;
    uint32_t* dl = reinterpret_cast<uint32_t*>(d + nx + no);
    if (host_reachable(X, 4 * nx)) {
      TC_HIP(hipMemcpyAsync(d, X, 4 * nx, hipMemcpyHostToDevice, s));
    } else {
      // pageable rows: parallel copy into pinned staging, then one transfer
      float* h = ens_host_buf(e, nx);
      copy_pool().copy(split_copy(h, X, 4 * nx));
      TC_HIP(hipMemcpyAsync(d, h, 4 * nx, hipMemcpyHostToDevice, s));
    }
    launch_predict(e, d, false, n, F, 0, transform, dout, s);
    if (leaf_idx && e->T) launch_leaf_index(e, d, n, F, dl, s);
    TC_HIP(hipMemcpyAsync(out, dout, 4 * no, hipMemcpyDeviceToHost, s));
    if (leaf_idx && e->T) TC_HIP(hipMemcpyAsync(leaf_idx, dl, 4 * nl, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
  });
}

int32_t tc_predict_noalloc(tc_ensemble* e, const float* X, uint32_t n, uint32_t F,
                           int32_t transform, float* out) {
  return guard([&] {
    TC_REQUIRE(e && (X || n == 0) && (out || n == 0), "tc_predict_noalloc: NULL argument");
    TC_REQUIRE(F == e->F, "feature count differs from the model's");
    if (n == 0) return;
    // the rows, then the outputs, each at a 256-byte boundary of the scratch
    const size_t xb = (4ull * n * F + 255) & ~size_t(255), ob = 4ull * n * e->K;
    if (xb + ob > kScratchBytes)
      throw Error(TC_EINVAL, "tc_predict_noalloc: " + std::to_string(n) + " rows of " +
                                 std::to_string(F) + " features need " + std::to_string(xb + ob) +
                                 " bytes of scratch, more than its " +
                                 std::to_string(kScratchBytes) + " bytes");
    unsigned char* base = scratch_of(e->ctx->device);
    float* dx = reinterpret_cast<float*>(base);
    float* dout = reinterpret_cast<float*>(base + xb);
    hipStream_t s = e->ctx->stream;
    TC_HIP(hipMemcpyAsync(dx, X, 4ull * n * F, hipMemcpyHostToDevice, s));
    launch_predict(e, dx, false, n, F, 0, transform, dout, s);
    TC_HIP(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
  });
}

int32_t tc_predict_binned(tc_ensemble* e, tc_dataset* ds, float* out) {
  return guard([&] {
    TC_REQUIRE(e && ds && out, "tc_predict_binned: NULL argument");
    TC_REQUIRE(ds->F == e->F, "feature count differs from the model's");
    if (ds->N == 0) return;
    std::lock_guard<std::mutex> lk(e->buf_mu);
    hipStream_t s = e->ctx->stream;
    float* d = ens_device_buf(e, static_cast<size_t>(ds->N) * e->K);
    launch_predict(e, ds->bins_cm, true, ds->N, ds->F, ds->N, 0, d, s);
    TC_HIP(hipMemcpyAsync(out, d, 4ull * ds->N * e->K, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
  });
}

int32_t tc_host_alloc(size_t bytes, void** out) {
  return guard([&] {
    TC_REQUIRE(out, "tc_host_alloc: NULL argument");
    *out = nullptr;
    void* p = nullptr;
    const hipError_t e = hipHostMalloc(&p, std::max<size_t>(1, bytes), hipHostMallocDefault);
    if (e != hipSuccess) throw Error(TC_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(g_host_mu);
    g_host_ranges[reinterpret_cast<uintptr_t>(p)] = std::max<size_t>(1, bytes);
    *out = p;
  });
}

int32_t tc_host_free(void* p) {
  return guard([&] {
    if (!p) return;
    {
      std::lock_guard<std::mutex> lk(g_host_mu);
      auto it = g_host_ranges.find(reinterpret_cast<uintptr_t>(p));
      TC_REQUIRE(it != g_host_ranges.end(), "tc_host_free: not memory from tc_host_alloc");
      g_host_ranges.erase(it);
    }
    TC_HIP(free_or_defer(p, true));
  });
}

int32_t tc_serve_start(tc_ensemble* e, const tc_serve_params* p, tc_server** out) {
  return guard([&] {
    TC_REQUIRE(e && p && out, "tc_serve_start: NULL argument");
    TC_REQUIRE(p->max_batch_rows > 0 && p->max_batch_rows <= (1u << 24) && p->n_streams > 0 &&
                   p->n_streams <= kServeRing,
               "tc_serve_start: bad parameters (max_batch_rows in [1, 2^24], n_streams in [1, 8])");
    *out = nullptr;
    ServeConfig c;
    c.F = e->F;
    c.K = e->K;
    c.max_batch_rows = p->max_batch_rows;
    c.memory_budget = p->memory_budget;
    c.n_slots = p->n_streams;
    c.seq_group = 8;
    c.seq_all = !e->any_split;
    c.fail_batch = p->fail_batch;
    auto be = std::make_unique<GpuServeBackend>(e, c, p->transform);
    auto* S = new tc_server();
    S->core = std::make_unique<ServeCore>(c, std::move(be));
    *out = S;
  });
}

static int32_t serve_status(int32_t rc, const std::string& err) {
  return rc == TC_OK ? TC_OK : fail(rc, err);
}

int32_t tc_serve_submit(tc_server* S, const float* X, uint32_t n, float* out, uint64_t* ticket) {
  if (!S || !ticket) return fail(TC_EINVAL, "tc_serve_submit: bad arguments");
  std::string err;
  return serve_status(S->core->submit(X, n, out, ticket, &err), "tc_serve_submit: " + err);
}

int32_t tc_serve_wait(tc_server* S, uint64_t ticket, int64_t timeout_us) {
  if (!S) return fail(TC_EINVAL, "NULL server");
  std::string err;
  return serve_status(S->core->wait(ticket, timeout_us, &err), "tc_serve_wait: " + err);
}

int32_t tc_serve_cancel(tc_server* S, uint64_t ticket) {
  if (!S) return fail(TC_EINVAL, "NULL server");
  std::string err;
  return serve_status(S->core->cancel(ticket, &err), "tc_serve_cancel: " + err);
}

int32_t tc_serve_stats(tc_server* S, uint64_t* batches, uint64_t* rows) {
  return guard([&] {
    TC_REQUIRE(S, "NULL server");
    const ServeCore::Stats st = S->core->stats();
    if (batches) *batches = st.batches;
    if (rows) *rows = st.rows;
  });
}

int32_t tc_serve_counters(tc_server* S, tc_serve_counters_t* c) {
  return guard([&] {
    TC_REQUIRE(S && c, "tc_serve_counters: NULL argument");
    const ServeCore::Stats st = S->core->stats();
    c->batches = st.batches;
    c->rows = st.rows;
    c->peak_admitted_bytes = st.peak_admitted_bytes;
    c->max_batches_in_flight = st.max_inflight;
    c->cancelled = st.cancelled;
    c->failed = st.failed;
    c->kernel_launches = st.launches;
  });
}

int32_t tc_serve_latencies(tc_server* S, float* us, uint64_t cap, uint64_t* count) {
  return guard([&] {
    TC_REQUIRE(S && count, "tc_serve_latencies: NULL argument");
    *count = S->core->latencies(us, cap);
  });
}

int32_t tc_serve_stop(tc_server* S) {
  return guard([&] {
    if (!S) return;
    S->core->stop();
    delete S;
  });
}

}  // extern "C"
