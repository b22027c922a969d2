// Host-side internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/turbocat_hip.h"
#include "tc_common.h"

// Handles are reference counted (refs): an object keeps what it uses alive
// -- every object its context, a trainer its datasets, a server its
// ensemble -- so the caller may destroy handles in any order.
struct tc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double last_loop_ms = 0.0, last_hist_ms = 0.0;
  uint64_t last_hist_rows = 0;     // regular-tree fits (tc_last_train_hist_rows)
  uint32_t last_hist_launches = 0;
  std::atomic<int> refs{1};
  // two pinned host buffers for streaming rows to the device, made on first use
  static constexpr size_t kStageBytes = size_t(32) << 20;
  void* stage[2] = {nullptr, nullptr};
  std::mutex stage_mu;
};

struct tc_dataset {
  std::atomic<int> refs{1};
  tc_ctx* ctx = nullptr;
  uint32_t N = 0, F = 0, Fp = 0, max_bins = 255;
  uint8_t* bins_rm = nullptr;   // [N][Fp], Fp = rm_stride(F)
  uint8_t* bins_cm = nullptr;   // [F][N]
  float* edges = nullptr;       // [F][255]
  uint32_t* n_edges = nullptr;  // [F]
  std::vector<float> h_edges;
  std::vector<uint32_t> h_n_edges;
};

struct tc_ensemble {
  std::atomic<int> refs{1};
  tc_ctx* ctx = nullptr;
  uint32_t T = 0, D = 0, F = 0;
  uint32_t K = 1;               // output channels (classes of a softmax ensemble)
  float base = 0.f;
  bool any_split = false;       // some tree has depth > 0
  int dmax = 8;                 // depth bound of the tree records (6, 8 or 16)
  uint32_t* rec = nullptr;      // [T] tree records (tc_predict_kernels.h, TreeRec)
  float* leaves = nullptr;      // [T][2^D][K]
  float* weight = nullptr;      // [T]
  // buffers of the host-pointer entry points, grown on demand
  float* d_buf = nullptr;
  size_t d_cap = 0;
  float* h_buf = nullptr;       // pinned
  size_t h_cap = 0;
  std::mutex buf_mu;            // one host-pointer call at a time uses the buffers
};

namespace tc {

// Error plumbing: internal code throws, the C-ABI boundary catches.
struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& msg);
int32_t fail(int32_t code, const std::string& msg);

#define TC_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::tc::Error(TC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

#define TC_REQUIRE(cond, msg)                            \
  do {                                                   \
    if (!(cond)) throw ::tc::Error(TC_EINVAL, (msg));    \
  } while (0)

template <typename F>
int32_t guard(F&& f) {
  try {
    f();
    return TC_OK;
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(TC_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(TC_EDEVICE, e.what());
  }
}

template <typename T>
T* dmalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  const hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) throw Error(TC_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return static_cast<T*>(p);
}

// hipFree and hipHostFree wait until the whole device is idle, and a device
// running a server's persistent kernel never is.  Every free of the library
// goes through these two: while any server runs they are queued, and done
// when the last server has stopped (its kernel finished).
hipError_t free_or_defer(void* p, bool host);
inline void dfree(void* p) { (void)free_or_defer(p, false); }  // device memory from hipMalloc
inline void hfree(void* p) { (void)free_or_defer(p, true); }   // pinned host memory
// A blocking copy ordered on stream s only (never on the null stream, so it
// waits for nothing a server runs).
inline void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  TC_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  TC_HIP(hipStreamSynchronize(s));
}
// Reference counts of the handles: ref takes one more, unref drops one and
// destroys the object with the last.
template <class H>
H* ref(H* h) {
  h->refs.fetch_add(1, std::memory_order_relaxed);
  return h;
}
void unref(tc_ctx* c);
void unref(tc_dataset* d);
void unref(tc_ensemble* e);
// A server's kernel is about to start / has finished.
void serving_begin();
void serving_end();

// Address of the allocation-free prediction's device scratch on `device`,
// taken once per device (the device must be current).
void scratch_setup(int device);

// Pinned staging buffer b (0 or 1) of the context, tc_ctx::kStageBytes.
void* ctx_stage(tc_ctx* c, int b);

// Edges (dataset.cpp:278-334) and bins of X (host, row-major) on the device,
// into ds (allocates its buffers; ds->h_edges / h_n_edges get the edges).
void quantize_on_device(tc_dataset* ds, const float* X, bool wrap);
// Bin X (host, row-major) with ds->h_edges on the device (allocates ds buffers).
void bin_on_device(tc_dataset* ds, const float* X);

}  // namespace tc
