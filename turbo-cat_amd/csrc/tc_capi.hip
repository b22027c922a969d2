// C-ABI: contexts, errors, device memory plumbing, datasets (A2/A3).
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "tc_internal.h"

namespace tc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int32_t fail(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {
// Frees are made under this lock, so that no server kernel can start while
// one is waiting for the device to go idle.
std::mutex g_free_mu;
int g_servers = 0;
std::vector<std::pair<void*, bool>> g_deferred;  // (pointer, pinned host memory)
}  // namespace

hipError_t free_or_defer(void* p, bool host) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_free_mu);
  if (g_servers > 0) {
    g_deferred.emplace_back(p, host);
    return hipSuccess;
  }
  return host ? hipHostFree(p) : hipFree(p);
}

void unref(tc_ctx* c) {
  if (!c || c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamDestroy(c->stream);
  for (void* p : c->stage) hfree(p);
  delete c;
}

void* ctx_stage(tc_ctx* c, int b) {
  if (!c->stage[b]) TC_HIP(hipHostMalloc(&c->stage[b], tc_ctx::kStageBytes, hipHostMallocDefault));
  return c->stage[b];
}

void unref(tc_dataset* ds) {
  if (!ds || ds->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  if (ds->ctx) (void)hipStreamSynchronize(ds->ctx->stream);
  dfree(ds->bins_rm);
  dfree(ds->bins_cm);
  dfree(ds->edges);
  dfree(ds->n_edges);
  unref(ds->ctx);
  delete ds;
}

void serving_begin() {
  std::lock_guard<std::mutex> lk(g_free_mu);
  ++g_servers;
}

void serving_end() {
  std::lock_guard<std::mutex> lk(g_free_mu);
  if (--g_servers > 0) return;
  for (const auto& d : g_deferred) (void)(d.second ? hipHostFree(d.first) : hipFree(d.first));
  g_deferred.clear();
}

}  // namespace tc

using namespace tc;

extern "C" {

const char* tc_last_error(void) { return g_last_error.c_str(); }

int32_t tc_ctx_create(int32_t device_id, tc_ctx** out) {
  return guard([&] {
    TC_REQUIRE(out, "tc_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
      throw Error(TC_EDEVICE, "no HIP device available");
    TC_REQUIRE(device_id >= 0 && device_id < n, "tc_ctx_create: device id out of range");
    hipDeviceProp_t prop;
    TC_HIP(hipGetDeviceProperties(&prop, device_id));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      throw Error(TC_EDEVICE, std::string("device is ") + prop.gcnArchName +
                                  "; this library is built for gfx950 (MI355X)");
    TC_HIP(hipSetDevice(device_id));
    scratch_setup(device_id);  // the allocation-free entry point's scratch, once per device
    auto* c = new tc_ctx();
    c->device = device_id;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      throw Error(TC_EDEVICE, "hipStreamCreate failed");
    }
    *out = c;
  });
}

int32_t tc_ctx_destroy(tc_ctx* ctx) {
  return guard([&] { unref(ctx); });
}

int32_t tc_ctx_synchronize(tc_ctx* ctx) {
  return guard([&] {
    TC_REQUIRE(ctx, "NULL ctx");
    TC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int32_t tc_device_alloc(tc_ctx* ctx, size_t bytes, void** out) {
  return guard([&] {
    TC_REQUIRE(ctx && out, "tc_device_alloc: NULL argument");
    TC_HIP(hipSetDevice(ctx->device));
    *out = dmalloc<uint8_t>(bytes);
  });
}

int32_t tc_device_free(tc_ctx* ctx, void* p) {
  return guard([&] {
    TC_REQUIRE(ctx, "NULL ctx");
    TC_HIP(free_or_defer(p, false));
  });
}

int32_t tc_memcpy_h2d(tc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    TC_REQUIRE(ctx && dst && src, "tc_memcpy_h2d: NULL argument");
    TC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int32_t tc_memcpy_d2h(tc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    TC_REQUIRE(ctx && dst && src, "tc_memcpy_d2h: NULL argument");
    TC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

static tc_dataset* make_dataset(tc_ctx* ctx, const float* X, uint32_t n, uint32_t F,
                                uint32_t max_bins, std::vector<float>&& edges,
                                std::vector<uint32_t>&& n_edges) {
  auto* ds = new tc_dataset();
  ds->ctx = ref(ctx);
  ds->N = n;
  ds->F = F;
  ds->max_bins = max_bins;
  ds->h_edges = std::move(edges);
  ds->h_n_edges = std::move(n_edges);
  try {
    TC_HIP(hipSetDevice(ctx->device));
    bin_on_device(ds, X);
  } catch (...) {
    unref(ds);
    throw;
  }
  return ds;
}

int32_t tc_quantize_fit(tc_ctx* ctx, const float* X, uint32_t n, uint32_t F, uint32_t max_bins,
                        int32_t wrap, float* edges_out, uint32_t* n_edges_out, tc_dataset** out) {
  return guard([&] {
    TC_REQUIRE(ctx && X && out, "tc_quantize_fit: NULL argument");
    TC_REQUIRE(F >= 1, "tc_quantize_fit: no features");
    TC_REQUIRE(max_bins >= 2 && max_bins <= 255, "max_bins must be in [2, 255]");
    *out = nullptr;
    auto* ds = new tc_dataset();
    ds->ctx = ref(ctx);
    ds->N = n;
    ds->F = F;
    ds->max_bins = max_bins;
    try {
      TC_HIP(hipSetDevice(ctx->device));
      quantize_on_device(ds, X, wrap != 0);
    } catch (...) {
      unref(ds);
      throw;
    }
    if (edges_out) std::memcpy(edges_out, ds->h_edges.data(), sizeof(float) * ds->h_edges.size());
    if (n_edges_out)
      std::memcpy(n_edges_out, ds->h_n_edges.data(), sizeof(uint32_t) * ds->h_n_edges.size());
    *out = ds;
  });
}

int32_t tc_dataset_create(tc_ctx* ctx, const float* X, uint32_t n, uint32_t F,
                          const float* edges, const uint32_t* n_edges, uint32_t max_bins,
                          tc_dataset** out) {
  return guard([&] {
    TC_REQUIRE(ctx && X && edges && n_edges && out, "tc_dataset_create: NULL argument");
    TC_REQUIRE(F >= 1, "tc_dataset_create: no features");
    TC_REQUIRE(max_bins >= 2 && max_bins <= 255, "max_bins must be in [2, 255]");
    *out = nullptr;
    std::vector<float> e(edges, edges + static_cast<size_t>(F) * 255);
    std::vector<uint32_t> ne(n_edges, n_edges + F);
    for (uint32_t f = 0; f < F; ++f) TC_REQUIRE(ne[f] <= 255, "n_edges must be <= 255");
    *out = make_dataset(ctx, X, n, F, max_bins, std::move(e), std::move(ne));
  });
}

int32_t tc_dataset_bins(tc_dataset* ds, uint8_t* bins_out) {
  return guard([&] {
    TC_REQUIRE(ds && bins_out, "tc_dataset_bins: NULL argument");
    copy_sync(bins_out, ds->bins_cm, static_cast<size_t>(ds->N) * ds->F, hipMemcpyDeviceToHost,
              ds->ctx->stream);
  });
}

int32_t tc_dataset_shape(tc_dataset* ds, uint32_t* n, uint32_t* F) {
  return guard([&] {
    TC_REQUIRE(ds, "NULL dataset");
    if (n) *n = ds->N;
    if (F) *F = ds->F;
  });
}

int32_t tc_dataset_destroy(tc_dataset* ds) {
  return guard([&] { unref(ds); });
}

}  // extern "C"
