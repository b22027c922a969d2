// A2/A3: feature quantization, on the device.  Edges follow the reference's
// quantile rule (src/core/dataset.cpp:278-334): per feature the non-NaN
// values are sorted and de-duplicated (U unique), A = min(max_bins,
// uint8(U)) when reproducing the reference's uint8 wrap (:308), else
// min(max_bins, U); the edges are the unique values at b*U/A for
// b = 1..A-1; a feature with no value gets one 0.0 edge (:288-295).  Bin =
// number of edges <= value (upper_bound, :336-344); NaN -> max_bins.
//
// X crosses PCIe once, in row blocks through pinned staging; each block is
// transposed into column-major order-preserving 32-bit keys (k_keys).  Groups
// of features are sorted at once as 64-bit keys (feature in the high word,
// rocPRIM's radix sort), k_edges selects each feature's edges from its sorted
// run, and k_bin_keys bins every value from the unsorted keys into both bin
// layouts the training kernels read.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "tc_copy_pool.h"
#include "tc_device.h"
#include "tc_internal.h"

namespace tc {

namespace {

// Order-preserving key of a float: two non-NaN values compare as their keys,
// with -0 taken as +0 (they compare equal in the reference's sort, unique and
// upper_bound); NaN is kNanKey, above every value.
constexpr uint32_t kNanKey = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t float_key(float v) {
  if (isnan(v)) return kNanKey;
  uint32_t u = __float_as_uint(v);
  if (u == 0x80000000u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Row block [nr][F] -> keys [f * ldk + row], through a 64 x 64 LDS tile.
__global__ void __launch_bounds__(256) k_keys(const float* __restrict__ X, uint32_t nr, uint32_t F,
                                              uint32_t* __restrict__ keys, size_t ldk) {
  __shared__ uint32_t t[64][65];
  const uint32_t r0 = blockIdx.x * 64u, f0 = blockIdx.y * 64u;
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  for (uint32_t i = ty; i < 64u; i += 4u) {
    const uint32_t r = r0 + i, f = f0 + tx;
    t[i][tx] = (r < nr && f < F) ? float_key(X[static_cast<size_t>(r) * F + f]) : 0u;
  }
  __syncthreads();
  for (uint32_t i = ty; i < 64u; i += 4u) {
    const uint32_t f = f0 + i, r = r0 + tx;
    if (f < F && r < nr) keys[static_cast<size_t>(f) * ldk + r] = t[tx][i];
  }
}

// keys of features [0, g) (f * N + i) -> (f << 32) | key, one run per feature.
__global__ void __launch_bounds__(256) k_compose(const uint32_t* __restrict__ keys, size_t total,
                                                 uint32_t N, unsigned long long* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * 256u + threadIdx.x;
  if (i >= total) return;
  out[i] = (static_cast<unsigned long long>(i / N) << 32) | keys[i];
}

constexpr uint32_t kEdgeThreads = 1024, kEdgeWaves = kEdgeThreads / 64;

// Edges of feature f0 + blockIdx.x from its sorted run (N keys, NaN last).
// Wave w owns the contiguous range [w R, (w + 1) R) of the run.  Pass 1
// counts each range's unique values (heads: a non-NaN key other than its
// predecessor) and NaNs; their prefix gives every wave the rank of its first
// head, and U.  The selected ranks floor(b U / A), b = 1..A-1, are distinct
// (A <= U) and increasing; pass 2 walks each range 64 keys at a time, ranks
// the heads with a ballot and stores the value of each selected rank as edge
// b - 1.  So a feature has A - 1 edges (none for A <= 1).
__global__ void __launch_bounds__(kEdgeThreads)
    k_edges(const unsigned long long* __restrict__ sorted, uint32_t N, uint32_t f0, uint32_t mb,
            int wrap, float* __restrict__ edges, uint32_t* __restrict__ n_edges) {
  __shared__ uint32_t s_heads[kEdgeWaves], s_nans[kEdgeWaves];
  __shared__ uint32_t s_sel[256];  // selected ranks
  const uint32_t g = blockIdx.x, f = f0 + g, tid = threadIdx.x;
  const uint32_t lane = tid & 63u, w = tid >> 6;
  const unsigned long long* seg = sorted + static_cast<size_t>(g) * N;
  auto key = [&](uint32_t i) { return static_cast<uint32_t>(seg[i]); };
  const uint32_t R = (N + kEdgeWaves - 1) / kEdgeWaves;
  const uint32_t lo = min(N, w * R), hi = min(N, lo + R);
  // eight steps of 64 keys (and their predecessors) in flight per wave
  constexpr uint32_t kSteps = 8;
  uint32_t heads = 0, nans = 0;
  for (uint32_t b0 = lo; b0 < hi; b0 += 64 * kSteps) {
    uint32_t k[kSteps], kp[kSteps];
#pragma unroll
    for (uint32_t q = 0; q < kSteps; ++q) {
      const uint32_t i = b0 + 64 * q + lane;
      k[q] = i < hi ? key(i) : kNanKey;
      kp[q] = (i < hi && i > 0) ? key(i - 1) : kNanKey;
    }
#pragma unroll
    for (uint32_t q = 0; q < kSteps; ++q) {
      const uint32_t i = b0 + 64 * q + lane;
      if (i < hi) {
        if (k[q] == kNanKey) ++nans;
        else if (i == 0 || kp[q] != k[q]) ++heads;
      }
    }
  }
  heads = wave_sum_u32(heads);
  nans = wave_sum_u32(nans);
  if (lane == 0) {
    s_heads[w] = heads;
    s_nans[w] = nans;
  }
  __syncthreads();
  uint32_t U = 0, nn = 0, first = 0;
  for (uint32_t v = 0; v < kEdgeWaves; ++v) {
    if (v == w) first = U;
    U += s_heads[v];
    nn += s_nans[v];
  }
  float* e = edges + static_cast<size_t>(f) * 255u;
  if (nn == N) {  // no value: one 0.0 edge, every row in the NaN bin
    if (tid == 0) {
      e[0] = 0.f;
      n_edges[f] = 1;
    }
    return;
  }
  const uint32_t ucast = wrap ? (U & 0xFFu) : min(U, 255u);
  const uint32_t A = min(mb, ucast);
  if (tid == 0) n_edges[f] = A > 0 ? A - 1 : 0;
  if (A <= 1) return;
  if (tid + 1 < A)
    s_sel[tid] = static_cast<uint32_t>(static_cast<unsigned long long>(tid + 1) * U / A);
  __syncthreads();
  // the first selected rank at or above the wave's first head
  uint32_t bc = 0;
  while (bc + 1 < A && s_sel[bc] < first) ++bc;
  uint32_t r0 = first;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t b0 = lo; b0 < hi && bc + 1 < A; b0 += 64 * kSteps) {
    uint32_t k[kSteps], kp[kSteps];
#pragma unroll
    for (uint32_t q = 0; q < kSteps; ++q) {
      const uint32_t i = b0 + 64 * q + lane;
      k[q] = i < hi ? key(i) : kNanKey;
      kp[q] = (i < hi && i > 0) ? key(i - 1) : kNanKey;
    }
#pragma unroll
    for (uint32_t q = 0; q < kSteps; ++q) {
      const uint32_t i = b0 + 64 * q + lane;
      const bool h = i < hi && k[q] != kNanKey && (i == 0 || kp[q] != k[q]);
      const unsigned long long m = __ballot(h);
      const uint32_t r = r0 + static_cast<uint32_t>(__popcll(m & below));
      const uint32_t r1 = r0 + static_cast<uint32_t>(__popcll(m));
      while (bc + 1 < A && s_sel[bc] < r1) {  // selected ranks in this step (uniform)
        if (h && r == s_sel[bc]) e[bc] = key_float(k[q]);
        ++bc;
      }
      r0 = r1;
    }
  }
}

// One thread per row, 16 features per workgroup column: the features' edges
// as keys in LDS, upper_bound of each key, one 16-byte store into the
// row-major bins and 16 coalesced byte stores into the column-major bins.
__global__ void __launch_bounds__(256) k_bin_keys(const uint32_t* __restrict__ keys, uint32_t N,
                                                  uint32_t F, uint32_t Fp,
                                                  const float* __restrict__ edges,
                                                  const uint32_t* __restrict__ n_edges,
                                                  uint8_t nan_bin, uint8_t* __restrict__ bins_rm,
                                                  uint8_t* __restrict__ bins_cm) {
  __shared__ uint32_t s_e[kTileF][256];
  __shared__ uint32_t s_n[kTileF];
  const uint32_t f0 = blockIdx.y * kTileF;
  for (uint32_t i = threadIdx.x; i < kTileF * 255; i += blockDim.x) {
    const uint32_t k = i / 255, j = i % 255;
    s_e[k][j] = (f0 + k < F) ? float_key(edges[static_cast<size_t>(f0 + k) * 255 + j]) : 0u;
  }
  if (threadIdx.x < kTileF) s_n[threadIdx.x] = (f0 + threadIdx.x < F) ? n_edges[f0 + threadIdx.x] : 0;
  __syncthreads();
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= N) return;
  uint32_t packed[4] = {0, 0, 0, 0};
  for (uint32_t k = 0; k < kTileF; ++k) {
    const uint32_t f = f0 + k;
    if (f >= F) break;
    const uint32_t v = keys[static_cast<size_t>(f) * N + row];
    uint32_t b;
    if (v == kNanKey) {
      b = nan_bin;
    } else {
      uint32_t lo = 0, hi = s_n[k];
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v < s_e[k][mid]) hi = mid; else lo = mid + 1;
      }
      b = lo;
    }
    packed[k >> 2] |= b << (8 * (k & 3));
    bins_cm[static_cast<size_t>(f) * N + row] = static_cast<uint8_t>(b);
  }
  *reinterpret_cast<uint4*>(bins_rm + static_cast<size_t>(row) * Fp + f0) =
      make_uint4(packed[0], packed[1], packed[2], packed[3]);
}

// Features sorted together: about 2^26 keys (512 MB of 64-bit keys and as
// much again for the sort's output).
constexpr size_t kSortKeys = size_t(1) << 26;

uint32_t bits_for(uint32_t v) {  // bits to hold 0..v
  uint32_t b = 0;
  while (b < 32 && (v >> b) != 0) ++b;
  return b;
}

}  // namespace

void quantize_on_device(tc_dataset* ds, const float* X, bool wrap) {
  const uint32_t N = ds->N, F = ds->F;
  hipStream_t s = ds->ctx->stream;
  std::lock_guard<std::mutex> stage_lock(ds->ctx->stage_mu);  // one stream of rows at a time
  ds->Fp = rm_stride(F);
  ds->bins_rm = dmalloc<uint8_t>(static_cast<size_t>(N) * ds->Fp);
  ds->bins_cm = dmalloc<uint8_t>(static_cast<size_t>(N) * F);
  ds->edges = dmalloc<float>(static_cast<size_t>(F) * 255);
  ds->n_edges = dmalloc<uint32_t>(F);
  ds->h_edges.assign(static_cast<size_t>(F) * 255, 0.f);
  ds->h_n_edges.assign(F, 0);
  if (N == 0) {  // no rows: every feature is empty (one 0.0 edge)
    std::fill(ds->h_n_edges.begin(), ds->h_n_edges.end(), 1u);
    copy_sync(ds->edges, ds->h_edges.data(), 4ull * F * 255, hipMemcpyHostToDevice, s);
    copy_sync(ds->n_edges, ds->h_n_edges.data(), 4ull * F, hipMemcpyHostToDevice, s);
    return;
  }
  uint32_t* keys = nullptr;
  unsigned long long *kin = nullptr, *kout = nullptr;
  void* temp = nullptr;
  float* dev[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto release = [&]() {
    (void)hipStreamSynchronize(s);
    for (int b = 0; b < 2; ++b) {
      if (ev[b]) (void)hipEventDestroy(ev[b]);
      dfree(dev[b]);
    }
    dfree(keys);
    dfree(kin);
    dfree(kout);
    dfree(temp);
  };
  try {
    keys = dmalloc<uint32_t>(static_cast<size_t>(N) * F);
    // stream X in row blocks: the host copy of block k + 1 into pinned
    // staging overlaps the transfer and transposition of block k
    const size_t row_bytes = 4ull * F;
    const uint32_t blk = static_cast<uint32_t>(
        std::max<size_t>(1, std::min<size_t>(N, tc_ctx::kStageBytes / row_bytes)));
    TC_REQUIRE(row_bytes <= tc_ctx::kStageBytes, "quantize: a row exceeds the staging buffer");
    float* pin[2];
    for (int b = 0; b < 2; ++b) {
      pin[b] = reinterpret_cast<float*>(ctx_stage(ds->ctx, b));
      dev[b] = dmalloc<float>(static_cast<size_t>(blk) * F);
      TC_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    }
    uint32_t k = 0;
    for (uint32_t r0 = 0; r0 < N; r0 += blk, ++k) {
      const uint32_t b = k & 1u, nr = std::min(blk, N - r0);
      if (k >= 2) TC_HIP(hipEventSynchronize(ev[b]));  // block k - 2 is done with the buffers
      copy_pool().copy(split_copy(pin[b], X + static_cast<size_t>(r0) * F, row_bytes * nr));
      TC_HIP(hipMemcpyAsync(dev[b], pin[b], row_bytes * nr, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_keys, dim3((nr + 63) / 64, (F + 63) / 64), dim3(256), 0, s, dev[b], nr, F,
                         keys + r0, static_cast<size_t>(N));
      TC_HIP(hipGetLastError());
      TC_HIP(hipEventRecord(ev[b], s));
    }
    // edges, a group of features at a time
    const uint32_t G = static_cast<uint32_t>(
        std::max<size_t>(1, std::min<size_t>(F, kSortKeys / std::max<uint32_t>(N, 1))));
    const size_t cap = static_cast<size_t>(G) * N;
    kin = dmalloc<unsigned long long>(cap);
    kout = dmalloc<unsigned long long>(cap);
    size_t temp_bytes = 0;
    TC_HIP(rocprim::radix_sort_keys(nullptr, temp_bytes, kin, kout, cap, 0u, 32u + bits_for(G - 1),
                                    s));
    temp = dmalloc<uint8_t>(std::max<size_t>(1, temp_bytes));
    TC_HIP(hipMemsetAsync(ds->edges, 0, 4ull * F * 255, s));
    for (uint32_t f0 = 0; f0 < F; f0 += G) {
      const uint32_t g = std::min(G, F - f0);
      const size_t total = static_cast<size_t>(g) * N;
      hipLaunchKernelGGL(k_compose, dim3(static_cast<uint32_t>((total + 255) / 256)), dim3(256), 0, s,
                         keys + static_cast<size_t>(f0) * N, total, N, kin);
      TC_HIP(hipGetLastError());
      size_t need = 0;
      const uint32_t end_bit = 32u + bits_for(g - 1);
      TC_HIP(rocprim::radix_sort_keys(nullptr, need, kin, kout, total, 0u, end_bit, s));
      TC_REQUIRE(need <= temp_bytes, "quantize: sort scratch smaller than the sort needs");
      TC_HIP(rocprim::radix_sort_keys(temp, need, kin, kout, total, 0u, end_bit, s));
      hipLaunchKernelGGL(k_edges, dim3(g), dim3(kEdgeThreads), 0, s, kout, N, f0, ds->max_bins & 0xFFu,
                         wrap ? 1 : 0, ds->edges, ds->n_edges);
      TC_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_bin_keys, dim3((N + 255) / 256, feature_tiles(F)), dim3(256), 0, s, keys, N,
                       F, ds->Fp, ds->edges, ds->n_edges, static_cast<uint8_t>(ds->max_bins),
                       ds->bins_rm, ds->bins_cm);
    TC_HIP(hipGetLastError());
    TC_HIP(hipMemcpyAsync(ds->h_edges.data(), ds->edges, 4ull * F * 255, hipMemcpyDeviceToHost, s));
    TC_HIP(hipMemcpyAsync(ds->h_n_edges.data(), ds->n_edges, 4ull * F, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
  } catch (...) {
    release();
    throw;
  }
  release();
}

// One thread per row, 16 features per workgroup column (blockIdx.y): edges of
// the 16 features in LDS, binary search per value, one 16-byte store into the
// row-major bins and 16 coalesced byte stores into the column-major bins.
__global__ void __launch_bounds__(256) k_bin(const float* __restrict__ X, uint32_t N, uint32_t F,
                                             uint32_t Fp, const float* __restrict__ edges,
                                             const uint32_t* __restrict__ n_edges, uint8_t nan_bin,
                                             uint8_t* __restrict__ bins_rm,
                                             uint8_t* __restrict__ bins_cm, uint32_t ldc) {
  __shared__ float s_e[kTileF][256];
  __shared__ uint32_t s_n[kTileF];
  const uint32_t f0 = blockIdx.y * kTileF;
  for (uint32_t i = threadIdx.x; i < kTileF * 255; i += blockDim.x) {
    const uint32_t k = i / 255, j = i % 255;
    s_e[k][j] = (f0 + k < F) ? edges[static_cast<size_t>(f0 + k) * 255 + j] : 0.f;
  }
  if (threadIdx.x < kTileF) s_n[threadIdx.x] = (f0 + threadIdx.x < F) ? n_edges[f0 + threadIdx.x] : 0;
  __syncthreads();
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= N) return;
  uint32_t packed[4] = {0, 0, 0, 0};
  const float* xr = X + static_cast<size_t>(row) * F;
  for (uint32_t k = 0; k < kTileF; ++k) {
    const uint32_t f = f0 + k;
    if (f >= F) break;
    const float v = xr[f];
    uint32_t b;
    if (isnan(v)) {
      b = nan_bin;
    } else {
      uint32_t lo = 0, hi = s_n[k];
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v < s_e[k][mid]) hi = mid; else lo = mid + 1;
      }
      b = lo;
    }
    packed[k >> 2] |= b << (8 * (k & 3));
    bins_cm[static_cast<size_t>(f) * ldc + row] = static_cast<uint8_t>(b);
  }
  *reinterpret_cast<uint4*>(bins_rm + static_cast<size_t>(row) * Fp + f0) =
      make_uint4(packed[0], packed[1], packed[2], packed[3]);
}

void bin_on_device(tc_dataset* ds, const float* X) {
  const uint32_t N = ds->N, F = ds->F;
  hipStream_t s = ds->ctx->stream;
  ds->Fp = rm_stride(F);
  ds->bins_rm = dmalloc<uint8_t>(static_cast<size_t>(N) * ds->Fp);
  ds->bins_cm = dmalloc<uint8_t>(static_cast<size_t>(N) * F);
  ds->edges = dmalloc<float>(static_cast<size_t>(F) * 255);
  ds->n_edges = dmalloc<uint32_t>(F);
  TC_HIP(hipMemcpyAsync(ds->edges, ds->h_edges.data(), sizeof(float) * F * 255,
                        hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(ds->n_edges, ds->h_n_edges.data(), sizeof(uint32_t) * F,
                        hipMemcpyHostToDevice, s));
  // stream X through the device in row blocks so X never needs to be resident
  const uint32_t block_rows = std::max<uint32_t>(1024, static_cast<uint32_t>(
      std::min<size_t>(N, (size_t(256) << 20) / (sizeof(float) * std::max(1u, F)))));
  float* xd = dmalloc<float>(static_cast<size_t>(std::min(block_rows, std::max(N, 1u))) * F);
  try {
    for (uint32_t r0 = 0; r0 < N; r0 += block_rows) {
      const uint32_t nr = std::min(block_rows, N - r0);
      TC_HIP(hipMemcpyAsync(xd, X + static_cast<size_t>(r0) * F, sizeof(float) * nr * F,
                            hipMemcpyHostToDevice, s));
      // rows r0.. of this block: offset the output pointers, keep full N stride
      dim3 grid((nr + 255) / 256, feature_tiles(F));
      hipLaunchKernelGGL(k_bin, grid, dim3(256), 0, s, xd, nr, F, ds->Fp, ds->edges, ds->n_edges,
                         static_cast<uint8_t>(ds->max_bins), ds->bins_rm + static_cast<size_t>(r0) * ds->Fp,
                         ds->bins_cm + r0, N);
      TC_HIP(hipGetLastError());
      TC_HIP(hipStreamSynchronize(s));
    }
  } catch (...) {
    dfree(xd);
    throw;
  }
  dfree(xd);
}

}  // namespace tc
