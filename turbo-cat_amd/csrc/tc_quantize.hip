// A2/A3: feature quantization.  Edges are the reference's quantile rule
// (src/core/dataset.cpp:278-334): per feature the non-NaN values are sorted
// and de-duplicated (U unique), A = min(max_bins, uint8(U)) when reproducing
// the reference's uint8 wrap (:308), else min(max_bins, U); edges are the
// unique values at b*U/A for b = 1..A-1.  Bin = number of edges <= value
// (upper_bound, :336-344); NaN -> max_bins.  Edge selection runs on the host
// (one sort per feature, parallel over features); binning every value runs
// on the GPU and writes both layouts the training kernels read.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "tc_device.h"
#include "tc_internal.h"

namespace tc {

void compute_edges(const float* X, uint32_t n, uint32_t F, uint32_t max_bins, bool wrap,
                   std::vector<float>& edges, std::vector<uint32_t>& n_edges) {
  edges.assign(static_cast<size_t>(F) * 255, 0.f);
  n_edges.assign(F, 0);
  const uint8_t nb = static_cast<uint8_t>(max_bins);
  auto one = [&](uint32_t f, std::vector<float>& vals) {
    vals.clear();
    for (uint32_t i = 0; i < n; ++i) {
      const float v = X[static_cast<size_t>(i) * F + f];
      if (!std::isnan(v)) vals.push_back(v);
    }
    float* e = edges.data() + static_cast<size_t>(f) * 255;
    if (vals.empty()) {  // all NaN: one 0.0 edge, rows in the NaN bin (:288-295)
      e[0] = 0.f;
      n_edges[f] = 1;
      return;
    }
    std::sort(vals.begin(), vals.end());
    vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
    const size_t U = vals.size();
    const uint32_t ucast = wrap ? static_cast<uint32_t>(static_cast<uint8_t>(U))
                                : static_cast<uint32_t>(std::min<size_t>(U, 255));
    const uint32_t A = std::min<uint32_t>(nb, ucast);
    uint32_t ne = 0;
    for (uint32_t b = 1; b < A; ++b) {
      const size_t idx = static_cast<size_t>(b) * U / A;
      if (ne == 0 || vals[idx] > e[ne - 1]) e[ne++] = vals[idx];
    }
    n_edges[f] = ne;
  };
  const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t nt = std::min(hw, std::max(1u, F));
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < nt; ++t) {
    pool.emplace_back([&, t]() {
      std::vector<float> vals;
      vals.reserve(n);
      for (uint32_t f = t; f < F; f += nt) one(f, vals);
    });
  }
  for (auto& th : pool) th.join();
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
