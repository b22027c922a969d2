// Symmetric-tree training kernels for gfx950.  One launch of each kernel
// serves every job of a batch (a single-job fit is a batch of one): per-job
// kernels put the job on blockIdx.y, row-level kernels use a flattened
// workgroup space, and histogram / partition kernels consume work lists.
//
// Per tree the launch sequence is fixed (so it can be captured in a graph);
// data-dependent control (level stops, tree count) lives in TreeState and
// makes later kernels of a stopped job return at once.
//
//   reset_counters, root_partials, tree_begin
//   for level d < D:
//     [d>0] zero_smaller(d), hist(d, smaller children), subtract(d)
//     [d=0] hist(0, root)
//     split_scan(d), finalize(d)
//     part_count(d), part_plan(d), part_scatter(d)
//   leaves, grad_update(apply)
#include "tc_device.h"

namespace tc {

struct Lists {
  HistWork* hist[kMaxDepth];
  uint32_t* hist_count[kMaxDepth];
  PartWork* part[kMaxDepth];
  uint32_t* part_count[kMaxDepth];
};

// ---------------------------------------------------------------------------
// Work-list helpers (called by one workgroup per job)

// Append the histogram work for rows [start, start+len) of `node` (all
// feature tiles of each row chunk, contiguous so the tiles share an XCD).
__device__ inline void emit_hist(const JobDev& J, uint32_t job, HistWork* list, uint32_t* count,
                                 uint32_t node, uint32_t start, uint32_t len) {
  if (len == 0) return;
  if (J.hist_rows_total) atomicAdd(J.hist_rows_total, static_cast<unsigned long long>(len));
  const uint32_t chunks = (len + J.hist_chunk - 1) / J.hist_chunk;
  const uint32_t base = atomicAdd(count, chunks * J.T);
  for (uint32_t c = 0; c < chunks; ++c) {
    const uint32_t s = start + c * J.hist_chunk;
    const uint32_t l = min(J.hist_chunk, start + len - s);
    for (uint32_t t = 0; t < J.T; ++t) {
      HistWork w;
      w.job = job; w.node = node; w.start = s; w.len = l; w.tile = t;
      w.pad0 = w.pad1 = w.pad2 = 0;
      list[base + c * J.T + t] = w;
    }
  }
}

__device__ inline uint32_t part_slots(uint32_t len) {
  return (len + kPartChunk - 1) / kPartChunk;
}

// ---------------------------------------------------------------------------
// Gradients (A6/A7) fused with the training-prediction update (A15 + K9).
//
// mode 0: preds = base for every row (before the first tree).
// mode 1: apply tree st->tree to every row: leaf index from the binned
//         compares go_right = bin != 255 && bin > threshold, MSB first
//         (SymmetricTree::predict_batch, symmetric_tree.cpp:813-840), then
//         preds += lr * leaf as one fused multiply-add (the reference
//         compiler's contraction of booster.cpp:460-463), then g/h.
// Rows in natural order, 1024 per workgroup, every row of the dataset (also
// when the tree was grown on a row sample).
constexpr int kGradTile = 1024;

__device__ inline uint32_t job_of_block(const uint32_t* job_wg_off, uint32_t njobs) {
  uint32_t lo = 0, hi = njobs;  // last job with job_wg_off[j] <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (job_wg_off[mid] <= blockIdx.x) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_grad_update(const JobDev* __restrict__ jobs,
                                                     const uint32_t* __restrict__ job_wg_off,
                                                     uint32_t njobs, int mode) {
  __shared__ uint32_t s_job;
  __shared__ uint16_t s_f[kMaxDepth];
  __shared__ uint8_t s_b[kMaxDepth];
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_job = job_of_block(job_wg_off, njobs);
  __syncthreads();
  const JobDev& J = jobs[s_job];
  const uint32_t lw = blockIdx.x - J.grad_wg_off;
  const uint32_t N = J.N;
  const uint32_t vecN = (N / 8u) * 8u;
  int t = -1, depth = 0;
  if (mode == 1) {
    t = J.st->tree;
    if (t < 0 || t >= static_cast<int>(J.n_trees)) return;
    depth = J.m_depth[t];
    if (tid < static_cast<uint32_t>(depth)) {
      s_f[tid] = J.m_feat[static_cast<size_t>(t) * J.D + tid];
      s_b[tid] = J.m_bin[static_cast<size_t>(t) * J.D + tid];
    }
    __syncthreads();
  }
  const float* leaves = J.m_leaves + static_cast<size_t>(t < 0 ? 0 : t) * (1u << J.D);
  for (int k = 0; k < kGradTile / 256; ++k) {
    const uint32_t row = lw * kGradTile + k * 256 + tid;
    if (row >= N) break;
    float pred;
    if (mode == 1) {
      uint32_t li = 0;
      for (int d = 0; d < depth; ++d) {
        const uint8_t v = J.bins_cm[static_cast<size_t>(s_f[d]) * N + row];
        li = (li << 1) | ((v != 255 && v > s_b[d]) ? 1u : 0u);
      }
      pred = __fmaf_rn(J.lr, leaves[li], J.preds[row]);
    } else {
      pred = J.base;
    }
    J.preds[row] = pred;
    J.gh[row] = row_gradient(J.loss, pred, J.y[row], row < vecN);
  }
}

// Root totals: (sum g, sum h) over 1024-position tiles of the root list
// rows[0][0 .. n_root), double, fixed order within the tile; k_tree_begin
// adds the tiles in order.
__global__ void __launch_bounds__(256) k_root_partials(const JobDev* __restrict__ jobs,
                                                       const uint32_t* __restrict__ job_wg_off,
                                                       uint32_t njobs) {
  __shared__ uint32_t s_job;
  __shared__ double s_red[2][4];
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_job = job_of_block(job_wg_off, njobs);
  __syncthreads();
  const JobDev& J = jobs[s_job];
  const uint32_t lw = blockIdx.x - J.grad_wg_off;
  double sg = 0.0, sh = 0.0;
  if (J.st->tree + 1 < static_cast<int>(J.n_trees)) {
    for (int k = 0; k < kGradTile / 256; ++k) {
      const uint32_t p = lw * kGradTile + k * 256 + tid;
      if (p >= J.n_root) break;
      const float2 v = J.gh[J.rows[p]];
      sg += v.x;
      sh += v.y;
    }
  }
  sg = warp_sum_d(sg);
  sh = warp_sum_d(sh);
  if ((tid & 63) == 0) { s_red[0][tid >> 6] = sg; s_red[1][tid >> 6] = sh; }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < 4; ++w) { a += s_red[0][w]; b += s_red[1][w]; }
    J.grad_partials[2 * lw + 0] = a;
    J.grad_partials[2 * lw + 1] = b;
  }
}

// ---------------------------------------------------------------------------
__global__ void k_reset_counters(Lists L) {
  const uint32_t i = threadIdx.x;
  if (i < kMaxDepth) {
    *L.hist_count[i] = 0;
    *L.part_count[i] = 0;
  }
}

// Root of a new tree (symmetric_tree.cpp:97-124): totals from the gradient
// partials (in workgroup order, double), level-0 stop test, root work.
__global__ void __launch_bounds__(256) k_tree_begin(const JobDev* __restrict__ jobs, Lists L) {
  const uint32_t job = blockIdx.y;
  const JobDev& J = jobs[job];
  const uint32_t tid = threadIdx.x;
  __shared__ int s_go;
  if (tid == 0) {
    TreeState* st = J.st;
    const int t = st->tree + 1;
    st->tree = t;
    st->depth = 0;
    for (int d = 0; d < kMaxDepth; ++d) st->best[d] = kEmptyKey;
    int go = 0;
    if (t < static_cast<int>(J.n_trees)) {
      double G = 0.0, H = 0.0;
      for (uint32_t w = 0; w < J.grad_wg_cnt; ++w) {
        G += J.grad_partials[2 * w];
        H += J.grad_partials[2 * w + 1];
      }
      J.ng[0] = static_cast<float>(G);
      J.nh[0] = static_cast<float>(H);
      J.nc[0] = J.n_root;
      J.off[0] = 0;
      J.off[1] = J.n_root;
      go = (J.n_root >= 2u * J.min_samples_leaf && J.nh[0] >= J.min_child_weight * 2.0f) ? 1 : 0;
      if (J.n_root == 0) go = 0;  // empty sample: depth-0 tree, leaf 0 (:39-43)
    }
    st->active = go;
    s_go = go;
  }
  __syncthreads();
  if (!s_go) return;
  // zero the root histogram
  float4* h4 = reinterpret_cast<float4*>(J.hist);
  const uint32_t n4 = J.F * 256u / 2u;
  for (uint32_t i = tid; i < n4; i += blockDim.x) h4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (tid == 0) {
    emit_hist(J, job, L.hist[0], L.hist_count[0], 0, 0, J.n_root);
    const uint32_t ns = part_slots(J.n_root);
    const uint32_t base = atomicAdd(L.part_count[0], ns);
    for (uint32_t s = 0; s < ns; ++s) {
      PartWork w;
      w.job = job; w.node = 0; w.start = s * kPartChunk;
      w.len = min(static_cast<uint32_t>(kPartChunk), J.n_root - w.start);
      w.slot = s; w.pad0 = w.pad1 = w.pad2 = 0;
      L.part[0][base + s] = w;
    }
  }
}

// ---------------------------------------------------------------------------
// A9: per-level histogram.  One workgroup = one work item (row chunk of one
// node) x one 16-feature tile.  Each lane loads the tile's 16 bins of a row
// with one 16-byte load from the row-major bins, its (g, h) pair, and adds
// them into a workgroup histogram in LDS (g plane and h plane, 257-float
// feature stride).  Lane l starts at feature l % 16 so that the 64 lanes of
// one LDS atomic instruction spread over 16 feature histograms.  At the end
// the non-empty bins are added to the node's histogram in HBM.
template <int U>
__global__ void __launch_bounds__(kHistThreads) k_hist(const JobDev* __restrict__ jobs,
                                                       const HistWork* __restrict__ work,
                                                       const uint32_t* __restrict__ count,
                                                       uint32_t level) {
  __shared__ float s_g[kTileF * 257];
  __shared__ float s_h[kTileF * 257];
  uint32_t li;
  if (!xcd_rank(blockIdx.x, *count, &li)) return;
  const HistWork w = work[li];
  const JobDev& J = jobs[w.job];
  if (!J.st->active) return;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kTileF * 257; i += kHistThreads) { s_g[i] = 0.f; s_h[i] = 0.f; }
  __syncthreads();

  const uint8_t* __restrict__ brm = J.bins_rm + w.tile * kTileF;
  const uint32_t* __restrict__ rows = J.rows + static_cast<size_t>(level) * J.N + w.start;
  const float2* __restrict__ gh = J.gh;
  const uint32_t Fp = J.Fp;
  const uint32_t s = tid & 15u;        // this lane's starting feature
  const uint32_t q = s >> 2, r = s & 3u;

  for (uint32_t j0 = 0; j0 < w.len; j0 += kHistThreads * U) {
    uint4 v[U];
    float2 g[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = j0 + u * kHistThreads + tid;
      ok[u] = j < w.len;
      const uint32_t row = ok[u] ? rows[j] : 0u;
      v[u] = ok[u] ? *reinterpret_cast<const uint4*>(brm + static_cast<size_t>(row) * Fp)
                   : make_uint4(0, 0, 0, 0);
      g[u] = ok[u] ? gh[row] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      // rotate the 16 bytes by s: byte k of wr = feature (k + s) % 16
      const uint32_t a0 = v[u].x, a1 = v[u].y, a2 = v[u].z, a3 = v[u].w;
      const uint32_t d0 = q == 0 ? a0 : q == 1 ? a1 : q == 2 ? a2 : a3;
      const uint32_t d1 = q == 0 ? a1 : q == 1 ? a2 : q == 2 ? a3 : a0;
      const uint32_t d2 = q == 0 ? a2 : q == 1 ? a3 : q == 2 ? a0 : a1;
      const uint32_t d3 = q == 0 ? a3 : q == 1 ? a0 : q == 2 ? a1 : a2;
      uint32_t wr[4];
      wr[0] = __builtin_amdgcn_alignbyte(d1, d0, r);
      wr[1] = __builtin_amdgcn_alignbyte(d2, d1, r);
      wr[2] = __builtin_amdgcn_alignbyte(d3, d2, r);
      wr[3] = __builtin_amdgcn_alignbyte(d0, d3, r);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint32_t f = (static_cast<uint32_t>(k) + s) & 15u;
        const uint32_t b = (wr[k >> 2] >> (8 * (k & 3))) & 0xFFu;
        atomicAdd(&s_g[f * 257 + b], g[u].x);
        atomicAdd(&s_h[f * 257 + b], g[u].y);
      }
    }
  }
  __syncthreads();
  const uint32_t f0 = w.tile * kTileF;
  float2* out = J.hist + (static_cast<size_t>(heap_index(level, w.node)) * J.F + f0) * 256u;
  const uint32_t nf = min(static_cast<uint32_t>(kTileF), J.F - f0);
  for (uint32_t i = tid; i < nf * 256u; i += kHistThreads) {
    const uint32_t f = i >> 8, b = i & 255u;
    const float hg = s_g[f * 257 + b], hh = s_h[f * 257 + b];
    if (hh != 0.f || hg != 0.f) {
      unsafeAtomicAdd(&out[i].x, hg);
      unsafeAtomicAdd(&out[i].y, hh);
    }
  }
}

// Zero the histograms of the smaller children at `level` (built by k_hist).
__global__ void __launch_bounds__(256) k_zero_smaller(const JobDev* __restrict__ jobs,
                                                      uint32_t level, uint32_t fmax) {
  const JobDev& J = jobs[blockIdx.y];
  if (!J.st->active) return;
  const uint32_t pair = blockIdx.x / fmax, f = blockIdx.x % fmax;
  if (f >= J.F) return;
  const uint32_t lc = 2 * pair, rc = lc + 1;
  const uint32_t nl = J.nc[heap_index(level, lc)], nr = J.nc[heap_index(level, rc)];
  const uint32_t sm = nl <= nr ? lc : rc;
  float2* h = J.hist + (static_cast<size_t>(heap_index(level, sm)) * J.F + f) * 256u;
  h[threadIdx.x] = make_float2(0.f, 0.f);
}

// A13: larger child = parent - smaller child, over every (feature, slot)
// (Histogram::subtract_from, histogram.cpp:39-42 via symmetric_tree.cpp:346).
__global__ void __launch_bounds__(256) k_subtract(const JobDev* __restrict__ jobs,
                                                  uint32_t level, uint32_t fmax) {
  const JobDev& J = jobs[blockIdx.y];
  if (!J.st->active) return;
  const uint32_t pair = blockIdx.x / fmax, f = blockIdx.x % fmax;
  if (f >= J.F) return;
  const uint32_t lc = 2 * pair, rc = lc + 1;
  const uint32_t nl = J.nc[heap_index(level, lc)], nr = J.nc[heap_index(level, rc)];
  const uint32_t sm = nl <= nr ? lc : rc, lg = nl <= nr ? rc : lc;
  const size_t F = J.F;
  const float2 p = J.hist[(heap_index(level - 1, pair) * F + f) * 256u + threadIdx.x];
  const float2 c = J.hist[(heap_index(level, sm) * F + f) * 256u + threadIdx.x];
  J.hist[(heap_index(level, lg) * F + f) * 256u + threadIdx.x] =
      make_float2(__fsub_rn(p.x, c.x), __fsub_rn(p.y, c.y));
}

// ---------------------------------------------------------------------------
// A10: level split scan.  Workgroup = (feature, job), thread = candidate bin.
// Per node the left sums are the sequential prefix over bins (one lane per
// node, in LDS), so for a given histogram every candidate's left sum, gain
// term and node-order total is the reference's float value
// (symmetric_tree.cpp:153-193, :214-262, with the deeper levels' total +=
// 0.5 * (...) as the fused multiply-add its compiler emits).
__global__ void __launch_bounds__(256) k_split_scan(const JobDev* __restrict__ jobs,
                                                    uint32_t level) {
  __shared__ float2 s_pref[kScanNodesPerPass][256];
  __shared__ unsigned long long s_red[4];
  const uint32_t job = blockIdx.y, f = blockIdx.x, tid = threadIdx.x;
  const JobDev& J = jobs[job];
  if (!J.st->active || f >= J.F) return;
  const uint32_t nn = 1u << level;
  const uint32_t mb = J.max_bins & 0xFFu;
  const float lambda = J.lambda_l2;
  float total = 0.f;
  bool any = false;
  for (uint32_t c0 = 0; c0 < nn; c0 += kScanNodesPerPass) {
    const uint32_t cs = min(static_cast<uint32_t>(kScanNodesPerPass), nn - c0);
    if (tid < cs) {
      const float2* hp = J.hist + (static_cast<size_t>(heap_index(level, c0 + tid)) * J.F + f) * 256u;
      float lg = 0.f, lh = 0.f;
      for (uint32_t b = 0; b + 1 < mb; ++b) {
        const float2 v = hp[b];
        lg = __fadd_rn(lg, v.x);
        lh = __fadd_rn(lh, v.y);
        s_pref[tid][b] = make_float2(lg, lh);
      }
    }
    __syncthreads();
    if (tid + 1 < mb) {
      for (uint32_t k = 0; k < cs; ++k) {
        const uint32_t hn = heap_index(level, c0 + k);
        if (level > 0 && J.nc[hn] == 0) continue;
        const float G = J.ng[hn], H = J.nh[hn];
        const float2 Lf = s_pref[k][tid];
        const float rg = __fsub_rn(G, Lf.x), rh = __fsub_rn(H, Lf.y);
        if (Lf.y < 1e-10f || rh < 1e-10f) continue;
        const float pg = H > 1e-10f ? __fdiv_rn(__fmul_rn(G, G), __fadd_rn(H, lambda)) : 0.f;
        const float lgn = __fdiv_rn(__fmul_rn(Lf.x, Lf.x), __fadd_rn(Lf.y, lambda));
        const float rgn = __fdiv_rn(__fmul_rn(rg, rg), __fadd_rn(rh, lambda));
        total = __fmaf_rn(0.5f, __fsub_rn(__fadd_rn(lgn, rgn), pg), total);
        any = true;
      }
    }
    __syncthreads();
  }
  unsigned long long key = kEmptyKey;
  if (tid + 1 < mb && (level > 0 || any) && !isnan(total)) {
    key = split_key(total == 0.f ? 0.f : total, f, tid);
  }
  key = warp_max_u64(key);
  if ((tid & 63) == 0) s_red[tid >> 6] = key;
  __syncthreads();
  if (tid == 0) {
    unsigned long long m = s_red[0];
    for (int i = 1; i < 4; ++i) m = s_red[i] > m ? s_red[i] : m;
    if (m != kEmptyKey) atomicMax(&J.st->best[level], m);
  }
}

// A11: decode the level's argmax, threshold lookup, min_split_gain stop
// (symmetric_tree.cpp:264-282).
__global__ void k_finalize(const JobDev* __restrict__ jobs, uint32_t level) {
  const JobDev& J = jobs[blockIdx.y];
  if (threadIdx.x != 0 || !J.st->active) return;
  TreeState* st = J.st;
  const unsigned long long key = st->best[level];
  if (key == kEmptyKey) { st->active = 0; return; }
  const float gain = float_unorder(static_cast<uint32_t>(key >> 32));
  const uint32_t fb = 0xFFFFFFFFu - static_cast<uint32_t>(key & 0xFFFFFFFFu);
  const uint32_t f = fb >> 8, b = fb & 255u;
  if (!(gain > -1e29f) || gain < J.min_split_gain) { st->active = 0; return; }
  float fthr = 0.f;
  const uint32_t ne = J.n_edges[f];
  if (b < ne) fthr = J.edges[static_cast<size_t>(f) * 255 + b];
  else if (ne > 0) fthr = J.edges[static_cast<size_t>(f) * 255 + ne - 1];
  const size_t o = static_cast<size_t>(st->tree) * J.D + level;
  J.m_feat[o] = static_cast<uint16_t>(f);
  J.m_bin[o] = static_cast<uint8_t>(b);
  J.m_fthr[o] = fthr;
  J.m_gain[o] = gain;
  st->depth = static_cast<int32_t>(level) + 1;
}

// ---------------------------------------------------------------------------
// A12: partition.  Pass 1 counts left rows and sums child statistics per work
// item (double, fixed reduction order); the plan kernel scans them per node;
// pass 3 writes each row to its child segment, stable.  Routing: NaN bin 255
// or bin <= threshold goes left (symmetric_tree.cpp:309-311).
__device__ inline bool part_active(const JobDev& J, uint32_t level) {
  return J.st->depth == static_cast<int32_t>(level) + 1 && J.st->tree < static_cast<int>(J.n_trees);
}

__global__ void __launch_bounds__(256) k_part_count(const JobDev* __restrict__ jobs,
                                                    const PartWork* __restrict__ work,
                                                    const uint32_t* __restrict__ count,
                                                    uint32_t level) {
  if (blockIdx.x >= *count) return;
  const PartWork w = work[blockIdx.x];
  const JobDev& J = jobs[w.job];
  if (!part_active(J, level)) return;
  const size_t o = static_cast<size_t>(J.st->tree) * J.D + level;
  const uint8_t* col = J.bins_cm + static_cast<size_t>(J.m_feat[o]) * J.N;
  const uint8_t thr = J.m_bin[o];
  const uint32_t* rows = J.rows + static_cast<size_t>(level) * J.N + w.start;
  const uint32_t tid = threadIdx.x;
  uint32_t nl = 0;
  double gl = 0.0, hl = 0.0, gr = 0.0, hr = 0.0;
  for (uint32_t j = tid; j < w.len; j += 256) {
    const uint32_t row = rows[j];
    const uint8_t b = col[row];
    const float2 v = J.gh[row];
    if (b == 255 || b <= thr) { ++nl; gl += v.x; hl += v.y; }
    else { gr += v.x; hr += v.y; }
  }
  __shared__ double s_d[4][4];
  __shared__ uint32_t s_u[4];
  nl = warp_sum_u(nl);
  gl = warp_sum_d(gl); hl = warp_sum_d(hl); gr = warp_sum_d(gr); hr = warp_sum_d(hr);
  const uint32_t wv = tid >> 6;
  if ((tid & 63) == 0) { s_u[wv] = nl; s_d[wv][0] = gl; s_d[wv][1] = hl; s_d[wv][2] = gr; s_d[wv][3] = hr; }
  __syncthreads();
  if (tid == 0) {
    uint32_t a = 0;
    double st[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
      a += s_u[i];
      for (int k = 0; k < 4; ++k) st[k] += s_d[i][k];
    }
    J.part_left[w.slot] = a;
    for (int k = 0; k < 4; ++k) J.part_stats[4 * static_cast<size_t>(w.slot) + k] = st[k];
  }
}

// Per job: child offsets and totals, level stop test, next level's work.
__global__ void __launch_bounds__(256) k_part_plan(const JobDev* __restrict__ jobs, Lists L,
                                                   uint32_t level) {
  const uint32_t job = blockIdx.y, tid = threadIdx.x;
  const JobDev& J = jobs[job];
  if (!part_active(J, level)) return;
  const uint32_t nn = 1u << level;
  const uint32_t stride = (1u << J.D) + 1u;
  const uint32_t* offd = J.off + static_cast<size_t>(level) * stride;
  uint32_t* offn = J.off + static_cast<size_t>(level + 1) * stride;
  __shared__ uint32_t s_base[(1 << kMaxDepth) + 1];
  __shared__ int s_any;
  if (tid == 0) {  // slot base per node: node order, ceil(len / chunk) slots each
    uint32_t acc = 0;
    for (uint32_t n = 0; n < nn; ++n) {
      s_base[n] = acc;
      acc += part_slots(offd[n + 1] - offd[n]);
    }
    s_base[nn] = acc;
    s_any = 0;
  }
  __syncthreads();
  const float mcw2 = J.min_child_weight * 2.0f;
  for (uint32_t n = tid; n < nn; n += 256) {
    const uint32_t b0 = s_base[n], b1 = s_base[n + 1];
    uint32_t acc = 0;
    double gl = 0, hl = 0, gr = 0, hr = 0;
    for (uint32_t s = b0; s < b1; ++s) {
      J.part_loff[s] = acc;
      acc += J.part_left[s];
      gl += J.part_stats[4 * static_cast<size_t>(s) + 0];
      hl += J.part_stats[4 * static_cast<size_t>(s) + 1];
      gr += J.part_stats[4 * static_cast<size_t>(s) + 2];
      hr += J.part_stats[4 * static_cast<size_t>(s) + 3];
    }
    const uint32_t len = offd[n + 1] - offd[n];
    const uint32_t hl_i = heap_index(level + 1, 2 * n), hr_i = hl_i + 1;
    J.nc[hl_i] = acc;
    J.nc[hr_i] = len - acc;
    J.ng[hl_i] = static_cast<float>(gl);
    J.nh[hl_i] = static_cast<float>(hl);
    J.ng[hr_i] = static_cast<float>(gr);
    J.nh[hr_i] = static_cast<float>(hr);
    offn[2 * n] = offd[n];
    offn[2 * n + 1] = offd[n] + acc;
    // stop test of the next level (symmetric_tree.cpp:115-124)
    if ((acc >= 2u * J.min_samples_leaf && J.nh[hl_i] >= mcw2) ||
        (len - acc >= 2u * J.min_samples_leaf && J.nh[hr_i] >= mcw2))
      s_any = 1;
  }
  if (tid == 0) offn[2 * nn] = offd[nn];
  __syncthreads();
  if (level + 1 >= J.D || !s_any) {
    if (tid == 0) J.st->active = 0;
    return;
  }
  if (tid == 0) {
    // next level: histograms of the smaller child of every pair (ties: left)
    for (uint32_t n = 0; n < nn; ++n) {
      const uint32_t nl = J.nc[heap_index(level + 1, 2 * n)];
      const uint32_t nr = J.nc[heap_index(level + 1, 2 * n + 1)];
      const uint32_t sm = nl <= nr ? 2 * n : 2 * n + 1;
      emit_hist(J, job, L.hist[level + 1], L.hist_count[level + 1], sm, offn[sm],
                offn[sm + 1] - offn[sm]);
    }
    // next level: partition work over every child, slots in node order
    uint32_t total = 0;
    for (uint32_t c = 0; c < 2 * nn; ++c) total += part_slots(offn[c + 1] - offn[c]);
    const uint32_t base = atomicAdd(L.part_count[level + 1], total);
    uint32_t slot = 0;
    for (uint32_t c = 0; c < 2 * nn; ++c) {
      const uint32_t len = offn[c + 1] - offn[c];
      for (uint32_t s = 0; s < part_slots(len); ++s, ++slot) {
        PartWork w;
        w.job = job; w.node = c; w.start = offn[c] + s * kPartChunk;
        w.len = min(static_cast<uint32_t>(kPartChunk), len - s * kPartChunk);
        w.slot = slot; w.pad0 = w.pad1 = w.pad2 = 0;
        L.part[level + 1][base + slot] = w;
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_part_scatter(const JobDev* __restrict__ jobs,
                                                      const PartWork* __restrict__ work,
                                                      const uint32_t* __restrict__ count,
                                                      uint32_t level) {
  if (blockIdx.x >= *count) return;
  const PartWork w = work[blockIdx.x];
  const JobDev& J = jobs[w.job];
  if (!part_active(J, level)) return;
  const size_t o = static_cast<size_t>(J.st->tree) * J.D + level;
  const uint8_t* col = J.bins_cm + static_cast<size_t>(J.m_feat[o]) * J.N;
  const uint8_t thr = J.m_bin[o];
  const uint32_t stride = (1u << J.D) + 1u;
  const uint32_t* offd = J.off + static_cast<size_t>(level) * stride;
  const uint32_t* offn = J.off + static_cast<size_t>(level + 1) * stride;
  const uint32_t* rows = J.rows + static_cast<size_t>(level) * J.N;
  uint32_t* rows_n = J.rows + static_cast<size_t>(level + 1) * J.N;
  const uint32_t loff = J.part_loff[w.slot];
  const uint32_t before = w.start - offd[w.node];
  uint32_t lpos = offn[2 * w.node] + loff;
  uint32_t rpos = offn[2 * w.node + 1] + (before - loff);
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ uint32_t s_cnt[2][4];
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t j0 = 0; j0 < w.len; j0 += 256) {
    const uint32_t j = j0 + tid;
    const bool valid = j < w.len;
    const uint32_t row = valid ? rows[w.start + j] : 0u;
    const uint8_t b = valid ? col[row] : 0;
    const bool left = valid && (b == 255 || b <= thr);
    const bool right = valid && !left;
    const unsigned long long ml = __ballot(left), mr = __ballot(right);
    if (lane == 0) { s_cnt[0][wv] = __popcll(ml); s_cnt[1][wv] = __popcll(mr); }
    __syncthreads();
    uint32_t pl = 0, pr = 0, tl = 0, tr = 0;
    for (uint32_t k = 0; k < 4; ++k) {
      if (k < wv) { pl += s_cnt[0][k]; pr += s_cnt[1][k]; }
      tl += s_cnt[0][k];
      tr += s_cnt[1][k];
    }
    if (left) rows_n[lpos + pl + __popcll(ml & lt)] = row;
    if (right) rows_n[rpos + pr + __popcll(mr & lt)] = row;
    lpos += tl;
    rpos += tr;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// A14: leaf values of the finished tree (compute_leaf_value,
// symmetric_tree.cpp:753-775).
__global__ void __launch_bounds__(256) k_leaves(const JobDev* __restrict__ jobs) {
  const JobDev& J = jobs[blockIdx.y];
  const int t = J.st->tree;
  if (t < 0 || t >= static_cast<int>(J.n_trees)) return;
  const int depth = J.st->depth;
  const uint32_t nl = 1u << depth;
  float* out = J.m_leaves + static_cast<size_t>(t) * (1u << J.D);
  for (uint32_t i = threadIdx.x; i < (1u << J.D); i += blockDim.x) {
    float v = 0.f;
    if (i < nl) {
      const uint32_t hi = heap_index(depth, i);
      const float G = J.ng[hi], H = J.nh[hi];
      if (!(H < 1e-10f)) {
        v = __fdiv_rn(-G, __fadd_rn(H, J.lambda_l2));
        if (J.lambda_l1 > 0.f) {
          if (v > J.lambda_l1) v = __fsub_rn(v, J.lambda_l1);
          else if (v < -J.lambda_l1) v = __fadd_rn(v, J.lambda_l1);
          else v = 0.f;
        }
        if (J.max_delta_step > 0.f) v = fmaxf(-J.max_delta_step, fminf(J.max_delta_step, v));
      }
    }
    out[i] = v;
  }
  if (threadIdx.x == 0) J.m_depth[t] = static_cast<uint16_t>(depth);
}

// Copy externally supplied gradients (teacher forcing) into gh.
__global__ void __launch_bounds__(256) k_set_gh(const JobDev* __restrict__ jobs, const float* g,
                                                const float* h) {
  const JobDev& J = jobs[0];
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p < J.N) J.gh[p] = make_float2(g[p], h[p]);
}

}  // namespace tc
