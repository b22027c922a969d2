// Symmetric-tree training kernels for gfx950.  One launch of each kernel
// serves every job of a batch (a single-job fit is a batch of one): per-job
// kernels put the job on blockIdx.y, row-level kernels use a flattened
// workgroup space, and histogram / partition kernels consume work lists.
//
// Per tree the launch sequence is fixed (so it can be captured in a graph);
// data-dependent control (level stops, tree count) lives in TreeState and
// makes later kernels of a stopped job return at once.
//
//   [root_partials]           (only for jobs whose gradient update cannot write them)
//   tree_begin
//   for level d < D:
//     [d=0] hist(0, root)
//     [d>0] hist(d, smaller children)  (one job: plans level d itself)
//     split_scan(d)           (partial slots summed, sibling = parent - smaller)
//     [d<D-1] partition(d)    (finalizes the level: split record, child totals)
//             [several jobs] part_plan(d)
//     [d=D-1] finalize(d)     (or, keeping the last rows, partition + part_plan)
//   grad_update(apply)        (leaf values, prediction update, gradients,
//                              validation rows, next root partials)
//   [eval]                    (losses and early stopping)
//
// Histograms are exact fixed-point sums: a row's g and h enter as int64
// round(v * 2^e), with e chosen per tree from max |g|, max |h| and the row
// count so that no sum overflows.  They are accumulated in LDS with integer
// atomics (on gfx950 about 20x the rate of LDS float atomics); each work item
// stores its cells (a node's last chunk into the node's histogram, the others
// into partial slots the split scan adds).  Integer addition is order-independent, so every
// histogram, and every split, is the same run to run and the same for a job
// trained alone or in a batch.  The same exactness makes the order of rows
// inside a node irrelevant, so the partition is a single unstable pass, and
// child totals come from the parent's histogram instead of a row pass.
#include "tc_device.h"

namespace tc {

struct Lists {
  HistWork* hist[kMaxDepth];
  uint32_t* hist_count[kMaxDepth];
  PartWork* part[kMaxDepth];
  uint32_t* part_count[kMaxDepth];
  uint32_t single;  // one job: its items start at 0 and the counts are stored, not reserved
};

// Room for n items in a level list: with one job the list is the job's
// (store the count, no round trip); otherwise reserve with an atomic.
__device__ inline uint32_t list_reserve(const Lists& L, uint32_t* count, uint32_t n) {
  if (L.single) {
    *count = n;
    return 0u;
  }
  return atomicAdd(count, n);
}

__device__ inline uint32_t part_slots(uint32_t len) {
  return (len + kPartChunk - 1) / kPartChunk;
}

// In-place exclusive scan of a[0..n) (LDS, n <= 1024) by a 256-thread
// workgroup; returns the total to every thread.  tmp: 256 words of LDS.
__device__ inline uint32_t block_scan_excl(uint32_t* a, uint32_t n, uint32_t* tmp) {
  // four elements per thread, an inclusive shuffle scan of the thread sums
  // per wave, then the 4 wave totals through LDS: two barriers in all
  const uint32_t tid = threadIdx.x, k0 = tid * 4, lane = tid & 63u, wave = tid >> 6;
  uint32_t loc[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t v = (k0 + i < n) ? a[k0 + i] : 0u;
    loc[i] = sum;
    sum += v;
  }
  const uint32_t incl = wave_incl_scan_u32(sum);
  if (lane == 63u) tmp[wave] = incl;
  __syncthreads();
  uint32_t wbase = 0, total = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; ++w) {
    const uint32_t t = tmp[w];
    wbase += w < wave ? t : 0u;
    total += t;
  }
  const uint32_t excl = wbase + incl - sum;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (k0 + i < n) a[k0 + i] = excl + loc[i];
  __syncthreads();
  return total;
}

// Item owning element e of a scanned list: the last i in [0, n) with
// off[i] <= e (off non-decreasing, e < total).
__device__ inline uint32_t find_item(const uint32_t* off, uint32_t n, uint32_t e) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ inline void put_hist_work(HistWork* dst, uint32_t job, uint32_t node, uint32_t start,
                                     uint32_t len, uint32_t tile, uint32_t cls, bool last) {
  HistWork w;
  w.job = job; w.node = node; w.start = start; w.len = len; w.tile = tile; w.cls = cls;
  w.last = last ? 1u : 0u;
  w.pad = 0;
  *dst = w;
}

__device__ inline void put_part_work(PartWork* dst, uint32_t job, uint32_t node, uint32_t start,
                                     uint32_t len, uint32_t slot) {
  PartWork w;
  w.job = job; w.node = node; w.start = start; w.len = len; w.slot = slot;
  w.pad0 = w.pad1 = w.pad2 = 0;
  *dst = w;
}

// Workgroup max of two non-negative floats, stored by thread 0 as float bits
// into dst[0..1] (the gradient update's per-tile maxima; the next tree start
// takes their max -- order-free, and no two workgroups touch one address).
__device__ inline void block_max2_store(float a, float b, uint32_t* dst) {
  __shared__ uint32_t s_m[2][4];
  a = wave_max_f32(a);
  b = wave_max_f32(b);
  if ((threadIdx.x & 63) == 0) {
    s_m[0][threadIdx.x >> 6] = __float_as_uint(a);
    s_m[1][threadIdx.x >> 6] = __float_as_uint(b);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t x = s_m[0][0], y = s_m[1][0];
    for (int w = 1; w < 4; ++w) {
      x = s_m[0][w] > x ? s_m[0][w] : x;
      y = s_m[1][w] > y ? s_m[1][w] : y;
    }
    dst[0] = x;
    dst[1] = y;
  }
}

// Workgroup max of two non-negative floats, merged into st (float bits of
// non-negative floats order like the floats).
__device__ inline void block_max2_to(TreeState* st, float a, float b) {
  __shared__ uint32_t s_m[2];
  if (threadIdx.x < 2) s_m[threadIdx.x] = 0u;
  __syncthreads();
  a = wave_max_f32(a);
  b = wave_max_f32(b);
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&s_m[0], __float_as_uint(a));
    atomicMax(&s_m[1], __float_as_uint(b));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_m[0]) atomicMax(&st->gmax_next, s_m[0]);
    if (s_m[1]) atomicMax(&st->hmax_next, s_m[1]);
  }
}

// ---------------------------------------------------------------------------
// Gradients (A6/A7) fused with the leaf values (A14) and the training-
// prediction update (A15 + K9).
//
// mode 0: preds = base for every row (before the first tree).
// mode 1: leaf values of tree st->tree from its final level's totals
//         (compute_leaf_value, symmetric_tree.cpp:753-775; every workgroup
//         computes them, workgroup 0 stores them), then apply the tree to
//         every row: leaf index from the binned
//         compares go_right = bin != 255 && bin > threshold, MSB first
//         (SymmetricTree::predict_batch, symmetric_tree.cpp:813-840), then
//         preds += lr * leaf as one fused multiply-add (the reference
//         compiler's contraction of booster.cpp:460-463), then g/h.
// Rows in natural order, 1024 per workgroup, every row of the dataset (also
// when the tree was grown on a row sample).  Also records max |g|, max |h|
// for the next tree's histogram scales, and empties the work lists.
constexpr int kGradTile = 1024;

// Mean per-row loss terms of Loss::compute_loss (A19): MSE (loss.cpp:336-346)
// and LogLoss (loss.cpp:110-121, sigmoid and 1e-7 clip of loss.hpp:99-105).
__device__ inline float row_loss(int loss, float pred, float label) {
  if (loss == kMSE) {
    const float d = __fsub_rn(pred, label);
    return __fmul_rn(d, d);
  }
  float q = __fdiv_rn(1.0f, __fadd_rn(1.0f, ref_expf(-pred)));
  q = fmaxf(1e-7f, fminf(1.0f - 1e-7f, q));
  return __fsub_rn(__fmul_rn(-label, logf(q)), __fmul_rn(__fsub_rn(1.0f, label), logf(__fsub_rn(1.0f, q))));
}

// A 256-thread workgroup's sum of one double per thread, in a fixed order
// (shuffle tree per wave, waves in order), stored by thread 0.
__device__ inline void tile_sum_store(double v, double* dst) {
  __shared__ double s_w[4];
  v = warp_sum_d(v);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *dst = ((s_w[0] + s_w[1]) + s_w[2]) + s_w[3];
  __syncthreads();
}

// One 256-thread workgroup's root partial (sum g, sum h of its 1024-row tile,
// channel c): every thread's double sum of its rows in order, a fixed
// shuffle tree per wave, the 4 waves added in order.
__device__ inline void root_partial_store(const JobDev& J, uint32_t lw, uint32_t c, double sg,
                                          double sh) {
  __shared__ double s_red[2][4];
  const uint32_t tid = threadIdx.x;
  sg = warp_sum_d(sg);
  sh = warp_sum_d(sh);
  if ((tid & 63) == 0) { s_red[0][tid >> 6] = sg; s_red[1][tid >> 6] = sh; }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < 4; ++w) { a += s_red[0][w]; b += s_red[1][w]; }
    J.grad_partials[2 * (lw * J.K + c)] = a;
    J.grad_partials[2 * (lw * J.K + c) + 1] = b;
  }
  __syncthreads();
}

__device__ inline uint32_t job_of_block(const uint32_t* job_wg_off, uint32_t njobs) {
  uint32_t lo = 0, hi = njobs;  // last job with job_wg_off[j] <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (job_wg_off[mid] <= blockIdx.x) lo = mid; else hi = mid;
  }
  return lo;
}

// K-class rows of k_grad_update: every channel's prediction += lr * leaf
// (fused), then A8 (CrossEntropyLoss::compute_multiclass_gradients +
// softmax, loss.cpp:273-318): max shift, exp, sum in class order, divide,
// g = p - [c == y], h = max(p (1 - p), 1e-6).
__device__ void grad_update_softmax(const JobDev& J, uint32_t lw, int mode, int depth,
                                    const uint16_t* s_f, const uint8_t* s_b, const float* s_leaf) {
  const uint32_t N = J.N, K = J.K, tid = threadIdx.x;
  float gmax = 0.f, hmax = 0.f;
  for (int k = 0; k < kGradTile / 256; ++k) {
    const uint32_t row = lw * kGradTile + k * 256 + tid;
    if (row < N) {
      uint32_t li = 0;
      if (mode == 1)
        for (int d = 0; d < depth; ++d) {
          const uint8_t v = J.bins_cm[static_cast<size_t>(s_f[d]) * N + row];
          li = (li << 1) | ((v != 255 && v > s_b[d]) ? 1u : 0u);
        }
      float e[kMaxClasses];  // static indices only (unrolled), so it stays in registers
      float m = -INFINITY;
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        float* pp = J.preds + static_cast<size_t>(c) * N + row;
        const float pr = mode == 1 ? __fmaf_rn(J.lr, s_leaf[li * K + c], *pp) : J.base;
        *pp = pr;
        e[c] = pr;
        m = fmaxf(m, pr);
      }
      float sum = 0.f;
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        e[c] = ref_expf(__fsub_rn(e[c], m));
        sum = __fadd_rn(sum, e[c]);
      }
      const uint32_t lab = static_cast<uint32_t>(J.y[row]);
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        const float p = __fdiv_rn(e[c], sum);
        const float g = __fsub_rn(p, c == lab ? 1.0f : 0.0f);
        const float hh = __fmul_rn(p, __fsub_rn(1.0f, p));
        const float h = hh > 1e-6f ? hh : 1e-6f;
        J.gh[static_cast<size_t>(c) * N + row] = make_float2(g, h);
        gmax = fmaxf(gmax, fabsf(g));
        hmax = fmaxf(hmax, h);
      }
    }
  }
  block_max2_store(gmax, hmax, J.ghmax_partials + 2 * lw);
}

__global__ void __launch_bounds__(256) k_grad_update(const JobDev* __restrict__ jobs,
                                                     const uint32_t* __restrict__ job_wg_off,
                                                     uint32_t njobs, int mode, Lists L) {
  __shared__ uint32_t s_job;
  __shared__ uint16_t s_f[kMaxDepth];
  __shared__ uint8_t s_b[kMaxDepth];
  extern __shared__ float s_leaf[];  // [2^D][K], sized by the launch
  const uint32_t tid = threadIdx.x;
  if (blockIdx.x == 0 && tid < kMaxDepth) {  // work lists of the next tree start empty
    *L.hist_count[tid] = 0;
    *L.part_count[tid] = 0;
  }
  if (tid == 0) s_job = job_of_block(job_wg_off, njobs);
  __syncthreads();
  const JobDev& J = jobs[s_job];
  const uint32_t lw = blockIdx.x - J.grad_wg_off;
  const uint32_t N = J.N;
  const uint32_t vecN = (N / 8u) * 8u;
  // each lane takes rows r + 256k (k < 4); every load -- predictions,
  // labels, and the split column of each level -- is issued before any use
  constexpr int kR = kGradTile / 256;
  uint32_t row[kR], li[kR];
  bool ok[kR];
  float prev[kR], lab[kR];
  if (J.K == 1) {
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      row[k] = lw * kGradTile + k * 256 + tid;
      ok[k] = row[k] < N;
      li[k] = 0;
      prev[k] = (ok[k] && mode == 1) ? J.preds[row[k]] : 0.f;
      lab[k] = ok[k] ? J.y[row[k]] : 0.f;
    }
  }
  int t = -1, depth = 0;
  if (mode == 1) {
    t = J.st->tree;
    if (t < 0 || t >= static_cast<int>(J.n_trees) || J.st->stopped) return;
    depth = J.st->depth;
    if (tid < static_cast<uint32_t>(depth)) {
      s_f[tid] = J.m_feat[static_cast<size_t>(t) * J.D + tid];
      s_b[tid] = J.m_bin[static_cast<size_t>(t) * J.D + tid];
    }
    const uint32_t K = J.K;
    float* out = J.m_leaves + (static_cast<size_t>(t) << J.D) * K;
    for (uint32_t i = tid; i < (K << J.D); i += 256) {  // i = leaf * K + channel
      float v = 0.f;
      if (i < (K << depth)) {
        const uint32_t hi = heap_index(depth, i / K) * K + i % K;
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
      s_leaf[i] = v;
      if (lw == 0) out[i] = v;
    }
    if (lw == 0 && tid == 0) J.m_depth[t] = static_cast<uint16_t>(depth);
    __syncthreads();
  }
  if (J.K > 1) {  // softmax channels
    grad_update_softmax(J, lw, mode, depth, s_f, s_b, s_leaf);
    return;
  }
  if (lw >= J.grad_wg_cnt) {  // validation rows: apply the tree, then the loss (booster.cpp:259-283)
    const uint32_t vw = lw - J.grad_wg_cnt, Nv = J.Nv;
    uint32_t vr[kR], vli[kR];
    bool vok[kR];
    float vp[kR], vy[kR];
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      vr[k] = vw * kGradTile + k * 256 + tid;
      vok[k] = vr[k] < Nv;
      vli[k] = 0;
      vp[k] = (vok[k] && mode == 1) ? J.v_preds[vr[k]] : 0.f;
      vy[k] = vok[k] ? J.v_y[vr[k]] : 0.f;
    }
    if (mode == 1) {
      uint32_t v[kMaxDepth][kR];
#pragma unroll
      for (int d = 0; d < kMaxDepth; ++d) {
        if (d < depth) {
          const uint8_t* col = J.v_bins_cm + static_cast<size_t>(s_f[d]) * Nv;
#pragma unroll
          for (int k = 0; k < kR; ++k) v[d][k] = vok[k] ? col[vr[k]] : 0u;
        }
      }
#pragma unroll
      for (int d = 0; d < kMaxDepth; ++d) {
        if (d < depth) {
          const uint32_t thr = s_b[d];
#pragma unroll
          for (int k = 0; k < kR; ++k) vli[k] = (vli[k] << 1) | ((v[d][k] != 255u && v[d][k] > thr) ? 1u : 0u);
        }
      }
    }
    double sl = 0.0;
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      if (!vok[k]) continue;
      // valid_preds[i] += lr * tree value, one fused multiply-add
      const float pred = mode == 1 ? __fmaf_rn(J.lr, s_leaf[vli[k]], vp[k]) : J.base;
      J.v_preds[vr[k]] = pred;
      sl += row_loss(J.loss, pred, vy[k]);
    }
    if (mode == 1) tile_sum_store(sl, J.loss_partials + lw);
    return;
  }
  if (mode == 1) {
    uint32_t v[kMaxDepth][kR];
#pragma unroll
    for (int d = 0; d < kMaxDepth; ++d) {
      if (d < depth) {
        const uint8_t* col = J.bins_cm + static_cast<size_t>(s_f[d]) * N;
#pragma unroll
        for (int k = 0; k < kR; ++k) v[d][k] = ok[k] ? col[row[k]] : 0u;
      }
    }
#pragma unroll
    for (int d = 0; d < kMaxDepth; ++d) {
      if (d < depth) {
        const uint32_t thr = s_b[d];
#pragma unroll
        for (int k = 0; k < kR; ++k) li[k] = (li[k] << 1) | ((v[d][k] != 255u && v[d][k] > thr) ? 1u : 0u);
      }
    }
  }
  float gmax = 0.f, hmax = 0.f;
  double sg = 0.0, sh = 0.0;  // the root partials of this tile (k_root_partials' order)
  double tl = 0.0;            // the training loss after this tree (A19)
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    if (!ok[k]) continue;
    const float pred = mode == 1 ? __fmaf_rn(J.lr, s_leaf[li[k]], prev[k]) : J.base;
    J.preds[row[k]] = pred;
    const float2 gh = row_gradient(J.loss, pred, lab[k], row[k] < vecN);
    J.gh[row[k]] = gh;
    gmax = fmaxf(gmax, fabsf(gh.x));
    hmax = fmaxf(hmax, fabsf(gh.y));
    sg += gh.x;
    sh += gh.y;
    tl += J.train_loss ? row_loss(J.loss, pred, lab[k]) : 0.0f;
  }
  if (mode == 1 && J.train_loss) tile_sum_store(tl, J.loss_partials + lw);
  if (J.fused_root) root_partial_store(J, lw, 0, sg, sh);
  block_max2_store(gmax, hmax, J.ghmax_partials + 2 * lw);
}

// A19 and early stopping (booster.cpp:252-352), one workgroup per job after
// the gradient update: the tiles' loss sums added in a fixed order, the mean
// losses of the tree just added recorded, then the stopping rule on the
// validation loss (improved = loss < best - 1e-10; stop once the tree index
// reaches es_min_iter and es_rounds trees in a row did not improve).
__global__ void __launch_bounds__(256) k_eval(const JobDev* __restrict__ jobs) {
  const JobDev& J = jobs[blockIdx.y];
  if (!J.loss_partials) return;
  TreeState* st = J.st;
  const int t = st->tree;
  if (st->stopped || t < 0 || t >= static_cast<int>(J.n_trees)) return;
  const uint32_t tid = threadIdx.x;
  __shared__ double s_w[2][4];
  double a = 0.0, b = 0.0;
  if (J.train_loss)
    for (uint32_t w = tid; w < J.grad_wg_cnt; w += 256) a += J.loss_partials[w];
  for (uint32_t w = tid; w < J.v_wg_cnt; w += 256) b += J.loss_partials[J.grad_wg_cnt + w];
  a = warp_sum_d(a);
  b = warp_sum_d(b);
  if ((tid & 63) == 0) {
    s_w[0][tid >> 6] = a;
    s_w[1][tid >> 6] = b;
  }
  __syncthreads();
  if (tid == 0) {
    const double ta = ((s_w[0][0] + s_w[0][1]) + s_w[0][2]) + s_w[0][3];
    const double va = ((s_w[1][0] + s_w[1][1]) + s_w[1][2]) + s_w[1][3];
    if (J.train_loss) J.train_loss[t] = static_cast<float>(ta / J.N);
    if (J.Nv) {
      const float vl = static_cast<float>(va / J.Nv);
      J.valid_loss[t] = vl;
      if (vl < __fsub_rn(st->es_best, 1e-10f)) {
        st->es_best = vl;
        st->es_best_iter = static_cast<uint32_t>(t);
        st->es_count = 0;
      } else {
        st->es_count += 1;
      }
      if (static_cast<uint32_t>(t) >= J.es_min_iter && st->es_count >= J.es_rounds) st->stopped = 1;
    }
  }
}

// Root of a new tree (symmetric_tree.cpp:97-124), one 256-thread workgroup
// per job: per-channel totals from the root partials (double, fixed order),
// level-0 stop test, the tree's histogram scales, root work lists.
// packed: histograms use the packed MSE cell (h == 2 for every row).
__device__ void tree_begin(const JobDev& J, uint32_t job, const Lists& L, int packed) {
  const uint32_t tid = threadIdx.x, K = J.K;
  __shared__ int s_go;
  __shared__ uint32_t s_hbase, s_pbase;
  __shared__ double s_tot[2][4];
  __shared__ double s_gh[2 * kMaxClasses];
  const uint32_t n = J.n_root;
  const uint32_t hcnt = ((n + J.hist_chunk - 1) / J.hist_chunk) * J.T * K, pcnt = part_slots(n);
  // the tree state, in flight while the partials are summed
  int prev_tree = 0;
  uint32_t gbits = 0, hbits = 0, stopped = 0;
  if (tid == 0) {
    prev_tree = J.st->tree;
    stopped = J.st->stopped;
  }
  {  // max |g|, max |h| of the gradients: the gradient update's tile maxima, and
     // what teacher forcing merged into the state
    __shared__ uint32_t s_mx[2][4];
    uint32_t x = tid == 0 ? J.st->gmax_next : 0u, y = tid == 0 ? J.st->hmax_next : 0u;
    for (uint32_t w = tid; w < J.grad_wg_cnt; w += 256) {
      x = max(x, J.ghmax_partials[2 * w]);
      y = max(y, J.ghmax_partials[2 * w + 1]);
    }
    x = wave_max_u32(x);
    y = wave_max_u32(y);
    if ((tid & 63) == 0) {
      s_mx[0][tid >> 6] = x;
      s_mx[1][tid >> 6] = y;
    }
    __syncthreads();
    if (tid == 0) {
      gbits = max(max(s_mx[0][0], s_mx[0][1]), max(s_mx[0][2], s_mx[0][3]));
      hbits = max(max(s_mx[1][0], s_mx[1][1]), max(s_mx[1][2], s_mx[1][3]));
    }
  }
  for (uint32_t c = 0; c < K; ++c) {  // thread t adds tiles t, t + 256, ... in order; then a fixed tree
    double a = 0.0, b = 0.0;
    for (uint32_t w = tid; w < J.grad_wg_cnt; w += 256) {
      a += J.grad_partials[2 * (w * K + c)];
      b += J.grad_partials[2 * (w * K + c) + 1];
    }
    a = warp_sum_d(a);
    b = warp_sum_d(b);
    if ((tid & 63) == 0) { s_tot[0][tid >> 6] = a; s_tot[1][tid >> 6] = b; }
    __syncthreads();
    if (tid == 0) {
      s_gh[2 * c] = (s_tot[0][0] + s_tot[0][1]) + (s_tot[0][2] + s_tot[0][3]);
      s_gh[2 * c + 1] = (s_tot[1][0] + s_tot[1][1]) + (s_tot[1][2] + s_tot[1][3]);
    }
    __syncthreads();
  }
  if (tid == 0 && stopped) {  // early stopping ended the fit: no more trees
    J.st->active = 0;
    s_go = 0;
  } else if (tid == 0) {
    TreeState* st = J.st;
    const int t = prev_tree + 1;
    st->tree = t;
    st->depth = 0;
    for (int d = 0; d < kMaxDepth; ++d) st->best[d] = kEmptyKey;
    const float gmax = __uint_as_float(gbits), hmax = __uint_as_float(hbits);
    st->gmax_next = 0u;
    st->hmax_next = 0u;
    int go = 0;
    if (t < static_cast<int>(J.n_trees)) {
      float Hsum = 0.f;  // the stop test's H sums the channels in class order
      for (uint32_t c = 0; c < K; ++c) {
        J.ng[c] = static_cast<float>(s_gh[2 * c]);
        J.nh[c] = static_cast<float>(s_gh[2 * c + 1]);
        Hsum = __fadd_rn(Hsum, J.nh[c]);
      }
      J.nc[0] = n;
      J.off[0] = 0;
      J.off[1] = n;
      go = (n >= 2u * J.min_samples_leaf && Hsum >= J.min_child_weight * 2.0f) ? 1 : 0;
      if (n == 0) go = 0;  // empty sample: depth-0 tree, leaf 0 (:39-43)
      // the packed bound uses the largest work item any chunking may use, so
      // the scales -- and every histogram -- do not depend on the chunking
      const int eg = fixed_exponent(gmax, n, kMaxHistChunk, packed != 0);
      const int eh = fixed_exponent(hmax, n, kMaxHistChunk, false);
      st->gscale = ldexp(1.0, eg);
      st->ginv = ldexp(1.0, -eg);
      st->hscale = ldexp(1.0, eh);
      st->hinv = ldexp(1.0, -eh);
      if (go) {
        s_hbase = list_reserve(L, L.hist_count[0], hcnt);
        s_pbase = list_reserve(L, L.part_count[0], pcnt);
        J.hfirst[0] = s_hbase;
        J.lvl_chunk[0] = J.hist_chunk;
        J.cursor[0] = 0;  // left position
        J.cursor[1] = 0;  // right rows taken
        if (J.hist_rows_total) atomicAdd(J.hist_rows_total, static_cast<unsigned long long>(n));
      }
    }
    st->active = go;
    s_go = go;
  }
  __syncthreads();
  if (!s_go) return;
  // chunk-major, then tile, then channel: the items reading the same row
  // segments are adjacent (and share an XCD)
  const uint32_t TK = J.T * K, nch = hcnt / TK;
  for (uint32_t e = tid; e < hcnt; e += blockDim.x) {
    const uint32_t start = (e / TK) * J.hist_chunk, r = e % TK;
    put_hist_work(L.hist[0] + s_hbase + e, job, 0, start, min(J.hist_chunk, n - start), r / K, r % K,
                  e / TK + 1 == nch);
  }
  for (uint32_t e = tid; e < pcnt; e += blockDim.x) {
    const uint32_t start = e * kPartChunk;
    put_part_work(L.part[0] + s_pbase + e, job, 0, start,
                  min(static_cast<uint32_t>(kPartChunk), n - start), e);
  }
}

// The root list of a plain row sample, ascending: rows[0 .. n_root) from the
// sample's bitmask sm (made on the host: the rows drawn before each chunk of
// kMaskChunkWords words, then the words; bit r of word r / 32 is row r).  A
// node's row order never changes a histogram (every one is an exact integer
// sum), and ascending rows turn the root's gathers of bins and gradients into
// sequential reads.  One workgroup per chunk c, one thread per 32-row word.
__device__ inline void rows_from_mask(const uint32_t* __restrict__ sm, uint32_t N, uint32_t n_root,
                                      uint32_t c, uint32_t* __restrict__ rows) {
  const uint32_t W = (N + 31u) / 32u, nch = (W + kMaskChunkWords - 1u) / kMaskChunkWords;
  if (c >= nch) return;  // uniform: the grid covers the largest job
  const uint32_t tid = threadIdx.x, w = c * kMaskChunkWords + tid;
  uint32_t bits = w < W ? sm[nch + w] : 0u;
  const uint32_t cnt = __popc(bits);
  __shared__ uint32_t s_w[kMaskChunkWords / 64u];
  const uint32_t incl = wave_incl_scan_u32(cnt);
  if ((tid & 63u) == 63u) s_w[tid >> 6] = incl;
  __syncthreads();
  uint32_t pos = sm[c] + incl - cnt;
  for (uint32_t i = 0; i < (tid >> 6); ++i) pos += s_w[i];
  while (bits) {
    if (pos >= n_root) break;  // a mask inconsistent with n_root writes nothing past it
    rows[pos++] = w * 32u + static_cast<uint32_t>(__builtin_ctz(bits));
    bits &= bits - 1u;
  }
}

// Grid (chunk, job): every job with a row-sample mask.
__global__ void __launch_bounds__(kMaskChunkWords) k_rows_from_mask(const JobDev* __restrict__ jobs) {
  const JobDev& J = jobs[blockIdx.y];
  if (J.smask) rows_from_mask(J.smask, J.N, J.n_root, blockIdx.x, J.rows);
}

// Grid (chunk): one sample (the regular-tree trainer's root rows).
__global__ void __launch_bounds__(kMaskChunkWords)
    k_rows_from_mask_one(const uint32_t* __restrict__ sm, uint32_t N, uint32_t n_root,
                         uint32_t* __restrict__ rows) {
  rows_from_mask(sm, N, n_root, blockIdx.x, rows);
}

// Start of a tree, over the row grid: per-channel root totals (sum g, sum h)
// over 1024-position tiles of the root list rows[0][0 .. n_root), double,
// fixed order within the tile (k_tree_begin adds the tiles in a fixed
// order); in the atomic histogram mode the root's histograms zeroed (deeper
// levels are zeroed by the split scan of the level above).
__global__ void __launch_bounds__(256) k_root_partials(const JobDev* __restrict__ jobs,
                                                       const uint32_t* __restrict__ job_wg_off,
                                                       uint32_t njobs) {
  __shared__ uint32_t s_job;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_job = job_of_block(job_wg_off, njobs);
  __syncthreads();
  const JobDev& J = jobs[s_job];
  const uint32_t lw = blockIdx.x - J.grad_wg_off, K = J.K;
  if (J.fused_root || lw >= J.grad_wg_cnt) return;  // written by the gradient update / validation rows
  const bool live = J.st->tree + 1 < static_cast<int>(J.n_trees);
  if (live && !J.store_hist) {
    const uint32_t cells = K * J.F * 256u;  // the root's histograms
    const uint32_t per = (cells + J.grad_wg_cnt - 1) / J.grad_wg_cnt;
    const uint32_t c1 = min(cells, (lw + 1) * per);
    for (uint32_t c = lw * per + tid; c < c1; c += 256) J.hist[c] = HistCell{0, 0};
  }
  for (uint32_t c = 0; c < K; ++c) {
    double sg = 0.0, sh = 0.0;
    if (live) {
      const float2* gh = J.gh + static_cast<size_t>(c) * J.N;
      for (int k = 0; k < kGradTile / 256; ++k) {
        const uint32_t p = lw * kGradTile + k * 256 + tid;
        if (p >= J.n_root) break;
        const float2 v = gh[J.root_identity ? p : J.rows[p]];
        sg += v.x;
        sh += v.y;
      }
    }
    root_partial_store(J, lw, c, sg, sh);
  }
}

__global__ void __launch_bounds__(256) k_tree_begin(const JobDev* __restrict__ jobs, Lists L,
                                                    int packed) {
  tree_begin(jobs[blockIdx.y], blockIdx.y, L, packed);
}

// ---------------------------------------------------------------------------
// Level plan: what the partition of a level leaves for the level below.
__device__ inline bool part_active(const JobDev& J, uint32_t level) {
  return J.st->depth == static_cast<int32_t>(level) + 1 && J.st->tree < static_cast<int>(J.n_trees);
}

// Partition cursors of a level: two sets, by the level's parity, so that the
// plan of the next level can reset one set while the other is still read.
__device__ inline uint32_t* cursor_set(const JobDev& J, uint32_t level) {
  return J.cursor + ((level & 1u) << J.D);
}

// LDS of a level plan (about 12 KB).
struct PlanShared {
  uint32_t hw[1 << (kMaxDepth - 1)];   // per pair: smaller child's rows -> items -> first item
  uint32_t pl[1 << (kMaxDepth - 1)];   // per pair: rows per piece
  uint32_t pw[1 << kMaxDepth];         // per child: partition slots -> first slot
  uint32_t off[(1 << kMaxDepth) + 1];  // child offsets
  uint32_t tmp[16];
  uint32_t cnt[kChunkCands];
  uint32_t chunk, htot, hbase, pbase;
  int any;
  unsigned long long rows;
};

// Histogram work item e of a planned level: pair n = the last with first
// item <= e, its smaller child (ties: left), piece (e - first) / (tiles x K),
// then tile and channel.  nn: nodes of the partitioned level.
__device__ inline HistWork plan_item(const JobDev& J, const PlanShared& S, uint32_t nn, uint32_t e,
                                     uint32_t job) {
  const uint32_t TK = J.T * J.K;
  const uint32_t n = find_item(S.hw, nn, e);
  const uint32_t lc = 2 * n, rc = lc + 1;
  const uint32_t nl = S.off[lc + 1] - S.off[lc], nr = S.off[rc + 1] - S.off[rc];
  const uint32_t sm = nl <= nr ? lc : rc;
  const uint32_t le = e - S.hw[n], r = le % TK;
  const uint32_t end = S.off[sm + 1], pl = S.pl[n];
  const uint32_t start = S.off[sm] + (le / TK) * pl;
  HistWork w;
  w.job = job; w.node = sm; w.start = start; w.len = min(pl, end - start);
  w.tile = r / J.K; w.cls = r % J.K; w.last = start + pl >= end ? 1u : 0u; w.pad = 0;
  return w;
}

// Plan of the level below `level` for job J, by one workgroup of NT threads,
// from level `level`'s partition (its offsets and left cursors): the
// children's counts and offsets (children of an empty node get zero
// totals), the next level's stop test (H summed over channels), its rows
// per histogram work item, and its work items (block scans over pairs).
// Returns whether the next level is histogrammed; then S.hw holds each
// pair's first item, S.pl its rows per piece, S.off the child offsets,
// S.chunk and S.htot.  emit 0: the plan only (a histogram workgroup's own
// item); 1: also store what later launches read (counts, offsets, stop flag,
// chunk, first items, the next partition's cursors and work list); 2: also
// the histogram work list.
template <uint32_t NT>
__device__ bool level_plan(const JobDev& J, const Lists& L, uint32_t job, uint32_t level,
                           PlanShared& S, int emit) {
  constexpr uint32_t PER = (512u + NT - 1) / NT;  // pairs per thread (nn <= 512)
  const uint32_t tid = threadIdx.x;
  const uint32_t nn = 1u << level;
  const uint32_t stride = (1u << J.D) + 1u;
  const uint32_t* offd = J.off + static_cast<size_t>(level) * stride;
  uint32_t* offn = J.off + static_cast<size_t>(level + 1) * stride;
  const uint32_t* curd = cursor_set(J, level);
  // the nodes' segments and left cursors, in flight while the state is checked
  uint32_t p_lo[PER], p_hi[PER], p_cur[PER];
  float p_hl[PER], p_hr[PER];  // single channel: the children's H
#pragma unroll
  for (uint32_t i = 0; i < PER; ++i) {
    const uint32_t n = tid + NT * i;
    p_lo[i] = p_hi[i] = p_cur[i] = 0;
    p_hl[i] = p_hr[i] = 0.f;
    if (n < nn) {
      p_lo[i] = offd[n];
      p_hi[i] = offd[n + 1];
      p_cur[i] = curd[2 * n];
      if (J.K == 1) {
        p_hl[i] = J.nh[heap_index(level + 1, 2 * n)];
        p_hr[i] = J.nh[heap_index(level + 1, 2 * n) + 1];
      }
    }
  }
  const uint32_t p_end = tid == 0 ? offd[nn] : 0u;
  if (!part_active(J, level)) return false;
  if (tid == 0) { S.any = 0; S.rows = 0ull; }
  if (tid < kChunkCands) S.cnt[tid] = 0;
  __syncthreads();
  const float mcw2 = J.min_child_weight * 2.0f;
#pragma unroll
  for (uint32_t i = 0; i < PER; ++i) {
    const uint32_t n = tid + NT * i;
    if (n >= nn) break;
    const uint32_t len = p_hi[i] - p_lo[i];
    const uint32_t acc = p_cur[i] - p_lo[i];  // rows sent left
    const uint32_t cl = heap_index(level + 1, 2 * n);
    if (emit) {
      J.nc[cl] = acc;
      J.nc[cl + 1] = len - acc;
      if (len == 0) {  // an empty node's children (no partition item computed them)
        for (uint32_t c = 0; c < J.K; ++c) {
          J.ng[cl * J.K + c] = 0.f;
          J.nh[cl * J.K + c] = 0.f;
          J.ng[(cl + 1) * J.K + c] = 0.f;
          J.nh[(cl + 1) * J.K + c] = 0.f;
        }
      }
      offn[2 * n] = p_lo[i];
      offn[2 * n + 1] = p_lo[i] + acc;
    }
    S.off[2 * n] = p_lo[i];
    S.off[2 * n + 1] = p_lo[i] + acc;
    // stop test of the next level (symmetric_tree.cpp:115-124; H sums the
    // channels in class order)
    float hl = 0.f, hr = 0.f;
    if (len == 0) {
      // an empty node's children: zero totals
    } else if (J.K == 1) {
      hl = p_hl[i];
      hr = p_hr[i];
    } else {
      for (uint32_t c = 0; c < J.K; ++c) {
        hl = __fadd_rn(hl, J.nh[cl * J.K + c]);
        hr = __fadd_rn(hr, J.nh[(cl + 1) * J.K + c]);
      }
    }
    if ((acc >= 2u * J.min_samples_leaf && hl >= mcw2) ||
        (len - acc >= 2u * J.min_samples_leaf && hr >= mcw2))
      S.any = 1;
    // next level: histogram of the smaller child of the pair (ties: left),
    // partition slots of both children
    const uint32_t sm_len = min(acc, len - acc);
    S.hw[n] = sm_len;
    S.pw[2 * n] = part_slots(acc);
    S.pw[2 * n + 1] = part_slots(len - acc);
    if (sm_len) atomicAdd(&S.rows, static_cast<unsigned long long>(sm_len));
  }
  if (tid == 0) {
    if (emit) offn[2 * nn] = p_end;
    S.off[2 * nn] = p_end;
  }
  __syncthreads();
  if (level + 1 >= J.D || !S.any) {
    if (emit && tid == 0) J.st->active = 0;
    return false;
  }
  // rows per work item of the next level: the smallest candidate whose
  // pieces (ceil(rows / chunk) per histogrammed child) fit J.deep_items
  // items, so that each CU gets one item of about the same size; or the
  // fixed chunk
  const uint32_t TK = J.T * J.K;
  uint32_t chunk = J.hist_chunk;
  if (J.deep_items) {
    // thread: candidate tid % 8 over pairs tid / 8, tid / 8 + NT / 8, ...
    const uint32_t budget = max(1u, J.deep_items / TK);  // pieces
    const uint32_t base = root_chunk(S.rows, 1u, budget);
    const uint32_t k = tid % kChunkCands, c = chunk_cand(base, k);
    uint32_t cnt = 0;
    for (uint32_t n = tid / kChunkCands; n < nn; n += NT / kChunkCands) cnt += (S.hw[n] + c - 1) / c;
    if (cnt) atomicAdd(&S.cnt[k], cnt);
    __syncthreads();
    chunk = chunk_cand(base, kChunkCands - 1);
    for (int q = kChunkCands - 1; q >= 0; --q)
      if (S.cnt[q] <= budget) chunk = chunk_cand(base, q);
  }
  if (emit && tid == 0) J.lvl_chunk[level + 1] = chunk;
  for (uint32_t n = tid; n < nn; n += NT) {
    S.pl[n] = piece_rows(S.hw[n], chunk);
    S.hw[n] = ((S.hw[n] + chunk - 1) / chunk) * TK;
  }
  __syncthreads();
  const uint32_t htot = block_scan_excl(S.hw, nn, S.tmp);
  const uint32_t ptot = emit ? block_scan_excl(S.pw, 2 * nn, S.tmp) : 0u;  // emit is uniform
  if (tid == 0) {
    S.chunk = chunk;
    S.htot = htot;
    S.hbase = 0;
    S.pbase = 0;
    if (emit) {
      if (emit >= 2) S.hbase = list_reserve(L, L.hist_count[level + 1], htot);
      S.pbase = list_reserve(L, L.part_count[level + 1], ptot);
      if (J.hist_rows_total) atomicAdd(J.hist_rows_total, S.rows);
    }
  }
  __syncthreads();
  if (!emit) return true;
  uint32_t* curn = cursor_set(J, level + 1);  // the other set: every cursor of this level was read
  for (uint32_t n = tid; n < nn; n += NT) {
    // first item of each pair's smaller child
    const uint32_t lc = 2 * n, rc = lc + 1;
    const uint32_t sm = S.off[lc + 1] - S.off[lc] <= S.off[rc + 1] - S.off[rc] ? lc : rc;
    J.hfirst[heap_index(level + 1, sm)] = S.hbase + S.hw[n];
    // partition cursors of the next level
    curn[2 * lc] = S.off[lc];
    curn[2 * lc + 1] = 0;
    curn[2 * rc] = S.off[rc];
    curn[2 * rc + 1] = 0;
  }
  if (emit >= 2) {
    HistWork* hout = L.hist[level + 1] + S.hbase;
    for (uint32_t e = tid; e < htot; e += NT) hout[e] = plan_item(J, S, nn, e, job);
  }
  PartWork* pout = L.part[level + 1] + S.pbase;
  for (uint32_t e = tid; e < ptot; e += NT) {
    const uint32_t c = find_item(S.pw, 2 * nn, e);
    const uint32_t start = S.off[c] + (e - S.pw[c]) * kPartChunk;
    put_part_work(pout + e, job, c, start,
                  min(static_cast<uint32_t>(kPartChunk), S.off[c + 1] - start), e);
  }
  return true;
}

// ---------------------------------------------------------------------------
// A9: per-level histogram.  One workgroup = one work item (row chunk of one
// node, one gradient channel) x one 16-feature tile.  Each lane loads the tile's 16 bins of a row
// with one 16-byte load from the row-major bins and the row's (g, h), turns
// them into fixed point once, and adds them into the workgroup's LDS
// histogram with 64-bit integer atomics.  The LDS cells are bin-major,
// feature-minor (cell b * 16 + f), so a cell's bank is set by its feature
// alone; lane l starts at feature l % 16, so every 16 lanes of one atomic
// instruction hit 16 different banks whatever their bins (the feature-major
// layout put the bank at (f + b) mod 16 and measured 63% of LDS cycles in
// bank-conflict stalls).
//
// MODE 1 (MSE, h == 2 for every row): one atomic per value on a cell holding
// g * 2^16 + 1 -- a signed 48-bit g sum above a 16-bit row count (a work
// item has fewer than 2^16 rows, so the count never carries); h = 2 * count.
// MODE 0: one atomic on the g plane and one on the h plane.  MODE 2 (the
// regular trees' cells, which need exact row counts): the g plane, and an h
// plane whose cells hold h * 2^16 + 1 (h >= 0 summed over one item stays
// inside 47 bits under the packed scale); the flush stores that packed word
// as the cell's h, which the reader unpacks (store mode only).
//
// Training (`partials` given, J.store_hist): a node's last chunk stores its
// cells into the node's HBM histogram, every other chunk into its slot of
// `partials` (slot = list index), all with plain stores; the split scan adds
// a node's slots to its histogram.  Nothing has to be zeroed beforehand and
// no global atomics are issued: int64 atomics from every workgroup (2 x 4096
// per item) are served by the memory-side atomic units at ~1.5e11 per second
// for the whole chip, several times the cost of the stores.  Otherwise the
// non-empty cells are added into the zeroed node histogram with int64
// atomics.
// SINGLE: one job, so its table entry is read without waiting for the item.
// PLAN (one job, below the root): no work list; every workgroup plans the
// level from the partition above (level_plan, in the LDS of its cells) and
// takes its own item, and workgroup 0 stores the plan's outputs for the
// later launches (the plan costs less than the launch it replaces).
template <int U, int MODE, bool SINGLE, bool PLAN>
__device__ __forceinline__ void hist_item(const JobDev* __restrict__ jobs,
                                          const HistWork* __restrict__ work,
                                          const uint32_t* __restrict__ count, uint32_t level,
                                          HistCell* __restrict__ partials, const Lists& L) {
  constexpr int kCells = kTileF * 256;
  __shared__ unsigned long long s_a[kCells];
  constexpr bool PACKED = MODE == 1;
  __shared__ unsigned long long s_b[PACKED ? 1 : kCells];
  static_assert(!PLAN || SINGLE, "the in-kernel plan serves one job");
  static_assert(sizeof(PlanShared) <= sizeof(s_a), "the plan fits the cells' LDS");
  uint32_t li;
  HistWork w;
  if (PLAN && level > 0) {
    PlanShared& S = *reinterpret_cast<PlanShared*>(s_a);
    if (!level_plan<kHistThreads>(jobs[0], L, 0u, level - 1, S, blockIdx.x == 0 ? 1 : 0)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && S.htot > gridDim.x)  // the grid's bound broke
      atomicOr(&jobs[0].st->fault, 4u);
    if (!xcd_rank(blockIdx.x, S.htot, &li)) return;
    w = plan_item(jobs[0], S, 1u << (level - 1), li, 0u);
    __syncthreads();  // the plan's LDS becomes the cells
  } else {
    if (!xcd_rank(blockIdx.x, *count, &li)) return;
    w = work[li];
  }
  const JobDev& J = jobs[SINGLE ? 0u : w.job];  // items exist only while the job's tree grows
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kCells; i += kHistThreads) {
    s_a[i] = 0ull;
    if (!PACKED) s_b[i] = 0ull;
  }
  const double gsc = J.st->gscale, hsc = J.st->hscale;
  __syncthreads();

  const uint8_t* __restrict__ brm = J.bins_rm + w.tile * kTileF;
  const uint32_t* __restrict__ rows = J.rows + static_cast<size_t>(level) * J.N + w.start;
  const bool direct = level == 0 && J.root_identity;  // row = position, no list read
  // several channels below the root: the level's (g, h) in the order of its row list
  const bool gpos = J.ghl != nullptr && level > 0;
  const float2* __restrict__ gh =
      gpos ? J.ghl + (static_cast<size_t>(level & 1u) * J.K + w.cls) * J.N + w.start
           : J.gh + static_cast<size_t>(w.cls) * J.N;
  const uint32_t Fp = J.Fp;
  const uint32_t s = tid & 15u;  // this lane's starting feature
  const uint32_t q = s >> 2, r = s & 3u;
  const uint32_t f0 = w.tile * kTileF;
  const uint32_t nf = min(static_cast<uint32_t>(kTileF), J.F - f0);  // real features of the tile

  // Software pipeline over batches of U x 1024 rows: while batch i is added
  // into LDS, the bins and (g, h) of batch i+1 and the row ids of batch i+2
  // are in flight.
  constexpr uint32_t STEP = kHistThreads * U;
  const uint32_t len = w.len;
  auto row_at = [&](uint32_t j) { return j < len ? (direct ? w.start + j : rows[j]) : 0u; };
  uint32_t r1[U], r2[U];
  uint4 v[U];
  float2 g[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t j = u * kHistThreads + tid;
    const uint32_t row = row_at(j);
    v[u] = j < len ? *reinterpret_cast<const uint4*>(brm + static_cast<size_t>(row) * Fp)
                   : make_uint4(0, 0, 0, 0);
    g[u] = j < len ? gh[gpos ? j : row] : make_float2(0.f, 0.f);
    r1[u] = row_at(STEP + j);
  }
  for (uint32_t j0 = 0; j0 < len; j0 += STEP) {
    uint4 vn[U];
    float2 gn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = j0 + STEP + u * kHistThreads + tid;
      vn[u] = j < len ? *reinterpret_cast<const uint4*>(brm + static_cast<size_t>(r1[u]) * Fp)
                      : make_uint4(0, 0, 0, 0);
      gn[u] = j < len ? gh[gpos ? j : r1[u]] : make_float2(0.f, 0.f);
      r2[u] = row_at(j + STEP);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (j0 + u * kHistThreads + tid >= len) continue;
      const long long gq = __double2ll_rn(static_cast<double>(g[u].x) * gsc);
      const unsigned long long va = PACKED ? (static_cast<unsigned long long>(gq) << 16) + 1ull
                                           : static_cast<unsigned long long>(gq);
      const unsigned long long hq =
          static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(g[u].y) * hsc));
      const unsigned long long vb = PACKED ? 0ull : MODE == 2 ? (hq << 16) + 1ull : hq;
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
        if (f < nf) {  // padding features of the last tile cost no atomics
          atomicAdd(&s_a[b * 16 + f], va);
          if (!PACKED) atomicAdd(&s_b[b * 16 + f], vb);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      v[u] = vn[u];
      g[u] = gn[u];
      r1[u] = r2[u];
    }
  }
  __syncthreads();
  const bool store = partials != nullptr;
  HistCell* out =
      store && !w.last
          ? partials + static_cast<size_t>(li) * kCells
          : J.hist + ((static_cast<size_t>(heap_index(level, w.node)) * J.K + w.cls) * J.F + f0) * 256u;
  const long long h2 = PACKED ? __double2ll_rn(2.0 * hsc) : 0;
  for (uint32_t i = tid; i < static_cast<uint32_t>(kCells); i += kHistThreads) {
    const uint32_t f = i & 15u, b = i >> 4;  // LDS order: conflict-free reads
    if (f >= nf) continue;
    const unsigned long long a = s_a[i];
    long long gs, hs;
    if (PACKED) {
      gs = static_cast<long long>(a) >> 16;
      hs = static_cast<long long>(a & 0xFFFFull) * h2;
    } else {
      gs = static_cast<long long>(a);
      hs = static_cast<long long>(s_b[i]);
    }
    if (store) {
      out[f * 256u + b] = HistCell{gs, hs};
    } else if (gs != 0 || hs != 0) {
      HistCell* o = out + f * 256u + b;
      atomicAdd(reinterpret_cast<unsigned long long*>(&o->g), static_cast<unsigned long long>(gs));
      atomicAdd(reinterpret_cast<unsigned long long*>(&o->h), static_cast<unsigned long long>(hs));
    }
  }
}

// Listed items (the root, batches; several per CU).  Two 1024-thread
// workgroups share a CU only while a wave holds at most about 78 SGPRs (64
// VGPRs, 64 KB LDS each; measured on gfx950 -- above that one runs at a time
// although the compiler's occupancy estimate still reads eight waves per
// SIMD), so the registers are capped here.
template <int U, int MODE, bool SINGLE>
__global__ void __launch_bounds__(kHistThreads) __attribute__((amdgpu_num_sgpr(76), amdgpu_waves_per_eu(8)))
    k_hist(const JobDev* __restrict__ jobs, const HistWork* __restrict__ work,
           const uint32_t* __restrict__ count, uint32_t level, HistCell* __restrict__ partials,
           Lists L) {
  hist_item<U, MODE, SINGLE, false>(jobs, work, count, level, partials, L);
}

// One job below the root: the level planned in the kernel, about one item per
// CU (no cap: a capped variant spills in its main loop).
template <int U, int MODE>
__global__ void __launch_bounds__(kHistThreads)
    k_hist_plan(const JobDev* __restrict__ jobs, const HistWork* __restrict__ work,
                const uint32_t* __restrict__ count, uint32_t level, HistCell* __restrict__ partials,
                Lists L) {
  hist_item<U, MODE, true, true>(jobs, work, count, level, partials, L);
}

// ---------------------------------------------------------------------------
// A10: level split scan.  Workgroup = (feature, job), thread = candidate bin.
// Per node the left sums of every candidate are staged in LDS (up to 32
// nodes at a time), then each candidate's gain terms are added over nodes in
// node order (symmetric_tree.cpp:153-193, :214-262, with the deeper levels'
// total += 0.5 * (...) as the fused multiply-add its compiler emits).
// FIXED (training): the left sums are the exact fixed-point prefixes, each
// rounded once to float.  Otherwise (the tc_split_scan entry point, float
// bins): the reference's sequential float prefix, so for a given float
// histogram every left sum, gain and total is the reference's float value.
//
// FIXED, per pass: one task per histogrammed node and channel (the root's
// channels at level 0; the smaller child of each pair below).  The task's
// cells start from the node's HBM histogram (its last chunk's cells, or in
// the atomic mode the whole sum; zero for an empty node) and every wave adds
// partial-histogram slots of the tasks into LDS (int64 atomics: exact in any
// order), two slots per wave in flight.  Then one wave per task stores the
// node's histogram, forms the larger sibling as parent - smaller (A13:
// Histogram::subtract_from, histogram.cpp:39-42 via symmetric_tree.cpp:346)
// and writes both prefixes; in the atomic mode it also zeroes the children's
// next-level histograms.
// TH threads: FIXED 1024 (one workgroup per CU: a lone job's few features
// finish sooner) or 512 (two per CU: a batch's thousands of workgroups
// overlap their memory waits); 256 otherwise.
template <bool FIXED, uint32_t TH>
__global__ void __launch_bounds__(TH) k_split_scan(const JobDev* __restrict__ jobs,
                                                   const float2* __restrict__ fhist,
                                                   uint32_t level,
                                                   const HistCell* __restrict__ partials) {
  constexpr uint32_t NW = TH / 64u;              // waves
  constexpr uint32_t Q = TH / 256u;              // threads per candidate bin
  // units staged per pass: two per wave (one task per wave: the smaller
  // child of a pair, and its sibling), 32 without tasks
  constexpr uint32_t NP = FIXED ? 2u * NW : static_cast<uint32_t>(kScanNodesPerPass);
  constexpr uint32_t kTasks = NP / 2;  // histogrammed units per pass
  static_assert(!FIXED || kTasks <= NW, "one wave per task");
  static_assert(FIXED ? (TH == 1024u || TH == 512u) : TH == 256u, "thread counts");
  // s_pref: the pass's left sums per unit and candidate.  FIXED: s_accg/h (the
  // tasks' summed cells) shares its 64 KB; each task's cells are in
  // registers before the first prefix is written.
  __shared__ __align__(16) unsigned char s_raw[NP * 256 * sizeof(float2)];
  float2(*s_pref)[256] = reinterpret_cast<float2(*)[256]>(s_raw);
  // the tasks' summed cells, one plane for g and one for h ([task][bin] each):
  // lanes add consecutive bins, so a wave's 64-bit atomics fall on 32
  // different bank pairs (interleaved (g, h) cells at four bins per lane put
  // them on 4 and stalled 70% of the LDS cycles in bank conflicts)
  unsigned long long* s_accg = reinterpret_cast<unsigned long long*>(s_raw);
  unsigned long long* s_acch = s_accg + kTasks * 256;
  static_assert(kTasks * 256 * 2 * sizeof(unsigned long long) <= sizeof(s_raw), "task cells fit");
  __shared__ unsigned long long s_red[4];
  __shared__ float s_chain[256];
  __shared__ uint8_t s_chain_any[256];
  __shared__ uint32_t s_sm[kTasks], s_first[kTasks], s_cpre[kTasks + 1], s_empty[kTasks];
  __shared__ float s_G[NP], s_H[NP], s_P[NP];  // unit totals; s_P: the unit's parent term
  __shared__ uint32_t s_N[NP];
  const uint32_t job = blockIdx.y, f = blockIdx.x, tid = threadIdx.x;
  const JobDev& J = jobs[job];
  if (f >= J.F) return;
  // FIXED waits for the tree state only after its first loads are issued
  const int32_t active = J.st->active;
  if (!FIXED && !active) return;
  const uint32_t K = J.K;
  const uint32_t nu = K << level;  // units (node, channel), node-major
  const uint32_t mb = J.max_bins & 0xFFu;
  const float lambda = J.lambda_l2;
  const double gi = FIXED ? J.st->ginv : 1.0, hi = FIXED ? J.st->hinv : 1.0;
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  const uint32_t bin = tid & 255u, part = tid >> 8;  // candidate, unit subset
  const size_t F = J.F;
  // units per pass: whole pairs of nodes (all channels) below the root
  const uint32_t P = !FIXED ? NP : level == 0 ? K : (NP / (2 * K)) * (2 * K);
  if (FIXED && (P == 0 || (level == 0 && K > kTasks))) {  // more channels than a pass holds
    if (tid == 0) atomicOr(&J.st->fault, 2u);
    return;
  }
  const bool store = FIXED && partials != nullptr && J.store_hist;
  float total = 0.f;
  bool any = false;
  for (uint32_t c0 = 0; c0 < nu; c0 += P) {
    const uint32_t cs = min(P, nu - c0);
    if (FIXED) {
      auto cells = [&](uint32_t lvl, uint32_t nd, uint32_t ch) {
        return J.hist + ((static_cast<size_t>(heap_index(lvl, nd)) * K + ch) * F + f) * 256u;
      };
      const uint32_t nt = level == 0 ? cs : cs / 2;  // tasks of this pass
      // task t: channel t % K of (level 0) the root / (deeper) pair t / K
      auto task_left = [&](uint32_t t) { return level == 0 ? 0u : c0 / K + 2u * (t / K); };
      if (tid < cs) {  // the pass's unit totals, for the gain terms
        const uint32_t hn = heap_index(level, (c0 + tid) / K);
        const float G = J.ng[hn * K + (c0 + tid) % K], H = J.nh[hn * K + (c0 + tid) % K];
        s_G[tid] = G;
        s_H[tid] = H;
        // the parent term G^2 / (H + lambda), the same for every candidate
        s_P[tid] = H > 1e-10f ? __fdiv_rn(__fmul_rn(G, G), __fadd_rn(H, lambda)) : 0.f;
        s_N[tid] = J.nc[hn];
      }
      if (tid < nt) {
        const uint32_t nl = task_left(tid), hl = heap_index(level, nl);
        // both siblings' counts and first items in one round trip
        const uint32_t cl = J.nc[hl], fl = J.hfirst[hl];
        const uint32_t cr = level > 0 ? J.nc[hl + 1] : 0u, fr = level > 0 ? J.hfirst[hl + 1] : 0u;
        const bool right = level > 0 && cl > cr;
        const uint32_t sm = right ? nl + 1 : nl;
        const uint32_t rows = right ? cr : cl, chunk = J.lvl_chunk[level];
        // sources: the node's HBM histogram, then (store mode) one partial
        // slot per chunk but the last; none for an empty node
        const bool empty = store && rows == 0;
        s_sm[tid] = sm;
        s_first[tid] = right ? fr : fl;
        s_cpre[tid + 1] = empty ? 0u : 1u + (store ? (rows - 1) / chunk : 0u);
        s_empty[tid] = empty ? 1u : 0u;
      }
      for (uint32_t i = tid; i < nt * 256u; i += TH) {
        s_accg[i] = 0ull;
        s_acch[i] = 0ull;
      }
      __syncthreads();
      if (!active) return;  // the job's tree is done (uniform)
      if (tid == 0) {
        s_cpre[0] = 0;
        for (uint32_t t = 0; t < nt; ++t) s_cpre[t + 1] += s_cpre[t];
      }
      // the wave that finishes task `wave` (nt <= NW) loads its parent now
      HistCell par[4];
      if (level > 0 && wave < nt) {
        const HistCell* pp = cells(level - 1, task_left(wave) >> 1, wave % K) + 4u * lane;
#pragma unroll
        for (int i = 0; i < 4; ++i) par[i] = pp[i];
      }
      __syncthreads();
      const uint32_t ne = s_cpre[nt];
      const uint32_t tile = f / kTileF, fi = f % kTileF, TK = J.T * K;
      auto source = [&](uint32_t e, uint32_t* t_out) {  // lane's first cell of source e
        uint32_t t = 0;
        while (s_cpre[t + 1] <= e) ++t;
        *t_out = t;
        const uint32_t j = e - s_cpre[t];
        if (j == 0) return static_cast<const HistCell*>(cells(level, s_sm[t], t % K) + lane);
        const uint32_t item = s_first[t] + (j - 1) * TK + tile * K + t % K;
        return partials + (static_cast<size_t>(item) * kTileF + fi) * 256u + lane;
      };
      constexpr uint32_t kInFlight = 2;  // sources loaded per wave before adding
      for (uint32_t e0 = wave; e0 < ne; e0 += kInFlight * NW) {
        HistCell v[kInFlight][4];
        uint32_t tt[kInFlight];
#pragma unroll
        for (uint32_t d = 0; d < kInFlight; ++d) {
          if (e0 + d * NW < ne) {
            const HistCell* p = source(e0 + d * NW, &tt[d]);
#pragma unroll
            for (int i = 0; i < 4; ++i) v[d][i] = p[64 * i];  // bins lane + 64 i
          }
        }
#pragma unroll
        for (uint32_t d = 0; d < kInFlight; ++d) {
          if (e0 + d * NW < ne) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const uint32_t c = tt[d] * 256u + lane + 64u * i;
              atomicAdd(&s_accg[c], static_cast<unsigned long long>(v[d][i].g));
              atomicAdd(&s_acch[c], static_cast<unsigned long long>(v[d][i].h));
            }
          }
        }
      }
      __syncthreads();
      // exact prefix of one unit's cells (lane l holds bins 4l .. 4l+3),
      // rounded once to float, into s_pref[k]
      auto prefix = [&](const HistCell* cell, uint32_t k) {
        long long pg[4], ph[4], ag = 0, ah = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          ag += cell[i].g;
          ah += cell[i].h;
          pg[i] = ag;
          ph[i] = ah;
        }
        // inclusive scan of the lane totals, then this lane's start
        long long sg = wave_incl_scan_i64(ag) - ag, sh = wave_incl_scan_i64(ah) - ah;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          s_pref[k][4 * lane + i] =
              make_float2(static_cast<float>(static_cast<double>(sg + pg[i]) * gi),
                          static_cast<float>(static_cast<double>(sh + ph[i]) * hi));
      };
      const bool has_task = wave < nt;  // nt <= NW: at most one task per wave
      const uint32_t t = wave;
      HistCell cs_[4];
      if (has_task) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          cs_[i] = HistCell{static_cast<long long>(s_accg[t * 256u + 4 * lane + i]),
                            static_cast<long long>(s_acch[t * 256u + 4 * lane + i])};
      }
      __syncthreads();  // the cells' space becomes s_pref
      if (has_task) {
        const uint32_t ch = t % K, nl = task_left(t), sm = s_sm[t];
        HistCell* hs = cells(level, sm, ch);
        if (s_cpre[t + 1] - s_cpre[t] > 1u || s_empty[t]) {  // a parent at the next level
#pragma unroll
          for (int i = 0; i < 4; ++i) hs[4 * lane + i] = cs_[i];
        }
        prefix(cs_, sm * K + ch - c0);
        if (level > 0) {
          const uint32_t lg = sm == nl ? nl + 1 : nl;
          HistCell* hl = cells(level, lg, ch);
          HistCell cl[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const HistCell a = par[i];
            cl[i] = HistCell{a.g - cs_[i].g, a.h - cs_[i].h};
            hl[4 * lane + i] = cl[i];
          }
          prefix(cl, lg * K + ch - c0);
        }
        if (!store && level + 1 < J.D) {  // the next level's histograms of this task's nodes
          const uint32_t n1 = level == 0 ? nl : nl + 1;  // the root, or the pair
          for (uint32_t nd = nl; nd <= n1; ++nd) {
            HistCell* z0 = cells(level + 1, 2 * nd, ch);
            HistCell* z1 = cells(level + 1, 2 * nd + 1, ch);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              z0[4 * lane + i] = HistCell{0, 0};
              z1[4 * lane + i] = HistCell{0, 0};
            }
          }
        }
      }
      __syncthreads();
    } else {
      // the reference's sequential float prefix over the given float bins
      for (uint32_t k = 0; k < cs; ++k)  // K == 1 here
        s_pref[k][tid] = fhist[(static_cast<size_t>(heap_index(level, c0 + k)) * F + f) * 256u + tid];
      __syncthreads();
      if (tid < cs) {
        float lg = 0.f, lh = 0.f;
        for (uint32_t b = 0; b + 1 < mb; ++b) {
          const float2 v = s_pref[tid][b];
          lg = __fadd_rn(lg, v.x);
          lh = __fadd_rn(lh, v.y);
          s_pref[tid][b] = make_float2(lg, lh);
        }
      }
      __syncthreads();
    }
    // gain of candidate `bin`: the unit terms added in (node, channel) order,
    // one sequence per candidate as in the reference.  FIXED: all threads
    // form the terms (two divisions each; the unit's parent term is staged
    // once), then the candidate's sequence runs through the parts in order.
    if (FIXED) {
      // part p forms the terms of units [8p, 8p + 8) in registers, then the
      // parts extend the candidate's sequence in order
      constexpr uint32_t kPer = NP / Q;
      float term[kPer];
      uint32_t use = 0;
#pragma unroll
      for (uint32_t u = 0; u < kPer; ++u) {
        const uint32_t k = part * kPer + u;
        term[u] = 0.f;
        if (k < cs && bin + 1 < mb && !(level > 0 && s_N[k] == 0)) {
          const float G = s_G[k], H = s_H[k];
          const float2 Lf = s_pref[k][bin];
          const float rg = __fsub_rn(G, Lf.x), rh = __fsub_rn(H, Lf.y);
          if (!(Lf.y < 1e-10f || rh < 1e-10f)) {
            const float lgn = __fdiv_rn(__fmul_rn(Lf.x, Lf.x), __fadd_rn(Lf.y, lambda));
            const float rgn = __fdiv_rn(__fmul_rn(rg, rg), __fadd_rn(rh, lambda));
            term[u] = __fsub_rn(__fadd_rn(lgn, rgn), s_P[k]);
            use |= 1u << u;
          }
        }
      }
      for (uint32_t p = 0; p < Q; ++p) {
        if (part == p && p * kPer < cs) {
          float tt = p == 0 ? total : s_chain[bin];
          bool aa = p == 0 ? any : s_chain_any[bin] != 0;
#pragma unroll
          for (uint32_t u = 0; u < kPer; ++u) {
            if ((use >> u) & 1u) {
              tt = __fmaf_rn(0.5f, term[u], tt);
              aa = true;
            }
          }
          s_chain[bin] = tt;
          s_chain_any[bin] = aa ? 1 : 0;
        }
        __syncthreads();
      }
      if (part == 0) {
        total = s_chain[bin];
        any = s_chain_any[bin] != 0;
      }
    } else if (bin + 1 < mb) {
      for (uint32_t k = part; k < cs; k += Q) {
        const uint32_t hn = heap_index(level, (c0 + k) / K);
        if (level > 0 && J.nc[hn] == 0) continue;
        const uint32_t hu = hn * K + (c0 + k) % K;
        const float G = J.ng[hu], H = J.nh[hu];
        const float2 Lf = s_pref[k][bin];
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
  if (part == 0) {
    unsigned long long key = kEmptyKey;
    if (bin + 1 < mb && (level > 0 || any) && !isnan(total)) {
      key = split_key(total == 0.f ? 0.f : total, f, bin);
    }
    key = wave_max_u64(key);
    if (lane == 0) s_red[wave] = key;
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long m = s_red[0];
    for (int i = 1; i < 4; ++i) m = s_red[i] > m ? s_red[i] : m;
    if (J.best_keys)  // every feature's key, read back by the level's decode
      J.best_keys[static_cast<size_t>(level) * F + f] = m;
    else if (m != kEmptyKey)
      atomicMax(&J.st->best[level], m);
  }
}

// A11: decode the level's argmax, threshold lookup, min_split_gain stop
// (symmetric_tree.cpp:264-282), recorded by one thread.
__device__ inline bool decode_split(const JobDev& J, uint32_t level, bool record, uint32_t* f_out,
                                    uint32_t* thr_out) {
  TreeState* st = J.st;
  unsigned long long key;
  if (J.best_keys) {  // the max over the features' keys, each wave for itself
    const unsigned long long* keys = J.best_keys + static_cast<size_t>(level) * J.F;
    unsigned long long m = kEmptyKey;
    for (uint32_t f = threadIdx.x & 63u; f < J.F; f += 64) m = keys[f] > m ? keys[f] : m;
    key = wave_max_u64(m);
  } else {
    key = st->best[level];
  }
  const float gain = float_unorder(static_cast<uint32_t>(key >> 32));
  const bool accept = key != kEmptyKey && gain > -1e29f && !(gain < J.min_split_gain);
  const uint32_t fb = 0xFFFFFFFFu - static_cast<uint32_t>(key & 0xFFFFFFFFu);
  const uint32_t f = fb >> 8, thr = fb & 255u;
  if (record) {
    if (!accept) {
      st->active = 0;
    } else {
      float fthr = 0.f;
      const uint32_t ne = J.n_edges[f];
      if (thr < ne) fthr = J.edges[static_cast<size_t>(f) * 255 + thr];
      else if (ne > 0) fthr = J.edges[static_cast<size_t>(f) * 255 + ne - 1];
      const size_t o = static_cast<size_t>(st->tree) * J.D + level;
      J.m_feat[o] = static_cast<uint16_t>(f);
      J.m_bin[o] = static_cast<uint8_t>(thr);
      J.m_fthr[o] = fthr;
      J.m_gain[o] = gain;
      st->depth = static_cast<int32_t>(level) + 1;
    }
  }
  *f_out = f;
  *thr_out = thr;
  return accept;
}

// Child totals of node n, channel c under split (f, thr), by one wave: left
// = bins <= thr plus the NaN slot 255, right = the other value bins, exact;
// lane 0 stores them (float).
__device__ inline void child_totals(const JobDev& J, uint32_t level, uint32_t n, uint32_t c,
                                    uint32_t f, uint32_t thr) {
  const uint32_t lane = threadIdx.x & 63u, K = J.K;
  const HistCell* hp =
      J.hist + ((static_cast<size_t>(heap_index(level, n)) * K + c) * J.F + f) * 256u;
  long long v[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t b = 4 * lane + i;
    const HistCell c = hp[b];
    const int side = (b <= thr || b == 255u) ? 0 : 2;
    v[side] += c.g;
    v[side + 1] += c.h;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = wave_sum_i64(v[k]);
  if (lane == 0) {
    const double gi = J.st->ginv, hi = J.st->hinv;
    const uint32_t cl = heap_index(level + 1, 2 * n);
    J.ng[cl * K + c] = static_cast<float>(static_cast<double>(v[0]) * gi);
    J.nh[cl * K + c] = static_cast<float>(static_cast<double>(v[1]) * hi);
    J.ng[(cl + 1) * K + c] = static_cast<float>(static_cast<double>(v[2]) * gi);
    J.nh[(cl + 1) * K + c] = static_cast<float>(static_cast<double>(v[3]) * hi);
  }
}

// Finalize of the last level (the levels above are finalized by their
// partition): workgroup (group of 4 units, job), one wave per unit (node,
// channel); workgroup 0 records the split.
__global__ void __launch_bounds__(256) k_finalize(const JobDev* __restrict__ jobs, uint32_t level) {
  const JobDev& J = jobs[blockIdx.y];
  if (!J.st->active) return;
  uint32_t f, thr;
  const bool accept = decode_split(J, level, blockIdx.x == 0 && threadIdx.x == 0, &f, &thr);
  const uint32_t u = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (accept && u < (J.K << level)) child_totals(J, level, u / J.K, u % J.K, f, thr);
}

// ---------------------------------------------------------------------------
// A12: partition, one pass.  Every workgroup decodes the level's split
// itself; the job's first work item records it (the level's finalize).
// Workgroups past the work list's capacity, one per (job, node), compute the
// node's child totals (one wave per channel) beside the row routing; inside
// a routing workgroup they lengthened the launch by 2-4 us.  Each work item (<= 4096 rows of one
// node) routes its rows -- NaN bin 255 or bin <= threshold goes left
// (symmetric_tree.cpp:309-311) -- keeping them in registers, reserves room
// with one 64-bit atomic on the node's cursor pair (left fills up from the
// node's start, right fills down from its end; set by the level's plan) and
// writes the rows, placed by wave ballots.
__device__ inline uint32_t level_row(const JobDev& J, uint32_t level, uint32_t pos) {
  return (level == 0 && J.root_identity) ? pos : J.rows[static_cast<size_t>(level) * J.N + pos];
}

template <bool SINGLE>
__global__ void __launch_bounds__(kPartThreads) k_partition(const JobDev* __restrict__ jobs,
                                                   const PartWork* __restrict__ work,
                                                   const uint32_t* __restrict__ count,
                                                   uint32_t level, uint32_t items_cap) {
  constexpr int kIt = kPartChunk / kPartThreads;  // rows per lane
  constexpr int kW = kPartThreads / 64;           // waves
  if (blockIdx.x >= items_cap) {  // one workgroup per (job, node): the node's child totals
    const uint32_t b = blockIdx.x - items_cap, node = b & ((1u << level) - 1u);
    const JobDev& J = jobs[SINGLE ? 0u : b >> level];
    if (!J.st->active || J.st->tree >= static_cast<int>(J.n_trees)) return;
    uint32_t feat, split_bin;
    if (!decode_split(J, level, false, &feat, &split_bin)) return;
    const uint32_t* offd = J.off + static_cast<size_t>(level) * ((1u << J.D) + 1u);
    // an empty node's children get zero totals from the plan
    if ((threadIdx.x >> 6) < J.K && offd[node + 1] > offd[node])
      child_totals(J, level, node, threadIdx.x >> 6, feat, split_bin);
    return;
  }
  // the list is read with its count (the grid never exceeds the list's capacity)
  const uint32_t cnt = *count;
  const PartWork w = work[blockIdx.x];
  if (blockIdx.x >= cnt) return;
  const JobDev& J = jobs[SINGLE ? 0u : w.job];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the item's row ids do not depend on the split: in flight while it is decoded
  uint32_t rv[kIt];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const uint32_t j = k * kPartThreads + tid;
    rv[k] = j < w.len ? level_row(J, level, w.start + j) : 0u;
  }
  if (!J.st->active || J.st->tree >= static_cast<int>(J.n_trees)) return;
  uint32_t feat, split_bin;
  if (!decode_split(J, level, w.slot == 0 && tid == 0, &feat, &split_bin)) return;
  const uint32_t* offd = J.off + static_cast<size_t>(level) * ((1u << J.D) + 1u);
  const uint8_t* col = J.bins_cm + static_cast<size_t>(feat) * J.N;
  const uint8_t thr = static_cast<uint8_t>(split_bin);
  uint32_t* rows_n = J.rows + static_cast<size_t>(level + 1) * J.N;
  uint32_t* cur = cursor_set(J, level) + 2 * w.node;
  __shared__ uint32_t s_cnt[2][kIt * kW];  // per (step, wave): count -> offset in the item
  __shared__ uint32_t s_base[2];
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // route every row of the item, keeping row ids and directions in registers
  uint32_t lbits = 0, rbits = 0;
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const uint32_t j = k * kPartThreads + tid;
    const bool valid = j < w.len;
    const uint8_t b = valid ? col[rv[k]] : 0;
    const bool left = valid && (b == 255 || b <= thr);
    const bool right = valid && !left;
    lbits |= static_cast<uint32_t>(left) << k;
    rbits |= static_cast<uint32_t>(right) << k;
    const unsigned long long ml = __ballot(left), mr = __ballot(right);
    if (lane == 0) {
      s_cnt[0][k * kW + wv] = __popcll(ml);
      s_cnt[1][k * kW + wv] = __popcll(mr);
    }
  }
  __syncthreads();
  __shared__ uint32_t s_tot[2];
  static_assert(kIt * kW == 64, "one wave scans one side's (step, wave) counts");
  if (wv < 2) {  // offsets in (step, wave) order: wave s scans side s, lane = entry
    const uint32_t v = s_cnt[wv][lane];
    const uint32_t incl = wave_incl_scan_u32(v);
    s_cnt[wv][lane] = incl - v;
    if (lane == 63) s_tot[wv] = incl;
  }
  __syncthreads();
  if (tid == 0) {  // one 64-bit reservation for both sides: left position, right rows taken
    const unsigned long long add =
        static_cast<unsigned long long>(s_tot[0]) | (static_cast<unsigned long long>(s_tot[1]) << 32);
    const unsigned long long old =
        add == 0 ? 0ull : atomicAdd(reinterpret_cast<unsigned long long*>(cur), add);
    s_base[0] = static_cast<uint32_t>(old);
    s_base[1] = offd[w.node + 1] - static_cast<uint32_t>(old >> 32) - s_tot[1];
  }
  __syncthreads();
  {  // both reservations must lie inside the node's segment; otherwise write nothing
    const uint32_t lo = offd[w.node], hi = offd[w.node + 1];
    const bool bad = (s_tot[0] && (s_base[0] < lo || s_base[0] > hi || s_tot[0] > hi - s_base[0])) ||
                     (s_tot[1] && (s_base[1] < lo || s_base[1] > hi || s_tot[1] > hi - s_base[1]));
    if (bad) {
      if (tid == 0) atomicOr(&J.st->fault, 1u);
      return;
    }
  }
  uint32_t dst[kIt];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const bool left = (lbits >> k) & 1u, right = (rbits >> k) & 1u;
    const unsigned long long ml = __ballot(left), mr = __ballot(right);
    dst[k] = left ? s_base[0] + s_cnt[0][k * kW + wv] + __popcll(ml & lt)
                  : s_base[1] + s_cnt[1][k * kW + wv] + __popcll(mr & lt);
    if (left || right) rows_n[dst[k]] = rv[k];
  }
  if (J.ghl) {  // several channels: carry the rows' (g, h) into the next level's row order
    const size_t N = J.N;
    const float2* src = level == 0 ? J.gh : J.ghl + static_cast<size_t>(level & 1u) * J.K * N;
    float2* out = J.ghl + static_cast<size_t>((level + 1) & 1u) * J.K * N;
    const uint32_t moved = lbits | rbits;
    for (uint32_t c = 0; c < J.K; ++c) {
      float2 v[kIt];
#pragma unroll
      for (int k = 0; k < kIt; ++k) {
        const uint32_t at = level == 0 ? rv[k] : w.start + k * kPartThreads + tid;
        v[k] = ((moved >> k) & 1u) ? src[c * N + at] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int k = 0; k < kIt; ++k)
        if ((moved >> k) & 1u) out[c * N + dst[k]] = v[k];
    }
  }
}

// Per job, after the partition: the plan of the next level with its stored
// outputs and both work lists (several jobs, or a partition of the last
// level; a one-job fit plans inside its next histogram launch).
__global__ void __launch_bounds__(256) k_part_plan(const JobDev* __restrict__ jobs, Lists L,
                                                   uint32_t level) {
  __shared__ PlanShared S;
  (void)level_plan<256>(jobs[blockIdx.y], L, blockIdx.y, level, S, 2);
}

// Copy externally supplied gradients (teacher forcing, [K][N]) into gh,
// recording max |g| and max |h| for the histogram scales.
__global__ void __launch_bounds__(256) k_set_gh(const JobDev* __restrict__ jobs, const float* g,
                                                const float* h) {
  const JobDev& J = jobs[0];
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;  // over K * N, class-major
  float a = 0.f, b = 0.f;
  if (p < J.N * J.K) {
    J.gh[p] = make_float2(g[p], h[p]);
    a = fabsf(g[p]);
    b = fabsf(h[p]);
  }
  block_max2_to(J.st, a, b);
}

// Fixed-point histogram cells -> float pairs (tc_hist_build).
__global__ void __launch_bounds__(256) k_hist_to_float(const HistCell* __restrict__ h,
                                                       const TreeState* __restrict__ st, uint32_t n,
                                                       float2* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    out[i] = make_float2(static_cast<float>(static_cast<double>(h[i].g) * st->ginv),
                         static_cast<float>(static_cast<double>(h[i].h) * st->hinv));
}

}  // namespace tc
