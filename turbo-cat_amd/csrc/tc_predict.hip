// A15-A17: symmetric-ensemble inference on gfx950, and the continuous-
// batching server built on it.
//
// Kernel layout: a workgroup takes R rows (R = 64 for F <= 620) and stages
// them in LDS with an odd row stride (conflict-free column reads).  Its 256
// threads are (q, r) = (tree lane, row): in each pass tree lane q evaluates
// one group of four trees for row r -- leaf index built MSB-first from depth
// comparisons (v >= threshold, NaN left; fast_float_ensemble.hpp:122-137)
// -- and the group value w0*l0 + w1*l1 + w2*l2 + w3*l3 (as the reference
// compiler contracts it: w1*l1 rounded, the other three fused onto it).  Then lane 0 of the row adds
// the groups to the row's sum in tree order, exactly the reference's
// 8-row-block order (fast_float_ensemble.hpp:480-541): groups added one by
// one, leftover trees fused in one by one (sum = fma(w, leaf, sum)).
//
// The reference sums some rows tree by tree instead, sum = fma(w_t, l_t, sum)
// for every tree in order: the rows of a call past its last full 8-row block
// (fast_float_ensemble.hpp:279-287, :543-560; fast_ensemble.hpp:428-446),
// every row of a binned call under 16 rows (symmetric_tree.cpp:954-961) and
// every row when no tree has a split (the fast ensembles are then empty,
// symmetric_tree.cpp:1144-1148).  Such a row is marked by seq_flag[row] (a
// served batch, whose rows come from several requests) or by row >= seq_from
// (one call); lane 0 then fuses the group's four leaf values one by one.  So a
// row's value depends only on the row, the model and its place in its own
// request, never on the batch the request was coalesced into.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "tc_device.h"
#include "tc_internal.h"

namespace tc {

struct EnsDev {
  const uint16_t* feat;
  const float* thr;
  const uint8_t* bin;
  const uint16_t* depth;
  const float* leaves;
  const float* weight;
  uint32_t T, D, K;
  float base;
};

template <bool BINNED>
__global__ void __launch_bounds__(256) k_predict(EnsDev E, const void* __restrict__ Xv, uint32_t n,
                                                 uint32_t F, uint32_t ldc, uint32_t R,
                                                 int transform, float* __restrict__ out,
                                                 uint32_t* __restrict__ leaf_idx, uint32_t seq_from,
                                                 const uint8_t* __restrict__ seq_flag) {
  extern __shared__ unsigned char s_raw[];
  const uint32_t Q = 256u / R;                 // tree lanes per row
  const uint32_t S = BINNED ? (F | 3u) + 1u : (F | 1u);  // row stride in LDS elements
  float* s_gv = reinterpret_cast<float*>(s_raw);           // [Q][R] group values
  float* s_lv = s_gv + 256;                                // [Q][4][R] leaf values (tree-by-tree rows)
  unsigned char* s_x = s_raw + sizeof(float) * 256 * 5;
  const uint32_t tid = threadIdx.x, q = tid / R, r = tid % R;
  const uint32_t row0 = blockIdx.x * R;
  const uint32_t nr = min(R, n - row0);
  // stage the rows
  if (BINNED) {
    const uint8_t* B = static_cast<const uint8_t*>(Xv);  // column-major [F][ldc]
    for (uint32_t i = tid; i < nr * F; i += 256) {
      const uint32_t f = i / nr, rr = i % nr;
      s_x[rr * S + f] = B[static_cast<size_t>(f) * ldc + row0 + rr];
    }
  } else {
    const float* X = static_cast<const float*>(Xv);
    float* sx = reinterpret_cast<float*>(s_x);
    const float* src = X + static_cast<size_t>(row0) * F;
    for (uint32_t i = tid; i < nr * F; i += 256) sx[(i / F) * S + (i % F)] = src[i];
  }
  __syncthreads();
  const uint32_t LP = 1u << E.D;
  const uint32_t groups = E.T / 4u;
  const bool seq = r < nr && (seq_flag ? seq_flag[row0 + r] != 0 : row0 + r >= seq_from);
  float sum = 0.f;
  for (uint32_t g0 = 0; g0 < groups; g0 += Q) {
    const uint32_t g = g0 + q;
    if (g < groups && r < nr) {
      float lv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t t = 4 * g + k;
        const uint32_t dep = E.depth[t];
        uint32_t li = 0;
        for (uint32_t d = 0; d < dep; ++d) {
          const uint32_t f = E.feat[t * E.D + d];
          uint32_t right;
          if (BINNED) {
            const uint8_t v = s_x[r * S + f];
            right = (v != 255 && v > E.bin[t * E.D + d]) ? 1u : 0u;
          } else {
            const float v = reinterpret_cast<const float*>(s_x)[r * S + f];
            right = (!isnan(v) && v >= E.thr[t * E.D + d]) ? 1u : 0u;
          }
          li = (li << 1) | right;
        }
        if (leaf_idx) leaf_idx[static_cast<size_t>(row0 + r) * E.T + t] = li;
        lv[k] = E.leaves[static_cast<size_t>(t) * LP + li];
      }
      if (seq) {
#pragma unroll
        for (int k = 0; k < 4; ++k) s_lv[(q * 4 + k) * R + r] = lv[k];
      } else {
        // the reference compiler's contraction of w0*l0 + w1*l1 + w2*l2 + w3*l3:
        // w1*l1 rounded, then w0*l0, w2*l2, w3*l3 fused onto it in that order
        const float* w = E.weight + 4 * g;
        float gv = __fmul_rn(w[1], lv[1]);
        gv = __fmaf_rn(w[0], lv[0], gv);
        gv = __fmaf_rn(w[2], lv[2], gv);
        gv = __fmaf_rn(w[3], lv[3], gv);
        s_gv[q * R + r] = gv;
      }
    }
    __syncthreads();
    if (q == 0 && r < nr) {
      const uint32_t ng = min(Q, groups - g0);
      if (seq) {
        for (uint32_t k = 0; k < ng; ++k)
          for (uint32_t j = 0; j < 4; ++j)
            sum = __fmaf_rn(E.weight[4 * (g0 + k) + j], s_lv[(k * 4 + j) * R + r], sum);
      } else {
        for (uint32_t k = 0; k < ng; ++k) sum = __fadd_rn(sum, s_gv[k * R + r]);
      }
    }
    __syncthreads();
  }
  if (q == 0 && r < nr) {
    for (uint32_t t = groups * 4u; t < E.T; ++t) {
      const uint32_t dep = E.depth[t];
      uint32_t li = 0;
      for (uint32_t d = 0; d < dep; ++d) {
        const uint32_t f = E.feat[t * E.D + d];
        uint32_t right;
        if (BINNED) {
          const uint8_t v = s_x[r * S + f];
          right = (v != 255 && v > E.bin[t * E.D + d]) ? 1u : 0u;
        } else {
          const float v = reinterpret_cast<const float*>(s_x)[r * S + f];
          right = (!isnan(v) && v >= E.thr[t * E.D + d]) ? 1u : 0u;
        }
        li = (li << 1) | right;
      }
      if (leaf_idx) leaf_idx[static_cast<size_t>(row0 + r) * E.T + t] = li;
      sum = __fmaf_rn(E.weight[t], E.leaves[static_cast<size_t>(t) * LP + li], sum);
    }
    float v = __fadd_rn(sum, E.base);
    if (transform == 1) v = __fdiv_rn(1.0f, __fadd_rn(1.0f, expf(-v)));
    out[row0 + r] = v;
  }
}

// A18 inference: the layout of k_predict, with K outputs per row.  Tree
// lane q of row r sums w_t * leaf_t[idx][c] over its trees t = q, q + Q, ...
// (fused multiply-adds, tree order) into K partial sums; lane 0 then adds
// base + the Q partials in lane order, and optionally applies softmax (max
// shift, exp, class-order sum, divide: CrossEntropyLoss::softmax,
// loss.cpp:273-285).  As in k_predict, a row's value depends only on the
// row and the model.
template <bool BINNED>
__global__ void __launch_bounds__(256) k_predict_multi(EnsDev E, const void* __restrict__ Xv,
                                                       uint32_t n, uint32_t F, uint32_t ldc,
                                                       uint32_t R, int transform,
                                                       float* __restrict__ out,
                                                       uint32_t* __restrict__ leaf_idx) {
  extern __shared__ unsigned char s_raw[];
  const uint32_t Q = 256u / R, K = E.K;
  const uint32_t S = BINNED ? (F | 3u) + 1u : (F | 1u);
  float* s_acc = reinterpret_cast<float*>(s_raw);  // [Q][R][K] partial sums
  unsigned char* s_x = s_raw + sizeof(float) * 256 * K;
  const uint32_t tid = threadIdx.x, q = tid / R, r = tid % R;
  const uint32_t row0 = blockIdx.x * R;
  const uint32_t nr = min(R, n - row0);
  if (BINNED) {
    const uint8_t* B = static_cast<const uint8_t*>(Xv);  // column-major [F][ldc]
    for (uint32_t i = tid; i < nr * F; i += 256) {
      const uint32_t f = i / nr, rr = i % nr;
      s_x[rr * S + f] = B[static_cast<size_t>(f) * ldc + row0 + rr];
    }
  } else {
    const float* X = static_cast<const float*>(Xv);
    float* sx = reinterpret_cast<float*>(s_x);
    const float* src = X + static_cast<size_t>(row0) * F;
    for (uint32_t i = tid; i < nr * F; i += 256) sx[(i / F) * S + (i % F)] = src[i];
  }
  __syncthreads();
  const uint32_t LP = 1u << E.D;
  float acc[kMaxClasses];
#pragma unroll
  for (int c = 0; c < kMaxClasses; ++c) acc[c] = 0.f;
  if (r < nr) {
    for (uint32_t t = q; t < E.T; t += Q) {
      const uint32_t dep = E.depth[t];
      uint32_t li = 0;
      for (uint32_t d = 0; d < dep; ++d) {
        const uint32_t f = E.feat[t * E.D + d];
        uint32_t right;
        if (BINNED) {
          const uint8_t v = s_x[r * S + f];
          right = (v != 255 && v > E.bin[t * E.D + d]) ? 1u : 0u;
        } else {
          const float v = reinterpret_cast<const float*>(s_x)[r * S + f];
          right = (!isnan(v) && v >= E.thr[t * E.D + d]) ? 1u : 0u;
        }
        li = (li << 1) | right;
      }
      if (leaf_idx) leaf_idx[static_cast<size_t>(row0 + r) * E.T + t] = li;
      const float w = E.weight[t];
      const float* lv = E.leaves + (static_cast<size_t>(t) * LP + li) * K;
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        acc[c] = __fmaf_rn(w, lv[c], acc[c]);
      }
    }
  }
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    s_acc[(q * R + r) * K + c] = acc[c];
  }
  __syncthreads();
  if (q == 0 && r < nr) {
    float v[kMaxClasses];
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) v[c] = 0.f;
    for (uint32_t qq = 0; qq < Q; ++qq) {  // lane partials in lane order
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        v[c] = __fadd_rn(v[c], s_acc[(qq * R + r) * K + c]);
      }
    }
    float m = -INFINITY;
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      v[c] = __fadd_rn(E.base, v[c]);
      m = fmaxf(m, v[c]);
    }
    if (transform == 1) {
      float z = 0.f;
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        v[c] = expf(__fsub_rn(v[c], m));
        z = __fadd_rn(z, v[c]);
      }
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= K) break;
        v[c] = __fdiv_rn(v[c], z);
      }
    }
    float* o = out + static_cast<size_t>(row0 + r) * K;
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      o[c] = v[c];
    }
  }
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

EnsDev ens_dev(const tc_ensemble* e) {
  EnsDev E;
  E.feat = e->feat; E.thr = e->thr; E.bin = e->bin; E.depth = e->depth; E.leaves = e->leaves;
  E.weight = e->weight; E.T = e->T; E.D = e->D; E.K = e->K; E.base = e->base;
  return E;
}

// Rows per workgroup and LDS bytes for a feature count and output channels.
void predict_shape(uint32_t F, uint32_t K, bool binned, uint32_t* R, size_t* lds) {
  const uint32_t S = binned ? (F | 3u) + 1u : (F | 1u);
  const size_t elem = binned ? 1 : 4;
  // K = 1: group values and the tree-by-tree rows' leaf values (5 x 256 floats)
  const size_t head = sizeof(float) * 256 * (K > 1 ? K : 5u);
  for (uint32_t r : {64u, 32u, 16u, 8u, 4u, 2u, 1u}) {
    const size_t bytes = head + static_cast<size_t>(r) * S * elem;
    if (bytes <= 160 * 1024 || r == 1) {
      *R = r;
      *lds = bytes;
      return;
    }
  }
}

// First row of a call of n rows that the reference sums tree by tree (see
// k_predict): past the last full 8-row block; every row of a binned call
// under 16 rows, or of an ensemble without splits.
uint32_t seq_rows_from(const tc_ensemble* e, bool binned, uint32_t n) {
  if (!e->any_split || (binned && n < 16)) return 0u;
  return n / 8u * 8u;
}

void launch_predict(const tc_ensemble* e, const void* X, bool binned, uint32_t n, uint32_t F,
                    uint32_t ldc, int transform, float* out, uint32_t* leaf_idx, hipStream_t s,
                    const uint8_t* seq_flag = nullptr) {
  if (n == 0) return;
  const uint32_t seq_from = seq_rows_from(e, binned, n);
  uint32_t R;
  size_t lds;
  predict_shape(F, e->K, binned, &R, &lds);
  TC_REQUIRE(lds <= 160 * 1024, "too many features for one LDS row tile");
  const dim3 grid((n + R - 1) / R);
  if (e->K > 1) {
    auto kern = binned ? &k_predict_multi<true> : &k_predict_multi<false>;
    if (lds > 64 * 1024)
      TC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, ens_dev(e), X, n, F, ldc, R, transform, out,
                       leaf_idx);
  } else if (binned) {
    if (lds > 64 * 1024)
      TC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_predict<true>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k_predict<true>, grid, dim3(256), lds, s, ens_dev(e), X, n, F, ldc, R,
                       transform, out, leaf_idx, seq_from, seq_flag);
  } else {
    if (lds > 64 * 1024)
      TC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_predict<false>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k_predict<false>, grid, dim3(256), lds, s, ens_dev(e), X, n, F, ldc, R,
                       transform, out, leaf_idx, seq_from, seq_flag);
  }
  TC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Continuous-batching server.
//
// Requests enter a queue; a scheduler thread admits them while the rows of
// admitted-but-unfinished requests stay inside the memory budget, and packs
// the head of the admitted queue (whole or partial requests) into the next
// free batch slot.  Each slot owns a stream, pinned staging and device
// buffers; its batch is copied in, predicted and copied out asynchronously,
// so up to n_streams batches overlap on the device while new arrivals queue
// for the next batch.  Completion scatters the slot's rows back into the
// requests' output buffers.
struct Request {
  uint64_t id;
  const float* X;
  float* out;
  uint32_t n;
  uint32_t next = 0;       // next row to schedule
  uint32_t done = 0;       // rows completed
  bool finished = false;
};

struct Segment {
  std::shared_ptr<Request> req;
  uint32_t row0, count, slot_row;
};

// Host copies of one batch split over a few threads (the caller works too):
// one thread's memcpy bandwidth, not PCIe, would otherwise bound the server.
class CopyPool {
 public:
  struct Piece {
    void* dst;
    const void* src;
    size_t bytes;
  };
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
  // Copies every piece; returns when all are done.
  void copy(const std::vector<Piece>& pieces) {
    if (pieces.empty()) return;
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
      std::memcpy(pc->dst, pc->src, pc->bytes);
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
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::vector<Piece>* jobs_ = nullptr;
  size_t next_ = 0, left_ = 0;
  bool stop_ = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  float* h_x = nullptr;    // pinned
  float* h_out = nullptr;  // pinned
  uint8_t* h_flag = nullptr;  // pinned: rows the reference sums tree by tree
  float* d_x = nullptr;
  float* d_out = nullptr;
  uint8_t* d_flag = nullptr;
  bool busy = false;
  std::vector<Segment> segs;
};

}  // namespace tc

struct tc_server {
  tc_ensemble* ens = nullptr;
  tc_serve_params p{};
  std::vector<tc::Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<tc::Request>> pending;   // not yet admitted
  std::deque<std::shared_ptr<tc::Request>> admitted;  // admitted, rows left to schedule
  std::unordered_map<uint64_t, std::shared_ptr<tc::Request>> live;
  uint64_t admitted_bytes = 0;
  uint64_t next_id = 1;
  uint64_t n_batches = 0, n_rows = 0;
  bool stop = false;
  std::string error;
  std::unique_ptr<tc::CopyPool> copier;
  std::thread worker;
};

namespace tc {

static void serve_loop(tc_server* S) {
  const uint32_t F = S->ens->F;
  const uint64_t row_bytes = 4ull * F + 4ull;
  (void)hipSetDevice(S->ens->ctx->device);
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    // admission: move pending requests while the budget allows (a request
    // larger than the whole budget is admitted alone when nothing else is)
    while (!S->pending.empty()) {
      const uint64_t b = row_bytes * S->pending.front()->n;
      if (S->admitted_bytes + b > S->p.memory_budget && S->admitted_bytes > 0) break;
      S->admitted_bytes += b;
      S->admitted.push_back(S->pending.front());
      S->pending.pop_front();
    }
    bool progressed = false;
    // retire finished slots
    for (auto& sl : S->slots) {
      if (!sl.busy) continue;
      if (hipEventQuery(sl.done) != hipSuccess) continue;
      for (auto& sg : sl.segs) {
        std::memcpy(sg.req->out + sg.row0, sl.h_out + sg.slot_row, 4ull * sg.count);
        sg.req->done += sg.count;
        if (sg.req->done == sg.req->n) {
          sg.req->finished = true;
          S->admitted_bytes -= row_bytes * sg.req->n;
        }
      }
      sl.segs.clear();
      sl.busy = false;
      progressed = true;
      S->cv_done.notify_all();
    }
    // fill free slots from the admitted queue; the rows are staged with the
    // queue unlocked (submitters keep going), in parallel pieces
    for (auto& sl : S->slots) {
      if (sl.busy || S->admitted.empty()) continue;
      uint32_t rows = 0;
      while (!S->admitted.empty() && rows < S->p.max_batch_rows) {
        auto& rq = S->admitted.front();
        const uint32_t take = std::min(rq->n - rq->next, S->p.max_batch_rows - rows);
        sl.segs.push_back(Segment{rq, rq->next, take, rows});
        rq->next += take;
        rows += take;
        if (rq->next == rq->n) S->admitted.pop_front();
      }
      if (rows == 0) continue;
      lk.unlock();
      {
        constexpr size_t kPiece = 1u << 20;
        std::vector<CopyPool::Piece> pieces;
        for (const auto& sg : sl.segs) {
          // a request's own tree-by-tree rows, wherever it lands in the batch
          const uint32_t from = seq_rows_from(S->ens, false, sg.req->n);
          for (uint32_t k = 0; k < sg.count; ++k) sl.h_flag[sg.slot_row + k] = sg.row0 + k >= from ? 1 : 0;
          const char* src = reinterpret_cast<const char*>(sg.req->X + static_cast<size_t>(sg.row0) * F);
          char* dst = reinterpret_cast<char*>(sl.h_x + static_cast<size_t>(sg.slot_row) * F);
          const size_t bytes = 4ull * sg.count * F;
          for (size_t o = 0; o < bytes; o += kPiece)
            pieces.push_back(CopyPool::Piece{dst + o, src + o, std::min(kPiece, bytes - o)});
        }
        S->copier->copy(pieces);
      }
      lk.lock();
      try {
        TC_HIP(hipMemcpyAsync(sl.d_x, sl.h_x, 4ull * rows * F, hipMemcpyHostToDevice, sl.stream));
        TC_HIP(hipMemcpyAsync(sl.d_flag, sl.h_flag, rows, hipMemcpyHostToDevice, sl.stream));
        launch_predict(S->ens, sl.d_x, false, rows, F, 0, S->p.transform, sl.d_out, nullptr, sl.stream,
                       sl.d_flag);
        TC_HIP(hipMemcpyAsync(sl.h_out, sl.d_out, 4ull * rows, hipMemcpyDeviceToHost, sl.stream));
        TC_HIP(hipEventRecord(sl.done, sl.stream));
      } catch (const std::exception& e) {
        S->error = e.what();
        S->stop = true;
      }
      sl.busy = true;
      ++S->n_batches;
      S->n_rows += rows;
      progressed = true;
    }
    bool any_busy = false;
    for (auto& sl : S->slots) any_busy |= sl.busy;
    // on stop, drain what was submitted unless the device failed
    if (S->stop && !any_busy &&
        ((S->admitted.empty() && S->pending.empty()) || !S->error.empty()))
      break;
    if (!progressed) {
      if (any_busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        S->cv_work.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
  }
  S->cv_done.notify_all();
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
    const uint32_t K = n_classes;
    *out = nullptr;
    for (uint32_t t = 0; t < n_trees; ++t) {
      TC_REQUIRE(depth[t] <= max_depth, "tree depth exceeds max_depth");
      for (uint32_t d = 0; d < depth[t]; ++d)
        TC_REQUIRE(feature[t * max_depth + d] < n_features, "split feature out of range");
    }
    auto* e = new tc_ensemble();
    e->ctx = ctx; e->T = n_trees; e->D = max_depth; e->F = n_features; e->K = K; e->base = base;
    for (uint32_t t = 0; t < n_trees; ++t) e->any_split = e->any_split || depth[t] > 0;
    try {
      const size_t TD = std::max<size_t>(1, static_cast<size_t>(n_trees) * max_depth);
      const size_t TL = std::max<size_t>(1, (static_cast<size_t>(n_trees) << max_depth) * K);
      e->feat = dmalloc<uint16_t>(TD);
      e->thr = dmalloc<float>(TD);
      e->bin = dmalloc<uint8_t>(TD);
      e->depth = dmalloc<uint16_t>(std::max(1u, n_trees));
      e->leaves = dmalloc<float>(TL);
      e->weight = dmalloc<float>(std::max(1u, n_trees));
      if (n_trees) {
        const size_t td = static_cast<size_t>(n_trees) * max_depth;
        std::vector<uint8_t> b(td, 0);
        if (bin) std::memcpy(b.data(), bin, td);
        TC_HIP(hipMemcpy(e->feat, feature, td * 2, hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(e->thr, fthr, td * 4, hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(e->bin, b.data(), td, hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(e->depth, depth, n_trees * 2ull, hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(e->leaves, leaves, (static_cast<size_t>(n_trees) << max_depth) * K * 4,
                         hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(e->weight, weights, n_trees * 4ull, hipMemcpyHostToDevice));
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
  return guard([&] {
    if (!e) return;
    dfree(e->feat); dfree(e->thr); dfree(e->bin); dfree(e->depth); dfree(e->leaves); dfree(e->weight);
    delete e;
  });
}

int32_t tc_predict_dev(tc_ensemble* e, const float* X_dev, uint32_t n, uint32_t F, int32_t transform,
                       float* out_dev) {
  return guard([&] {
    TC_REQUIRE(e && (X_dev || n == 0) && (out_dev || n == 0), "tc_predict_dev: NULL argument");
    TC_REQUIRE(F == e->F, "feature count differs from the model's");
    launch_predict(e, X_dev, false, n, F, 0, transform, out_dev, nullptr, e->ctx->stream);
  });
}

int32_t tc_predict(tc_ensemble* e, const float* X, uint32_t n, uint32_t F, int32_t transform,
                   float* out, uint32_t* leaf_idx) {
  return guard([&] {
    TC_REQUIRE(e && (X || n == 0) && (out || n == 0), "tc_predict: NULL argument");
    TC_REQUIRE(F == e->F, "feature count differs from the model's");
    if (n == 0) return;
    hipStream_t s = e->ctx->stream;
    float* dx = dmalloc<float>(static_cast<size_t>(n) * F + static_cast<size_t>(n) * e->K);
    uint32_t* dl = nullptr;
    try {
      if (leaf_idx) dl = dmalloc<uint32_t>(static_cast<size_t>(n) * std::max(1u, e->T));
      TC_HIP(hipMemcpyAsync(dx, X, 4ull * n * F, hipMemcpyHostToDevice, s));
      launch_predict(e, dx, false, n, F, 0, transform, dx + static_cast<size_t>(n) * F, dl, s);
      TC_HIP(hipMemcpyAsync(out, dx + static_cast<size_t>(n) * F, 4ull * n * e->K,
                            hipMemcpyDeviceToHost, s));
      if (leaf_idx && e->T)
        TC_HIP(hipMemcpyAsync(leaf_idx, dl, 4ull * n * e->T, hipMemcpyDeviceToHost, s));
      TC_HIP(hipStreamSynchronize(s));
    } catch (...) {
      dfree(dx);
      dfree(dl);
      throw;
    }
    dfree(dx);
    dfree(dl);
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
    launch_predict(e, dx, false, n, F, 0, transform, dout, nullptr, s);
    TC_HIP(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
  });
}

int32_t tc_predict_binned(tc_ensemble* e, tc_dataset* ds, float* out) {
  return guard([&] {
    TC_REQUIRE(e && ds && out, "tc_predict_binned: NULL argument");
    TC_REQUIRE(ds->F == e->F, "feature count differs from the model's");
    if (ds->N == 0) return;
    hipStream_t s = e->ctx->stream;
    float* d = dmalloc<float>(static_cast<size_t>(ds->N) * e->K);
    try {
      launch_predict(e, ds->bins_cm, true, ds->N, ds->F, ds->N, 0, d, nullptr, s);
      TC_HIP(hipMemcpyAsync(out, d, 4ull * ds->N * e->K, hipMemcpyDeviceToHost, s));
      TC_HIP(hipStreamSynchronize(s));
    } catch (...) {
      dfree(d);
      throw;
    }
    dfree(d);
  });
}

int32_t tc_serve_start(tc_ensemble* e, const tc_serve_params* p, tc_server** out) {
  return guard([&] {
    TC_REQUIRE(e && p && out, "tc_serve_start: NULL argument");
    TC_REQUIRE(e->K == 1, "tc_serve_start: serving takes single-output ensembles");
    TC_REQUIRE(p->max_batch_rows > 0 && p->n_streams > 0 && p->n_streams <= 8,
               "tc_serve_start: bad parameters");
    *out = nullptr;
    auto* S = new tc_server();
    S->ens = e;
    S->p = *p;
    S->slots.resize(p->n_streams);
    try {
      for (auto& sl : S->slots) {
        TC_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        TC_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_x), 4ull * p->max_batch_rows * e->F,
                             hipHostMallocDefault));
        TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_out), 4ull * p->max_batch_rows,
                             hipHostMallocDefault));
        TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_flag), p->max_batch_rows,
                             hipHostMallocDefault));
        sl.d_x = dmalloc<float>(static_cast<size_t>(p->max_batch_rows) * e->F);
        sl.d_out = dmalloc<float>(p->max_batch_rows);
        sl.d_flag = dmalloc<uint8_t>(p->max_batch_rows);
      }
    } catch (...) {
      for (auto& sl : S->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_x) (void)hipHostFree(sl.h_x);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.h_flag) (void)hipHostFree(sl.h_flag);
        dfree(sl.d_x);
        dfree(sl.d_out);
        dfree(sl.d_flag);
      }
      delete S;
      throw;
    }
    S->copier = std::make_unique<CopyPool>(7);
    S->worker = std::thread(serve_loop, S);
    *out = S;
  });
}

int32_t tc_serve_submit(tc_server* S, const float* X, uint32_t n, float* out, uint64_t* ticket) {
  return guard([&] {
    TC_REQUIRE(S && X && out && ticket && n > 0, "tc_serve_submit: bad arguments");
    auto rq = std::make_shared<Request>();
    rq->X = X;
    rq->out = out;
    rq->n = n;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->stop) throw Error(TC_ESTATE, "server stopped: " + S->error);
      rq->id = S->next_id++;
      S->pending.push_back(rq);
      S->live[rq->id] = rq;
    }
    S->cv_work.notify_one();
    *ticket = rq->id;
  });
}

int32_t tc_serve_wait(tc_server* S, uint64_t ticket, int64_t timeout_us) {
  int32_t code = TC_OK;
  const int32_t r = guard([&] {
    TC_REQUIRE(S, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    TC_REQUIRE(it != S->live.end(), "unknown ticket");
    auto rq = it->second;
    auto pred = [&] { return rq->finished || (S->stop && !S->error.empty()); };
    if (timeout_us < 0) {
      S->cv_done.wait(lk, pred);
    } else if (!S->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), pred)) {
      code = TC_ETIMEOUT;
      return;
    }
    if (!rq->finished) throw Error(TC_EDEVICE, "server failed: " + S->error);
    S->live.erase(ticket);
  });
  if (r != TC_OK) return r;
  if (code != TC_OK) return fail(code, "tc_serve_wait: timed out");
  return TC_OK;
}

int32_t tc_serve_stats(tc_server* S, uint64_t* batches, uint64_t* rows) {
  return guard([&] {
    TC_REQUIRE(S, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (batches) *batches = S->n_batches;
    if (rows) *rows = S->n_rows;
  });
}

int32_t tc_serve_stop(tc_server* S) {
  return guard([&] {
    if (!S) return;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    for (auto& sl : S->slots) {
      (void)hipStreamSynchronize(sl.stream);
      (void)hipStreamDestroy(sl.stream);
      (void)hipEventDestroy(sl.done);
      (void)hipHostFree(sl.h_x);
      (void)hipHostFree(sl.h_out);
      (void)hipHostFree(sl.h_flag);
      dfree(sl.d_x);
      dfree(sl.d_out);
      dfree(sl.d_flag);
    }
    delete S;
  });
}

}  // extern "C"
