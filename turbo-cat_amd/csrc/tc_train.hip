// Host side of training: device state for a batch of jobs, the fixed
// per-tree launch sequence (optionally replayed as one hipGraph per tree),
// and the C-ABI entry points built on it (tc_train, tc_train_batch,
// tc_tree_build, tc_hist_build, tc_split_scan, tc_hist_subtract,
// tc_gradients).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <future>
#include <memory>
#include <numeric>
#include <random>

#include "tc_device.h"
#include "tc_internal.h"
#include "tc_train_kernels.hip"

namespace tc {

namespace {

constexpr int kHistUnroll = 2;  // rows per lane in flight in k_hist (U = 3 needs more than 64 VGPRs)
// k_hist_plan (one job below the root): one row per lane in flight.  Its items
// hold a few rows per lane, so a deeper pipeline buys little (49 VGPRs instead
// of 65).  Measured: 1M x 100 LogLoss levels 1-5 144.5 -> 135.5 us per tree in
// a kernel trace (U = 4: 157 us); 50K x 20 MSE 6,276 -> 6,325 trees/s.
constexpr int kPlanUnroll = 1;

uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }
uint32_t round8(uint32_t x) { return std::max(8u, (x + 7u) & ~7u); }

__global__ void k_iota(uint32_t* rows, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rows[i] = i;
}

__global__ void k_init_state(TreeState* st) {
  st->active = 0;
  st->depth = 0;
  st->tree = -1;
  st->gmax_next = 0u;
  st->hmax_next = 0u;
  st->fault = 0u;
  st->stopped = 0u;
  st->es_best = 1e30f;  // booster.cpp:119: lower is better for the loss
  st->es_best_iter = 0u;
  st->es_count = 0u;
  st->pad = 0u;
  for (int d = 0; d < kMaxDepth; ++d) st->best[d] = kEmptyKey;
  st->gscale = st->hscale = st->ginv = st->hinv = 1.0;
}

__global__ void k_sub_elementwise(const float* a, const float* b, float* c, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = __fsub_rn(a[i], b[i]);
}

// A8 over row-major n x K scores (CrossEntropyLoss::compute_multiclass_gradients,
// loss.cpp:287-318), the same arithmetic as the trainer's softmax rows.
__global__ void k_gradients_multi(const float* y, const float* p, float* g, float* h, uint32_t n,
                                  uint32_t K) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = p + static_cast<size_t>(i) * K;
  float e[kMaxClasses];
  float m = -INFINITY;
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    e[c] = r[c];
    m = fmaxf(m, e[c]);
  }
  float sum = 0.f;
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    e[c] = ref_expf(__fsub_rn(e[c], m));
    sum = __fadd_rn(sum, e[c]);
  }
  const uint32_t lab = static_cast<uint32_t>(y[i]);
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    const float pc = __fdiv_rn(e[c], sum);
    g[static_cast<size_t>(i) * K + c] = __fsub_rn(pc, c == lab ? 1.0f : 0.0f);
    const float hh = __fmul_rn(pc, __fsub_rn(1.0f, pc));
    h[static_cast<size_t>(i) * K + c] = hh > 1e-6f ? hh : 1e-6f;
  }
}

__global__ void k_gradients(int loss, const float* y, const float* p, float* g, float* h, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = row_gradient(loss, p[i], y[i], i < (n / 8u) * 8u);
  g[i] = v.x;
  h[i] = v.y;
}

}  // namespace

// Booster::next_random, xorshift64* (booster.cpp:1219-1225).
uint64_t next_random(uint64_t& state) {
  state ^= state >> 12;
  state ^= state << 25;
  state ^= state >> 27;
  return state * 0x2545F4914F6CDD1DULL;
}

// Dataset::random_subsample (dataset.cpp:552-562): the same engine and the
// same std::shuffle, so the same rows in the same order as the reference.
// The shuffle runs in place in out (n slots); the sample is its first k.
uint32_t random_subsample(uint32_t n, float ratio, uint64_t seed, uint32_t* out) {
  const uint32_t k = static_cast<uint32_t>(static_cast<float>(n) * ratio);
  std::iota(out, out + n, 0u);
  std::mt19937_64 rng(seed);
  std::shuffle(out, out + n, rng);
  return std::min(k, n);
}

// Plain row samples (Dataset::random_subsample) of the coming trees, drawn
// ahead: a ring of slots, each tree's shuffle on a host thread of its own
// (the seeds are taken in tree order, xorshift64* as Booster::next_random,
// with the feature-subset draw after each when colsample is on, so every tree
// draws the reference's rows).  A slot holds the sample as a bitmask of the
// drawn rows (bit r of word r / 32) preceded by the rows drawn before each
// chunk of kMaskChunkWords words; the device makes the tree's root list from
// it (k_rows_from_mask*).  Only the set of rows is the reference's: the list
// comes out in ascending row order, which changes no histogram (exact integer
// sums) and makes the root's reads sequential.  The draw threads make no HIP
// call: one made while the trainer captures a graph would void the capture.
class AheadSampler {
 public:
  AheadSampler() = default;
  AheadSampler(const AheadSampler&) = delete;
  AheadSampler& operator=(const AheadSampler&) = delete;
  ~AheadSampler() { release(); }

  static size_t mask_words(uint32_t N) {
    const size_t W = (N + 31u) / 32u;
    return W + (W + kMaskChunkWords - 1u) / kMaskChunkWords;
  }
  static uint32_t mask_chunks(uint32_t N) {
    return ((N + 31u) / 32u + kMaskChunkWords - 1u) / kMaskChunkWords;
  }

  // ring: slots (trees drawn ahead, one host thread each)
  void init(uint32_t N, float ratio, uint64_t seed, bool colsample, size_t ring) {
    N_ = N;
    ratio_ = ratio;
    state_ = seed;
    colsample_ = colsample;
    host_.assign(ring, nullptr);
    ev_.assign(ring, nullptr);
    draws_.resize(ring);
    scratch_.assign(ring, std::vector<uint32_t>(std::max(1u, N)));
    for (size_t b = 0; b < ring; ++b) {
      TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_[b]), 4ull * mask_words(N),
                           hipHostMallocDefault));
      TC_HIP(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming));
    }
  }

  // The next tree's bitmask, copied to d_mask on stream s.
  void next(uint32_t* d_mask, hipStream_t s) {
    while (submitted_ < prepared_ + host_.size()) submit();
    const size_t b = prepared_ % host_.size();
    draws_[b].get();  // rethrows a failed draw
    ++prepared_;
    TC_HIP(hipMemcpyAsync(d_mask, host_[b], 4ull * mask_words(N_), hipMemcpyHostToDevice, s));
    TC_HIP(hipEventRecord(ev_[b], s));
  }

  void release() {
    for (auto& d : draws_)  // draws still running write into the buffers
      if (d.valid()) d.wait();
    draws_.clear();
    for (size_t b = 0; b < host_.size(); ++b) {
      if (host_[b]) hfree(host_[b]);
      if (ev_[b]) (void)hipEventDestroy(ev_[b]);
    }
    host_.clear();
    ev_.clear();
    scratch_.clear();
  }

 private:
  // Starts the next tree's shuffle into its slot, once the copy of the tree
  // that used the slot before (a ring's length of trees ago) has left it.
  void submit() {
    const uint64_t seed = next_random(state_);
    if (colsample_) (void)next_random(state_);  // feature subset: drawn, unused
    const size_t b = submitted_ % host_.size();
    TC_HIP(hipEventSynchronize(ev_[b]));
    uint32_t* out = host_[b];
    uint32_t* sc = scratch_[b].data();
    const uint32_t N = N_;
    const float ratio = ratio_;
    draws_[b] = std::async(std::launch::async, [=] {
      const uint32_t k = random_subsample(N, ratio, seed, sc);
      const uint32_t W = (N + 31u) / 32u, nch = mask_chunks(N);
      uint32_t* mask = out + nch;
      std::memset(mask, 0, 4ull * W);
      for (uint32_t i = 0; i < k; ++i) mask[sc[i] >> 5] |= 1u << (sc[i] & 31u);
      uint32_t run = 0;  // rows drawn before each chunk
      for (uint32_t c = 0; c < nch; ++c) {
        out[c] = run;
        const uint32_t w1 = std::min(W, (c + 1u) * kMaskChunkWords);
        for (uint32_t w = c * kMaskChunkWords; w < w1; ++w)
          run += static_cast<uint32_t>(__builtin_popcount(mask[w]));
      }
    });
    ++submitted_;
  }

  uint32_t N_ = 0;
  float ratio_ = 1.f;
  uint64_t state_ = 0;
  bool colsample_ = false;
  std::vector<uint32_t*> host_;
  std::vector<hipEvent_t> ev_;
  std::vector<std::future<void>> draws_;
  std::vector<std::vector<uint32_t>> scratch_;
  uint64_t submitted_ = 0, prepared_ = 0;  // trees drawn / copied
};

// Slots of an AheadSampler: up to 12 trees ahead (each shuffle and bitmask
// about 3 ms for 1M rows; 1M x 100, subsample 0.8: 4 slots 1,388, 8
// 2,239-2,513, 12 2,984 trees/s), at most 256 MB of shuffle scratch over
// `jobs` samplers of N rows.
inline size_t ahead_ring(uint32_t N, uint32_t jobs) {
  const uint64_t per = 4ull * std::max(1u, N) * std::max(1u, jobs);
  return static_cast<size_t>(std::clamp<uint64_t>((uint64_t(256) << 20) / per, 2, 12));
}

// Rows Dataset::goss_subsample (dataset.cpp:494-550) keeps for one tree: the
// uint32(n * top_rate) rows of largest |g| -- std::nth_element over (|g|, row)
// pairs ordered by |g| alone, so equal |g| fall where the library puts them --
// then min(uint32(n * other_rate), n - top) rows of the rest by a partial
// Fisher-Yates draw from std::mt19937_64.  The same library calls on the same
// array in the same order as the reference: the same rows in the same order.
// gi: the (|g|, row) pairs' storage, kept across trees.
uint32_t goss_select(const float* g, uint32_t n, float top_rate, float other_rate, uint64_t seed,
                     uint32_t* out, std::vector<std::pair<float, uint32_t>>& gi) {
  const uint32_t top_n = static_cast<uint32_t>(static_cast<float>(n) * top_rate);
  uint32_t other_n = static_cast<uint32_t>(static_cast<float>(n) * other_rate);
  if (top_n == 0 && other_n == 0) return 0;
  gi.resize(n);
  for (uint32_t i = 0; i < n; ++i) gi[i] = {std::abs(g[i]), i};
  if (top_n > 0 && top_n < n)
    std::nth_element(gi.begin(), gi.begin() + top_n, gi.end(),
                     [](const std::pair<float, uint32_t>& a, const std::pair<float, uint32_t>& b) {
                       return a.first > b.first;
                     });
  uint32_t k = 0;
  for (uint32_t i = 0; i < top_n; ++i) out[k++] = gi[i].second;
  if (other_n > 0 && top_n < n) {
    const uint32_t rest = n - top_n;
    other_n = std::min(other_n, rest);
    std::mt19937_64 rng(seed);
    for (uint32_t i = 0; i < other_n; ++i) {
      std::uniform_int_distribution<uint32_t> dist(i, rest - 1);
      const uint32_t j = dist(rng);
      if (i != j) std::swap(gi[top_n + i], gi[top_n + j]);
      out[k++] = gi[top_n + i].second;
    }
  }
  return k;
}

uint32_t goss_subsample(const float* g, uint32_t n, float top_rate, float other_rate,
                        uint64_t seed, uint32_t* out) {
  std::vector<std::pair<float, uint32_t>> gi;
  return goss_select(g, n, top_rate, other_rate, seed, out, gi);
}

// Rows goss_subsample returns for n rows (independent of the gradients).
uint32_t goss_rows(uint32_t n, float top_rate, float other_rate) {
  const uint32_t top_n = static_cast<uint32_t>(static_cast<float>(n) * top_rate);
  const uint32_t other_n = static_cast<uint32_t>(static_cast<float>(n) * other_rate);
  if (top_n == 0 && other_n == 0) return 0;
  return top_n + (top_n < n ? std::min(other_n, n - top_n) : 0u);
}

// Rows Dataset::mvs_subsample (dataset.cpp:576-673) keeps for one tree:
// target = max(1, uint32(n * ratio)); a floor t = max(the |g| of rank
// uint32((1 - ratio) * n) in ascending order, max |g| / 100, 1e-10); row i
// kept when a uniform float draw is below min(1, s * max(|g_i|, t) / S * n),
// S the float sum of the floored scores and s = target / S * S / n, in the
// reference's float operation order; if fewer than uint32(target * 0.8f)
// rows were kept, the rest are shuffled with the same engine and the first
// ones appended up to target.  The same library calls (std::mt19937_64,
// std::uniform_real_distribution<float>, std::shuffle) on the same values in
// the same order give the same rows in the same order.  S is summed in row
// order: the reference's own OpenMP reduction adds per-thread partial sums,
// which equals this order on one thread.
//
// The draws depend on the seed alone (one per row, in row order), so they are
// made apart from the selection (MvsDraws), ahead of the tree that uses them
// (MvsAhead); the rank's value is the one std::nth_element places there,
// found here by a radix select (the value of a rank does not depend on how
// it is found).
struct MvsDraws {
  std::vector<float> u;  // row i's uniform draw
  std::mt19937_64 after;  // the engine once the n draws are made
};

inline void mvs_draw(uint64_t seed, uint32_t n, MvsDraws* d) {
  d->u.resize(n);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> uniform(0.0f, 1.0f);
  for (uint32_t i = 0; i < n; ++i) d->u[i] = uniform(rng);
  d->after = rng;
}

// f(begin, end, part) over `parts` contiguous row ranges, on host threads
// (the calling thread takes part 0).
template <class Fn>
void row_parts(uint32_t n, uint32_t parts, Fn f) {
  auto lo = [&](uint32_t t) { return static_cast<uint32_t>(static_cast<uint64_t>(n) * t / parts); };
  std::vector<std::future<void>> fs;
  for (uint32_t t = 1; t < parts; ++t) fs.push_back(std::async(std::launch::async, f, lo(t), lo(t + 1), t));
  f(lo(0), lo(1), 0u);
  for (auto& x : fs) x.get();
}

struct MvsScratch {
  std::vector<float> a, sc;
  std::vector<uint32_t> cnt;               // [part][65536] radix counts
  std::vector<std::vector<uint32_t>> kept;  // per part: its kept rows
  std::vector<float> gmax;
};

// The value of ascending rank k (0-based) among n >= k + 1 non-negative
// floats: two 16-bit counting passes over their bit patterns, which order as
// the values do (counts per row range, summed).
inline float kth_nonneg(const float* a, uint32_t n, uint32_t k, uint32_t parts, MvsScratch& s) {
  s.cnt.assign(static_cast<size_t>(parts) * 65536u, 0u);
  row_parts(n, parts, [&](uint32_t b, uint32_t e, uint32_t t) {
    uint32_t* c = s.cnt.data() + static_cast<size_t>(t) * 65536u;
    for (uint32_t i = b; i < e; ++i) ++c[__builtin_bit_cast(uint32_t, a[i]) >> 16];
  });
  auto total = [&](uint32_t v) {
    uint32_t x = 0;
    for (uint32_t t = 0; t < parts; ++t) x += s.cnt[static_cast<size_t>(t) * 65536u + v];
    return x;
  };
  uint32_t hi = 0, below = 0;
  for (uint32_t c = total(0); below + c <= k; c = total(hi)) {
    below += c;
    ++hi;
  }
  s.cnt.assign(static_cast<size_t>(parts) * 65536u, 0u);
  row_parts(n, parts, [&](uint32_t b, uint32_t e, uint32_t t) {
    uint32_t* c = s.cnt.data() + static_cast<size_t>(t) * 65536u;
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t v = __builtin_bit_cast(uint32_t, a[i]);
      if ((v >> 16) == hi) ++c[v & 0xFFFFu];
    }
  });
  uint32_t lo = 0;
  for (uint32_t c = total(0); below + c <= k; c = total(lo)) {
    below += c;
    ++lo;
  }
  return __builtin_bit_cast(float, (hi << 16) | lo);
}

// The passes whose result does not depend on the order rows are visited in
// (|g| and its maximum, the rank's value, the floored scores, each row's
// keep test) run over row ranges on host threads; the score sum is one
// pass in row order, as the reference's float sum is taken.
uint32_t mvs_select(const float* g, uint32_t n, float ratio, const MvsDraws& d, uint32_t* out,
                    MvsScratch& s) {
  if (n == 0 || !(ratio > 0.0f)) return 0;
  const uint32_t parts = std::clamp(n / 131072u, 1u, 8u);
  const uint32_t target = std::max(static_cast<uint32_t>(static_cast<float>(n) * ratio), 1u);
  s.a.resize(n);
  s.sc.resize(n);
  s.gmax.assign(parts, 0.0f);
  float* a = s.a.data();
  float* sc = s.sc.data();
  row_parts(n, parts, [&](uint32_t b, uint32_t e, uint32_t t) {
    float m = 0.0f;
    for (uint32_t i = b; i < e; ++i) {
      a[i] = std::abs(g[i]);
      m = std::max(m, a[i]);
    }
    s.gmax[t] = m;
  });
  float gmax = 0.0f;
  for (float m : s.gmax) gmax = std::max(gmax, m);
  const uint32_t ti =
      std::min(static_cast<uint32_t>((1.0f - ratio) * static_cast<float>(n)), n - 1);
  float thr = std::max(kth_nonneg(a, n, ti, parts, s), gmax * 0.01f);
  if (thr < 1e-10f) thr = 1e-10f;
  row_parts(n, parts, [&](uint32_t b, uint32_t e, uint32_t) {
    for (uint32_t i = b; i < e; ++i) sc[i] = std::max(a[i], thr);
  });
  float ssum = 0.0f;
  for (uint32_t i = 0; i < n; ++i) ssum += sc[i];
  const float fn = static_cast<float>(n);
  const float scale = static_cast<float>(target) / ssum * ssum / fn;
  const float* u = d.u.data();
  s.kept.resize(parts);
  row_parts(n, parts, [&](uint32_t b, uint32_t e, uint32_t t) {
    std::vector<uint32_t>& kt = s.kept[t];
    kt.clear();
    for (uint32_t i = b; i < e; ++i) {
      const float prob = std::min(1.0f, scale * sc[i] / ssum * fn);
      if (u[i] < prob) kt.push_back(i);
    }
  });
  uint32_t k = 0;
  for (const auto& kt : s.kept) {  // ranges in row order: rows ascending
    std::copy(kt.begin(), kt.end(), out + k);
    k += static_cast<uint32_t>(kt.size());
  }
  if (k < static_cast<uint32_t>(static_cast<float>(target) * 0.8f)) {
    std::mt19937_64 rng = d.after;
    std::vector<bool> kept(n, false);
    for (uint32_t i = 0; i < k; ++i) kept[out[i]] = true;
    std::vector<uint32_t> rest;
    rest.reserve(n - k);
    for (uint32_t i = 0; i < n; ++i)
      if (!kept[i]) rest.push_back(i);
    std::shuffle(rest.begin(), rest.end(), rng);
    const uint32_t need = target - k;
    for (uint32_t i = 0; i < need && i < rest.size(); ++i) out[k++] = rest[i];
  }
  return k;
}

uint32_t mvs_subsample(const float* g, uint32_t n, float ratio, uint64_t seed, uint32_t* out) {
  if (n == 0 || !(ratio > 0.0f)) return 0;
  MvsDraws d;
  mvs_draw(seed, n, &d);
  MvsScratch s;
  return mvs_select(g, n, ratio, d, out, s);
}

// MVS draws made ahead on host threads (no HIP call there), a ring of trees
// at a time; each tree's seed is taken in tree order from the booster's
// engine state, with the feature-subset draw after it as the training loop
// takes it (booster.cpp:200-226).
class MvsAhead {
 public:
  MvsAhead() = default;
  MvsAhead(const MvsAhead&) = delete;
  MvsAhead& operator=(const MvsAhead&) = delete;
  ~MvsAhead() { release(); }

  void init(uint32_t N, uint64_t state, bool colsample, size_t ring) {
    N_ = N;
    state_ = state;
    colsample_ = colsample;
    slots_ = std::vector<MvsDraws>(ring);
    futs_.resize(ring);
  }

  // The next tree's draws (valid until the call after next).
  const MvsDraws& next() {
    while (submitted_ < prepared_ + slots_.size()) submit();
    const size_t b = prepared_ % slots_.size();
    futs_[b].get();  // rethrows a failed draw
    ++prepared_;
    return slots_[b];
  }

  void release() {
    for (auto& f : futs_)
      if (f.valid()) f.wait();
    futs_.clear();
    slots_.clear();
  }

 private:
  // the slot handed out by the previous next() is free again
  void submit() {
    const uint64_t seed = next_random(state_);
    if (colsample_) (void)next_random(state_);  // feature subset: drawn, unused
    const size_t b = submitted_ % slots_.size();
    MvsDraws* d = &slots_[b];
    const uint32_t N = N_;
    futs_[b] = std::async(std::launch::async, [=] { mvs_draw(seed, N, d); });
    ++submitted_;
  }

  uint32_t N_ = 0;
  uint64_t state_ = 0;
  bool colsample_ = false;
  std::vector<MvsDraws> slots_;
  std::vector<std::future<void>> futs_;
  size_t submitted_ = 0, prepared_ = 0;
};

struct JobSpec {
  const tc_dataset* ds;
  const float* y_host;   // labels (may be null for teacher forcing)
  tc_train_params p;
  float base;
  const tc_dataset* valid = nullptr;  // validation rows, binned with the training edges
  const float* y_valid = nullptr;     // their labels (host)
  uint32_t es_rounds = 0;             // early stopping rounds (with a validation set)
  bool history = false;               // keep the training loss after every tree
};

// Gradient channels of a job: K classes for softmax, else 1.
uint32_t channels(const tc_train_params& p) {
  return p.loss == TC_LOSS_SOFTMAX ? p.n_classes : 1u;
}

class Trainer {
 public:
  Trainer(tc_ctx* ctx, const std::vector<JobSpec>& specs) : ctx_(ctx), specs_(specs) {
    njobs_ = static_cast<uint32_t>(specs.size());
    TC_REQUIRE(njobs_ > 0, "no jobs");
    D_ = specs[0].p.tree.max_depth;
    TC_REQUIRE(D_ >= 1 && D_ <= static_cast<uint32_t>(kMaxDepth), "max_depth must be in [1, 10]");
    n_trees_ = specs[0].p.n_trees;
    for (const auto& s : specs) {
      TC_REQUIRE(s.p.tree.max_depth == D_, "all jobs of a batch share max_depth");
      TC_REQUIRE(s.p.n_trees == n_trees_, "all jobs of a batch share n_trees");
      TC_REQUIRE(s.p.loss == TC_LOSS_MSE || s.p.loss == TC_LOSS_LOGLOSS ||
                     s.p.loss == TC_LOSS_SOFTMAX, "unsupported loss");
      TC_REQUIRE((s.p.tree.max_bins & 0xFFu) >= 2, "max_bins must be in [2, 255]");
      fmax_ = std::max(fmax_, s.ds->F);
      kmax_ = std::max(kmax_, channels(s.p));
    }
    try {
      setup();
    } catch (...) {
      release();
      throw;
    }
  }
  ~Trainer() { release(); }

  // Row sample of the next tree for every sampling job, drawn on the host.  A
  // plain row sample is shuffled ahead of time (a ring of up to 12 trees) and
  // copied as a bitmask that the tree's first launch turns into rows[0]; GOSS
  // and MVS select on the last tree's gradients into two alternating pinned
  // buffers, copied into rows[0] on the stream.
  void prepare_samples() {
    for (uint32_t j = 0; j < njobs_; ++j) {
      Sampler& S = samplers_[j];
      if (!S.active && !S.colsample) continue;
      const uint32_t N = h_jobs_[j].N;
      if (S.plain) {  // the shuffle needs no gradients: drawn ahead on host threads
        // the sample's bitmask; the tree's first launch makes the root list from it
        S.plain->next(const_cast<uint32_t*>(h_jobs_[j].smask), ctx_->stream);
        continue;
      }
      uint64_t seed = 0;
      const MvsDraws* md = nullptr;
      if (S.mvs) {  // drawn ahead: the seed and the feature-subset draw are taken there
        md = &S.mvs_ahead->next();
      } else {
        if (S.active) seed = next_random(S.state);
        if (S.colsample) (void)next_random(S.state);  // feature subset: drawn, unused
      }
      if (!S.active) continue;
      const int b = S.next;
      S.next ^= 1;
      TC_HIP(hipEventSynchronize(S.ev[b]));
      // GOSS / MVS: the gradients the next tree is grown on (every queued tree first)
      TC_HIP(hipMemcpyAsync(S.gh_host, h_jobs_[j].gh, sizeof(float2) * N, hipMemcpyDeviceToHost,
                            ctx_->stream));
      TC_HIP(hipStreamSynchronize(ctx_->stream));
      for (uint32_t i = 0; i < N; ++i) S.g[i] = S.gh_host[i].x;
      if (S.goss) {
        goss_select(S.g.data(), N, S.top_rate, S.other_rate, seed, S.host[b], S.goss_pairs);
      } else {  // a sample of its own size each tree: the device reads n_root from the job table
        h_jobs_[j].n_root = S.n_root_host[b] = mvs_select(S.g.data(), N, S.ratio, *md, S.host[b],
                                                          S.mvs_scratch);
        TC_HIP(hipMemcpyAsync(reinterpret_cast<char*>(d_jobs_ + j) + offsetof(JobDev, n_root),
                              &S.n_root_host[b], sizeof(uint32_t), hipMemcpyHostToDevice,
                              ctx_->stream));
      }
      TC_HIP(hipMemcpyAsync(h_jobs_[j].rows, S.host[b], 4ull * h_jobs_[j].n_root,
                            hipMemcpyHostToDevice, ctx_->stream));
      TC_HIP(hipEventRecord(S.ev[b], ctx_->stream));
    }
  }

  bool sampling() const {
    for (const auto& S : samplers_) if (S.active || S.colsample) return true;
    return false;
  }

  void init_gradients() {
    hipLaunchKernelGGL(k_grad_update, dim3(total_grad_wg_), dim3(256), leaf_lds(), ctx_->stream,
                       d_jobs_, d_job_wg_off_, njobs_, 0, lists_);
    TC_HIP(hipGetLastError());
  }

  // Teacher forcing: job 0's gradients, class-major K x N.
  void set_gradients(const float* g, const float* h) {
    const uint32_t N = h_jobs_[0].N * h_jobs_[0].K;
    float* dg = dmalloc<float>(N);
    float* dh = dmalloc<float>(N);
    TC_HIP(hipMemcpyAsync(dg, g, sizeof(float) * N, hipMemcpyHostToDevice, ctx_->stream));
    TC_HIP(hipMemcpyAsync(dh, h, sizeof(float) * N, hipMemcpyHostToDevice, ctx_->stream));
    // the scales come from these gradients alone: clear the gradient update's tile maxima
    TC_HIP(hipMemsetAsync(h_jobs_[0].ghmax_partials, 0, 8ull * h_jobs_[0].grad_wg_cnt, ctx_->stream));
    hipLaunchKernelGGL(k_set_gh, dim3(std::max(1u, ceil_div(N, 256))), dim3(256), 0,
                       ctx_->stream, d_jobs_, dg, dh);
    forced_ = true;  // arbitrary h: no packed histogram cells
    // the root partials now come from these gradients, not from a gradient update
    h_jobs_[0].fused_root = 0;
    root_partials_ = true;
    TC_HIP(hipMemcpyAsync(d_jobs_, h_jobs_.data(), sizeof(JobDev) * njobs_, hipMemcpyHostToDevice,
                          ctx_->stream));
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(ctx_->stream));
    dfree(dg);
    dfree(dh);
  }

  // The per-tree launch sequence.  hist_events (optional) brackets every
  // histogram launch for live timing.
  void enqueue_tree(std::vector<std::pair<hipEvent_t, hipEvent_t>>* hist_events) {
    hipStream_t s = ctx_->stream;
    const bool packed = packed_hist();
    if (mask_chunks_)  // plain row samples: the root lists from the copied bitmasks
      hipLaunchKernelGGL(k_rows_from_mask, dim3(mask_chunks_, njobs_), dim3(kMaskChunkWords), 0, s,
                         d_jobs_);
    if (root_partials_)  // otherwise the gradient update wrote every job's root partials
      hipLaunchKernelGGL(k_root_partials, dim3(total_grad_wg_), dim3(256), 0, s, d_jobs_,
                         d_job_wg_off_, njobs_);
    hipLaunchKernelGGL(k_tree_begin, dim3(1, njobs_), dim3(256), 0, s, d_jobs_, lists_,
                       packed ? 1 : 0);
    for (uint32_t d = 0; d < D_; ++d) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (hist_events) {
        TC_HIP(hipEventCreate(&e0));
        TC_HIP(hipEventCreate(&e1));
        TC_HIP(hipEventRecord(e0, s));
      }
      {
        // one job: below the root the histogram launch plans its own level
        const dim3 grid(round8(hist_cap_[d]));
        const bool one = njobs_ == 1, plan = one && d > 0;
        auto* k = packed ? (plan ? k_hist_plan<kPlanUnroll, 1>
                                 : one ? k_hist<kHistUnroll, 1, true>
                                       : k_hist<kHistUnroll, 1, false>)
                         : (plan ? k_hist_plan<kPlanUnroll, 0>
                                 : one ? k_hist<kHistUnroll, 0, true>
                                       : k_hist<kHistUnroll, 0, false>);
        hipLaunchKernelGGL(k, grid, dim3(kHistThreads), 0, s, d_jobs_, lists_.hist[d],
                           lists_.hist_count[d], d, partials_, lists_);
      }
      if (hist_events) {
        TC_HIP(hipEventRecord(e1, s));
        hist_events->emplace_back(e0, e1);
      }
      // a batch: 512 threads, two workgroups per CU (a pass then holds 16 units: at most 8
      // channels)
      if (njobs_ == 1 || kmax_ > 8u)
        hipLaunchKernelGGL((k_split_scan<true, 1024>), dim3(fmax_, njobs_), dim3(1024), 0, s,
                           d_jobs_, static_cast<const float2*>(nullptr), d, partials_);
      else
        hipLaunchKernelGGL((k_split_scan<true, 512>), dim3(fmax_, njobs_), dim3(512), 0, s,
                           d_jobs_, static_cast<const float2*>(nullptr), d, partials_);
      if (d + 1 < D_ || keep_rows_) {  // training never reads the last level's rows
        // the partition also finalizes the level (split record, child totals)
        // the work items, then one workgroup per (job, node) for the child totals
        hipLaunchKernelGGL(njobs_ == 1 ? k_partition<true> : k_partition<false>,
                           dim3(part_cap_[d] + (njobs_ << d)), dim3(kPartThreads), 0, s, d_jobs_,
                           lists_.part[d], lists_.part_count[d], d, part_cap_[d]);
        if (njobs_ > 1 || d + 1 == D_)  // otherwise the next histogram launch plans the level
          hipLaunchKernelGGL(k_part_plan, dim3(1, njobs_), dim3(256), 0, s, d_jobs_, lists_, d);
      } else {
        // one wave per (node, channel) unit of the widest job
        hipLaunchKernelGGL(k_finalize, dim3(ceil_div(static_cast<uint64_t>(kmax_) << d, 4), njobs_),
                           dim3(256), 0, s, d_jobs_, d);
      }
    }
    hipLaunchKernelGGL(k_grad_update, dim3(total_grad_wg_), dim3(256), leaf_lds(), s, d_jobs_,
                       d_job_wg_off_, njobs_, 1, lists_);
    if (eval_) hipLaunchKernelGGL(k_eval, dim3(1, njobs_), dim3(256), 0, s, d_jobs_);
    TC_HIP(hipGetLastError());
  }

  // Enqueue n more trees as graph replays (captured on first use).
  void step_graph(uint32_t n) {
    hipStream_t s = ctx_->stream;
    uint32_t built = 0;
    if (!graph_exec_) {
      built = autotune(n);
      hipGraph_t g;
      TC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      enqueue_tree(nullptr);
      TC_HIP(hipStreamEndCapture(s, &g));
      const hipError_t e = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      TC_HIP(e);
    }
    const bool smp = sampling();
    for (uint32_t t = built; t < n; ++t) {
      if (smp) prepare_samples();
      TC_HIP(hipGraphLaunch(graph_exec_, s));
    }
  }

  // Enqueue n trees without a graph, timing every histogram launch.
  void step_timed(uint32_t n, double* hist_ms, uint32_t* launches) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> hev;
    for (uint32_t t = 0; t < n; ++t) {
      if (sampling()) prepare_samples();
      enqueue_tree(&hev);
    }
    TC_HIP(hipStreamSynchronize(ctx_->stream));
    double hs = 0.0;
    for (auto& e : hev) {
      float x = 0.f;
      TC_HIP(hipEventElapsedTime(&x, e.first, e.second));
      hs += x;
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    if (hist_ms) *hist_ms = hs;
    if (launches) *launches = static_cast<uint32_t>(hev.size());
  }

  // Run all trees; returns loop time (ms) and histogram time (ms).
  // Early stopping: every 16 trees the jobs' stop flags are copied to pinned
  // memory; once a copy shows every job stopped, no more trees are enqueued
  // (the trees enqueued meanwhile do nothing on the device).
  struct StopPoll {
    uint32_t* flags = nullptr;
    hipEvent_t ev = nullptr;
    bool pending = false;
    StopPoll() = default;
    StopPoll(const StopPoll&) = delete;
    StopPoll& operator=(const StopPoll&) = delete;
    ~StopPoll() {
      if (ev) (void)hipEventSynchronize(ev);
      if (flags) hfree(flags);
      if (ev) (void)hipEventDestroy(ev);
    }
  };
  bool validating() const {
    for (const JobDev& J : h_jobs_) if (J.Nv == 0) return false;
    return !h_jobs_.empty();
  }
  bool poll_stopped(StopPoll& P, uint32_t t) {
    if (!validating() || t % 16 != 0) return false;
    hipStream_t s = ctx_->stream;
    if (!P.flags) {
      TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&P.flags), 4ull * njobs_, hipHostMallocDefault));
      TC_HIP(hipEventCreateWithFlags(&P.ev, hipEventDisableTiming));
    }
    if (P.pending && hipEventQuery(P.ev) == hipSuccess) {
      P.pending = false;
      bool all = true;
      for (uint32_t j = 0; j < njobs_; ++j) all = all && P.flags[j] != 0;
      if (all) return true;
    }
    if (!P.pending) {
      for (uint32_t j = 0; j < njobs_; ++j)
        TC_HIP(hipMemcpyAsync(P.flags + j,
                              reinterpret_cast<const char*>(h_jobs_[j].st) + offsetof(TreeState, stopped),
                              4, hipMemcpyDeviceToHost, s));
      TC_HIP(hipEventRecord(P.ev, s));
      P.pending = true;
    }
    return false;
  }
  void run(bool use_graph, bool time_hist, double* loop_ms, double* hist_ms) {
    hipStream_t s = ctx_->stream;
    StopPoll poll;
    hipEvent_t a, b;
    TC_HIP(hipEventCreate(&a));
    TC_HIP(hipEventCreate(&b));
    std::vector<std::pair<hipEvent_t, hipEvent_t>> hev;
    if (use_graph && !time_hist) {
      TC_HIP(hipEventRecord(a, s));
      const uint32_t built = autotune(n_trees_);
      hipGraph_t g;
      TC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      enqueue_tree(nullptr);
      TC_HIP(hipStreamEndCapture(s, &g));
      hipGraphExec_t ge;
      TC_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      TC_HIP(hipGraphDestroy(g));
      const bool smp = sampling();
      for (uint32_t t = built; t < n_trees_; ++t) {
        if (poll_stopped(poll, t)) break;
        if (smp) prepare_samples();
        TC_HIP(hipGraphLaunch(ge, s));
      }
      TC_HIP(hipEventRecord(b, s));
      TC_HIP(hipStreamSynchronize(s));
      TC_HIP(hipGraphExecDestroy(ge));
    } else {
      TC_HIP(hipEventRecord(a, s));
      for (uint32_t t = 0; t < n_trees_; ++t) {
        if (poll_stopped(poll, t)) break;
        if (sampling()) prepare_samples();
        enqueue_tree(time_hist ? &hev : nullptr);
      }
      TC_HIP(hipEventRecord(b, s));
      TC_HIP(hipStreamSynchronize(s));
    }
    float ms = 0.f;
    TC_HIP(hipEventElapsedTime(&ms, a, b));
    if (loop_ms) *loop_ms = ms;
    double hs = 0.0;
    for (auto& e : hev) {
      float x = 0.f;
      TC_HIP(hipEventElapsedTime(&x, e.first, e.second));
      hs += x;
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    if (hist_ms) *hist_ms = hs;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    TC_HIP(hipGetLastError());
  }

  // ---- histogram work-item shapes and their autotuner.
  // Candidate shapes: root items per job (each of N x tiles x channels rows
  // of work), and the items each deeper level is cut into (k_part_plan sizes
  // that level's items from its rows).  The rows of a work item never change
  // a result (every histogram is an exact integer sum and the fixed-point
  // scale does not depend on the item size), so the tuner times the fit's
  // own first trees.
  // The first candidate is the shape used when a fit has too few trees to
  // tune: two root items per CU and one item per CU below the root (256 CUs),
  // the fastest shape measured on the 1M x 100 LogLoss workload (a level waits
  // for its CUs that take a second item).
  struct TuneChoice {
    uint32_t root_items, deep_items;
  };
  static constexpr TuneChoice kTune[8] = {{512, 256},  {512, 512},  {1024, 512}, {512, 128},
                                          {1024, 256}, {512, 1024}, {1024, 1024}, {2048, 1024}};

  // Upper bound of job J's histogram work items at level d.  Below the root
  // the histogrammed (smaller) children hold at most N/2 rows over 2^(d-1)
  // pairs, and each pair adds at most one partial chunk per tile-channel.
  static uint32_t hist_items(const JobDev& J, uint32_t d) {
    const uint32_t TK = J.T * J.K;
    if (d == 0) return ceil_div(J.N, J.hist_chunk) * TK;
    const uint32_t pairs = 1u << (d - 1);
    if (!J.deep_items) return (ceil_div(J.N / 2, J.hist_chunk) + pairs) * TK;
    // k_part_plan's chunk: at most deep_items / TK pieces when a candidate
    // fits; otherwise a chunk >= rows / (deep_items / TK) or 16384 (rows <=
    // N/2), with at most rows / chunk + 1 pieces per pair
    const uint32_t full = std::max((J.deep_items + 1) / TK + 2, ceil_div(J.N / 2, 16384) + 1);
    return (full + pairs) * TK;
  }

  void set_chunks(uint32_t j, const TuneChoice& c) {
    JobDev& J = h_jobs_[j];
    if (fixed_chunk_[j]) {
      J.hist_chunk = fixed_chunk_[j];
      J.deep_items = specs_[j].p.hist_deep_items;
      return;
    }
    J.hist_chunk = root_chunk(J.N, J.T * J.K, c.root_items);
    J.deep_items = c.deep_items;
  }

  // Launch grids of k_hist for the current shapes (<= the allocated lists).
  void recompute_caps() {
    for (uint32_t d = 0; d < D_; ++d) {
      uint32_t c = 0;
      for (uint32_t j = 0; j < njobs_; ++j) c += hist_items(h_jobs_[j], d);
      hist_cap_[d] = c;
    }
  }

  void choose(uint32_t c) {
    for (uint32_t j = 0; j < njobs_; ++j) set_chunks(j, kTune[c]);
    recompute_caps();
    TC_HIP(hipMemcpyAsync(d_jobs_, h_jobs_.data(), sizeof(JobDev) * njobs_, hipMemcpyHostToDevice,
                          ctx_->stream));
    choice_ = c;
  }

  // Times one tree per candidate shape (two each when the budget allows),
  // after one warm-up tree, and keeps the fastest per histogrammed row (trees
  // grown on samples of different sizes, as under MVS, compare fairly).
  // Returns the trees built.
  uint32_t autotune(uint32_t budget) {
    constexpr uint32_t kN = sizeof(kTune) / sizeof(kTune[0]);
    if (!tune_ || tuned_) return 0;
    const uint32_t reps = budget >= 2 * kN + 1 ? 2u : 1u;
    if (budget < kN * reps + 1) return 0;  // too few trees: keep the default shape
    hipStream_t s = ctx_->stream;
    std::vector<hipEvent_t> ev(2 * kN * reps);
    for (auto& e : ev) TC_HIP(hipEventCreate(&e));
    // every job's histogrammed-row counter after each timed tree (and before the first)
    const uint32_t nsnap = kN * reps + 1;
    unsigned long long* snap = dmalloc<unsigned long long>(static_cast<size_t>(nsnap) * njobs_);
    auto snapshot = [&](uint32_t k) {
      for (uint32_t j = 0; j < njobs_; ++j)
        TC_HIP(hipMemcpyAsync(snap + static_cast<size_t>(k) * njobs_ + j, hist_rows_[j], 8,
                              hipMemcpyDeviceToDevice, s));
    };
    try {
      if (sampling()) prepare_samples();
      enqueue_tree(nullptr);
      snapshot(0);
      for (uint32_t r = 0, k = 0; r < reps; ++r)
        for (uint32_t c = 0; c < kN; ++c, k += 2) {
          choose(c);
          if (sampling()) prepare_samples();
          TC_HIP(hipEventRecord(ev[k], s));
          enqueue_tree(nullptr);
          TC_HIP(hipEventRecord(ev[k + 1], s));
          snapshot(r * kN + c + 1);
        }
      TC_HIP(hipStreamSynchronize(s));
      std::vector<unsigned long long> rows(static_cast<size_t>(nsnap) * njobs_);
      copy_sync(rows.data(), snap, 8 * rows.size(), hipMemcpyDeviceToHost, s);
      uint32_t best = 0;
      for (uint32_t c = 0; c < kN; ++c) {
        double t = 0.0, nr = 0.0;
        for (uint32_t r = 0; r < reps; ++r) {
          const uint32_t k = r * kN + c;
          float x = 0.f;
          TC_HIP(hipEventElapsedTime(&x, ev[2 * k], ev[2 * k + 1]));
          t += x;
          for (uint32_t j = 0; j < njobs_; ++j)
            nr += static_cast<double>(rows[static_cast<size_t>(k + 1) * njobs_ + j] -
                                      rows[static_cast<size_t>(k) * njobs_ + j]);
        }
        tune_ms_[c] = static_cast<float>(t / reps);
        tune_cost_[c] = t / std::max(1.0, nr);
        if (tune_cost_[c] < tune_cost_[best]) best = c;
      }
      choose(best);
    } catch (...) {
      dfree(snap);
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    dfree(snap);
    for (auto& e : ev) (void)hipEventDestroy(e);
    tuned_ = true;
    return 1 + kN * reps;
  }

  void hist_chunks(uint32_t j, uint32_t* root, uint32_t* deep) const {
    *root = h_jobs_[j].hist_chunk;
    *deep = h_jobs_[j].deep_items;
  }

  // Device-side consistency flags (TreeState::fault) of every job.
  void check_faults() {
    for (uint32_t j = 0; j < njobs_; ++j) {
      uint32_t f = 0;
      copy_sync(&f, reinterpret_cast<const char*>(h_jobs_[j].st) + offsetof(TreeState, fault), 4,
                hipMemcpyDeviceToHost, ctx_->stream);
      if (f & 2u)
        throw Error(TC_EDEVICE, "internal consistency check failed in the split scan of job " +
                                    std::to_string(j) + " (more channels than a pass holds)");
      if (f & 4u)
        throw Error(TC_EDEVICE, "internal consistency check failed in the histogram of job " +
                                    std::to_string(j) + " (more work items than the grid)");
      if (f) throw Error(TC_EDEVICE, "internal consistency check failed in the partition of job " +
                                         std::to_string(j));
    }
  }

  // Per-tree losses and early-stopping outcome of job j.
  void history(uint32_t j, tc_train_history* h) {
    const JobDev& J = h_jobs_[j];
    TreeState st{};
    TC_HIP(hipMemcpy(&st, J.st, sizeof(st), hipMemcpyDeviceToHost));
    const uint32_t built = st.tree < 0 ? 0u : std::min(n_trees_, static_cast<uint32_t>(st.tree) + 1u);
    h->n_trees_built = built;
    h->stopped_early = st.stopped ? 1 : 0;
    h->best_iteration = st.es_best_iter;
    h->best_valid_loss = J.Nv ? st.es_best : 0.0f;
    if (h->train_loss) {
      std::fill(h->train_loss, h->train_loss + n_trees_, 0.0f);
      if (J.train_loss) TC_HIP(hipMemcpy(h->train_loss, J.train_loss, 4ull * built, hipMemcpyDeviceToHost));
    }
    if (h->valid_loss) {
      std::fill(h->valid_loss, h->valid_loss + n_trees_, 0.0f);
      if (J.valid_loss) TC_HIP(hipMemcpy(h->valid_loss, J.valid_loss, 4ull * built, hipMemcpyDeviceToHost));
    }
  }

  void download(uint32_t j, tc_model* m) {
    check_faults();
    const JobDev& J = h_jobs_[j];
    const size_t TD = static_cast<size_t>(n_trees_) * D_;
    hipStream_t s = ctx_->stream;
    if (m->feature) TC_HIP(hipMemcpyAsync(m->feature, J.m_feat, TD * 2, hipMemcpyDeviceToHost, s));
    if (m->bin) TC_HIP(hipMemcpyAsync(m->bin, J.m_bin, TD, hipMemcpyDeviceToHost, s));
    if (m->fthr) TC_HIP(hipMemcpyAsync(m->fthr, J.m_fthr, TD * 4, hipMemcpyDeviceToHost, s));
    if (m->gain) TC_HIP(hipMemcpyAsync(m->gain, J.m_gain, TD * 4, hipMemcpyDeviceToHost, s));
    if (m->depth) TC_HIP(hipMemcpyAsync(m->depth, J.m_depth, n_trees_ * 2, hipMemcpyDeviceToHost, s));
    if (m->leaves)
      TC_HIP(hipMemcpyAsync(m->leaves, J.m_leaves, (static_cast<size_t>(n_trees_) << D_) * J.K * 4,
                            hipMemcpyDeviceToHost, s));
    if (m->train_preds)
      TC_HIP(hipMemcpyAsync(m->train_preds, J.preds, sizeof(float) * J.N * J.K,
                            hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    m->base = J.base;
  }

  // Partition the last level as well, so that leaf_index can read it.
  void keep_final_rows() { keep_rows_ = true; }

  // Per-row leaf index of tree 0 from the final row lists (teacher-forcing
  // checks of the partition; needs keep_final_rows before the tree).
  void leaf_index(uint32_t* out) {
    const JobDev& J = h_jobs_[0];
    uint16_t depth = 0;
    TC_HIP(hipMemcpy(&depth, J.m_depth, 2, hipMemcpyDeviceToHost));
    std::vector<uint32_t> rows(J.N), off((1u << depth) + 1);
    TC_HIP(hipMemcpy(rows.data(), J.rows + static_cast<size_t>(depth) * J.N, 4ull * J.N,
                     hipMemcpyDeviceToHost));
    TC_HIP(hipMemcpy(off.data(), J.off + static_cast<size_t>(depth) * ((1u << D_) + 1),
                     4ull * off.size(), hipMemcpyDeviceToHost));
    for (uint32_t n = 0; n < (1u << depth); ++n)
      for (uint32_t p = off[n]; p < off[n + 1]; ++p) out[rows[p]] = n;
  }

  uint64_t hist_rows(uint32_t j) const {
    uint64_t v = 0;
    TC_HIP(hipMemcpy(&v, hist_rows_[j], 8, hipMemcpyDeviceToHost));
    return v;
  }

  uint32_t njobs() const { return njobs_; }
  const JobDev& job(uint32_t j) const { return h_jobs_[j]; }

 private:
  template <typename T>
  T* alloc(size_t n) {
    T* p = dmalloc<T>(n);
    allocs_.push_back(p);
    return p;
  }

  void setup() {
    hipStream_t s = ctx_->stream;
    h_jobs_.resize(njobs_);
    fixed_chunk_.assign(njobs_, 0);
    std::fill(list_cap_, list_cap_ + kMaxDepth, 0u);
    hist_rows_.resize(njobs_);
    samplers_.resize(njobs_);
    std::vector<uint32_t> wg_off(njobs_ + 1, 0);
    std::fill(hist_cap_, hist_cap_ + kMaxDepth, 0u);
    std::fill(part_cap_, part_cap_ + kMaxDepth, 0u);
    for (uint32_t j = 0; j < njobs_; ++j) {
      const JobSpec& S = specs_[j];
      const tc_dataset* ds = S.ds;
      JobDev& J = h_jobs_[j];
      std::memset(&J, 0, sizeof(J));
      const uint32_t N = ds->N;
      J.bins_rm = ds->bins_rm; J.bins_cm = ds->bins_cm; J.edges = ds->edges; J.n_edges = ds->n_edges;
      J.N = N; J.F = ds->F; J.Fp = ds->Fp; J.T = feature_tiles(ds->F);
      Sampler& SM = samplers_[j];
      SM.state = S.p.seed;
      SM.ratio = S.p.subsample;
      // 0 (unset) or >= 1 means no sampling
      SM.active = S.p.subsample > 0.0f && S.p.subsample < 1.0f;
      SM.colsample = S.p.colsample_bytree > 0.0f && S.p.colsample_bytree < 1.0f;
      J.n_root = SM.active ? std::min(N, static_cast<uint32_t>(static_cast<float>(N) * S.p.subsample))
                           : N;
      // booster.cpp:201-212: MVS, else GOSS, takes the place of the row sample
      if (S.p.use_mvs) {
        SM.mvs = SM.active = true;
        SM.ratio = S.p.mvs_subsample;
        J.n_root = N;  // until the first tree's sample is drawn
      } else if (S.p.use_goss) {
        SM.goss = SM.active = true;
        SM.top_rate = S.p.goss_top_rate;
        SM.other_rate = S.p.goss_other_rate;
        J.n_root = goss_rows(N, SM.top_rate, SM.other_rate);
      }
      if (SM.goss || SM.mvs) {
        TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&SM.gh_host), sizeof(float2) * std::max(1u, N),
                             hipHostMallocDefault));
        SM.g.resize(N);
      }
      if (SM.mvs) {
        SM.mvs_ahead = std::make_unique<MvsAhead>();
        SM.mvs_ahead->init(N, SM.state, SM.colsample, ahead_ring(N, njobs_));
      }
      J.root_identity = SM.active ? 0u : 1u;
      if (SM.active && !SM.goss && !SM.mvs) {  // a plain row sample: drawn ahead
        SM.plain = std::make_unique<AheadSampler>();
        SM.plain->init(N, SM.ratio, SM.state, SM.colsample, ahead_ring(N, njobs_));
        J.smask = alloc<uint32_t>(AheadSampler::mask_words(N));
        mask_chunks_ = std::max(mask_chunks_, AheadSampler::mask_chunks(N));
      } else if (SM.active) {  // GOSS, MVS: each tree's gradients; two buffers alternate
        for (int b = 0; b < 2; ++b) {
          TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&SM.host[b]), 4ull * std::max(1u, N),
                               hipHostMallocDefault));
          TC_HIP(hipEventCreateWithFlags(&SM.ev[b], hipEventDisableTiming));
        }
      }
      const uint32_t K = channels(S.p);
      J.loss = S.p.loss; J.K = K; J.D = D_; J.max_bins = S.p.tree.max_bins;
      J.min_samples_leaf = S.p.tree.min_samples_leaf;
      J.lr = S.p.learning_rate; J.lambda_l2 = S.p.tree.lambda_l2; J.lambda_l1 = S.p.tree.lambda_l1;
      J.max_delta_step = S.p.tree.max_delta_step; J.min_child_weight = S.p.tree.min_child_weight;
      J.min_split_gain = S.p.tree.min_split_gain;
      J.base = S.base;
      fixed_chunk_[j] = S.p.hist_chunk ? std::min(S.p.hist_chunk, kMaxHistChunk) : 0u;
      if (!fixed_chunk_[j]) tune_ = true;
      set_chunks(j, kTune[0]);
      float* y = alloc<float>(N);
      if (S.y_host) TC_HIP(hipMemcpyAsync(y, S.y_host, 4ull * N, hipMemcpyHostToDevice, s));
      else TC_HIP(hipMemsetAsync(y, 0, 4ull * N, s));
      J.y = y;
      J.preds = alloc<float>(static_cast<size_t>(N) * K);
      J.gh = alloc<float2>(static_cast<size_t>(N) * K);
      J.ghl = K > 1 ? alloc<float2>(2ull * N * K) : nullptr;
      J.rows = alloc<uint32_t>(static_cast<size_t>(D_ + 1) * N);
      hipLaunchKernelGGL(k_iota, dim3(std::max(1u, ceil_div(N, 256))), dim3(256), 0, s, J.rows, N);
      J.off = alloc<uint32_t>(static_cast<size_t>(D_ + 1) * ((1u << D_) + 1));
      const size_t nheap = (2ull << D_) - 1;
      J.ng = alloc<float>(nheap * K);
      J.nh = alloc<float>(nheap * K);
      J.nc = alloc<uint32_t>(nheap);
      J.hist = alloc<HistCell>(((1ull << D_) - 1) * K * J.F * 256);
      J.hfirst = alloc<uint32_t>((1ull << D_) - 1);
      J.best_keys = alloc<unsigned long long>(static_cast<size_t>(D_) * ds->F);
      J.lvl_chunk = alloc<uint32_t>(D_);
      J.st = alloc<TreeState>(1);
      hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, J.st);
      J.cursor = alloc<uint32_t>(2u << D_);  // two sets (level parity) of 2^(D-1) nodes x 2 words
      J.grad_wg_cnt = std::max(1u, ceil_div(N, kGradTile));
      J.grad_wg_off = wg_off[j];
      if (S.valid) {  // validation rows ride on the gradient update's grid
        TC_REQUIRE(S.valid->F == ds->F, "validation set: feature count differs from training");
        TC_REQUIRE(K == 1, "validation sets need a single-output loss (MSE or LogLoss)");
        TC_REQUIRE(S.y_valid != nullptr, "validation set without labels");
        const uint32_t Nv = S.valid->N;
        J.Nv = Nv;
        J.v_bins_cm = S.valid->bins_cm;
        J.v_wg_cnt = ceil_div(Nv, kGradTile);
        float* vy = alloc<float>(std::max(1u, Nv));
        if (Nv) TC_HIP(hipMemcpyAsync(vy, S.y_valid, 4ull * Nv, hipMemcpyHostToDevice, s));
        J.v_y = vy;
        J.v_preds = alloc<float>(std::max(1u, Nv));
        J.valid_loss = alloc<float>(std::max(1u, n_trees_));
        TC_HIP(hipMemsetAsync(J.valid_loss, 0, 4ull * std::max(1u, n_trees_), s));
        J.es_rounds = S.es_rounds;
        J.es_min_iter = std::max(50u, n_trees_ / 10);  // booster.cpp:345
      }
      if (S.history) {
        J.train_loss = alloc<float>(std::max(1u, n_trees_));
        TC_HIP(hipMemsetAsync(J.train_loss, 0, 4ull * std::max(1u, n_trees_), s));
      }
      if (J.Nv || J.train_loss) {
        J.loss_partials = alloc<double>(J.grad_wg_cnt + J.v_wg_cnt);
        eval_ = true;
      }
      wg_off[j + 1] = wg_off[j] + J.grad_wg_cnt + J.v_wg_cnt;
      J.grad_partials = alloc<double>(2ull * J.grad_wg_cnt * K);
      J.ghmax_partials = alloc<uint32_t>(2ull * J.grad_wg_cnt);
      TC_HIP(hipMemsetAsync(J.ghmax_partials, 0, 8ull * J.grad_wg_cnt, s));
      J.n_trees = n_trees_;
      const size_t TD = static_cast<size_t>(n_trees_) * D_;
      J.m_feat = alloc<uint16_t>(TD);
      J.m_bin = alloc<uint8_t>(TD);
      J.m_fthr = alloc<float>(TD);
      J.m_gain = alloc<float>(TD);
      J.m_depth = alloc<uint16_t>(n_trees_);
      J.m_leaves = alloc<float>((static_cast<size_t>(n_trees_) << D_) * K);
      TC_HIP(hipMemsetAsync(J.m_feat, 0, TD * 2, s));
      TC_HIP(hipMemsetAsync(J.m_bin, 0, TD, s));
      TC_HIP(hipMemsetAsync(J.m_fthr, 0, TD * 4, s));
      TC_HIP(hipMemsetAsync(J.m_gain, 0, TD * 4, s));
      TC_HIP(hipMemsetAsync(J.m_depth, 0, 2ull * n_trees_, s));
      TC_HIP(hipMemsetAsync(J.m_leaves, 0, ((static_cast<size_t>(n_trees_) << D_) * K) * 4, s));
      hist_rows_[j] = alloc<unsigned long long>(1);
      TC_HIP(hipMemsetAsync(hist_rows_[j], 0, 8, s));
      J.hist_rows_total = hist_rows_[j];
      // work-list capacities: room for the most work items any choice
      // gives (tree_begin / k_part_plan emit at most this many)
      for (uint32_t d = 0; d < D_; ++d) {
        uint32_t most = 0;
        for (const TuneChoice& c : kTune) {
          set_chunks(j, c);
          most = std::max(most, hist_items(J, d));
        }
        list_cap_[d] += most;
        part_cap_[d] += ceil_div(N, kPartChunk) + (1u << d);
      }
      set_chunks(j, kTune[0]);
    }
    recompute_caps();
    // partial-histogram slots: one per work item of the longest list (reused
    // level after level), unless that would take more than a quarter of the
    // free device memory -- then every item adds into zeroed node histograms
    // with atomics
    {
      uint32_t cap = 1;
      for (uint32_t d = 0; d < D_; ++d) cap = std::max(cap, list_cap_[d]);
      const size_t bytes = static_cast<size_t>(cap) * kTileF * 256 * sizeof(HistCell);
      size_t free_b = 0, total_b = 0;
      TC_HIP(hipMemGetInfo(&free_b, &total_b));
      // TURBOCAT_HIST_ATOMIC=1 selects the atomic mode (same results; tests use it)
      const char* force = std::getenv("TURBOCAT_HIST_ATOMIC");
      const bool atomic = force != nullptr && force[0] == '1';
      if (!atomic && bytes <= free_b / 4)
        partials_ = alloc<HistCell>(static_cast<size_t>(cap) * kTileF * 256);
      for (JobDev& J : h_jobs_) {
        J.store_hist = partials_ ? 1u : 0u;
        J.fused_root = J.store_hist && J.root_identity && J.K == 1 ? 1u : 0u;
        if (!J.fused_root) root_partials_ = true;
      }
    }
    for (uint32_t d = 0; d < D_; ++d) {
      lists_.hist[d] = alloc<HistWork>(std::max(1u, list_cap_[d]));
      lists_.part[d] = alloc<PartWork>(std::max(1u, part_cap_[d]));
      part_cap_[d] = std::max(1u, part_cap_[d]);
    }
    for (uint32_t d = D_; d < static_cast<uint32_t>(kMaxDepth); ++d) {
      lists_.hist[d] = lists_.hist[0];
      lists_.part[d] = lists_.part[0];
    }
    uint32_t* counters = alloc<uint32_t>(2 * kMaxDepth);
    TC_HIP(hipMemsetAsync(counters, 0, 8 * kMaxDepth, s));
    for (int d = 0; d < kMaxDepth; ++d) {
      lists_.hist_count[d] = counters + d;
      lists_.part_count[d] = counters + kMaxDepth + d;
      lists_.single = njobs_ == 1 ? 1u : 0u;
    }
    total_grad_wg_ = wg_off[njobs_];
    d_job_wg_off_ = alloc<uint32_t>(njobs_ + 1);
    TC_HIP(hipMemcpyAsync(d_job_wg_off_, wg_off.data(), 4ull * (njobs_ + 1), hipMemcpyHostToDevice, s));
    d_jobs_ = alloc<JobDev>(njobs_);
    TC_HIP(hipMemcpyAsync(d_jobs_, h_jobs_.data(), sizeof(JobDev) * njobs_, hipMemcpyHostToDevice, s));
    TC_HIP(hipStreamSynchronize(s));
  }

  void release() {
    if (ctx_ && ctx_->stream) (void)hipStreamSynchronize(ctx_->stream);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
    for (void* p : allocs_) dfree(p);
    allocs_.clear();
    for (auto& S : samplers_) {
      S.plain.reset();  // waits for its draws
      for (int b = 0; b < 2; ++b) {
        if (S.host[b]) hfree(S.host[b]);
        if (S.ev[b]) (void)hipEventDestroy(S.ev[b]);
        S.host[b] = nullptr;
        S.ev[b] = nullptr;
      }
      if (S.gh_host) hfree(S.gh_host);
      S.gh_host = nullptr;
    }
  }

  struct Sampler {
    uint64_t state = 42;
    float ratio = 1.f;
    bool active = false, colsample = false;
    bool goss = false;                 // GOSS on the current gradients instead of a shuffle
    bool mvs = false;                  // MVS on the current gradients (ratio: mvs_subsample)
    float top_rate = 0.f, other_rate = 0.f;
    float2* gh_host = nullptr;         // pinned copy of the gradients (GOSS, MVS)
    uint32_t n_root_host[2] = {0, 0};  // MVS: each buffer's row count, the source of its copy
    std::vector<float> g;              // their g
    // GOSS, MVS: pinned row samples and the events of their copies, alternating
    uint32_t* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int next = 0;
    std::unique_ptr<AheadSampler> plain;  // a plain row sample, drawn ahead
    std::unique_ptr<MvsAhead> mvs_ahead;  // MVS: the uniform draws, made ahead
    MvsScratch mvs_scratch;
    std::vector<std::pair<float, uint32_t>> goss_pairs;  // GOSS: kept across trees
  };

  tc_ctx* ctx_;
  std::vector<JobSpec> specs_;
  uint32_t njobs_ = 0, D_ = 0, n_trees_ = 0, fmax_ = 0, kmax_ = 1;
  std::vector<JobDev> h_jobs_;
  std::vector<unsigned long long*> hist_rows_;
  JobDev* d_jobs_ = nullptr;
  uint32_t* d_job_wg_off_ = nullptr;
  uint32_t total_grad_wg_ = 0;
  Lists lists_{};
  HistCell* partials_ = nullptr;        // [item][16][256] of the level being histogrammed
  bool root_partials_ = false;          // some job needs k_root_partials
  uint32_t hist_cap_[kMaxDepth] = {}, part_cap_[kMaxDepth] = {};
  std::vector<void*> allocs_;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<Sampler> samplers_;
  uint32_t mask_chunks_ = 0;  // k_rows_from_mask's grid: chunks of the largest sampled job
  bool forced_ = false;
  bool keep_rows_ = false;  // partition the last level too (leaf_index)
  bool eval_ = false;       // some job records losses / early-stops (k_eval per tree)
  // histogram work-item shapes (rows per item at the root and below)
  std::vector<uint32_t> fixed_chunk_;   // per job: caller's hist_chunk, 0 = tuned
  uint32_t list_cap_[kMaxDepth] = {};   // allocated work-list entries per level
  bool tune_ = false, tuned_ = false;
  uint32_t choice_ = 0;
  float tune_ms_[8] = {};
  double tune_cost_[8] = {};  // ms per histogrammed row of each candidate

  // Dynamic LDS of k_grad_update: the leaf values of one tree, [2^D][K].
  size_t leaf_lds() const { return 4ull * (static_cast<size_t>(kmax_) << D_); }

  // Packed MSE histogram cells need h == 2 on every row of every job.
  bool packed_hist() const {
    if (forced_) return false;
    for (const auto& S : specs_) if (S.p.loss != TC_LOSS_MSE) return false;
    return true;
  }
};

}  // namespace tc

struct tc_trainer {
  tc_ctx* ctx = nullptr;              // held (a reference)
  std::vector<tc_dataset*> datasets;  // held, one reference each
  std::unique_ptr<tc::Trainer> tr;
  ~tc_trainer() {
    tr.reset();
    for (tc_dataset* d : datasets) tc::unref(d);
    tc::unref(ctx);
  }
};

namespace tc {

float init_prediction_host(int32_t loss, const float* y, uint32_t n) {
  // loss.cpp:217-226 (LogLoss) and :364-370 (MSE): sequential float sum;
  // softmax starts every class at 0 (loss.cpp:269-271)
  if (loss == TC_LOSS_SOFTMAX) return 0.f;
  float s = 0.f;
  for (uint32_t i = 0; i < n; ++i) s += y[i];
  if (loss == TC_LOSS_MSE) return s / static_cast<float>(n);
  float p = s / static_cast<float>(n);
  p = std::max(1e-7f, std::min(1.0f - 1e-7f, p));
  return std::log(p / (1.0f - p));
}

void check_params(const tc_train_params* p) {
  TC_REQUIRE(p != nullptr, "params is NULL");
  TC_REQUIRE(p->tree.max_depth >= 1 && p->tree.max_depth <= static_cast<uint32_t>(kMaxDepth),
             "max_depth must be in [1, 10]");
  TC_REQUIRE(p->learning_rate == p->learning_rate, "learning_rate is NaN");
  TC_REQUIRE(p->hist_deep_items <= 65536u, "hist_deep_items must be <= 65536");
  if (p->use_goss) {
    TC_REQUIRE(p->goss_top_rate >= 0.f && p->goss_top_rate <= 1.f &&
                   p->goss_other_rate >= 0.f && p->goss_other_rate <= 1.f,
               "goss_top_rate and goss_other_rate must be in [0, 1]");
    TC_REQUIRE(p->loss != TC_LOSS_SOFTMAX, "use_goss needs a single-channel loss");
  }
  if (p->use_mvs) {
    TC_REQUIRE(p->mvs_subsample > 0.f && p->mvs_subsample <= 1.f, "mvs_subsample must be in (0, 1]");
    TC_REQUIRE(p->loss != TC_LOSS_SOFTMAX, "use_mvs needs a single-channel loss");
  }
  if (p->loss == TC_LOSS_SOFTMAX) {
    TC_REQUIRE(p->n_classes >= 2 && p->n_classes <= static_cast<uint32_t>(kMaxClasses),
               "softmax needs n_classes in [2, 16]");
    TC_REQUIRE((p->n_classes << p->tree.max_depth) <= static_cast<uint32_t>(kMaxLeafValues),
               "softmax: n_classes * 2^max_depth must be <= 8192");
  }
}

}  // namespace tc

using namespace tc;

extern "C" {

int32_t tc_random_subsample(uint32_t n, float ratio, uint64_t seed, uint32_t* out, uint32_t* n_out) {
  return guard([&] {
    TC_REQUIRE(out && n_out, "tc_random_subsample: NULL argument");
    *n_out = random_subsample(n, ratio, seed, out);
  });
}

int32_t tc_goss_subsample(const float* g, uint32_t n, float top_rate, float other_rate,
                          uint64_t seed, uint32_t* out, uint32_t* n_out) {
  return guard([&] {
    TC_REQUIRE(g && out && n_out, "tc_goss_subsample: NULL argument");
    TC_REQUIRE(top_rate >= 0.f && top_rate <= 1.f && other_rate >= 0.f && other_rate <= 1.f,
               "tc_goss_subsample: rates must be in [0, 1]");
    *n_out = goss_subsample(g, n, top_rate, other_rate, seed, out);
  });
}

int32_t tc_mvs_subsample(const float* g, uint32_t n, float ratio, uint64_t seed, uint32_t* out,
                         uint32_t* n_out) {
  return guard([&] {
    TC_REQUIRE(g && out && n_out, "tc_mvs_subsample: NULL argument");
    TC_REQUIRE(ratio <= 1.f, "tc_mvs_subsample: ratio must be <= 1");
    *n_out = mvs_subsample(g, n, ratio, seed, out);
  });
}

int32_t tc_init_prediction(int32_t loss, const float* y, uint32_t n, float* base) {
  return guard([&] {
    TC_REQUIRE(y && base && n > 0, "tc_init_prediction: bad arguments");
    TC_REQUIRE(loss == TC_LOSS_MSE || loss == TC_LOSS_LOGLOSS, "unsupported loss");
    *base = init_prediction_host(loss, y, n);
  });
}

int32_t tc_train(tc_dataset* ds, const float* y, const tc_train_params* p, tc_model* model) {
  return guard([&] {
    TC_REQUIRE(ds && y && model, "tc_train: NULL argument");
    check_params(p);
    TC_REQUIRE(ds->N > 0, "tc_train: empty dataset");
    JobSpec s{ds, y, *p, init_prediction_host(p->loss, y, ds->N)};
    Trainer tr(ds->ctx, {s});
    tr.init_gradients();
    double loop_ms = 0, hist_ms = 0;
    tr.run(p->use_graph != 0, false, &loop_ms, &hist_ms);
    ds->ctx->last_loop_ms = loop_ms;
    ds->ctx->last_hist_ms = 0.0;
    tr.download(0, model);
  });
}

int32_t tc_train_valid(tc_dataset* ds, const float* y, const tc_train_params* p, tc_dataset* valid,
                       const float* y_valid, uint32_t early_stopping_rounds, tc_model* model,
                       tc_train_history* history) {
  return guard([&] {
    TC_REQUIRE(ds && y && model, "tc_train_valid: NULL argument");
    check_params(p);
    TC_REQUIRE(ds->N > 0, "tc_train_valid: empty dataset");
    TC_REQUIRE(p->loss == TC_LOSS_MSE || p->loss == TC_LOSS_LOGLOSS,
               "tc_train_valid: MSE or LogLoss");
    TC_REQUIRE(!valid || valid->ctx == ds->ctx, "tc_train_valid: datasets on different contexts");
    TC_REQUIRE(!valid || y_valid, "tc_train_valid: validation labels missing");
    JobSpec s{ds, y, *p, init_prediction_host(p->loss, y, ds->N)};
    s.valid = valid;
    s.y_valid = y_valid;
    s.es_rounds = early_stopping_rounds;
    s.history = history != nullptr && history->train_loss != nullptr;
    Trainer tr(ds->ctx, {s});
    tr.init_gradients();
    double loop_ms = 0, hist_ms = 0;
    tr.run(p->use_graph != 0, false, &loop_ms, &hist_ms);
    ds->ctx->last_loop_ms = loop_ms;
    ds->ctx->last_hist_ms = 0.0;
    tr.download(0, model);
    if (history) tr.history(0, history);
  });
}

int32_t tc_train_batch(tc_ctx* ctx, uint32_t n_jobs, tc_dataset* const* ds, const float* const* y,
                       const tc_train_params* p, tc_model* models) {
  return guard([&] {
    TC_REQUIRE(ctx && ds && y && p && models && n_jobs > 0, "tc_train_batch: NULL argument");
    std::vector<JobSpec> specs;
    for (uint32_t j = 0; j < n_jobs; ++j) {
      check_params(&p[j]);
      TC_REQUIRE(ds[j] && ds[j]->ctx == ctx && ds[j]->N > 0, "tc_train_batch: bad dataset");
      specs.push_back(JobSpec{ds[j], y[j], p[j], init_prediction_host(p[j].loss, y[j], ds[j]->N)});
    }
    Trainer tr(ctx, specs);
    tr.init_gradients();
    double loop_ms = 0, hist_ms = 0;
    tr.run(p[0].use_graph != 0, false, &loop_ms, &hist_ms);
    ctx->last_loop_ms = loop_ms;
    ctx->last_hist_ms = 0.0;
    for (uint32_t j = 0; j < n_jobs; ++j) tr.download(j, &models[j]);
  });
}

int32_t tc_trainer_create(tc_ctx* ctx, uint32_t n_jobs, tc_dataset* const* ds, const float* const* y,
                          const tc_train_params* p, tc_trainer** out) {
  return guard([&] {
    TC_REQUIRE(ctx && ds && y && p && out && n_jobs > 0, "tc_trainer_create: NULL argument");
    *out = nullptr;
    std::vector<JobSpec> specs;
    for (uint32_t j = 0; j < n_jobs; ++j) {
      check_params(&p[j]);
      TC_REQUIRE(ds[j] && ds[j]->ctx == ctx && ds[j]->N > 0, "tc_trainer_create: bad dataset");
      specs.push_back(JobSpec{ds[j], y[j], p[j], init_prediction_host(p[j].loss, y[j], ds[j]->N)});
    }
    auto h = std::make_unique<tc_trainer>();
    h->ctx = ref(ctx);
    for (uint32_t j = 0; j < n_jobs; ++j) h->datasets.push_back(ref(ds[j]));
    h->tr = std::make_unique<Trainer>(ctx, specs);
    h->tr->init_gradients();
    TC_HIP(hipStreamSynchronize(ctx->stream));
    *out = h.release();
  });
}

int32_t tc_trainer_step(tc_trainer* t, uint32_t n_trees) {
  return guard([&] {
    TC_REQUIRE(t, "NULL trainer");
    t->tr->step_graph(n_trees);
  });
}

int32_t tc_trainer_step_timed(tc_trainer* t, uint32_t n_trees, double* hist_ms, uint32_t* launches) {
  return guard([&] {
    TC_REQUIRE(t, "NULL trainer");
    t->tr->step_timed(n_trees, hist_ms, launches);
  });
}

int32_t tc_trainer_sync(tc_trainer* t) {
  return guard([&] {
    TC_REQUIRE(t, "NULL trainer");
    TC_HIP(hipStreamSynchronize(t->ctx->stream));
    t->tr->check_faults();
  });
}

int32_t tc_trainer_model(tc_trainer* t, uint32_t job, tc_model* m) {
  return guard([&] {
    TC_REQUIRE(t && m, "tc_trainer_model: NULL argument");
    TC_REQUIRE(job < t->tr->njobs(), "job index out of range");
    t->tr->download(job, m);
  });
}

int32_t tc_trainer_hist_rows(tc_trainer* t, uint32_t job, uint64_t* rows) {
  return guard([&] {
    TC_REQUIRE(t && rows, "tc_trainer_hist_rows: NULL argument");
    TC_REQUIRE(job < t->tr->njobs(), "job index out of range");
    *rows = t->tr->hist_rows(job);
  });
}

int32_t tc_trainer_hist_chunks(tc_trainer* t, uint32_t job, uint32_t* root_rows,
                               uint32_t* deep_rows) {
  return guard([&] {
    TC_REQUIRE(t && root_rows && deep_rows, "tc_trainer_hist_chunks: NULL argument");
    TC_REQUIRE(job < t->tr->njobs(), "job index out of range");
    t->tr->hist_chunks(job, root_rows, deep_rows);
  });
}

int32_t tc_trainer_destroy(tc_trainer* t) {
  return guard([&] { delete t; });
}

int32_t tc_last_train_timing(tc_ctx* ctx, double* loop_ms, double* hist_ms) {
  return guard([&] {
    TC_REQUIRE(ctx, "NULL ctx");
    if (loop_ms) *loop_ms = ctx->last_loop_ms;
    if (hist_ms) *hist_ms = ctx->last_hist_ms;
  });
}

// Teacher-forced tree of K channels (class-major g, h).
static void tree_build_k(tc_dataset* ds, const float* g, const float* h, uint32_t K,
                         const tc_tree_params* tp, uint16_t* feature, uint8_t* bin, float* fthr,
                         float* gain, float* leaves, uint32_t* leaf_idx, int32_t* depth) {
  TC_REQUIRE(ds && g && h && tp && feature && bin && fthr && gain && leaves && depth,
             "tc_tree_build: NULL argument");
  tc_train_params p{};
  p.loss = K > 1 ? TC_LOSS_SOFTMAX : TC_LOSS_MSE;  // only the channel count matters here
  p.n_classes = K;
  p.n_trees = 1;
  p.learning_rate = 0.f;
  p.tree = *tp;
  p.subsample = 1.0f;
  p.colsample_bytree = 1.0f;
  check_params(&p);
  TC_REQUIRE(ds->N > 0, "tc_tree_build: empty dataset");
  JobSpec s{ds, nullptr, p, 0.f};
  Trainer tr(ds->ctx, {s});
  tr.set_gradients(g, h);
  if (leaf_idx) tr.keep_final_rows();
  tr.enqueue_tree(nullptr);
  TC_HIP(hipStreamSynchronize(ds->ctx->stream));
  const uint32_t D = tp->max_depth;
  std::vector<uint16_t> dep(1);
  tc_model m{};
  m.feature = feature; m.bin = bin; m.fthr = fthr; m.gain = gain; m.depth = dep.data();
  std::vector<float> lv((static_cast<size_t>(1) << D) * K);
  m.leaves = lv.data();
  tr.download(0, &m);
  *depth = dep[0];
  std::memcpy(leaves, lv.data(), sizeof(float) * (static_cast<size_t>(1) << dep[0]) * K);
  if (leaf_idx) tr.leaf_index(leaf_idx);
}

int32_t tc_tree_build(tc_dataset* ds, const float* g, const float* h, const tc_tree_params* tp,
                      uint16_t* feature, uint8_t* bin, float* fthr, float* gain, float* leaves,
                      uint32_t* leaf_idx, int32_t* depth) {
  return guard([&] { tree_build_k(ds, g, h, 1, tp, feature, bin, fthr, gain, leaves, leaf_idx, depth); });
}

int32_t tc_tree_build_multi(tc_dataset* ds, const float* g, const float* h, uint32_t K,
                            const tc_tree_params* tp, uint16_t* feature, uint8_t* bin, float* fthr,
                            float* gain, float* leaves, uint32_t* leaf_idx, int32_t* depth) {
  return guard([&] {
    TC_REQUIRE(K >= 1 && K <= static_cast<uint32_t>(kMaxClasses), "K must be in [1, 16]");
    tree_build_k(ds, g, h, K, tp, feature, bin, fthr, gain, leaves, leaf_idx, depth);
  });
}

int32_t tc_hist_build(tc_dataset* ds, const float* g, const float* h, const uint32_t* rows,
                      uint32_t n_rows, float* hist_out, uint32_t* count_out) {
  return guard([&] {
    TC_REQUIRE(ds && g && h && hist_out, "tc_hist_build: NULL argument");
    tc_ctx* ctx = ds->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t N = ds->N, F = ds->F;
    const uint32_t nr = rows ? n_rows : N;
    std::vector<uint32_t> r(nr);
    for (uint32_t i = 0; i < nr; ++i) {
      r[i] = rows ? rows[i] : i;
      TC_REQUIRE(r[i] < N, "tc_hist_build: row index out of range");
    }
    std::vector<void*> owned;
    auto al = [&](size_t bytes) { void* p = dmalloc<uint8_t>(bytes); owned.push_back(p); return p; };
    try {
      JobDev J;
      std::memset(&J, 0, sizeof(J));
      J.bins_rm = ds->bins_rm; J.bins_cm = ds->bins_cm; J.N = N; J.F = F; J.Fp = ds->Fp;
      J.T = feature_tiles(F);
      J.K = 1;
      std::vector<float2> ghh(N);
      for (uint32_t i = 0; i < N; ++i) ghh[i] = make_float2(g[i], h[i]);
      J.gh = static_cast<float2*>(al(8ull * N));
      TC_HIP(hipMemcpyAsync(J.gh, ghh.data(), 8ull * N, hipMemcpyHostToDevice, s));
      J.rows = static_cast<uint32_t*>(al(4ull * std::max(1u, nr)));
      if (nr) TC_HIP(hipMemcpyAsync(J.rows, r.data(), 4ull * nr, hipMemcpyHostToDevice, s));
      J.hist = static_cast<HistCell*>(al(sizeof(HistCell) * F * 256));
      TC_HIP(hipMemsetAsync(J.hist, 0, sizeof(HistCell) * F * 256, s));
      float2* fh = static_cast<float2*>(al(sizeof(float2) * F * 256));
      // fixed-point scales from the selected rows, as k_tree_begin sets them
      float gmax = 0.f, hmax = 0.f;
      for (uint32_t i = 0; i < nr; ++i) {
        gmax = std::max(gmax, std::fabs(g[r[i]]));
        hmax = std::max(hmax, std::fabs(h[r[i]]));
      }
      const uint32_t ch = 2048;
      const int eg = fixed_exponent(gmax, nr, ch, false), eh = fixed_exponent(hmax, nr, ch, false);
      TreeState st{};
      st.active = 1;
      st.gscale = std::ldexp(1.0, eg);
      st.ginv = std::ldexp(1.0, -eg);
      st.hscale = std::ldexp(1.0, eh);
      st.hinv = std::ldexp(1.0, -eh);
      J.st = static_cast<TreeState*>(al(sizeof(TreeState)));
      TC_HIP(hipMemcpyAsync(J.st, &st, sizeof(st), hipMemcpyHostToDevice, s));
      std::vector<HistWork> wl;
      for (uint32_t s0 = 0; s0 < nr; s0 += ch)
        for (uint32_t t = 0; t < J.T; ++t) {
          HistWork w{};
          w.job = 0; w.node = 0; w.start = s0; w.len = std::min(ch, nr - s0); w.tile = t;
          wl.push_back(w);
        }
      const uint32_t cnt = static_cast<uint32_t>(wl.size());
      HistWork* dw = static_cast<HistWork*>(al(sizeof(HistWork) * std::max<size_t>(1, wl.size())));
      if (cnt) TC_HIP(hipMemcpyAsync(dw, wl.data(), sizeof(HistWork) * cnt, hipMemcpyHostToDevice, s));
      uint32_t* dc = static_cast<uint32_t*>(al(4));
      TC_HIP(hipMemcpyAsync(dc, &cnt, 4, hipMemcpyHostToDevice, s));
      JobDev* dj = static_cast<JobDev*>(al(sizeof(JobDev)));
      TC_HIP(hipMemcpyAsync(dj, &J, sizeof(J), hipMemcpyHostToDevice, s));
      if (cnt)
        hipLaunchKernelGGL((k_hist<kHistUnroll, 0, true>), dim3(round8(cnt)),
                           dim3(kHistThreads), 0, s, dj, dw, dc, 0u, static_cast<HistCell*>(nullptr),
                           Lists{});
      hipLaunchKernelGGL(k_hist_to_float, dim3(ceil_div(F * 256ull, 256)), dim3(256), 0, s, J.hist,
                         J.st, F * 256u, fh);
      TC_HIP(hipGetLastError());
      TC_HIP(hipMemcpyAsync(hist_out, fh, sizeof(float2) * F * 256, hipMemcpyDeviceToHost, s));
      if (count_out) {
        // GradientPair::count (types.hpp:123-146): the same rows once more with
        // every (g, h) = (1, 1) at scale 1, so each cell's g sum is its row count
        std::fill(ghh.begin(), ghh.end(), make_float2(1.f, 1.f));
        TC_HIP(hipMemcpyAsync(J.gh, ghh.data(), 8ull * N, hipMemcpyHostToDevice, s));
        TC_HIP(hipMemsetAsync(J.hist, 0, sizeof(HistCell) * F * 256, s));
        TreeState one{};
        one.active = 1;
        one.gscale = one.ginv = one.hscale = one.hinv = 1.0;
        TC_HIP(hipMemcpyAsync(J.st, &one, sizeof(one), hipMemcpyHostToDevice, s));
        if (cnt)
          hipLaunchKernelGGL((k_hist<kHistUnroll, 0, true>), dim3(round8(cnt)),
                             dim3(kHistThreads), 0, s, dj, dw, dc, 0u,
                             static_cast<HistCell*>(nullptr), Lists{});
        TC_HIP(hipGetLastError());
        std::vector<HistCell> cells(static_cast<size_t>(F) * 256);
        TC_HIP(hipMemcpyAsync(cells.data(), J.hist, sizeof(HistCell) * cells.size(),
                              hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < cells.size(); ++i) count_out[i] = static_cast<uint32_t>(cells[i].g);
      }
      TC_HIP(hipStreamSynchronize(s));
    } catch (...) {
      for (void* p : owned) dfree(p);
      throw;
    }
    for (void* p : owned) dfree(p);
  });
}

int32_t tc_hist_subtract(tc_ctx* ctx, const float* parent, const float* child, float* out,
                         uint32_t F) {
  return guard([&] {
    TC_REQUIRE(ctx && parent && child && out, "tc_hist_subtract: NULL argument");
    const uint32_t n = F * 512;
    float *a = dmalloc<float>(n), *b = nullptr, *c = nullptr;
    try {
      b = dmalloc<float>(n);
      c = dmalloc<float>(n);
      TC_HIP(hipMemcpy(a, parent, 4ull * n, hipMemcpyHostToDevice));
      TC_HIP(hipMemcpy(b, child, 4ull * n, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_sub_elementwise, dim3(std::max(1u, ceil_div(n, 256))), dim3(256), 0,
                         ctx->stream, a, b, c, n);
      TC_HIP(hipGetLastError());
      TC_HIP(hipStreamSynchronize(ctx->stream));
      TC_HIP(hipMemcpy(out, c, 4ull * n, hipMemcpyDeviceToHost));
    } catch (...) {
      dfree(a); dfree(b); dfree(c);
      throw;
    }
    dfree(a); dfree(b); dfree(c);
  });
}

int32_t tc_split_scan(tc_ctx* ctx, const float* hist, const float* node_g, const float* node_h,
                      const uint32_t* node_count, uint32_t n_nodes, uint32_t F, uint32_t max_bins,
                      float lambda_l2, int32_t is_root, int32_t* feature, int32_t* bin,
                      float* gain) {
  return guard([&] {
    TC_REQUIRE(ctx && hist && node_g && node_h && node_count && feature && bin && gain,
               "tc_split_scan: NULL argument");
    TC_REQUIRE(n_nodes >= 1 && (n_nodes & (n_nodes - 1)) == 0 && n_nodes <= 512,
               "tc_split_scan: n_nodes must be a power of two <= 512");
    uint32_t level = 0;
    while ((1u << level) < n_nodes) ++level;
    TC_REQUIRE((level == 0) == (is_root != 0), "tc_split_scan: is_root must equal (n_nodes == 1)");
    hipStream_t s = ctx->stream;
    std::vector<void*> owned;
    auto al = [&](size_t bytes) { void* p = dmalloc<uint8_t>(bytes); owned.push_back(p); return p; };
    try {
      const size_t nheap = (2ull << level) - 1;
      const size_t hb = static_cast<size_t>(F) * 256 * 8;
      JobDev J;
      std::memset(&J, 0, sizeof(J));
      J.F = F; J.K = 1; J.max_bins = max_bins; J.lambda_l2 = lambda_l2; J.D = level + 1;
      float2* fh = static_cast<float2*>(al(hb * nheap));
      TC_HIP(hipMemcpyAsync(fh + ((1ull << level) - 1) * F * 256, hist, hb * n_nodes,
                            hipMemcpyHostToDevice, s));
      J.ng = static_cast<float*>(al(4 * nheap));
      J.nh = static_cast<float*>(al(4 * nheap));
      J.nc = static_cast<uint32_t*>(al(4 * nheap));
      const size_t o = (1ull << level) - 1;
      TC_HIP(hipMemcpyAsync(J.ng + o, node_g, 4ull * n_nodes, hipMemcpyHostToDevice, s));
      TC_HIP(hipMemcpyAsync(J.nh + o, node_h, 4ull * n_nodes, hipMemcpyHostToDevice, s));
      TC_HIP(hipMemcpyAsync(J.nc + o, node_count, 4ull * n_nodes, hipMemcpyHostToDevice, s));
      TreeState st{};
      st.active = 1;
      J.st = static_cast<TreeState*>(al(sizeof(TreeState)));
      TC_HIP(hipMemcpyAsync(J.st, &st, sizeof(st), hipMemcpyHostToDevice, s));
      JobDev* dj = static_cast<JobDev*>(al(sizeof(JobDev)));
      TC_HIP(hipMemcpyAsync(dj, &J, sizeof(J), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL((k_split_scan<false, 256>), dim3(F, 1), dim3(256), 0, s, dj,
                         static_cast<const float2*>(fh), level, static_cast<const HistCell*>(nullptr));
      TC_HIP(hipGetLastError());
      TC_HIP(hipMemcpyAsync(&st, J.st, sizeof(st), hipMemcpyDeviceToHost, s));
      TC_HIP(hipStreamSynchronize(s));
      const unsigned long long key = st.best[level];
      if (key == kEmptyKey) {
        *feature = -1; *bin = -1; *gain = -1e30f;
      } else {
        const uint32_t k = static_cast<uint32_t>(key >> 32);
        const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
        float gv;
        std::memcpy(&gv, &u, 4);
        const uint32_t fb = 0xFFFFFFFFu - static_cast<uint32_t>(key & 0xFFFFFFFFu);
        *feature = static_cast<int32_t>(fb >> 8);
        *bin = static_cast<int32_t>(fb & 255u);
        *gain = gv;
      }
    } catch (...) {
      for (void* p : owned) dfree(p);
      throw;
    }
    for (void* p : owned) dfree(p);
  });
}

int32_t tc_gradients_multi(tc_ctx* ctx, const float* y, const float* preds, float* g, float* h,
                           uint32_t n, uint32_t K) {
  return guard([&] {
    TC_REQUIRE(ctx && y && preds && g && h, "tc_gradients_multi: NULL argument");
    TC_REQUIRE(K >= 1 && K <= static_cast<uint32_t>(kMaxClasses), "K must be in [1, 16]");
    if (n == 0) return;
    const size_t nk = static_cast<size_t>(n) * K;
    float* buf = dmalloc<float>(n + 3 * nk);
    try {
      float *dy = buf, *dp = buf + n, *dg = dp + nk, *dh = dg + nk;
      TC_HIP(hipMemcpy(dy, y, 4ull * n, hipMemcpyHostToDevice));
      TC_HIP(hipMemcpy(dp, preds, 4 * nk, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_gradients_multi, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, dy,
                         dp, dg, dh, n, K);
      TC_HIP(hipGetLastError());
      TC_HIP(hipStreamSynchronize(ctx->stream));
      TC_HIP(hipMemcpy(g, dg, 4 * nk, hipMemcpyDeviceToHost));
      TC_HIP(hipMemcpy(h, dh, 4 * nk, hipMemcpyDeviceToHost));
    } catch (...) {
      dfree(buf);
      throw;
    }
    dfree(buf);
  });
}

int32_t tc_gradients(tc_ctx* ctx, int32_t loss, const float* y, const float* preds, float* g,
                     float* h, uint32_t n) {
  return guard([&] {
    TC_REQUIRE(ctx && y && preds && g && h, "tc_gradients: NULL argument");
    TC_REQUIRE(loss == TC_LOSS_MSE || loss == TC_LOSS_LOGLOSS, "unsupported loss");
    if (n == 0) return;
    float* buf = dmalloc<float>(4ull * n);
    try {
      TC_HIP(hipMemcpy(buf, y, 4ull * n, hipMemcpyHostToDevice));
      TC_HIP(hipMemcpy(buf + n, preds, 4ull * n, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_gradients, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, loss, buf,
                         buf + n, buf + 2ull * n, buf + 3ull * n, n);
      TC_HIP(hipGetLastError());
      TC_HIP(hipStreamSynchronize(ctx->stream));
      TC_HIP(hipMemcpy(g, buf + 2ull * n, 4ull * n, hipMemcpyDeviceToHost));
      TC_HIP(hipMemcpy(h, buf + 3ull * n, 4ull * n, hipMemcpyDeviceToHost));
    } catch (...) {
      dfree(buf);
      throw;
    }
    dfree(buf);
  });
}

}  // extern "C"
#include "tc_rtree.hip"
