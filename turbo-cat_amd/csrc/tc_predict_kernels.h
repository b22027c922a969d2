// A15-A18 inference kernels for gfx950, and the persistent serving kernel.
//
// Model layout.  Each tree is one record of TreeRec<DMAX>::kDwords dwords:
// DMAX float thresholds, DMAX uint16 features (two per dword), DMAX uint8
// bin thresholds (four per dword) and the depth; unused levels hold feature
// 0.  A workgroup's four waves each evaluate their own trees, so a tree index
// is uniform across a wave: its record is read through the constant address
// space into scalar registers (s_load), and every level's feature and
// threshold are scalar operands of the row compares.  All DMAX levels of a
// tree are compared without branches (the row's values are independent LDS
// reads) and the bits of the levels past the tree's depth are shifted out.
//
// Tile.  R <= 64 rows (lane r of every wave = row r) are staged in LDS with an
// odd row stride in dwords (conflict-free reads).  K = 1 sums in the
// reference's orders (fast_float_ensemble.hpp:480-560, fast_ensemble.hpp:
// 256-446): per pass each wave forms G groups of four trees, a group being
// w0*l0 + w1*l1 + w2*l2 + w3*l3 as the reference compiler contracts it (w1*l1
// rounded, the other three fused onto it), and wave 0 adds the pass's groups
// to the row's sum in tree order; leftover trees (T mod 4) are fused in one by
// one.  A row the reference sums tree by tree (the rows past a call's last
// full 8-row block, every row of a binned call under 16 rows, every row when
// no tree has a split) gets the four leaf values of each group instead and
// wave 0 fuses them one by one.  K > 1 (no reference counterpart): wave w
// sums w_t * leaf_t[c] over its trees t = w, w + 4, ... (fused, tree order)
// into K partials; wave 0 adds the four partials in wave order, then the base,
// then optionally the softmax.  A row's value depends only on the row, the
// model and its place in its own call, never on the tile or batch.
#pragma once
#include "tc_device.h"

#define TC_CONST_AS __attribute__((address_space(4)))

namespace tc {

constexpr uint32_t kPredWaves = 4;     // waves per workgroup
constexpr uint32_t kPredThreads = 256;
constexpr uint32_t kPredRowsMax = 64;  // rows per tile (one lane per row)

template <int DMAX>
struct TreeRec {
  static constexpr int kThr = 0;
  static constexpr int kFeat = DMAX;
  static constexpr int kBin = kFeat + (DMAX + 1) / 2;
  static constexpr int kDepth = kBin + (DMAX + 3) / 4;
  static constexpr int kDwords = (kDepth + 1 + 15) / 16 * 16;
};

struct EnsDev {
  const uint32_t* rec;     // [T][kDwords] tree records
  const float* leaves;     // [T][2^D][K]
  const float* weight;     // [T]
  uint32_t T, D, K;
  float base;
};

template <typename T>
__device__ __forceinline__ const TC_CONST_AS T* as_const(const T* p) {
  return (const TC_CONST_AS T*)p;
}

// Leaf index of a tree for the row at `xrow` (LDS; floats, or bins).
template <bool BINNED, int DMAX>
__device__ __forceinline__ uint32_t leaf_of(const TC_CONST_AS uint32_t* rec, const void* xrow) {
  using L = TreeRec<DMAX>;
  uint32_t raw = 0;
#pragma unroll
  for (int d = 0; d < DMAX; ++d) {
    const uint32_t f = (rec[L::kFeat + d / 2] >> (16 * (d & 1))) & 0xFFFFu;
    uint32_t bit;
    if (BINNED) {
      // go right = bin != 255 && bin > threshold (fast_ensemble.hpp:293-301)
      const uint32_t v = static_cast<const uint8_t*>(xrow)[f];
      const uint32_t b = (rec[L::kBin + d / 4] >> (8 * (d & 3))) & 0xFFu;
      bit = (v != 255u && v > b) ? 1u : 0u;
    } else {
      // go right = !isnan(v) && v >= threshold (fast_float_ensemble.hpp:122-137):
      // an ordered compare is false for NaN
      const float v = static_cast<const float*>(xrow)[f];
      bit = (v >= __uint_as_float(rec[L::kThr + d])) ? 1u : 0u;
    }
    raw |= bit << (DMAX - 1 - d);
  }
  return raw >> (DMAX - rec[L::kDepth]);
}

// LDS row stride in elements: an odd number of dwords per row.
__host__ __device__ inline uint32_t pred_stride(uint32_t F, bool binned) {
  return binned ? 4u * (((F + 3u) / 4u) | 1u) : (F | 1u);
}

// Stages nr contiguous rows of F floats (src, global or host memory) into
// LDS rows of stride S: a flat pass over the nr * F values, eight loads in
// flight per thread before their stores.
__device__ __forceinline__ void stage_rows(const float* src_generic, uint32_t nr, uint32_t F,
                                           float* sx, uint32_t S) {
  typedef const __attribute__((address_space(1))) float* gptr;
  const gptr src = (gptr)src_generic;
  const uint32_t total = nr * F, tid = threadIdx.x;
  const uint32_t dq = kPredThreads / F, dm = kPredThreads % F;  // the stride in (row, column)
  uint32_t i = tid, row = tid / F, col = tid % F;
  while (i < total) {
    float v[8];
    uint32_t rr[8], cc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      rr[u] = row;
      cc[u] = col;
      v[u] = i < total ? src[i] : 0.f;
      i += kPredThreads;
      row += dq;
      col += dm;
      if (col >= F) {
        col -= F;
        ++row;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (rr[u] < nr) sx[rr[u] * S + cc[u]] = v[u];
  }
}

// LDS of the group values / leaf values: [kPredWaves * G][4][64] floats.
template <int G>
constexpr uint32_t pred_val_floats() {
  return kPredWaves * G * 4 * kPredRowsMax;
}

// K = 1: the value of row r (meaningful in wave 0) of a staged tile.
template <bool BINNED, int DMAX, int G>
__device__ float tile_sum(const EnsDev& E, const void* xrow, uint32_t r, bool active, bool seq,
                          float* s_val) {
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const TC_CONST_AS uint32_t* rec = as_const(E.rec);
  const TC_CONST_AS float* wgt = as_const(E.weight);
  const uint32_t T = E.T, groups = T / 4u, LP = 1u << E.D;
  constexpr uint32_t kPass = kPredWaves * G;  // groups per pass
  constexpr int RD = TreeRec<DMAX>::kDwords;
  float sum = 0.f;
  for (uint32_t g0 = 0; g0 < groups; g0 += kPass) {
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const uint32_t slot = w * G + j;
      const uint32_t g = g0 + slot;
      if (g < groups) {
        float lv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t t = 4 * g + k;
          const uint32_t li = leaf_of<BINNED, DMAX>(rec + t * RD, xrow);
          lv[k] = E.leaves[static_cast<size_t>(t) * LP + li];
        }
        float* sv = s_val + slot * 4 * kPredRowsMax + r;
        if (seq) {
#pragma unroll
          for (int k = 0; k < 4; ++k) sv[k * kPredRowsMax] = lv[k];
        } else {
          float gv = __fmul_rn(wgt[4 * g + 1], lv[1]);
          gv = __fmaf_rn(wgt[4 * g], lv[0], gv);
          gv = __fmaf_rn(wgt[4 * g + 2], lv[2], gv);
          gv = __fmaf_rn(wgt[4 * g + 3], lv[3], gv);
          sv[0] = gv;
        }
      }
    }
    __syncthreads();
    if (w == 0 && active) {
      const uint32_t ng = min(kPass, groups - g0);
      for (uint32_t q = 0; q < ng; ++q) {
        const float* sv = s_val + q * 4 * kPredRowsMax + r;
        if (seq) {
          const uint32_t t0 = 4 * (g0 + q);
#pragma unroll
          for (int k = 0; k < 4; ++k) sum = __fmaf_rn(wgt[t0 + k], sv[k * kPredRowsMax], sum);
        } else {
          sum = __fadd_rn(sum, sv[0]);
        }
      }
    }
    __syncthreads();
  }
  if (w == 0 && active) {
    for (uint32_t t = groups * 4u; t < T; ++t) {
      const uint32_t li = leaf_of<BINNED, DMAX>(rec + t * RD, xrow);
      sum = __fmaf_rn(wgt[t], E.leaves[static_cast<size_t>(t) * LP + li], sum);
    }
  }
  return sum;
}

// K > 1: wave w's partial sums of row r into s_acc[w][c][64]; after the
// barrier, wave 0 forms the row's K outputs into v[].
template <bool BINNED, int DMAX>
__device__ void tile_sum_multi(const EnsDev& E, const void* xrow, uint32_t r, float* s_acc,
                               float* v, int transform) {
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const TC_CONST_AS uint32_t* rec = as_const(E.rec);
  const TC_CONST_AS float* wgt = as_const(E.weight);
  const uint32_t K = E.K, LP = 1u << E.D;
  constexpr int RD = TreeRec<DMAX>::kDwords;
  float acc[kMaxClasses];
#pragma unroll
  for (int c = 0; c < kMaxClasses; ++c) acc[c] = 0.f;
  for (uint32_t t = w; t < E.T; t += kPredWaves) {
    const uint32_t li = leaf_of<BINNED, DMAX>(rec + t * RD, xrow);
    const float wt = wgt[t];
    const float* lv = E.leaves + (static_cast<size_t>(t) * LP + li) * K;
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      acc[c] = __fmaf_rn(wt, lv[c], acc[c]);
    }
  }
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
      s_acc[(w * K + c) * kPredRowsMax + r] = acc[c];
  }
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) v[c] = 0.f;
  for (uint32_t q = 0; q < kPredWaves; ++q) {  // partials in wave order
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      v[c] = __fadd_rn(v[c], s_acc[(q * K + c) * kPredRowsMax + r]);
    }
  }
  float m = -INFINITY;
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    v[c] = __fadd_rn(E.base, v[c]);
    m = fmaxf(m, v[c]);
  }
  if (transform == 1) {  // CrossEntropyLoss::softmax (loss.cpp:273-285)
    float z = 0.f;
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      v[c] = ref_expf(__fsub_rn(v[c], m));
      z = __fadd_rn(z, v[c]);
    }
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      v[c] = __fdiv_rn(v[c], z);
    }
  }
}

__device__ __forceinline__ float finish_k1(const EnsDev& E, float sum, int transform) {
  float v = __fadd_rn(sum, E.base);  // booster.cpp:918-924
  if (transform == 1) v = __fdiv_rn(1.0f, __fadd_rn(1.0f, ref_expf(-v)));  // loss.cpp:213-215
  return v;
}

constexpr int kPredG = 1;  // groups per wave per pass

// LDS of a tile: the values area (K = 1: group / leaf values; K > 1: the
// waves' partial sums), then the rows.
__host__ __device__ inline uint32_t pred_head_bytes(bool multi, uint32_t K) {
  return multi ? 4u * kPredWaves * K * kPredRowsMax : 4u * pred_val_floats<kPredG>();
}

// One call over rows [0, n) of a device matrix: float X row-major n x F, or
// bins column-major F x ldc.  seq_from: first row the reference sums tree by
// tree.  leaf_idx (optional): n x T.
template <bool BINNED, int DMAX, bool MULTI>
__global__ void __launch_bounds__(kPredThreads)
    k_predict(EnsDev E, const void* __restrict__ Xv, uint32_t n, uint32_t F, uint32_t ldc,
              uint32_t R, int transform, float* __restrict__ out, uint32_t seq_from) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  float* s_val = reinterpret_cast<float*>(s_raw);
  unsigned char* s_x = s_raw + pred_head_bytes(MULTI, E.K);
  const uint32_t S = pred_stride(F, BINNED);
  const uint32_t tid = threadIdx.x, w = tid >> 6, r = tid & 63u;
  const uint32_t row0 = blockIdx.x * R;
  const uint32_t nr = min(R, n - row0);
  if (BINNED) {
    const uint8_t* B = static_cast<const uint8_t*>(Xv);
    for (uint32_t f = w; f < F; f += kPredWaves)
      if (r < nr) s_x[r * S + f] = B[static_cast<size_t>(f) * ldc + row0 + r];
  } else {
    stage_rows(static_cast<const float*>(Xv) + static_cast<size_t>(row0) * F, nr, F,
               reinterpret_cast<float*>(s_x), S);
  }
  __syncthreads();
  const bool active = r < nr;
  const void* xrow = BINNED ? static_cast<const void*>(s_x + (active ? r : 0) * S)
                            : static_cast<const void*>(reinterpret_cast<float*>(s_x) + (active ? r : 0) * S);
  if (MULTI) {
    float v[kMaxClasses];
    tile_sum_multi<BINNED, DMAX>(E, xrow, r, s_val, v, transform);
    if (w == 0 && active) {
      float* o = out + static_cast<size_t>(row0 + r) * E.K;
#pragma unroll
      for (uint32_t c = 0; c < kMaxClasses; ++c) {
        if (c >= E.K) break;
        o[c] = v[c];
      }
    }
  } else {
    const bool seq = active && row0 + r >= seq_from;
    const float sum = tile_sum<BINNED, DMAX, kPredG>(E, xrow, r, active, seq, s_val);
    if (w == 0 && active) out[row0 + r] = finish_k1(E, sum, transform);
  }
}

// Per-row, per-tree leaf indices (tests and debugging): n x T.
template <int DMAX>
__global__ void __launch_bounds__(256) k_leaf_index(EnsDev E, const float* __restrict__ X,
                                                    uint32_t n, uint32_t F,
                                                    uint32_t* __restrict__ leaf_idx) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (uint32_t t = 0; t < E.T; ++t) {
    const TC_CONST_AS uint32_t* rec = as_const(E.rec) + t * TreeRec<DMAX>::kDwords;
    leaf_idx[static_cast<size_t>(i) * E.T + t] =
        leaf_of<false, DMAX>(rec, X + static_cast<size_t>(i) * F);
  }
}

// ---------------------------------------------------------------------------
// Serving kernel: resident while there is work, bounded in lifetime.
//
// The host stages each batch in a slot of server-owned pinned host memory
// (rows, one tree-by-tree flag per row, outputs; slot s holds rows
// [s * max_rows, (s + 1) * max_rows) of each array) and publishes it in a
// ring of descriptors in uncached pinned host memory: it writes the batch's
// first tile, slot and row count, then the batch number (release).
// Workgroup 0, thread 0 is the dispatcher: it waits for the descriptor of the
// next batch number, checks it (the expected first tile, slot < n_slots,
// 1 <= rows <= max_rows; anything else stops the kernel with a fault flag),
// copies it into device memory, advances the count of published tiles and
// reports the batches taken in host memory (ServeHost::taken).  Every other
// workgroup claims tiles (R rows) in order with one atomic, waits for its
// tile to be published, finds its batch (rechecking the tile range, slot and
// rows), predicts the tile's rows from the slot's staging and writes the
// outputs there; the last tile of a batch sets the batch's done word in host
// memory.  The device dereferences no pointer written by the host: every
// host memory access is an index into the fixed staging arrays, checked
// against their bounds.  No copy call per batch, and no launch per batch
// while batches keep coming.
//
// Lifetime: a hardware queue runs its packets in order, and HIP may place
// another stream's work on this kernel's queue; so the kernel never stays
// for good.  The dispatcher stops taking batches when no batch came for
// idle_ticks, when it has run for slice_ticks, on the host's stop word, or
// when the host's heartbeat has not changed for hb_ticks; it then raises
// ServeCtl::stop, every published tile is processed, and the kernel ends.
// The host launches it again, from the first batch not taken, whenever a
// published batch is not taken and no launch is running.  Counters in
// ServeCtl are relative to the launch's first batch and tile.
constexpr uint32_t kServeRing = 8;
constexpr unsigned long long kServeNoBatch = ~0ull;  // a descriptor not yet written

struct ServeHostDesc {
  unsigned long long seq;    // batch number; written last
  unsigned long long tile0;  // global index of the batch's first tile
  uint32_t slot, rows;
};

struct ServeHost {
  unsigned long long heartbeat;            // changed by the host while it lives
  unsigned long long taken;                // batches taken by the device (all launches)
  uint32_t stop;
  uint32_t pad;
  unsigned long long done[kServeRing];     // seq + 1 of the last batch finished in each position
  ServeHostDesc ring[kServeRing];
};

struct ServeDevDesc {
  unsigned long long seq, tile0;
  uint32_t slot, rows, ntiles, pad;
};

// Device-side state of one launch, zeroed before it.
struct ServeCtl {
  unsigned long long next_tile;      // tiles claimed in this launch
  unsigned long long pub_tiles;      // tiles published in this launch
  unsigned long long pub_batches;    // batches published in this launch
  uint32_t stop;
  uint32_t fault;                    // 1: a worker found no batch for its tile; 2: a bad descriptor
  uint32_t done_cnt[kServeRing];     // tiles finished of each position's batch
  ServeDevDesc ring[kServeRing];
};

// Where a launch starts and when its dispatcher stops (wall-clock ticks).
struct ServeLaunch {
  unsigned long long batch0, tile0;  // first batch number / tile index of the launch
  unsigned long long hb_ticks, idle_ticks, slice_ticks;
};

// Staging of a server: n_slots x max_rows rows (host memory).
struct ServeStage {
  const float* x;       // [slot][row][F]
  const uint8_t* flag;  // [slot][row]
  float* out;           // [slot][row][K]
  uint32_t n_slots, max_rows;
};

// Polling uses relaxed atomic loads (read at the scope's coherence point,
// no cache invalidation per poll); the acquire is one fence once the polled
// value says the data is there.
__device__ __forceinline__ unsigned long long rx_sys(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t rx_sys32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ unsigned long long rx_dev(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t rx_dev32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void acquire_system() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, ""); }
__device__ __forceinline__ void release_system() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, ""); }

__device__ void serve_dispatch(ServeCtl* ctl, ServeHost* host, ServeStage st, uint32_t R,
                               const ServeLaunch L) {
  unsigned long long pubb = L.batch0, pubt = L.tile0;
  const unsigned long long t_start = wall_clock64();
  unsigned long long hb = rx_sys(&host->heartbeat), t_hb = t_start, t_act = t_start;
  bool bad = false, spent = false;
  while (!bad && !spent) {
    while (true) {  // every batch the host has published, in order, within the slice
      if (wall_clock64() - t_start > L.slice_ticks) {
        spent = true;
        break;
      }
      const ServeHostDesc* hd = &host->ring[pubb % kServeRing];
      if (rx_sys(&hd->seq) != pubb) break;  // not written yet
      acquire_system();                      // the fields written before the batch number
      const unsigned long long tile0 = rx_sys(&hd->tile0);
      const uint32_t slot = rx_sys32(&hd->slot), rows = rx_sys32(&hd->rows);
      if (tile0 != pubt || slot >= st.n_slots || rows == 0 || rows > st.max_rows) {
        atomicOr(&ctl->fault, 2u);
        bad = true;
        break;
      }
      const uint32_t nt = (rows + R - 1) / R;
      ServeDevDesc* dd = &ctl->ring[pubb % kServeRing];
      dd->tile0 = tile0;
      dd->slot = slot;
      dd->rows = rows;
      dd->ntiles = nt;
      __hip_atomic_store(&dd->seq, pubb, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      pubt += nt;
      ++pubb;
      __hip_atomic_store(&ctl->pub_batches, pubb - L.batch0, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl->pub_tiles, pubt - L.tile0, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&host->taken, pubb, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      t_act = wall_clock64();
    }
    if (bad || spent || rx_sys32(&host->stop)) break;
    const unsigned long long h = rx_sys(&host->heartbeat), now = wall_clock64();
    if (now - t_act > L.idle_ticks || now - t_start > L.slice_ticks) break;
    if (h != hb) {
      hb = h;
      t_hb = now;
    } else if (now - t_hb > L.hb_ticks) {
      break;  // the host stopped beating: end rather than spin
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __hip_atomic_store(&ctl->stop, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int DMAX, bool MULTI>
__global__ void __launch_bounds__(kPredThreads)
    k_serve(EnsDev E, ServeCtl* ctl, ServeHost* host, ServeStage st, uint32_t F, uint32_t R,
            int transform, ServeLaunch L) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) serve_dispatch(ctl, host, st, R, L);
    return;
  }
  __shared__ unsigned long long s_t, s_seq_no;
  __shared__ uint32_t s_slot, s_rows, s_ntiles, s_k;
  __shared__ int s_exit;
  float* s_val = reinterpret_cast<float*>(s_raw);
  float* sx = reinterpret_cast<float*>(s_raw + pred_head_bytes(MULTI, E.K));
  const uint32_t S = pred_stride(F, false);
  const uint32_t tid = threadIdx.x, w = tid >> 6, r = tid & 63u, K = E.K;
  unsigned long long cur_b = L.batch0;  // thread 0: the first batch that may hold the next tile
  while (true) {
    if (tid == 0) {
      const unsigned long long t = L.tile0 + atomicAdd(&ctl->next_tile, 1ull);
      int ex = 0;
      while (true) {
        if (t < L.tile0 + rx_dev(&ctl->pub_tiles)) break;
        if (rx_dev32(&ctl->stop)) {
          // the dispatcher publishes its last tiles before it raises stop:
          // look again after synchronizing with the stop
          acquire_agent();
          if (t < L.tile0 + rx_dev(&ctl->pub_tiles)) break;
          ex = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(20);
      }
      if (!ex) {
        // the batch's descriptor (dispatcher) and staged rows (host) are
        // visible from here on
        acquire_system();
        // batches are published in order with contiguous tile ranges; a
        // position holding a later batch means the earlier one is finished
        // (the position was reused), so it cannot hold this unprocessed tile.
        // The host reuses a position only after its batch is done, so every
        // batch before the last kServeRing published is finished: the scan
        // starts there (a few probes, however long this wave was idle).
        const unsigned long long nb = L.batch0 + rx_dev(&ctl->pub_batches);
        if (nb > kServeRing && cur_b < nb - kServeRing) cur_b = nb - kServeRing;
        while (true) {
          if (cur_b >= nb) {  // inconsistent: no published batch holds the tile
            atomicOr(&ctl->fault, 1u);
            ex = 1;
            break;
          }
          const ServeDevDesc* d = &ctl->ring[cur_b % kServeRing];
          const unsigned long long sq = rx_dev(&d->seq);
          const unsigned long long t0 = d->tile0;
          const uint32_t nt = d->ntiles, rows = d->rows, slot = d->slot;
          if (sq == cur_b && t >= t0 && t < t0 + nt) {
            if (slot >= st.n_slots || rows == 0 || rows > st.max_rows || nt != (rows + R - 1) / R) {
              atomicOr(&ctl->fault, 2u);
              ex = 1;
            } else {
              s_seq_no = sq;
              s_slot = slot;
              s_rows = rows;
              s_ntiles = nt;
              s_k = static_cast<uint32_t>(t - t0);
            }
            break;
          }
          ++cur_b;
        }
      }
      s_t = t;
      s_exit = ex;
    }
    __syncthreads();
    if (s_exit) return;
    // in bounds: slot < n_slots, rbase + nr <= rows <= max_rows
    const uint32_t rbase = s_k * R, nr = min(R, s_rows - rbase);
    const size_t row0 = static_cast<size_t>(s_slot) * st.max_rows + rbase;
    stage_rows(st.x + row0 * F, nr, F, sx, S);
    __syncthreads();
    const bool active = r < nr;
    const float* xrow = sx + (active ? r : 0) * S;
    if (MULTI) {
      float v[kMaxClasses];
      tile_sum_multi<false, DMAX>(E, xrow, r, s_val, v, transform);
      if (w == 0 && active) {
        float* o = st.out + (row0 + r) * K;
#pragma unroll
        for (uint32_t c = 0; c < kMaxClasses; ++c) {
          if (c >= K) break;
          o[c] = v[c];
        }
      }
    } else {
      const bool seq = active && st.flag[row0 + r] != 0;
      const float sum = tile_sum<false, DMAX, kPredG>(E, xrow, r, active, seq, s_val);
      if (w == 0 && active) st.out[row0 + r] = finish_k1(E, sum, transform);
    }
    // each wave's outputs reach host memory before the batch is marked done
    release_system();
    __syncthreads();
    if (tid == 0) {
      const uint32_t pos = static_cast<uint32_t>(s_seq_no % kServeRing);
      const uint32_t old = __hip_atomic_fetch_add(&ctl->done_cnt[pos], 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 == s_ntiles) {
        __hip_atomic_store(&ctl->done_cnt[pos], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&host->done[pos], s_seq_no + 1, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    __syncthreads();  // the tile's LDS and shared words are rewritten next round
  }
}

}  // namespace tc
