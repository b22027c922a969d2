// Shared device/host definitions for the TurboCat MI355X (gfx950) hot path.
//
// Data layout in HBM (one "job" = one boosting run over one dataset):
//   bins_rm   uint8  [N][Fp]     row-major (one 16-byte load = one 16-feature
//                                 tile of a row); the row stride Fp keeps a
//                                 row inside whole 128-byte lines (rm_stride)
//   bins_cm   uint8  [F][N]      column-major (split column reads)
//   gh        float2 [K][N]      gradient / hessian per channel and row (K = 1
//                                 except K-class softmax; class-major)
//   ghl       float2 [2][K][N]   K > 1: the same pairs in the order of a level's
//                                 row list (the histogram reads them by position)
//   preds     float  [K][N]      raw training predictions, class-major
//   rows      uint32 [D+1][N]    per level, row ids grouped by node (order
//                                 within a node is irrelevant: every sum
//                                 over a node is an exact integer sum)
//   off       uint32 [D+1][2^D+1] node segment offsets into rows[level]
//   hist      HistCell{int64 g, int64 h} [2^D-1][K][F][256]  fixed-point
//                                 per node and channel (heap order: level d
//                                 node n at (2^d - 1 + n)); slot 255 = NaN bin
//   ng/nh     float [2^(D+1)-1][K]  node totals per channel, heap order
//   nc        uint32 [2^(D+1)-1]    node row counts
//   partials  HistCell [item][16][256]  the level's work items but each
//                                 node's last chunk: each item's tile
//                                 histogram (plain stores), added to the
//                                 node's histogram by the split scan
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tc {

constexpr int kMaxDepth = 10;          // 2^10 leaves; u16 leaf index range
constexpr int kTileF = 16;             // features per histogram tile (16 B)
constexpr int kHistThreads = 1024;     // 16 waves; two workgroups per CU (64 KB LDS, <= 64 VGPRs)
constexpr int kPartChunk = 4096;       // rows per partition work item
constexpr int kPartThreads = 1024;
constexpr uint32_t kMaxHistChunk = 32768;  // rows per histogram work item (packed count < 2^16)
constexpr uint32_t kMaskChunkWords = 256;  // 32-row words of a row-sample mask per workgroup
constexpr int kScanNodesPerPass = 32;  // node prefixes staged in LDS at a time
constexpr uint32_t kEmptyKey = 0u;

enum Loss : int32_t { kMSE = 0, kLogLoss = 1, kSoftmax = 2 };
constexpr int kMaxClasses = 16;         // softmax channels
constexpr int kMaxLeafValues = 8192;    // 2^depth * K (leaf values staged in LDS)

// Per-tree state of one job, device resident.
struct TreeState {
  int32_t active;       // 1 while levels are still being grown
  int32_t depth;        // splits applied so far in the current tree
  int32_t tree;         // index of the tree being built
  uint32_t gmax_next;   // float bits of max |g| of the gradients being written
  uint32_t hmax_next;   // float bits of max |h| of the gradients being written
  uint32_t fault;       // nonzero: a kernel found inconsistent state and skipped its writes
                        // (1: a partition reservation outside its node; 2: a split-scan pass
                        // too small for the job's channels; 4: a level planned with more
                        // histogram items than its launch has workgroups)
  uint32_t stopped;     // early stopping ended the fit after tree `tree`
  float es_best;        // best validation loss so far (early stopping)
  uint32_t es_best_iter;
  uint32_t es_count;    // trees since the last improvement
  uint32_t pad;
  unsigned long long best[kMaxDepth];  // argmax key per level
  // Fixed-point scales of the current tree's histograms: a row's g enters as
  // round(g * gscale) (int64), h as round(h * hscale).  Powers of two chosen
  // from max |g| and the root row count so that no sum can overflow.
  double gscale, hscale;
  double ginv, hinv;
};

// Histogram slot in HBM: exact fixed-point sums of one (node, feature, bin).
struct HistCell {
  long long g, h;
};

// One unit of histogram work: rows [start, start+len) of rows[level] for
// node `node` (level-local index), feature tile `tile`, channel `cls`.
// `last`: the node's last chunk (its cells go into the node's histogram).
struct HistWork {
  uint32_t job, node, start, len, tile, cls, last, pad;
};

// One unit of partition work: rows [start, start+len) of rows[level],
// node `node`, result slot `slot` (per job, node order).
struct PartWork {
  uint32_t job, node, start, len, slot, pad0, pad1, pad2;
};

// Everything a kernel needs to know about one job.  Lives in device memory.
struct JobDev {
  // dataset
  const uint8_t* bins_rm;
  const uint8_t* bins_cm;
  const float* edges;       // [F][255]
  const uint32_t* n_edges;  // [F]
  uint32_t N, F, Fp, T;     // Fp = rm_stride(F), T = feature_tiles(F)
  uint32_t n_root;          // rows of the tree's root list (N, or the row sample)
  uint32_t root_identity;   // 1: the root list is 0..N-1 (no sampling), rows[0] not read
  const uint32_t* smask;    // a plain row sample: [chunks] rows drawn before each chunk of
                            // kMaskChunkWords words, then [ceil(N/32)] words (bit r of word
                            // r/32: row r drawn); k_rows_from_mask makes rows[0] from it.
                            // null: rows[0] is copied in (GOSS, MVS) or not used
  uint32_t fused_root;      // 1: the gradient update writes the root partials (K = 1, no
                            // sampling, stored histograms); k_root_partials skips the job
  // hyper-parameters
  int32_t loss;
  uint32_t K;               // gradient channels: classes for softmax, else 1
  uint32_t D, max_bins, min_samples_leaf;
  float lr, lambda_l2, lambda_l1, max_delta_step, min_child_weight, min_split_gain;
  float base;
  uint32_t hist_chunk;      // rows per histogram work item at the root
  uint32_t deep_items;      // below the root: work items a level is cut into (about; the
                            // rows per item follow the level's rows); 0: hist_chunk rows
  uint32_t* lvl_chunk;      // [D] rows per histogram work item at each level of the tree
  uint32_t store_hist;      // 1: each node's last chunk stores the node's histogram and the
                            // other chunks store partial slots (no zeroing, no atomics);
                            // 0: every chunk adds into the zeroed node histogram
  // training state
  const float* y;           // [N] (class index for softmax)
  float* preds;             // [K][N]
  float2* gh;               // [K][N]
  float2* ghl;              // K > 1: [2][K][N] each level's (g, h) in the order of its row list
                            // (by level parity, written by the partition); null: read gh by row
  uint32_t* rows;           // [D+1][N]
  uint32_t* off;            // [D+1][2^D + 1]
  float* ng;                // heap [2^(D+1) - 1][K]
  float* nh;
  uint32_t* nc;             // heap [2^(D+1) - 1]
  HistCell* hist;           // heap [2^D - 1][K][F][256]
  uint32_t* hfirst;         // heap [2^D - 1]: first work-list item of a histogrammed node
  TreeState* st;
  unsigned long long* best_keys;  // [D][F] each feature's best split key of the level
                                  // (null: merged with atomics into TreeState::best)
  // partition cursors, one 64-bit word per node of the level being split:
  // [node][0] next free position of the left child (counts up from the
  // node's start), [node][1] rows taken by the right child (filled down from
  // the node's end); two sets of 2^(D-1) nodes, by level parity
  uint32_t* cursor;
  // root-sum partials (sum g, sum h over 1024-row tiles of the root list)
  double* grad_partials;    // [wg][K][2]
  uint32_t* ghmax_partials; // [wg][2] float bits of each tile's max |g|, max |h| (all channels)
  uint32_t grad_wg_off;     // first workgroup of this job in the row grid
  uint32_t grad_wg_cnt;
  // model, tree-major
  uint32_t n_trees;
  uint16_t* m_feat;         // [n_trees][D]
  uint8_t* m_bin;
  float* m_fthr;
  float* m_gain;
  uint16_t* m_depth;        // [n_trees]
  float* m_leaves;          // [n_trees][2^D][K]
  unsigned long long* hist_rows_total;  // rows fed to histogram kernels (all trees)
  // validation set and per-tree losses (K = 1); Nv = 0: no validation set
  const uint8_t* v_bins_cm; // [F][Nv], binned with the training edges
  const float* v_y;         // [Nv]
  float* v_preds;           // [Nv] raw validation predictions
  uint32_t Nv, v_wg_cnt;    // validation rows; their gradient-update workgroups
  double* loss_partials;    // [grad_wg_cnt + v_wg_cnt] loss sums per 1024-row tile
  float* train_loss;        // [n_trees] mean training loss after each tree (null: not kept)
  float* valid_loss;        // [n_trees] mean validation loss after each tree
  uint32_t es_rounds;       // early stopping: trees without improvement
  uint32_t es_min_iter;     // first tree after which the fit may stop
};

// 16-feature tiles of a row.
__host__ __device__ inline uint32_t feature_tiles(uint32_t F) { return (F + kTileF - 1) / kTileF; }

// Row stride of the row-major bins: the tiles, padded so that no row
// straddles a 128-byte line (16, 32, 64, 128, then multiples of 128).  The
// histogram gathers the rows of a node; a 112-byte row would touch two lines
// about 7 times in 8.  Padding bytes are never written or read.
__host__ __device__ inline uint32_t rm_stride(uint32_t F) {
  const uint32_t t = feature_tiles(F) * kTileF;
  if (t > 128) return (t + 127) / 128 * 128;
  uint32_t p = kTileF;
  while (p < t) p <<= 1;
  return p;
}

// Rows per root work item when the root's `rows` rows of `tk` tile-channels
// are cut into at most `items` work items: rows / floor(items / tk), rounded
// up, in [1024, 16384].  All root items start at once when `items` is twice
// the CU count: 511 items of 13,699 rows for 1M rows x 7 tiles and 512 items
// run two per CU on 256 CUs, where 16,384-row items (434 of them) leave some
// CUs one item and most two.
__host__ __device__ inline uint32_t root_chunk(uint64_t rows, uint32_t tk, uint32_t items) {
  const uint64_t per = items / tk > 0 ? items / tk : 1u;
  const uint64_t ch = (rows + per - 1) / per;
  return ch < 1024u ? 1024u : (ch > 16384u ? 16384u : static_cast<uint32_t>(ch));
}

// Below the root a level's chunk is the smallest of kChunkCands candidates,
// base x (1 + k/4) with base = root_chunk(level rows, 1, budget), whose
// pieces (ceil(rows / chunk) per histogrammed node) fit the budget.  Each
// piece costs about its rows on one CU, so the level takes about its largest
// piece; powers of two overshot the fitted size by 1.4x on average.
constexpr uint32_t kChunkCands = 8;
__host__ __device__ inline uint32_t chunk_cand(uint32_t base, uint32_t k) {
  const uint32_t c = base + base * k / 4u;
  return c < 16384u ? c : 16384u;
}

// A node of n rows is cut into ceil(n / chunk) pieces of equal length (the
// last up to one piece shorter; none empty); the length of each but the last.
__host__ __device__ inline uint32_t piece_rows(uint32_t n, uint32_t chunk) {
  const uint32_t p = (n + chunk - 1) / chunk;
  return p ? (n + p - 1) / p : 0u;
}

__host__ __device__ inline uint32_t heap_index(uint32_t level, uint32_t node) {
  return (1u << level) - 1u + node;
}

// ceil(log2(n)) for n >= 1 (0 for n <= 1).
__host__ __device__ inline int ceil_log2(uint32_t n) {
  int b = 0;
  for (uint32_t v = n > 1u ? n - 1u : 0u; v != 0u; v >>= 1) ++b;
  return b;
}

// Exponent e of a histogram's fixed-point scale 2^e for values |v| <= vmax
// summed over at most n rows.  |round(v * 2^e)| <= 2^(62 - ceil_log2(n)), so
// any sum of n of them fits int64.  `packed`: also keep a sum over one work
// item's `chunk` rows inside the signed 48-bit field of the packed cell.
__host__ __device__ inline int fixed_exponent(float vmax, uint32_t n, uint32_t chunk, bool packed) {
  int e = 62 - ceil_log2(n);
  if (packed) {
    const int ec = 46 - ceil_log2(chunk);
    e = ec < e ? ec : e;
  }
  if (!(vmax > 0.f) || vmax > 3.4e38f) return e < 40 ? e : 40;  // all zero, or not finite
  int k = 0;
  (void)frexpf(vmax, &k);  // vmax <= 2^k
  e -= k;
  return e < -900 ? -900 : (e > 900 ? 900 : e);
}

}  // namespace tc
