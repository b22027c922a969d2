// Regular (non-oblivious) depthwise trees on gfx950: the reference's
// Tree::build Depthwise branch (src/tree/tree.cpp:162-217,
// build_recursive_optimized :297-390, make_leaf :655-681) with its
// SplitFinder (src/tree/histogram.cpp:569-761), the regular-tree boosting
// loops (Booster::train / build_tree, src/boosting/booster.cpp:160-232,
// :438-493; Booster::train_multiclass :1263-1400) and the regular-tree
// ensemble prediction (TreeEnsemble::predict_batch_optimized and
// predict_batch_multiclass_optimized, tree.cpp:1471-1709, over the bins of
// Booster::predict_raw_fast, booster.cpp:680-871, or Dataset::apply_bins).
//
// Part of the tc_train.hip translation unit (it launches k_hist).
//
// Growth.  The reference recurses depth first, smaller child first, and
// turns a node into a leaf when max_leaves leaves exist at its visit.  A
// node's split depends only on its own rows, so the tree is grown here one
// level at a time for every node at once -- per level one histogram launch
// (the smaller child of each pair; the larger is parent - smaller), one
// split scan per (feature, pair), one decision per node, one partition and
// one plan of the next level -- and a single thread then replays the
// reference's depth-first visit over the grown nodes: the max_leaves cut,
// the add_node numbering and the depth come out exactly as the recursion
// makes them.
//
// Node statistics follow the reference's GradientPair flow: the root's are
// the sums over its rows, a child's are the split's left prefix (bins <= b;
// NaN rows, bin 255, are not in it although the partition sends them left)
// or the parent's minus that prefix (float subtraction, count subtraction in
// uint32).  Histograms are exact fixed-point sums (as for the symmetric
// trees) with exact row counts: the h plane holds h * 2^16 + 1 per row
// (k_hist MODE 2) and MSE's packed cell carries the count (MODE 1).
#include "tc_device.h"

namespace tc {

constexpr uint32_t kRMaxDepth = 12;   // 2^12 nodes at the deepest level
constexpr uint32_t kRNone = 0xFFFFFFFFu;
constexpr uint32_t kRPlanThreads = 1024;

enum : uint8_t { kRPending = 0, kRSplit = 1, kRLeaf = 2 };

// A node of the tree being grown, stored in level order: level d's nodes are
// nodes[lvl_begin[d] .. lvl_begin[d] + lvl_count[d]).  Level d + 1 holds the
// two children of every node split at level d, in the order of their parents
// (children 2r, 2r + 1 of the r-th split node; left first).
struct RNode {
  uint32_t off, len;      // the node's rows: rows[off, off + len) of its level's row list
  float G, H;             // GradientPair sums (node_stats)
  uint32_t C;             // GradientPair count
  uint32_t parent;        // grown index of the parent (kRNone: the root)
  uint32_t path;          // one bit per level: 0 where the node is its parent's first-visited child
  uint16_t depth;
  uint8_t state;          // kRPending, kRSplit, kRLeaf
  uint8_t dl;             // default_left (learned after the partition, types.hpp:193)
  uint16_t feat;
  uint8_t bin;
  uint8_t right_first;    // the right child is the smaller one, so visited first
  float gain;
  uint32_t child;         // grown index of the left child (right = child + 1)
  uint32_t nleft;         // rows the partition sent left
  float GL, HL;           // the split's left sums
  uint32_t CL;
  uint32_t pad;
};

// One regular tree being grown (one job: a binary/regression fit, or one
// class of a multiclass iteration), device resident.
struct RJob {
  const uint8_t* bins_rm;   // [N][Fp]
  const uint8_t* bins_cm;   // [F][N]
  uint32_t N, F, Fp, T;     // T = feature tiles
  uint32_t D, max_leaves, max_bins, min_samples_leaf;  // D: levels grown
  uint32_t leafwise;        // 1: lossguide (grown to max_depth + 1, k_rleafwise replays)
  uint32_t max_depth;       // the tree's max_depth (D, or D - 1 for lossguide)
  float lambda_l2, min_child_weight, min_split_gain, gamma, max_delta_step, leaf_smooth;
  int32_t learn_missing;
  int32_t mode;             // k_hist cell mode: 1 packed g + count (MSE), 2 packed h + count
  uint32_t hist_chunk;      // rows per root work item
  uint32_t deep_items;      // below the root: work items a level is cut into (about)
  uint32_t n_root;          // rows of the root (N, or the row sample)
  const uint32_t* root_rows;  // the root's rows (null: 0 .. N-1)
  uint32_t* rows;           // [2][N] row lists of the levels below the root, by level parity
  float2* gh;               // this job's (g, h) per row [N]
  TreeState* st;            // tree index, scales, fault flags
  RNode* nodes;             // grown nodes [2^(D+1) - 1]
  uint32_t* lvl_begin;      // [D + 2]
  uint32_t* lvl_count;      // [D + 2]
  unsigned long long* keys; // [2^(D-1)][F]: each pending node's best key per feature
  HistCell* nhist;          // [2][2^(D-1)][F][256] node histograms by level parity
  uint32_t* ncnt;           // [2][2^(D-1)][F][256] their row counts
  HistCell* lastc;          // [2^(D-1)][F][256] each pair's last histogram chunk (k_hist output)
  uint32_t* pair_first;     // [2][2^(D-1)] first histogram item of each pair
  uint32_t* pair_nch;       // [2][2^(D-1)] chunks of each pair's smaller child
  unsigned long long* cursor;  // [2^(D-1)]: left position | right rows taken << 32
  uint32_t* ghmax_partials; // [grad_wg_cnt][2] per-tile max |g|, max |h|
  uint32_t grad_wg_cnt;
  // model, tree-major, NC = 2^(D+1) - 1 node slots per tree
  uint32_t n_trees, NC;
  uint16_t* m_feat;
  uint8_t* m_bin;
  uint8_t* m_flags;         // bit 0 is_leaf, bit 1 default_left (Tree::save, tree.cpp:1768)
  uint32_t* m_left;
  uint32_t* m_right;
  float* m_value;
  float* m_gain;
  uint32_t* m_nodes;        // [n_trees]
  uint32_t* m_leaves;       // [n_trees]
  uint16_t* m_depth;        // [n_trees]
  unsigned long long* hist_rows_total;
};

// Work lists of a level, shared by the jobs, two sets by level parity.
struct RLists {
  HistWork* hist[2];
  uint32_t* hist_count[2];
  PartWork* part[2];
  uint32_t* part_count[2];
};

__device__ inline uint32_t rpairs_cap(const RJob& J) { return J.D > 0 ? (1u << (J.D - 1)) : 1u; }

// The row at position pos of level d's row list.
__device__ inline uint32_t rlevel_row(const RJob& J, uint32_t d, uint32_t pos) {
  if (d == 0) return J.root_rows ? J.root_rows[pos] : pos;
  return J.rows[static_cast<size_t>(d & 1u) * J.N + pos];
}

// Tree::make_leaf (tree.cpp:655-681).
__device__ inline float rleaf_value(const RJob& J, float G, float H, uint32_t C) {
  float v = __fdiv_rn(-G, __fadd_rn(H, J.lambda_l2));
  if (J.leaf_smooth > 0.0f) {
    const float c = static_cast<float>(C);
    v = __fmul_rn(v, __fdiv_rn(c, __fadd_rn(c, J.leaf_smooth)));
  }
  if (J.max_delta_step > 0.0f) v = fmaxf(-J.max_delta_step, fminf(J.max_delta_step, v));
  return v;
}

// Histogram work items of one node: chunk-major, then tile.
__device__ inline void rput_hist_items(HistWork* out, uint32_t job, uint32_t slot, uint32_t off,
                                       uint32_t len, uint32_t chunk, uint32_t T, uint32_t e0,
                                       uint32_t e1) {
  const uint32_t nch = (len + chunk - 1) / chunk;
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t j = e / T, t = e % T;
    const uint32_t start = off + j * chunk;
    put_hist_work(out + e, job, slot, start, min(chunk, off + len - start), t, 0u, j + 1 == nch);
  }
}

// ---------------------------------------------------------------------------
// Start of a tree, one 256-thread workgroup per job: the histogram scales
// from the gradients' maxima, the root node, the root's work lists.
__global__ void __launch_bounds__(256) k_rbegin(RJob* __restrict__ jobs, RLists L) {
  RJob& J = jobs[blockIdx.y];
  const uint32_t tid = threadIdx.x;
  __shared__ uint32_t s_mx[2][4];
  __shared__ uint32_t s_hbase, s_pbase;
  __shared__ int s_go;
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
  const uint32_t n = J.n_root, T = J.T, chunk = J.hist_chunk;
  const uint32_t nch = (n + chunk - 1) / chunk, hcnt = nch * T, pcnt = part_slots(n);
  if (tid == 0) {
    TreeState* st = J.st;
    const uint32_t gbits = max(max(s_mx[0][0], s_mx[0][1]), max(s_mx[0][2], s_mx[0][3]));
    const uint32_t hbits = max(max(s_mx[1][0], s_mx[1][1]), max(s_mx[1][2], s_mx[1][3]));
    st->gmax_next = 0u;
    st->hmax_next = 0u;
    const int t = st->tree + 1;
    st->tree = t;
    int go = t < static_cast<int>(J.n_trees) && !st->stopped ? 1 : 0;
    for (uint32_t d = 0; d < J.D + 2; ++d) {
      J.lvl_begin[d] = 0;
      J.lvl_count[d] = 0;
    }
    if (go) {
      const float gmax = __uint_as_float(gbits), hmax = __uint_as_float(hbits);
      // the packed bound uses the largest work item any chunking may use
      const int eg = fixed_exponent(gmax, n, kMaxHistChunk, J.mode == 1);
      const int eh = fixed_exponent(hmax, n, kMaxHistChunk, J.mode == 2);
      st->gscale = ldexp(1.0, eg);
      st->ginv = ldexp(1.0, -eg);
      st->hscale = ldexp(1.0, eh);
      st->hinv = ldexp(1.0, -eh);
      if (n == 0) go = 0;  // no rows: a tree without nodes (tree.cpp:172-174)
    }
    if (go) {
      RNode r;
      r.off = 0; r.len = n; r.G = 0.f; r.H = 0.f; r.C = n; r.parent = kRNone; r.path = 0;
      r.depth = 0; r.state = kRPending; r.dl = 1; r.feat = 0; r.bin = 0; r.right_first = 0;
      r.gain = 0.f; r.child = 0; r.nleft = 0; r.GL = r.HL = 0.f; r.CL = 0; r.pad = 0;
      J.nodes[0] = r;
      J.lvl_count[0] = 1;
      s_hbase = atomicAdd(L.hist_count[0], hcnt);
      s_pbase = atomicAdd(L.part_count[0], pcnt);
      J.pair_first[0] = s_hbase;
      J.pair_nch[0] = nch;
      if (J.hist_rows_total) atomicAdd(J.hist_rows_total, static_cast<unsigned long long>(n));
    }
    st->active = go;
    s_go = go;
  }
  __syncthreads();
  if (!s_go) return;
  const uint32_t job = blockIdx.y;
  for (uint32_t e = tid; e < hcnt; e += 256) rput_hist_items(L.hist[0] + s_hbase, job, 0, 0, n, chunk, T, e, e + 1);
  for (uint32_t e = tid; e < pcnt; e += 256) {
    const uint32_t start = e * kPartChunk;
    put_part_work(L.part[0] + s_pbase + e, job, 0, start, min(static_cast<uint32_t>(kPartChunk), n - start), e);
  }
}

// ---------------------------------------------------------------------------
// Split scan of level d: workgroup (feature f, pair p, job), thread = bin.
// The pair's smaller child is summed from its histogram chunks (exact), the
// larger is parent - smaller (Histogram::subtract_from, histogram.cpp:39-42),
// both are kept for the next level, and every pending node of the pair gets
// its best bin on f: SplitFinder::find_best_split_feature (histogram.cpp:
// 617-679) -- left = the exact prefix over bins <= b (each sum rounded once
// to float), right = parent - left in float, both sides need count >=
// min_samples_leaf and hess >= min_child_weight, gain = 0.5 * (GL^2/(HL+l) +
// GR^2/(HR+l) - G^2/(H+l)) - gamma with its compiler's contraction, kept
// when gain >= min_split_gain; the best key per (node, feature) has ties to
// the lowest bin (and, over features, the lowest feature).
__global__ void __launch_bounds__(256) k_rsplit(const RJob* __restrict__ jobs, uint32_t d,
                                                const HistCell* __restrict__ partials) {
  const RJob& J = jobs[blockIdx.z];
  const uint32_t f = blockIdx.x, p = blockIdx.y, b = threadIdx.x;
  if (f >= J.F || !J.st->active) return;
  const uint32_t b0 = J.lvl_begin[d], cnt = J.lvl_count[d];
  if (d == 0 ? (p > 0 || cnt == 0) : (2 * p + 1 >= cnt)) return;
  const uint32_t na = b0 + (d == 0 ? 0u : 2u * p);
  const RNode A = J.nodes[na];
  RNode Cn = A;
  if (d > 0) Cn = J.nodes[na + 1];
  const bool pa = A.state == kRPending, pc = d > 0 && Cn.state == kRPending;
  if (!pa && !pc) return;
  const bool right_small = d > 0 && Cn.len < A.len;  // nl <= nr: the left child is the smaller
  const uint32_t par = d & 1u, PC = rpairs_cap(J);
  const uint32_t nch = J.pair_nch[par * PC + p], first = J.pair_first[par * PC + p];
  const uint32_t F = J.F, tile = f / kTileF, fi = f % kTileF;
  const double gi = J.st->ginv, hi = J.st->hinv;
  const long long h2 = J.mode == 1 ? __double2ll_rn(2.0 * J.st->hscale) : 1;
  // the smaller child's cells: its chunks' partial slots, then the last chunk
  long long sg = 0, sh = 0;
  uint32_t sc = 0;
  for (uint32_t j = 0; j < nch; ++j) {
    const HistCell c = j + 1 == nch
                           ? J.lastc[(static_cast<size_t>(p) * F + f) * 256u + b]
                           : partials[(static_cast<size_t>(first + j * J.T + tile) * kTileF + fi) * 256u + b];
    sg += c.g;
    if (J.mode == 1) {
      sh += c.h;
      sc += static_cast<uint32_t>(c.h / h2);
    } else {
      sh += c.h >> 16;
      sc += static_cast<uint32_t>(c.h & 0xFFFF);
    }
  }
  // both nodes' cells
  long long ag = sg, ah = sh, cg = 0, ch = 0;
  uint32_t ac = sc, cc = 0;
  if (d > 0) {
    const uint32_t ps = A.parent - J.lvl_begin[d - 1];
    const size_t pcell = ((static_cast<size_t>(par ^ 1u) * PC + ps) * F + f) * 256u + b;
    const HistCell P = J.nhist[pcell];
    const uint32_t pcn = J.ncnt[pcell];
    const long long lg = P.g - sg, lh = P.h - sh;
    const uint32_t lc = pcn - sc;
    if (right_small) {
      cg = sg; ch = sh; cc = sc;
      ag = lg; ah = lh; ac = lc;
    } else {
      cg = lg; ch = lh; cc = lc;
    }
  }
  __shared__ long long s_w[3][4];
  __shared__ unsigned long long s_key[4];
  const uint32_t lane = b & 63u, wave = b >> 6;
  for (int side = 0; side < 2; ++side) {
    if (side == 1 && d == 0) break;
    const bool pend = side == 0 ? pa : pc;
    const RNode& X = side == 0 ? A : Cn;
    const long long xg = side == 0 ? ag : cg, xh = side == 0 ? ah : ch;
    const uint32_t xc = side == 0 ? ac : cc;
    const uint32_t slot = (side == 0 ? na : na + 1) - b0;
    if (pend) {  // kept for the next level (the node may split)
      const size_t o = ((static_cast<size_t>(par) * PC + slot) * F + f) * 256u + b;
      // slot < 2^(D-1) for a pending node: level D holds leaves only
      J.nhist[o] = HistCell{xg, xh};
      J.ncnt[o] = xc;
    }
    // exact inclusive prefix over bins (all three sums), every thread
    long long pg = wave_incl_scan_i64(xg), ph = wave_incl_scan_i64(xh);
    uint32_t pcnt = wave_incl_scan_u32(xc);
    if (lane == 63) {
      s_w[0][wave] = pg;
      s_w[1][wave] = ph;
      s_w[2][wave] = static_cast<long long>(pcnt);
    }
    __syncthreads();
    long long tg = 0, th = 0, tc = 0;
    for (uint32_t w = 0; w < 4; ++w) {
      if (w < wave) {
        pg += s_w[0][w];
        ph += s_w[1][w];
        pcnt += static_cast<uint32_t>(s_w[2][w]);
      }
      tg += s_w[0][w];
      th += s_w[1][w];
      tc += s_w[2][w];
    }
    __syncthreads();
    float G = X.G, H = X.H;
    const uint32_t C = X.C;
    if (d == 0) {  // the root's sums over its rows (every feature's cells add up to them)
      G = static_cast<float>(static_cast<double>(tg) * gi);
      H = static_cast<float>(static_cast<double>(th) * hi);
      if (f == 0 && b == 0) {
        J.nodes[0].G = G;
        J.nodes[0].H = H;
      }
    }
    (void)tc;
    unsigned long long key = kEmptyKey;
    if (pend && b + 1 < (J.max_bins & 0xFFu)) {
      const float lam = J.lambda_l2;
      const float GL = static_cast<float>(static_cast<double>(pg) * gi);
      const float HL = static_cast<float>(static_cast<double>(ph) * hi);
      const uint32_t CL = pcnt;
      const float GR = __fsub_rn(G, GL), HR = __fsub_rn(H, HL);
      const uint32_t CR = C - CL;  // uint32, as GradientPair's count
      const float mcw = J.min_child_weight;
      const uint32_t msl = J.min_samples_leaf;
      if (!(CL < msl || HL < mcw || CR < msl || HR < mcw)) {
        const float pgain = __fdiv_rn(__fmul_rn(G, G), __fadd_rn(H, lam));
        const float gl = __fdiv_rn(__fmul_rn(GL, GL), __fadd_rn(HL, lam));
        const float gr = __fdiv_rn(__fmul_rn(GR, GR), __fadd_rn(HR, lam));
        const float gain = __fmaf_rn(0.5f, __fsub_rn(__fadd_rn(gl, gr), pgain), -J.gamma);
        if (gain >= J.min_split_gain && gain > -1e30f) key = split_key(gain == 0.f ? 0.f : gain, f, b);
      }
    }
    key = wave_max_u64(key);
    if (lane == 0) s_key[wave] = key;
    __syncthreads();
    if (b == 0 && pend) {
      unsigned long long m = s_key[0];
      for (int w = 1; w < 4; ++w) m = s_key[w] > m ? s_key[w] : m;
      J.keys[static_cast<size_t>(slot) * F + f] = m;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Decision of level d, one wave per node: the root's stop test (tree.cpp:
// 310-319; deeper nodes had theirs at the plan), the best key over features,
// split or leaf (tree.cpp:323-330), the split's exact left sums, the node's
// partition cursor.  Workgroup (0, 0) also empties the next level's lists.
__global__ void __launch_bounds__(256) k_rdecide(RJob* __restrict__ jobs, uint32_t d, RLists L) {
  RJob& J = jobs[blockIdx.y];
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    *L.hist_count[(d + 1) & 1u] = 0;
    *L.part_count[(d + 1) & 1u] = 0;
  }
  if (!J.st->active) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t b0 = J.lvl_begin[d];
  if (slot >= J.lvl_count[d]) return;
  RNode& X = J.nodes[b0 + slot];
  if (X.state != kRPending) return;
  if (d == 0 && (d >= J.D || X.len < 2u * J.min_samples_leaf || X.H < 2.0f * J.min_child_weight)) {
    if (lane == 0) X.state = kRLeaf;
    return;
  }
  const uint32_t F = J.F;
  unsigned long long m = kEmptyKey;
  for (uint32_t f = lane; f < F; f += 64) {
    const unsigned long long k = J.keys[static_cast<size_t>(slot) * F + f];
    m = k > m ? k : m;
  }
  m = wave_max_u64(m);
  const float gain = float_unorder(static_cast<uint32_t>(m >> 32));
  if (m == kEmptyKey || gain < J.min_split_gain) {
    if (lane == 0) X.state = kRLeaf;
    return;
  }
  const uint32_t fb = 0xFFFFFFFFu - static_cast<uint32_t>(m & 0xFFFFFFFFu);
  const uint32_t feat = fb >> 8, bin = fb & 255u;
  const uint32_t PC = rpairs_cap(J);
  const size_t o = ((static_cast<size_t>(d & 1u) * PC + slot) * F + feat) * 256u;
  long long g = 0, h = 0;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t bb = 4 * lane + i;
    if (bb <= bin) {
      const HistCell v = J.nhist[o + bb];
      g += v.g;
      h += v.h;
      c += J.ncnt[o + bb];
    }
  }
  g = wave_sum_i64(g);
  h = wave_sum_i64(h);
  c = wave_sum_u32(c);
  if (lane == 0) {
    X.state = kRSplit;
    X.feat = static_cast<uint16_t>(feat);
    X.bin = static_cast<uint8_t>(bin);
    X.gain = gain;
    X.GL = static_cast<float>(static_cast<double>(g) * J.st->ginv);
    X.HL = static_cast<float>(static_cast<double>(h) * J.st->hinv);
    X.CL = c;
    J.cursor[slot] = static_cast<unsigned long long>(X.off);
  }
}

// ---------------------------------------------------------------------------
// Partition of level d (tree.cpp:40-69, :341-359): every split node's rows,
// NaN (bin 255) to the left (a fresh node's default_left), bin <= split bin
// to the left, the rest to the right, into the next level's row list inside
// the node's own segment (left filled up from its start, right down from its
// end), placed by wave ballots after one 64-bit reservation per item.
__global__ void __launch_bounds__(kPartThreads) k_rpartition(RJob* __restrict__ jobs,
                                                             const PartWork* __restrict__ work,
                                                             const uint32_t* __restrict__ count,
                                                             uint32_t d) {
  constexpr int kIt = kPartChunk / kPartThreads;
  constexpr int kW = kPartThreads / 64;
  const uint32_t cnt = *count;
  if (blockIdx.x >= cnt) return;
  const PartWork w = work[blockIdx.x];
  RJob& J = jobs[w.job];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint32_t rv[kIt];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const uint32_t j = k * kPartThreads + tid;
    rv[k] = j < w.len ? rlevel_row(J, d, w.start + j) : 0u;
  }
  if (!J.st->active) return;
  const RNode& X = J.nodes[J.lvl_begin[d] + w.node];
  if (X.state != kRSplit) return;
  const uint8_t* col = J.bins_cm + static_cast<size_t>(X.feat) * J.N;
  const uint8_t thr = X.bin;
  uint32_t* rows_n = J.rows + static_cast<size_t>((d + 1) & 1u) * J.N;
  unsigned long long* cur = J.cursor + w.node;
  __shared__ uint32_t s_cnt[2][kIt * kW];
  __shared__ uint32_t s_base[2], s_tot[2];
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t lbits = 0, rbits = 0;
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const uint32_t j = k * kPartThreads + tid;
    const bool valid = j < w.len;
    const uint8_t bb = valid ? col[rv[k]] : 0;
    const bool left = valid && (bb == 255 || bb <= thr);
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
  static_assert(kIt * kW == 64, "one wave scans one side's (step, wave) counts");
  if (wv < 2) {
    const uint32_t v = s_cnt[wv][lane];
    const uint32_t incl = wave_incl_scan_u32(v);
    s_cnt[wv][lane] = incl - v;
    if (lane == 63) s_tot[wv] = incl;
  }
  __syncthreads();
  if (tid == 0) {
    const unsigned long long add =
        static_cast<unsigned long long>(s_tot[0]) | (static_cast<unsigned long long>(s_tot[1]) << 32);
    const unsigned long long old = add == 0 ? 0ull : atomicAdd(cur, add);
    s_base[0] = static_cast<uint32_t>(old);
    s_base[1] = X.off + X.len - static_cast<uint32_t>(old >> 32) - s_tot[1];
  }
  __syncthreads();
  {  // both reservations inside the node's segment, or nothing is written
    const uint32_t lo = X.off, hi = X.off + X.len;
    const bool bad = (s_tot[0] && (s_base[0] < lo || s_base[0] > hi || s_tot[0] > hi - s_base[0])) ||
                     (s_tot[1] && (s_base[1] < lo || s_base[1] > hi || s_tot[1] > hi - s_base[1]));
    if (bad) {
      if (tid == 0) atomicOr(&J.st->fault, 1u);
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const bool left = (lbits >> k) & 1u, right = (rbits >> k) & 1u;
    const unsigned long long ml = __ballot(left), mr = __ballot(right);
    const uint32_t dst = left ? s_base[0] + s_cnt[0][k * kW + wv] + __popcll(ml & lt)
                              : s_base[1] + s_cnt[1][k * kW + wv] + __popcll(mr & lt);
    if (left || right) rows_n[dst] = rv[k];
  }
}

// In-place exclusive scan of a[0..n) in LDS by a kRPlanThreads workgroup
// (n <= 4 * kRPlanThreads); returns the total.  tmp: 16 words of LDS.
__device__ inline uint32_t rblock_scan_excl(uint32_t* a, uint32_t n, uint32_t* tmp) {
  constexpr uint32_t NW = kRPlanThreads / 64;
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
  for (uint32_t w = 0; w < NW; ++w) {
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

// ---------------------------------------------------------------------------
// Plan of level d + 1 from the partition of level d, one workgroup per job:
// the children of every split node (segments, GradientPair sums, stop test
// of tree.cpp:310-319), the split node's learned default_left (tree.cpp:
// 353-356) and visiting order (the smaller child first, :370-390), the
// histogram of each pair's smaller child (when one of the two may split),
// the split scan's sources and both work lists.
__global__ void __launch_bounds__(kRPlanThreads) k_rplan(RJob* __restrict__ jobs, uint32_t d,
                                                        RLists L) {
  RJob& J = jobs[blockIdx.y];
  const uint32_t job = blockIdx.y;
  if (!J.st->active) return;
  constexpr uint32_t kMaxLvl = 1u << (kRMaxDepth - 1);  // nodes of the widest split level
  __shared__ uint32_t s_rank[kMaxLvl];       // split nodes -> rank
  __shared__ uint32_t s_hw[kMaxLvl];         // pairs: histogram items -> first item
  __shared__ uint32_t s_pw[2 * kMaxLvl];     // children: partition items -> first item
  __shared__ uint32_t s_tmp[16];
  __shared__ uint32_t s_hbase, s_pbase, s_chunk;
  __shared__ unsigned long long s_rows;
  const uint32_t tid = threadIdx.x;
  const uint32_t b0 = J.lvl_begin[d], n = J.lvl_count[d], b1 = b0 + n;
  if (tid == 0) s_rows = 0ull;
  for (uint32_t s = tid; s < n; s += kRPlanThreads) s_rank[s] = J.nodes[b0 + s].state == kRSplit ? 1u : 0u;
  __syncthreads();
  const uint32_t nsplit = rblock_scan_excl(s_rank, n, s_tmp);
  if (tid == 0) {
    J.lvl_begin[d + 1] = b1;
    J.lvl_count[d + 1] = 2 * nsplit;
  }
  if (nsplit == 0) {
    if (tid == 0) J.st->active = 0;  // nothing left to grow
    return;
  }
  const bool last = d + 1 >= J.D;  // the children are leaves
  const float mcw2 = 2.0f * J.min_child_weight;
  const uint32_t msl2 = 2u * J.min_samples_leaf;
  // children, and each pair's smaller-child rows (0: no histogram needed)
  for (uint32_t s = tid; s < n; s += kRPlanThreads) {
    RNode& P = J.nodes[b0 + s];
    if (P.state != kRSplit) continue;
    const uint32_t r = s_rank[s];
    const unsigned long long cw = J.cursor[s];
    const uint32_t nl = static_cast<uint32_t>(cw) - P.off, nr = static_cast<uint32_t>(cw >> 32);
    if (nl + nr != P.len) atomicOr(&J.st->fault, 1u);  // the partition lost rows
    uint8_t dl = 1;
    if (J.learn_missing && nl != P.len && nr != P.len) dl = nl >= nr ? 1 : 0;
    const bool right_first = nl > nr;
    P.dl = dl;
    P.nleft = nl;
    P.right_first = right_first ? 1 : 0;
    P.child = b1 + 2 * r;
    RNode c[2];
    for (int k = 0; k < 2; ++k) {
      RNode& x = c[k];
      x.off = k == 0 ? P.off : P.off + nl;
      x.len = k == 0 ? nl : nr;
      x.G = k == 0 ? P.GL : __fsub_rn(P.G, P.GL);
      x.H = k == 0 ? P.HL : __fsub_rn(P.H, P.HL);
      x.C = k == 0 ? P.CL : P.C - P.CL;
      x.parent = b0 + s;
      x.path = (P.path << 1) | ((k == 0) == right_first ? 1u : 0u);
      x.depth = static_cast<uint16_t>(d + 1);
      x.state = (last || x.len < msl2 || x.H < mcw2) ? kRLeaf : kRPending;
      x.dl = 1; x.feat = 0; x.bin = 0; x.right_first = 0; x.gain = 0.f; x.child = 0; x.nleft = 0;
      x.GL = x.HL = 0.f; x.CL = 0; x.pad = 0;
      J.nodes[b1 + 2 * r + k] = x;
    }
    const bool need = c[0].state == kRPending || c[1].state == kRPending;
    const uint32_t sm = need ? (nl <= nr ? nl : nr) : 0u;
    s_hw[r] = sm;
    if (sm) atomicAdd(&s_rows, static_cast<unsigned long long>(sm));
    s_pw[2 * r] = c[0].state == kRPending ? part_slots(nl) : 0u;
    s_pw[2 * r + 1] = c[1].state == kRPending ? part_slots(nr) : 0u;
  }
  __syncthreads();
  if (last) return;
  if (tid == 0) {
    // rows per work item: the level's histogram rows cut into about deep_items items
    const uint64_t per = max(1u, J.deep_items / J.T);
    const uint64_t chn = (s_rows + per - 1) / per;
    s_chunk = chn < 1024u ? 1024u : (chn > 16384u ? 16384u : static_cast<uint32_t>(chn));
    if (J.hist_rows_total) atomicAdd(J.hist_rows_total, s_rows);
  }
  __syncthreads();
  const uint32_t chunk = s_chunk, T = J.T;
  for (uint32_t r = tid; r < nsplit; r += kRPlanThreads) s_hw[r] = ((s_hw[r] + chunk - 1) / chunk) * T;
  __syncthreads();
  const uint32_t htot = rblock_scan_excl(s_hw, nsplit, s_tmp);
  const uint32_t ptot = rblock_scan_excl(s_pw, 2 * nsplit, s_tmp);
  if (tid == 0) {
    s_hbase = atomicAdd(L.hist_count[(d + 1) & 1u], htot);
    s_pbase = atomicAdd(L.part_count[(d + 1) & 1u], ptot);
  }
  __syncthreads();
  const uint32_t par = (d + 1) & 1u, PC = rpairs_cap(J);
  for (uint32_t r = tid; r < nsplit; r += kRPlanThreads) {
    const uint32_t e0 = s_hw[r], e1 = r + 1 < nsplit ? s_hw[r + 1] : htot;
    J.pair_first[par * PC + r] = s_hbase + e0;
    J.pair_nch[par * PC + r] = (e1 - e0) / T;
  }
  // item e: pair r = the last with first item <= e, chunk-major then tile
  for (uint32_t e = tid; e < htot; e += kRPlanThreads) {
    const uint32_t r = find_item(s_hw, nsplit, e);
    const RNode& a = J.nodes[b1 + 2 * r];
    const RNode& c = J.nodes[b1 + 2 * r + 1];
    const RNode& sm = a.len <= c.len ? a : c;
    const uint32_t le = e - s_hw[r];
    rput_hist_items(L.hist[par] + s_hbase + s_hw[r], job, r, sm.off, sm.len, chunk, T, le, le + 1);
  }
  PartWork* pout = L.part[par] + s_pbase;
  for (uint32_t e = tid; e < ptot; e += kRPlanThreads) {
    const uint32_t c = find_item(s_pw, 2 * nsplit, e);
    const RNode& x = J.nodes[b1 + c];
    const uint32_t start = x.off + (e - s_pw[c]) * kPartChunk;
    put_part_work(pout + e, job, c, start, min(static_cast<uint32_t>(kPartChunk), x.off + x.len - start), e);
  }
}

// ---------------------------------------------------------------------------
// End of a tree, one workgroup per job: one thread replays the reference's
// depth-first build over the grown nodes -- a node is visited, becomes a
// leaf if it did not split or if max_leaves leaves exist at its visit
// (tree.cpp:310-319), else takes ids n, n + 1 for its children (add_node,
// :357-363, :648-653) whose smaller one is visited first (:370-390) -- and
// writes the tree in that numbering (leaf values: make_leaf, :655-681).
// The walk is a chain of dependent reads, so the workgroup first copies what
// it needs of every grown node into LDS (up to kRFinLds nodes; a larger tree
// walks the HBM records).  Also empties the level-0 lists for the next tree.
struct RFin {
  float G, H;
  uint32_t C, child;
  float gain;
  uint16_t feat, depth;
  uint8_t bin, split, dl, right_first;
};
constexpr uint32_t kRFinLds = 2048;

__device__ inline RFin rfin_of(const RNode& X) {
  RFin r;
  r.G = X.G; r.H = X.H; r.C = X.C; r.child = X.child; r.gain = X.gain;
  r.feat = X.feat; r.depth = X.depth; r.bin = X.bin;
  r.split = X.state == kRSplit ? 1 : 0;
  r.dl = X.dl; r.right_first = X.right_first;
  return r;
}

__global__ void __launch_bounds__(256) k_rfinish(RJob* __restrict__ jobs, RLists L) {
  RJob& J = jobs[blockIdx.y];
  __shared__ RFin s_n[kRFinLds];
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    *L.hist_count[0] = 0;
    *L.part_count[0] = 0;
  }
  TreeState* st = J.st;
  const int t = st->tree;
  if (t < 0 || t >= static_cast<int>(J.n_trees) || st->stopped) return;
  uint32_t ng = 0;  // grown nodes: every level's
  for (uint32_t d = 0; d <= J.D; ++d) ng = max(ng, J.lvl_begin[d] + J.lvl_count[d]);
  const bool staged = ng <= kRFinLds;
  if (staged)
    for (uint32_t i = tid; i < ng; i += 256) s_n[i] = rfin_of(J.nodes[i]);
  __syncthreads();
  if (tid != 0) return;
  // the job's fields in registers: the stores below could otherwise alias the
  // job record and force a reload (and a wait) after each of them
  const size_t o = static_cast<size_t>(t) * J.NC;
  uint16_t* __restrict__ m_feat = J.m_feat + o;
  uint8_t* __restrict__ m_bin = J.m_bin + o;
  uint8_t* __restrict__ m_flags = J.m_flags + o;
  uint32_t* __restrict__ m_left = J.m_left + o;
  uint32_t* __restrict__ m_right = J.m_right + o;
  float* __restrict__ m_value = J.m_value + o;
  float* __restrict__ m_gain = J.m_gain + o;
  const RNode* __restrict__ nodes = J.nodes;
  const uint32_t max_leaves = J.max_leaves;
  const float lam = J.lambda_l2, smooth = J.leaf_smooth, mds = J.max_delta_step;
  __shared__ uint32_t s_sg[2 * kRMaxDepth + 4], s_sid[2 * kRMaxDepth + 4];  // the DFS stack
  uint32_t next = 0, leaves = 0, depth = 0;
  if (J.lvl_count[0] > 0) {
    int sp = 1;
    s_sg[0] = 0;
    s_sid[0] = 0;
    next = 1;
    while (sp > 0) {
      --sp;
      const uint32_t g = s_sg[sp], id = s_sid[sp];
      const RFin X = staged ? s_n[g] : rfin_of(nodes[g]);
      depth = max(depth, static_cast<uint32_t>(X.depth));
      if (X.split && leaves < max_leaves) {
        const uint32_t l = next, r = next + 1;
        next += 2;
        m_feat[id] = X.feat;
        m_bin[id] = X.bin;
        m_flags[id] = static_cast<uint8_t>((X.dl & 1u) << 1);
        m_left[id] = l;
        m_right[id] = r;
        m_value[id] = 0.f;
        m_gain[id] = X.gain;
        // pushed larger first: the smaller child's subtree is built first
        const uint32_t a = X.right_first ? X.child : X.child + 1, b = X.right_first ? X.child + 1 : X.child;
        s_sg[sp] = a; s_sid[sp] = X.right_first ? l : r;
        s_sg[sp + 1] = b; s_sid[sp + 1] = X.right_first ? r : l;
        sp += 2;
      } else {
        m_feat[id] = 0;
        m_bin[id] = 0;
        m_flags[id] = 3u;  // is_leaf, default_left of a fresh node
        m_left[id] = 0;
        m_right[id] = 0;
        // Tree::make_leaf (tree.cpp:655-681)
        float v = __fdiv_rn(-X.G, __fadd_rn(X.H, lam));
        if (smooth > 0.0f) {
          const float c = static_cast<float>(X.C);
          v = __fmul_rn(v, __fdiv_rn(c, __fadd_rn(c, smooth)));
        }
        if (mds > 0.0f) v = fmaxf(-mds, fminf(mds, v));
        m_value[id] = v;
        m_gain[id] = 0.f;
        ++leaves;
      }
    }
  }
  J.m_nodes[t] = next;
  J.m_leaves[t] = leaves;
  J.m_depth[t] = static_cast<uint16_t>(depth);
}

// End of a lossguide tree (Tree::build_leafwise, tree.cpp:685-925), one
// workgroup per job.  Every node's split depends on its own rows only, so the
// tree was grown level by level to max_depth + 1 (every node the leaf-wise
// loop may queue has its split); one thread now replays the reference's loop
// over the grown nodes: the root is queued if its split is valid; while
// candidates are queued and fewer than max_leaves leaves exist, the one of
// largest gain is popped -- a leaf if it is at max_depth, else it splits and
// its children (ids n, n + 1; left first) are queued when they may split
// (rows >= 2 min_samples_leaf, hess >= 2 min_child_weight, a valid split: the
// grown node split) and leaves + 1 < max_leaves, or become leaves; what is
// left in the queue becomes leaves.  The queue is std::priority_queue's
// binary max-heap by gain, pushed and popped in libstdc++'s steps, so that
// equal gains leave it in the reference's order.
struct RCand {
  float gain;
  uint32_t g, id;  // grown node, id in the tree
};
constexpr uint32_t kRLeafLds = 1024;    // grown nodes staged in LDS
constexpr uint32_t kRHeap = 1u << (kRMaxDepth - 1);  // queued at once: at most 2^max_depth

__device__ inline void rq_up(RCand* q, uint32_t hole, RCand v) {
  while (hole > 0) {  // std::__push_heap: past every parent that compares less
    const uint32_t parent = (hole - 1) / 2;
    if (!(q[parent].gain < v.gain)) break;
    q[hole] = q[parent];
    hole = parent;
  }
  q[hole] = v;
}

__device__ inline RCand rq_pop(RCand* q, uint32_t* n) {
  const RCand top = q[0];
  const uint32_t len = *n - 1;  // the heap without its last slot
  if (len > 0) {  // std::__pop_heap + std::__adjust_heap
    const RCand v = q[len];
    uint32_t hole = 0, second = 0;
    while (second < (len - 1) / 2) {  // the hole to a leaf along the larger child
      second = 2 * (second + 1);
      if (q[second].gain < q[second - 1].gain) --second;
      q[hole] = q[second];
      hole = second;
    }
    if ((len & 1u) == 0 && second == (len - 2) / 2) {
      second = 2 * (second + 1);
      q[hole] = q[second - 1];
      hole = second - 1;
    }
    rq_up(q, hole, v);
  }
  *n = len;
  return top;
}

__global__ void __launch_bounds__(256) k_rleafwise(RJob* __restrict__ jobs, RLists L) {
  RJob& J = jobs[blockIdx.y];
  __shared__ RFin s_n[kRLeafLds];
  __shared__ RCand s_q[kRHeap];
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    *L.hist_count[0] = 0;
    *L.part_count[0] = 0;
  }
  TreeState* st = J.st;
  const int t = st->tree;
  if (t < 0 || t >= static_cast<int>(J.n_trees) || st->stopped) return;
  uint32_t ng = 0;  // grown nodes: every level's
  for (uint32_t d = 0; d <= J.D; ++d) ng = max(ng, J.lvl_begin[d] + J.lvl_count[d]);
  const bool staged = ng <= kRLeafLds;
  if (staged)
    for (uint32_t i = tid; i < ng; i += 256) s_n[i] = rfin_of(J.nodes[i]);
  __syncthreads();
  if (tid != 0) return;
  const size_t o = static_cast<size_t>(t) * J.NC;
  uint16_t* __restrict__ m_feat = J.m_feat + o;
  uint8_t* __restrict__ m_bin = J.m_bin + o;
  uint8_t* __restrict__ m_flags = J.m_flags + o;
  uint32_t* __restrict__ m_left = J.m_left + o;
  uint32_t* __restrict__ m_right = J.m_right + o;
  float* __restrict__ m_value = J.m_value + o;
  float* __restrict__ m_gain = J.m_gain + o;
  const RNode* __restrict__ nodes = J.nodes;
  const uint32_t max_leaves = J.max_leaves, max_depth = J.max_depth, NC = J.NC;
  const float lam = J.lambda_l2, smooth = J.leaf_smooth, mds = J.max_delta_step;
  uint32_t next = 0, leaves = 0, depth = 0, qn = 0;
  auto leaf = [&](uint32_t id, const RFin& X) {  // Tree::make_leaf (tree.cpp:655-681)
    m_feat[id] = 0;
    m_bin[id] = 0;
    m_flags[id] = 3u;  // is_leaf, default_left of a fresh node
    m_left[id] = 0;
    m_right[id] = 0;
    float v = __fdiv_rn(-X.G, __fadd_rn(X.H, lam));
    if (smooth > 0.0f) {
      const float c = static_cast<float>(X.C);
      v = __fmul_rn(v, __fdiv_rn(c, __fadd_rn(c, smooth)));
    }
    if (mds > 0.0f) v = fmaxf(-mds, fminf(mds, v));
    m_value[id] = v;
    m_gain[id] = 0.f;
    ++leaves;
  };
  if (J.lvl_count[0] > 0) {
    next = 1;
    const RFin R = staged ? s_n[0] : rfin_of(nodes[0]);
    if (R.split) {
      s_q[0] = RCand{R.gain, 0u, 0u};
      qn = 1;
    } else {
      leaf(0, R);
    }
    while (qn > 0 && leaves < max_leaves) {
      const RCand c = rq_pop(s_q, &qn);
      const RFin X = staged ? s_n[c.g] : rfin_of(nodes[c.g]);
      if (X.depth >= max_depth || next + 2 > NC) {  // (the slots fit a depth-max_depth tree)
        leaf(c.id, X);
        continue;
      }
      const uint32_t l = next, r = next + 1;
      next += 2;
      m_feat[c.id] = X.feat;
      m_bin[c.id] = X.bin;
      m_flags[c.id] = static_cast<uint8_t>((X.dl & 1u) << 1);
      m_left[c.id] = l;
      m_right[c.id] = r;
      m_value[c.id] = 0.f;
      m_gain[c.id] = X.gain;
      depth = max(depth, static_cast<uint32_t>(X.depth) + 1u);
      for (uint32_t side = 0; side < 2; ++side) {
        const uint32_t g = X.child + side, id = side ? r : l;
        const RFin Y = staged ? s_n[g] : rfin_of(nodes[g]);
        if (Y.split && leaves + 1 < max_leaves && qn < kRHeap) {
          rq_up(s_q, qn, RCand{Y.gain, g, id});
          ++qn;
        } else {
          leaf(id, Y);
        }
      }
    }
    while (qn > 0) {
      const RCand c = rq_pop(s_q, &qn);
      leaf(c.id, staged ? s_n[c.g] : rfin_of(nodes[c.g]));
    }
  }
  J.m_nodes[t] = next;
  J.m_leaves[t] = leaves;
  J.m_depth[t] = static_cast<uint16_t>(depth);
}

// Tree::predict(dataset, row) (tree.cpp:929-950): the binned walk of tree t
// of job J for one row, NaN (bin 255) by the node's default_left.
__device__ __forceinline__ float rtree_row_value(const RJob& J, size_t o, const uint8_t* br) {
  uint32_t k = 0;
  for (uint32_t step = 0; step <= kRMaxDepth; ++step) {
    const uint8_t fl = J.m_flags[o + k];
    if (fl & 1u) return J.m_value[o + k];
    const uint8_t b = br[J.m_feat[o + k]];
    if (b == 255) k = (fl & 2u) ? J.m_left[o + k] : J.m_right[o + k];
    else k = b > J.m_bin[o + k] ? J.m_right[o + k] : J.m_left[o + k];
  }
  return J.m_value[o + k];
}

// The last tree of every job staged in LDS for the row walks (12 bytes a
// node; u16 child ids), when the trees hold at most kRWalkLds nodes in all;
// otherwise the walks read the model arrays.  s_base[k]: job k's first node,
// s_base[njobs]: the total (0 when no job has a tree to apply).
struct RWalk {
  uint32_t a;   // feature | bin << 16 | flags << 24
  uint32_t lr;  // left | right << 16
  float v;      // leaf value
};
constexpr uint32_t kRWalkLds = 2560;

__device__ inline bool rwalk_stage(const RJob* jobs, uint32_t njobs, RWalk* s_w, uint32_t* s_base) {
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < njobs; ++k) {
      s_base[k] = tot;
      const RJob& J = jobs[k];
      const int t = J.st->tree;
      if (t >= 0 && t < static_cast<int>(J.n_trees) && !J.st->stopped) tot += J.m_nodes[t];
    }
    s_base[njobs] = tot;
  }
  __syncthreads();
  const bool staged = s_base[njobs] <= kRWalkLds;
  if (staged)
    for (uint32_t k = 0; k < njobs; ++k) {
      const RJob& J = jobs[k];
      const uint32_t nn = s_base[k + 1] - s_base[k];
      const size_t o = static_cast<size_t>(max(J.st->tree, 0)) * J.NC;
      for (uint32_t i = tid; i < nn; i += blockDim.x) {
        RWalk w;
        w.a = J.m_feat[o + i] | (static_cast<uint32_t>(J.m_bin[o + i]) << 16) |
              (static_cast<uint32_t>(J.m_flags[o + i]) << 24);
        w.lr = J.m_left[o + i] | (J.m_right[o + i] << 16);
        w.v = J.m_value[o + i];
        s_w[s_base[k] + i] = w;
      }
    }
  __syncthreads();
  return staged;
}

// Tree::predict(dataset, row) over a staged tree.
__device__ __forceinline__ float rwalk_lds(const RWalk* nd, const uint8_t* br) {
  uint32_t k = 0;
  for (uint32_t step = 0; step <= kRMaxDepth; ++step) {
    const RWalk x = nd[k];
    const uint32_t fl = x.a >> 24;
    if (fl & 1u) return x.v;
    const uint8_t b = br[x.a & 0xFFFFu];
    const bool right = b == 255 ? !(fl & 2u) : b > ((x.a >> 16) & 0xFFu);
    k = right ? (x.lr >> 16) : (x.lr & 0xFFFFu);
  }
  return nd[k].v;
}

// The prediction update and the next gradients, over rows (1024 per
// 256-thread workgroup), for every job at once.  mode 0: predictions = base
// (before the first tree).  mode 1: predictions += lr * value of the last
// tree of each job (booster.cpp:485-490, :1378-1383, one fused multiply-add
// as the reference's compiler contracts it), then the gradients of the loss
// (A6/A7; K > 1 jobs: the K classes' softmax gradients, A8) and each job's
// per-tile max |g|, max |h| for the next tree's scales.
__global__ void __launch_bounds__(256) k_rupdate(const RJob* __restrict__ jobs, uint32_t njobs,
                                                 int mode, int loss, float lr, const float* base,
                                                 const float* __restrict__ y, float* __restrict__ preds,
                                                 float2* __restrict__ gh) {
  const RJob& J0 = jobs[0];
  const uint32_t N = J0.N, tid = threadIdx.x, lw = blockIdx.x;
  const uint32_t vecN = (N / 8u) * 8u;
  constexpr int kQ = kGradTile / 256;  // rows per thread: lw * 1024 + q * 256 + tid
  static_assert(kQ == 4, "four rows per thread");
  __shared__ RWalk s_w[kRWalkLds];
  __shared__ uint32_t s_base[kMaxClasses + 1];
  uint32_t row[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) row[q] = lw * kGradTile + q * 256 + tid;
  // phase 1: every job's prediction of the thread's rows
  const bool staged = mode == 1 && rwalk_stage(jobs, njobs, s_w, s_base);
  const uint8_t* br[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) br[q] = J0.bins_rm + static_cast<size_t>(min(row[q], N - 1)) * J0.Fp;
  float p0[kQ];  // job 0's new predictions, kept for a single-output loss's gradients
  for (uint32_t k = 0; k < njobs; ++k) {
    const RJob& J = jobs[k];
    float* pk = preds + static_cast<size_t>(k) * N;
    float pr[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) pr[q] = mode == 1 && row[q] < N ? pk[row[q]] : base[k];
    if (mode == 1) {
      const int t = J.st->tree;
      if (t >= 0 && t < static_cast<int>(J.n_trees) && !J.st->stopped && J.m_nodes[t] > 0) {
        float v[kQ];
        if (staged) {
#pragma unroll
          for (int q = 0; q < kQ; ++q) v[q] = rwalk_lds(s_w + s_base[k], br[q]);
        } else {
#pragma unroll
          for (int q = 0; q < kQ; ++q) v[q] = rtree_row_value(J, static_cast<size_t>(t) * J.NC, br[q]);
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q) pr[q] = __fmaf_rn(lr, v[q], pr[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      if (row[q] < N) pk[row[q]] = pr[q];
      if (k == 0) p0[q] = pr[q];
    }
  }
  // phase 2: the gradients of each row from its (just written) predictions
  float gmax[kMaxClasses], hmax[kMaxClasses];
#pragma unroll
  for (uint32_t k = 0; k < kMaxClasses; ++k) gmax[k] = hmax[k] = 0.f;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const uint32_t r = row[q];
    if (r >= N) continue;
    if (loss != kSoftmax) {
      const float2 v = row_gradient(loss, p0[q], y[r], r < vecN);
      gh[r] = v;
      gmax[0] = fmaxf(gmax[0], fabsf(v.x));
      hmax[0] = fmaxf(hmax[0], fabsf(v.y));
      continue;
    }
    float e[kMaxClasses];
    float m = -INFINITY;
#pragma unroll
    for (uint32_t k = 0; k < kMaxClasses; ++k) {
      e[k] = k < njobs ? preds[static_cast<size_t>(k) * N + r] : -INFINITY;
      m = fmaxf(m, e[k]);
    }
    float sum = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < kMaxClasses; ++k) {
      if (k >= njobs) break;
      e[k] = ref_expf(__fsub_rn(e[k], m));
      sum = __fadd_rn(sum, e[k]);
    }
    const uint32_t lab = static_cast<uint32_t>(y[r]);
#pragma unroll
    for (uint32_t k = 0; k < kMaxClasses; ++k) {
      if (k >= njobs) break;
      const float p = __fdiv_rn(e[k], sum);
      const float g = __fsub_rn(p, k == lab ? 1.0f : 0.0f);
      const float hh = __fmul_rn(p, __fsub_rn(1.0f, p));
      const float h = hh > 1e-6f ? hh : 1e-6f;
      gh[static_cast<size_t>(k) * N + r] = make_float2(g, h);
      gmax[k] = fmaxf(gmax[k], fabsf(g));
      hmax[k] = fmaxf(hmax[k], h);
    }
  }
  // each job's tile maxima
  __shared__ float s_m[2][kMaxClasses][4];
#pragma unroll
  for (uint32_t k = 0; k < kMaxClasses; ++k) {
    const float a = wave_max_f32(gmax[k]), b = wave_max_f32(hmax[k]);
    if ((tid & 63) == 0) {
      s_m[0][k][tid >> 6] = a;
      s_m[1][k][tid >> 6] = b;
    }
  }
  __syncthreads();
  if (tid < njobs) {
    const float a = fmaxf(fmaxf(s_m[0][tid][0], s_m[0][tid][1]), fmaxf(s_m[0][tid][2], s_m[0][tid][3]));
    const float b = fmaxf(fmaxf(s_m[1][tid][0], s_m[1][tid][1]), fmaxf(s_m[1][tid][2], s_m[1][tid][3]));
    jobs[tid].ghmax_partials[2 * lw] = __float_as_uint(a);
    jobs[tid].ghmax_partials[2 * lw + 1] = __float_as_uint(b);
  }
}

// Validation rows after each iteration (booster.cpp:306-330, :1400-1409):
// every job's new tree applied to the validation rows' bins (Tree::predict
// over Dataset::apply_bins with the training edges), v_preds += lr * value
// (fused), then each row's loss term -- MSE / LogLoss as Loss::compute_loss
// (loss.cpp:110-121, :336-346), softmax rows -(log max(p_label, 1e-7))
// (CrossEntropyLoss::compute_loss, loss.cpp:232-244) -- summed per
// 1024-row tile in double, in a fixed order.
__global__ void __launch_bounds__(256) k_rvalid(const RJob* __restrict__ jobs, uint32_t njobs,
                                                int loss, float lr, const uint8_t* __restrict__ vbins,
                                                uint32_t vFp, uint32_t Nv, const float* __restrict__ vy,
                                                float* __restrict__ vpreds, double* __restrict__ part) {
  const RJob& J0 = jobs[0];
  const int t = J0.st->tree;
  if (t < 0 || t >= static_cast<int>(J0.n_trees) || J0.st->stopped) return;
  const uint32_t tid = threadIdx.x, lw = blockIdx.x;
  __shared__ float s_p[kMaxClasses][256];
  __shared__ RWalk s_w[kRWalkLds];
  __shared__ uint32_t s_base[kMaxClasses + 1];
  const bool staged = rwalk_stage(jobs, njobs, s_w, s_base);
  double acc = 0.0;
  for (int q = 0; q < kGradTile / 256; ++q) {
    const uint32_t row = lw * kGradTile + q * 256 + tid;
    if (row >= Nv) continue;
    const uint8_t* br = vbins + static_cast<size_t>(row) * vFp;
    for (uint32_t k = 0; k < njobs; ++k) {
      const RJob& J = jobs[k];
      float* pp = vpreds + static_cast<size_t>(k) * Nv + row;
      float pr = *pp;
      if (J.m_nodes[t] > 0)
        pr = __fmaf_rn(lr, staged ? rwalk_lds(s_w + s_base[k], br)
                                  : rtree_row_value(J, static_cast<size_t>(t) * J.NC, br), pr);
      *pp = pr;
      s_p[k][tid] = pr;
    }
    if (loss != kSoftmax) {
      acc += static_cast<double>(row_loss(loss, s_p[0][tid], vy[row]));
      continue;
    }
    float e[kMaxClasses];
    float m = -INFINITY;
#pragma unroll
    for (uint32_t k = 0; k < kMaxClasses; ++k) {
      e[k] = k < njobs ? s_p[k][tid] : -INFINITY;
      m = fmaxf(m, e[k]);
    }
    float sum = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < kMaxClasses; ++k) {
      if (k >= njobs) break;
      e[k] = ref_expf(__fsub_rn(e[k], m));
      sum = __fadd_rn(sum, e[k]);
    }
    const uint32_t lab = static_cast<uint32_t>(vy[row]);
    float pl = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < kMaxClasses; ++k)
      if (k == lab) pl = __fdiv_rn(e[k], sum);
    acc += static_cast<double>(-logf(fmaxf(pl, 1e-7f)));
  }
  tile_sum_store(acc, part + lw);
}

// After k_rvalid: the iteration's mean validation loss and the reference's
// early-stopping rule (booster.cpp:339-371 / :1411-1430: an improvement is a
// loss below the best by more than 1e-10; stop once iter >= max(50, n / 10)
// and es_rounds iterations in a row did not improve).  On a stop every job is
// stopped and *built records the iterations built (the stopping one
// included).
__global__ void __launch_bounds__(256) k_reval(RJob* __restrict__ jobs, uint32_t njobs,
                                               uint32_t Nv, const double* __restrict__ part,
                                               uint32_t ntiles, float* __restrict__ vloss,
                                               uint32_t es_rounds, uint32_t es_min_iter,
                                               uint32_t* __restrict__ built) {
  TreeState* st = jobs[0].st;
  const int t = st->tree;
  if (t < 0 || t >= static_cast<int>(jobs[0].n_trees) || st->stopped) return;
  const uint32_t tid = threadIdx.x;
  __shared__ double s_w[4];
  double a = 0.0;
  for (uint32_t w = tid; w < ntiles; w += 256) a += part[w];
  a = warp_sum_d(a);
  if ((tid & 63) == 0) s_w[tid >> 6] = a;
  __syncthreads();
  if (tid != 0) return;
  const float vl = static_cast<float>((((s_w[0] + s_w[1]) + s_w[2]) + s_w[3]) / Nv);
  vloss[t] = vl;
  if (vl < __fsub_rn(st->es_best, 1e-10f)) {
    st->es_best = vl;
    st->es_best_iter = static_cast<uint32_t>(t);
    st->es_count = 0;
  } else {
    st->es_count += 1;
  }
  *built = static_cast<uint32_t>(t) + 1u;
  if (static_cast<uint32_t>(t) >= es_min_iter && st->es_count >= es_rounds)
    for (uint32_t k = 0; k < njobs; ++k) jobs[k].st->stopped = 1;
}

// Teacher forcing: job 0's (g, h) from given arrays, their maxima merged into
// the tree state for the scales.
__global__ void __launch_bounds__(256) k_rset_gh(RJob* __restrict__ jobs, const float* g,
                                                 const float* h) {
  const RJob& J = jobs[0];
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  float a = 0.f, b = 0.f;
  if (p < J.N) {
    J.gh[p] = make_float2(g[p], h[p]);
    a = fabsf(g[p]);
    b = fabsf(h[p]);
  }
  block_max2_to(J.st, a, b);
}

}  // namespace tc

// ===========================================================================
// Host side.
namespace tc {
namespace {

struct RSpec {
  const tc_dataset* ds = nullptr;
  const float* y = nullptr;        // host labels (null: teacher forcing)
  int32_t loss = TC_LOSS_MSE;
  uint32_t K = 1;                  // jobs: the classes of a softmax fit, else 1
  uint32_t n_iter = 1;
  float lr = 0.1f;
  tc_rtree_params tree{};
  float subsample = 1.f, colsample = 1.f;
  bool use_goss = false, use_mvs = false;  // MVS, else GOSS, else the plain row sample
  float goss_top = 0.2f, goss_other = 0.1f, mvs_ratio = 0.5f;
  uint64_t seed = 42;
  float base = 0.f;
  const tc_dataset* valid = nullptr;  // validation rows, binned with the training edges
  const float* y_valid = nullptr;
  uint32_t es_rounds = 0;
};

void check_rtree_params(const tc_rtree_params* p) {
  TC_REQUIRE(p != nullptr, "tree params are NULL");
  TC_REQUIRE(p->grow_policy == 0 || p->grow_policy == 1, "grow_policy must be 0 or 1");
  TC_REQUIRE(p->max_depth >= 1 && p->max_depth + (p->grow_policy == 1 ? 1u : 0u) <= kRMaxDepth,
             p->grow_policy == 1 ? "max_depth must be in [1, 11] for lossguide"
                                 : "max_depth must be in [1, 12]");
  TC_REQUIRE(p->max_leaves >= 1, "max_leaves must be >= 1");
  TC_REQUIRE((p->max_bins & 0xFFu) >= 2 && p->max_bins <= 255, "max_bins must be in [2, 255]");
  TC_REQUIRE(p->lambda_l2 == p->lambda_l2 && p->min_child_weight == p->min_child_weight &&
                 p->min_split_gain == p->min_split_gain && p->gamma == p->gamma,
             "NaN tree parameter");
}

// Device state of one regular-tree fit: one job per tree grown at a time (one,
// or K for the classes of a multiclass iteration), the per-level launches,
// the per-iteration graph.
class RTrainer {
 public:
  RTrainer(tc_ctx* ctx, const RSpec& s) : ctx_(ctx), s_(s) {
    check_rtree_params(&s.tree);
    try {
      setup();
    } catch (...) {
      release();
      throw;
    }
  }
  ~RTrainer() { release(); }
  RTrainer(const RTrainer&) = delete;
  RTrainer& operator=(const RTrainer&) = delete;

  // predictions = base, first gradients
  void init_gradients() {
    hipLaunchKernelGGL(k_rupdate, dim3(grad_wg_), dim3(256), 0, ctx_->stream, d_jobs_, nj_, 0,
                       loss_, s_.lr, d_base_, y_, preds_, gh_);
    TC_HIP(hipGetLastError());
  }

  // Teacher forcing: job 0's (g, h) and, optionally, its root rows.
  void set_gradients(const float* g, const float* h, const uint32_t* rows, uint32_t n_rows) {
    const uint32_t N = N_;
    float* dg = dmalloc<float>(N);
    float* dh = nullptr;
    try {
      dh = dmalloc<float>(N);
      TC_HIP(hipMemcpyAsync(dg, g, 4ull * N, hipMemcpyHostToDevice, ctx_->stream));
      TC_HIP(hipMemcpyAsync(dh, h, 4ull * N, hipMemcpyHostToDevice, ctx_->stream));
      hipLaunchKernelGGL(k_rset_gh, dim3(std::max(1u, ceil_div(N, 256))), dim3(256), 0, ctx_->stream,
                         d_jobs_, dg, dh);
      TC_HIP(hipGetLastError());
      if (rows) {
        TC_HIP(hipMemcpyAsync(sample_, rows, 4ull * n_rows, hipMemcpyHostToDevice, ctx_->stream));
        set_root(n_rows, true);
      }
      TC_HIP(hipStreamSynchronize(ctx_->stream));
    } catch (...) {
      dfree(dg);
      dfree(dh);
      throw;
    }
    dfree(dg);
    dfree(dh);
  }

  // One iteration: a tree per job, then (update) predictions and gradients.
  // hev (optional) receives an event pair around every histogram launch.
  void enqueue_iteration(bool update, std::vector<std::pair<hipEvent_t, hipEvent_t>>* hev = nullptr) {
    hipStream_t s = ctx_->stream;
    if (sampling_ && !selective_)  // the root rows of the iteration's plain sample
      hipLaunchKernelGGL(k_rows_from_mask_one, dim3(AheadSampler::mask_chunks(N_)),
                         dim3(kMaskChunkWords), 0, s, mask_, N_, n_sample_, sample_);
    hipLaunchKernelGGL(k_rbegin, dim3(1, nj_), dim3(256), 0, s, d_jobs_, lists_);
    for (uint32_t d = 0; d < D_; ++d) {
      const uint32_t par = d & 1u;
      auto* kh = mode_ == 1 ? (nj_ == 1 ? k_hist<kHistUnroll, 1, true> : k_hist<kHistUnroll, 1, false>)
                            : (nj_ == 1 ? k_hist<kHistUnroll, 2, true> : k_hist<kHistUnroll, 2, false>);
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (hev) {
        TC_HIP(hipEventCreate(&e0));
        TC_HIP(hipEventCreate(&e1));
        TC_HIP(hipEventRecord(e0, s));
      }
      hipLaunchKernelGGL(kh, dim3(round8(hist_cap_[d])), dim3(kHistThreads), 0, s,
                         d_hjobs_ + static_cast<size_t>(d) * nj_, lists_.hist[par], lists_.hist_count[par],
                         0u, partials_, Lists{});
      if (hev) {
        TC_HIP(hipEventRecord(e1, s));
        hev->emplace_back(e0, e1);
      }
      hipLaunchKernelGGL(k_rsplit, dim3(F_, d == 0 ? 1u : (1u << (d - 1)), nj_), dim3(256), 0, s,
                         d_jobs_, d, partials_);
      hipLaunchKernelGGL(k_rdecide, dim3(std::max(1u, (1u << d) / 4u), nj_), dim3(256), 0, s, d_jobs_,
                         d, lists_);
      hipLaunchKernelGGL(k_rpartition, dim3(part_cap_[d]), dim3(kPartThreads), 0, s, d_jobs_,
                         lists_.part[par], lists_.part_count[par], d);
      hipLaunchKernelGGL(k_rplan, dim3(1, nj_), dim3(kRPlanThreads), 0, s, d_jobs_, d, lists_);
    }
    if (s_.tree.grow_policy == 1)
      hipLaunchKernelGGL(k_rleafwise, dim3(1, nj_), dim3(256), 0, s, d_jobs_, lists_);
    else
      hipLaunchKernelGGL(k_rfinish, dim3(1, nj_), dim3(256), 0, s, d_jobs_, lists_);
    if (update)
      hipLaunchKernelGGL(k_rupdate, dim3(grad_wg_), dim3(256), 0, s, d_jobs_, nj_, 1, loss_, s_.lr,
                         d_base_, y_, preds_, gh_);
    if (update && Nv_) {
      hipLaunchKernelGGL(k_rvalid, dim3(v_tiles_), dim3(256), 0, s, d_jobs_, nj_, loss_, s_.lr,
                         s_.valid->bins_rm, s_.valid->Fp, Nv_, vy_, vpreds_, vpart_);
      hipLaunchKernelGGL(k_reval, dim3(1), dim3(256), 0, s, d_jobs_, nj_, Nv_, vpart_, v_tiles_,
                         vloss_, s_.es_rounds, std::max(50u, s_.n_iter / 10), built_);
    }
    TC_HIP(hipGetLastError());
  }

  // Early-stopping outcome (valid only after the fit has run).
  void history(tc_train_history* h) {
    TreeState st{};
    TC_HIP(hipMemcpy(&st, h_jobs_[0].st, sizeof(st), hipMemcpyDeviceToHost));
    uint32_t built = s_.n_iter;
    if (Nv_) TC_HIP(hipMemcpy(&built, built_, 4, hipMemcpyDeviceToHost));
    h->n_trees_built = built;
    h->stopped_early = st.stopped ? 1 : 0;
    h->best_iteration = st.es_best_iter;
    h->best_valid_loss = Nv_ ? st.es_best : 0.0f;
    if (h->train_loss) std::fill(h->train_loss, h->train_loss + s_.n_iter, 0.0f);
    if (h->valid_loss) {
      std::fill(h->valid_loss, h->valid_loss + s_.n_iter, 0.0f);
      if (Nv_) TC_HIP(hipMemcpy(h->valid_loss, vloss_, 4ull * built, hipMemcpyDeviceToHost));
    }
  }

  // Booster::train / train_multiclass: per iteration the row sample (drawn on
  // the host with the reference's engine), then the iteration's trees.
  void run(bool graph, bool time_hist = false) {
    hipStream_t s = ctx_->stream;
    hipGraphExec_t ge = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> hev;
    if (time_hist) graph = false;
    if (graph) {
      hipGraph_t g;
      TC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      enqueue_iteration(true);
      TC_HIP(hipStreamEndCapture(s, &g));
      const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      TC_HIP(e);
    }
    hipEvent_t a, b;
    TC_HIP(hipEventCreate(&a));
    TC_HIP(hipEventCreate(&b));
    TC_HIP(hipEventRecord(a, s));
    try {
      for (uint32_t it = 0; it < s_.n_iter; ++it) {
        if (stop_seen(it)) break;
        prepare_sample();
        if (ge) TC_HIP(hipGraphLaunch(ge, s));
        else enqueue_iteration(true, time_hist ? &hev : nullptr);
      }
      TC_HIP(hipEventRecord(b, s));
      TC_HIP(hipStreamSynchronize(s));
    } catch (...) {
      if (ge) (void)hipGraphExecDestroy(ge);
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      for (auto& e : hev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
      throw;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    ctx_->last_loop_ms = ms;
    double hs = 0.0;
    for (auto& e : hev) {
      float x = 0.f;
      (void)hipEventElapsedTime(&x, e.first, e.second);
      hs += x;
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    ctx_->last_hist_ms = hs;
    ctx_->last_hist_launches = static_cast<uint32_t>(hev.size());
    ctx_->last_hist_rows = hist_rows();
    if (ge) (void)hipGraphExecDestroy(ge);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }

  // Trees of every job, iteration-major, class-minor, into out (node_cap =
  // 2^(D+1) - 1 slots per tree); train_preds K x N.
  void download(tc_rmodel* out, uint32_t n_iter) {
    hipStream_t s = ctx_->stream;
    TC_HIP(hipStreamSynchronize(s));
    uint32_t fault = 0;
    for (uint32_t j = 0; j < nj_; ++j) {
      TreeState st;
      TC_HIP(hipMemcpy(&st, h_jobs_[j].st, sizeof(st), hipMemcpyDeviceToHost));
      fault |= st.fault;
    }
    if (fault) throw Error(TC_EDEVICE, "regular-tree training: a kernel found inconsistent state (fault " +
                                           std::to_string(fault) + ")");
    const size_t NC = NC_;
    std::vector<uint16_t> f16(NC * n_iter), dep(n_iter);
    std::vector<uint8_t> u8a(NC * n_iter), u8b(NC * n_iter);
    std::vector<uint32_t> l32(NC * n_iter), r32(NC * n_iter), nn(n_iter), nl(n_iter);
    std::vector<float> v32(NC * n_iter), g32(NC * n_iter);
    for (uint32_t j = 0; j < nj_; ++j) {
      const RJob& J = h_jobs_[j];
      auto get = [&](void* dst, const void* src, size_t bytes) {
        TC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
      };
      get(f16.data(), J.m_feat, 2 * NC * n_iter);
      get(u8a.data(), J.m_bin, NC * n_iter);
      get(u8b.data(), J.m_flags, NC * n_iter);
      get(l32.data(), J.m_left, 4 * NC * n_iter);
      get(r32.data(), J.m_right, 4 * NC * n_iter);
      get(v32.data(), J.m_value, 4 * NC * n_iter);
      get(g32.data(), J.m_gain, 4 * NC * n_iter);
      get(nn.data(), J.m_nodes, 4ull * n_iter);
      get(nl.data(), J.m_leaves, 4ull * n_iter);
      get(dep.data(), J.m_depth, 2ull * n_iter);
      for (uint32_t t = 0; t < n_iter; ++t) {
        const size_t src = NC * t, dst = NC * (static_cast<size_t>(t) * nj_ + j);
        if (out->feature) std::memcpy(out->feature + dst, f16.data() + src, 2 * NC);
        if (out->bin) std::memcpy(out->bin + dst, u8a.data() + src, NC);
        if (out->flags) std::memcpy(out->flags + dst, u8b.data() + src, NC);
        if (out->left) std::memcpy(out->left + dst, l32.data() + src, 4 * NC);
        if (out->right) std::memcpy(out->right + dst, r32.data() + src, 4 * NC);
        if (out->value) std::memcpy(out->value + dst, v32.data() + src, 4 * NC);
        if (out->gain) std::memcpy(out->gain + dst, g32.data() + src, 4 * NC);
        const size_t ti = static_cast<size_t>(t) * nj_ + j;
        if (out->n_nodes) out->n_nodes[ti] = nn[t];
        if (out->n_leaves) out->n_leaves[ti] = nl[t];
        if (out->depth) out->depth[ti] = dep[t];
      }
    }
    if (out->train_preds)
      TC_HIP(hipMemcpy(out->train_preds, preds_, 4ull * N_ * nj_, hipMemcpyDeviceToHost));
    out->base = s_.base;
  }

  uint64_t hist_rows() const {
    unsigned long long v = 0;
    (void)hipMemcpy(&v, hist_rows_, 8, hipMemcpyDeviceToHost);
    return v;
  }

 private:
  template <typename T>
  T* alloc(size_t n) {
    T* p = dmalloc<T>(n);
    allocs_.push_back(p);
    return p;
  }
  template <typename T>
  T* alloc_zero(size_t n) {
    T* p = alloc<T>(n);
    TC_HIP(hipMemsetAsync(p, 0, sizeof(T) * std::max<size_t>(1, n), ctx_->stream));
    return p;
  }

  // With a validation set: every 16 iterations the stop flag is copied to
  // pinned memory; once a copy shows the fit stopped, nothing more is
  // enqueued (iterations enqueued meanwhile do nothing on the device).
  bool stop_seen(uint32_t it) {
    if (!Nv_ || it % 16 != 0) return false;
    if (stop_pending_ && hipEventQuery(stop_ev_) == hipSuccess) {
      stop_pending_ = false;
      if (*stop_host_) return true;
    }
    if (!stop_pending_) {
      TC_HIP(hipMemcpyAsync(stop_host_, reinterpret_cast<const char*>(h_jobs_[0].st) +
                                            offsetof(TreeState, stopped),
                            4, hipMemcpyDeviceToHost, ctx_->stream));
      TC_HIP(hipEventRecord(stop_ev_, ctx_->stream));
      stop_pending_ = true;
    }
    return false;
  }

  // The root's rows: all N (identity) or n rows of the sample buffer.
  void set_root(uint32_t n, bool sampled) {
    for (uint32_t j = 0; j < nj_; ++j) {
      h_jobs_[j].n_root = n;
      h_jobs_[j].root_rows = sampled ? sample_ : nullptr;
      h_jobs_[j].hist_chunk = root_chunk(std::max(1u, n), T_ * nj_, 512);
      JobDev& H = h_hjobs_[j];  // level 0's k_hist record
      H.rows = sampled ? sample_ : rows0_dummy_;
      H.root_identity = sampled ? 0u : 1u;
    }
    recompute_caps();
    TC_HIP(hipMemcpyAsync(d_jobs_, h_jobs_.data(), sizeof(RJob) * nj_, hipMemcpyHostToDevice,
                          ctx_->stream));
    TC_HIP(hipMemcpyAsync(d_hjobs_, h_hjobs_.data(), sizeof(JobDev) * nj_, hipMemcpyHostToDevice,
                          ctx_->stream));
  }

  void recompute_caps() {
    uint32_t root = 0;
    for (const RJob& J : h_jobs_) root += ceil_div(std::max(1u, J.n_root), J.hist_chunk) * T_;
    for (uint32_t d = 0; d < D_; ++d) {
      // below the root: chunks of >= 1024 rows over at most N/2 histogram
      // rows, plus one partial chunk per pair
      hist_cap_[d] = d == 0 ? root : nj_ * (ceil_div(N_ / 2 + 1, 1024) + (1u << (d - 1))) * T_;
      part_cap_[d] = nj_ * (ceil_div(N_, kPartChunk) + (1u << d));
    }
  }

  // The next iteration's row sample.  A plain one (Dataset::random_subsample)
  // is drawn ahead on host threads and copied as a bitmask that the
  // iteration's first launch turns into the root rows.  GOSS and MVS select
  // on the gradients the iteration's trees grow on (every queued iteration
  // first): Dataset::goss_subsample on |g| -- a softmax fit's largest |g|
  // over its classes (booster.cpp:1326-1347) -- or Dataset::mvs_subsample;
  // the rows and their count are copied to the device.
  void prepare_sample() {
    if (!sampling_ && !colsample_) return;
    if (sampling_ && !selective_) {
      plain_.next(mask_, ctx_->stream);
      return;
    }
    uint64_t seed = 0;
    const MvsDraws* md = nullptr;
    if (sampling_ && s_.use_mvs) {  // drawn ahead: the seed and the feature-subset draw are taken there
      md = &mvs_ahead_.next();
    } else {
      if (sampling_) seed = next_random(rng_);
      if (colsample_) (void)next_random(rng_);  // feature subset: drawn, unused (tree.cpp:193-196)
    }
    if (!sampling_) return;
    hipStream_t s = ctx_->stream;
    const int b = next_;
    next_ ^= 1;
    TC_HIP(hipEventSynchronize(sev_[b]));
    TC_HIP(hipMemcpyAsync(gh_host_, gh_, sizeof(float2) * N_ * nj_, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < N_; ++i) {
      float m = 0.0f;
      if (nj_ == 1) {
        m = gh_host_[i].x;
      } else {
        for (uint32_t c = 0; c < nj_; ++c)
          m = std::max(m, std::abs(gh_host_[static_cast<size_t>(c) * N_ + i].x));
      }
      gsel_[i] = m;
    }
    const uint32_t n = s_.use_mvs
                           ? mvs_select(gsel_.data(), N_, s_.mvs_ratio, *md, shost_[b], mvs_scratch_)
                           : goss_select(gsel_.data(), N_, s_.goss_top, s_.goss_other, seed, shost_[b],
                                         goss_pairs_);
    nroot_host_[b] = n;
    TC_HIP(hipMemcpyAsync(sample_, shost_[b], 4ull * std::max(1u, n), hipMemcpyHostToDevice, s));
    for (uint32_t j = 0; j < nj_; ++j)
      TC_HIP(hipMemcpyAsync(reinterpret_cast<char*>(d_jobs_ + j) + offsetof(RJob, n_root),
                            &nroot_host_[b], sizeof(uint32_t), hipMemcpyHostToDevice, s));
    TC_HIP(hipEventRecord(sev_[b], s));
  }

  void setup() {
    hipStream_t s = ctx_->stream;
    const tc_dataset* ds = s_.ds;
    N_ = ds->N;
    F_ = ds->F;
    T_ = feature_tiles(F_);
    // lossguide: one more level is grown, so that every candidate the
    // leaf-wise loop may queue (a node at max_depth included) has its split
    D_ = s_.tree.max_depth + (s_.tree.grow_policy == 1 ? 1u : 0u);
    nj_ = s_.K;
    loss_ = s_.loss;
    TC_REQUIRE(nj_ >= 1 && nj_ <= static_cast<uint32_t>(kMaxClasses), "classes must be in [1, 16]");
    TC_REQUIRE(loss_ == TC_LOSS_SOFTMAX ? nj_ >= 2 : nj_ == 1, "classes do not match the loss");
    TC_REQUIRE(N_ >= 1, "no rows");
    mode_ = loss_ == TC_LOSS_MSE ? 1 : 2;
    NC_ = (2u << s_.tree.max_depth) - 1u;  // node slots of a tree of the model
    NG_ = (2u << D_) - 1u;                 // grown nodes
    const uint32_t PC = 1u << (D_ - 1);
    grad_wg_ = std::max(1u, ceil_div(N_, kGradTile));
    rng_ = s_.seed;
    selective_ = s_.use_goss || s_.use_mvs;
    sampling_ = selective_ || (s_.subsample > 0.f && s_.subsample < 1.f);
    colsample_ = s_.colsample > 0.f && s_.colsample < 1.f;
    // the root's grids hold n_sample_ rows: GOSS's fixed count, every row for
    // MVS (whose sample size the device reads from the job record)
    n_sample_ = s_.use_mvs    ? N_
                : s_.use_goss ? std::min(N_, goss_rows(N_, s_.goss_top, s_.goss_other))
                : sampling_   ? std::min(N_, static_cast<uint32_t>(static_cast<float>(N_) * s_.subsample))
                              : N_;
    y_ = alloc_zero<float>(N_);
    if (s_.y) TC_HIP(hipMemcpyAsync(y_, s_.y, 4ull * N_, hipMemcpyHostToDevice, s));
    preds_ = alloc_zero<float>(static_cast<size_t>(N_) * nj_);
    gh_ = alloc_zero<float2>(static_cast<size_t>(N_) * nj_);
    std::vector<float> base(nj_, s_.base);
    d_base_ = alloc<float>(nj_);
    TC_HIP(hipMemcpyAsync(d_base_, base.data(), 4ull * nj_, hipMemcpyHostToDevice, s));
    sample_ = alloc<uint32_t>(N_);
    rows0_dummy_ = alloc<uint32_t>(1);
    hist_rows_ = alloc_zero<unsigned long long>(1);
    if (sampling_ && !selective_) {
      plain_.init(N_, s_.subsample, rng_, colsample_, ahead_ring(N_, 1));
      mask_ = alloc<uint32_t>(AheadSampler::mask_words(N_));
    } else if (selective_) {
      TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&gh_host_), sizeof(float2) * N_ * nj_,
                           hipHostMallocDefault));
      for (int b = 0; b < 2; ++b) {
        TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&shost_[b]), 4ull * N_, hipHostMallocDefault));
        TC_HIP(hipEventCreateWithFlags(&sev_[b], hipEventDisableTiming));
      }
      gsel_.resize(N_);
      if (s_.use_mvs) mvs_ahead_.init(N_, rng_, colsample_, ahead_ring(N_, 1));
    }
    if (s_.valid) {  // validation rows ride after each iteration's update
      TC_REQUIRE(s_.valid->F == F_, "validation set: feature count differs from training");
      TC_REQUIRE(s_.y_valid != nullptr, "validation set without labels");
      Nv_ = s_.valid->N;
    }
    if (Nv_) {
      v_tiles_ = ceil_div(Nv_, kGradTile);
      vy_ = alloc<float>(Nv_);
      TC_HIP(hipMemcpyAsync(vy_, s_.y_valid, 4ull * Nv_, hipMemcpyHostToDevice, s));
      std::vector<float> vp(static_cast<size_t>(Nv_) * nj_, s_.base);
      vpreds_ = alloc<float>(vp.size());
      TC_HIP(hipMemcpyAsync(vpreds_, vp.data(), 4ull * vp.size(), hipMemcpyHostToDevice, s));
      TC_HIP(hipStreamSynchronize(s));  // vp leaves scope
      vpart_ = alloc<double>(v_tiles_);
      vloss_ = alloc_zero<float>(std::max(1u, s_.n_iter));
      built_ = alloc_zero<uint32_t>(1);
      TC_HIP(hipHostMalloc(reinterpret_cast<void**>(&stop_host_), 4, hipHostMallocDefault));
      *stop_host_ = 0;
      TC_HIP(hipEventCreateWithFlags(&stop_ev_, hipEventDisableTiming));
    }
    h_jobs_.resize(nj_);
    h_hjobs_.resize(static_cast<size_t>(D_) * nj_);
    for (uint32_t j = 0; j < nj_; ++j) {
      RJob& J = h_jobs_[j];
      std::memset(&J, 0, sizeof(J));
      J.bins_rm = ds->bins_rm;
      J.bins_cm = ds->bins_cm;
      J.N = N_; J.F = F_; J.Fp = ds->Fp; J.T = T_;
      J.D = D_;
      J.leafwise = s_.tree.grow_policy == 1 ? 1u : 0u;
      J.max_depth = s_.tree.max_depth;
      J.max_leaves = s_.tree.max_leaves;
      J.max_bins = s_.tree.max_bins;
      J.min_samples_leaf = s_.tree.min_samples_leaf;
      J.lambda_l2 = s_.tree.lambda_l2;
      J.min_child_weight = s_.tree.min_child_weight;
      J.min_split_gain = s_.tree.min_split_gain;
      J.gamma = s_.tree.gamma;
      J.max_delta_step = s_.tree.max_delta_step;
      J.leaf_smooth = s_.tree.leaf_smooth;
      J.learn_missing = s_.tree.learn_missing_direction;
      J.mode = mode_;
      J.deep_items = std::max(T_, 256u / nj_);
      J.rows = alloc<uint32_t>(2ull * N_);
      J.gh = gh_ + static_cast<size_t>(j) * N_;
      J.st = alloc<TreeState>(1);
      hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, J.st);
      J.nodes = alloc<RNode>(NG_);
      J.lvl_begin = alloc_zero<uint32_t>(D_ + 2);
      J.lvl_count = alloc_zero<uint32_t>(D_ + 2);
      J.keys = alloc<unsigned long long>(static_cast<size_t>(PC) * F_);
      J.nhist = alloc<HistCell>(2ull * PC * F_ * 256);
      J.ncnt = alloc<uint32_t>(2ull * PC * F_ * 256);
      J.lastc = alloc<HistCell>(static_cast<size_t>(PC) * F_ * 256);
      J.pair_first = alloc_zero<uint32_t>(2ull * PC);
      J.pair_nch = alloc_zero<uint32_t>(2ull * PC);
      J.cursor = alloc_zero<unsigned long long>(PC);
      J.ghmax_partials = alloc_zero<uint32_t>(2ull * grad_wg_);
      J.grad_wg_cnt = grad_wg_;
      J.n_trees = s_.n_iter;
      J.NC = NC_;
      const size_t M = static_cast<size_t>(s_.n_iter) * NC_;
      J.m_feat = alloc_zero<uint16_t>(M);
      J.m_bin = alloc_zero<uint8_t>(M);
      J.m_flags = alloc_zero<uint8_t>(M);
      J.m_left = alloc_zero<uint32_t>(M);
      J.m_right = alloc_zero<uint32_t>(M);
      J.m_value = alloc_zero<float>(M);
      J.m_gain = alloc_zero<float>(M);
      J.m_nodes = alloc_zero<uint32_t>(s_.n_iter);
      J.m_leaves = alloc_zero<uint32_t>(s_.n_iter);
      J.m_depth = alloc_zero<uint16_t>(s_.n_iter);
      J.hist_rows_total = hist_rows_;  // every class's trees
      for (uint32_t d = 0; d < D_; ++d) {  // k_hist's record of level d
        JobDev& H = h_hjobs_[static_cast<size_t>(d) * nj_ + j];
        std::memset(&H, 0, sizeof(H));
        H.bins_rm = ds->bins_rm; H.bins_cm = ds->bins_cm; H.edges = ds->edges; H.n_edges = ds->n_edges;
        H.N = N_; H.F = F_; H.Fp = ds->Fp; H.T = T_;
        H.K = 1;
        H.D = D_;
        H.gh = J.gh;
        H.ghl = nullptr;
        H.rows = d == 0 ? rows0_dummy_ : J.rows + static_cast<size_t>(d & 1u) * N_;
        H.root_identity = d == 0 ? 1u : 0u;
        H.st = J.st;
        H.hist = J.lastc;
        H.store_hist = 1;
      }
    }
    d_jobs_ = alloc<RJob>(nj_);
    d_hjobs_ = alloc<JobDev>(static_cast<size_t>(D_) * nj_);
    TC_HIP(hipMemcpyAsync(d_hjobs_, h_hjobs_.data(), sizeof(JobDev) * h_hjobs_.size(),
                          hipMemcpyHostToDevice, s));
    set_root(n_sample_, sampling_);  // also uploads the job records
    // work lists and partial slots sized for the largest level
    uint32_t hmax = 0, pmax = 0;
    for (uint32_t d = 0; d < D_; ++d) {
      // the root's items at any row count (teacher forcing may choose fewer rows)
      hmax = std::max(hmax, hist_cap_[d]);
      pmax = std::max(pmax, part_cap_[d]);
    }
    hmax = std::max(hmax, nj_ * ceil_div(N_, root_chunk(1, T_ * nj_, 512)) * T_);
    for (int p = 0; p < 2; ++p) {
      lists_.hist[p] = alloc<HistWork>(hmax);
      lists_.part[p] = alloc<PartWork>(pmax);
      lists_.hist_count[p] = alloc_zero<uint32_t>(1);
      lists_.part_count[p] = alloc_zero<uint32_t>(1);
    }
    partials_ = alloc<HistCell>(static_cast<size_t>(hmax) * kTileF * 256);
    TC_HIP(hipStreamSynchronize(s));
  }

  void release() {
    if (ctx_ && ctx_->stream) (void)hipStreamSynchronize(ctx_->stream);
    for (void* p : allocs_) dfree(p);
    allocs_.clear();
    if (stop_host_) hfree(stop_host_);
    stop_host_ = nullptr;
    if (stop_ev_) (void)hipEventDestroy(stop_ev_);
    stop_ev_ = nullptr;
    plain_.release();  // waits for its draws
    mvs_ahead_.release();
    for (int b = 0; b < 2; ++b) {
      if (shost_[b]) hfree(shost_[b]);
      if (sev_[b]) (void)hipEventDestroy(sev_[b]);
      shost_[b] = nullptr;
      sev_[b] = nullptr;
    }
    if (gh_host_) hfree(gh_host_);
    gh_host_ = nullptr;
  }

  tc_ctx* ctx_;
  RSpec s_;
  uint32_t N_ = 0, F_ = 0, T_ = 0, D_ = 0, nj_ = 1, NC_ = 0, NG_ = 0, grad_wg_ = 1;
  int32_t loss_ = 0;
  int mode_ = 2;
  std::vector<RJob> h_jobs_;
  std::vector<JobDev> h_hjobs_;   // [level][job] k_hist records
  RJob* d_jobs_ = nullptr;
  JobDev* d_hjobs_ = nullptr;
  RLists lists_{};
  HistCell* partials_ = nullptr;
  uint32_t hist_cap_[kRMaxDepth] = {};
  uint32_t part_cap_[kRMaxDepth] = {};
  float* y_ = nullptr;
  float* preds_ = nullptr;
  float2* gh_ = nullptr;
  float* d_base_ = nullptr;
  uint32_t* sample_ = nullptr;
  uint32_t* rows0_dummy_ = nullptr;
  unsigned long long* hist_rows_ = nullptr;
  std::vector<void*> allocs_;
  // validation
  uint32_t Nv_ = 0, v_tiles_ = 0;
  float* vy_ = nullptr;
  float* vpreds_ = nullptr;
  double* vpart_ = nullptr;
  float* vloss_ = nullptr;
  uint32_t* built_ = nullptr;
  uint32_t* stop_host_ = nullptr;
  hipEvent_t stop_ev_ = nullptr;
  bool stop_pending_ = false;
  // row sampling
  bool sampling_ = false, colsample_ = false;
  uint32_t n_sample_ = 0;
  uint64_t rng_ = 42;
  AheadSampler plain_;        // the row samples, drawn ahead
  uint32_t* mask_ = nullptr;  // the iteration's sample as a bitmask (device)
  // GOSS / MVS: the selection on each iteration's gradients (host)
  bool selective_ = false;
  float2* gh_host_ = nullptr;           // pinned copy of every class's (g, h)
  std::vector<float> gsel_;             // the gradient the selection reads per row
  MvsAhead mvs_ahead_;                  // MVS: the uniform draws, made ahead
  MvsScratch mvs_scratch_;
  std::vector<std::pair<float, uint32_t>> goss_pairs_;  // GOSS: kept across iterations
  uint32_t* shost_[2] = {nullptr, nullptr};  // pinned row samples, alternating
  hipEvent_t sev_[2] = {nullptr, nullptr};
  uint32_t nroot_host_[2] = {0, 0};    // each buffer's row count, the source of its copy
  int next_ = 0;
};

}  // namespace

// ---------------------------------------------------------------------------
// Regular-tree ensemble prediction.  Each split node becomes a compare on the
// raw float that gives the same direction as the rule's bin: with edges e
// sorted and unique, bin > b <=> v >= e[b] when bins count the edges <= v,
// v > e[b] when they count the edges < v, never when b >= n_edges; NaN (bin
// 255) goes left.
struct RPNode {
  uint32_t a;       // feature | kind << 16; kind 0: right if v >= t, 1: right if v > t,
                    // 2: always left, 3: leaf
  float v;          // threshold, or the leaf value
  uint32_t left, right;  // ensemble-global node ids
};

// One row per lane, 64 rows per workgroup.  STAGED: the tile's rows are
// copied to LDS first (coalesced; an odd row stride keeps the lanes' reads of
// one feature on different banks); otherwise (rows too wide for 64 KB) each
// lane reads its row from HBM.
template <bool STAGED>
__global__ void __launch_bounds__(64) k_rpredict(const RPNode* __restrict__ nodes,
                                                 const uint32_t* __restrict__ root,
                                                 const float* __restrict__ weight,
                                                 const uint32_t* __restrict__ tree_class,
                                                 const float* __restrict__ base, uint32_t T,
                                                 uint32_t K, const float* __restrict__ X, uint32_t n,
                                                 uint32_t F, int transform, float* __restrict__ out) {
  extern __shared__ float s_x[];  // 64 rows x (F + 1) floats
  const uint32_t tid = threadIdx.x, r0 = blockIdx.x * 64u, S = F + 1;
  if (STAGED) {
    for (uint32_t i = tid; i < 64u * F; i += 64) {
      const uint32_t rr = i / F, f = i % F;
      s_x[rr * S + f] = r0 + rr < n ? X[static_cast<size_t>(r0 + rr) * F + f] : 0.f;
    }
    __syncthreads();
  }
  const uint32_t row = r0 + tid;
  if (row >= n) return;
  const float* x = STAGED ? s_x + tid * S : X + static_cast<size_t>(row) * F;
  float acc[kMaxClasses];
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) acc[c] = 0.f;
  for (uint32_t t = 0; t < T; ++t) {
    uint32_t k = root[t];
    if (k == kRNone) continue;  // a tree without nodes adds nothing
    RPNode nd = nodes[k];
    for (uint32_t step = 0; step <= kRMaxDepth + 1 && (nd.a >> 16) != 3u; ++step) {
      const uint32_t kind = nd.a >> 16;
      const float v = x[nd.a & 0xFFFFu];
      const bool right = kind == 0u ? v >= nd.v : (kind == 1u ? v > nd.v : false);  // NaN: false
      nd = nodes[right ? nd.right : nd.left];
    }
    const uint32_t c = tree_class ? tree_class[t] : 0u;
#pragma unroll
    for (uint32_t q = 0; q < kMaxClasses; ++q)
      if (q == c) acc[q] = __fmaf_rn(weight[t], nd.v, acc[q]);
  }
  float o[kMaxClasses];
  float m = -INFINITY;
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    o[c] = __fadd_rn(acc[c], base[c]);
    m = fmaxf(m, o[c]);
  }
  if (transform && K == 1) {
    o[0] = __fdiv_rn(1.0f, __fadd_rn(1.0f, ref_expf(-o[0])));  // LogLoss::transform_prediction
  } else if (transform) {  // CrossEntropyLoss::transform_to_proba: max shift, exp, sum, divide
    float sum = 0.f;
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      o[c] = ref_expf(__fsub_rn(o[c], m));
      sum = __fadd_rn(sum, o[c]);
    }
#pragma unroll
    for (uint32_t c = 0; c < kMaxClasses; ++c) {
      if (c >= K) break;
      o[c] = __fdiv_rn(o[c], sum);
    }
  }
#pragma unroll
  for (uint32_t c = 0; c < kMaxClasses; ++c) {
    if (c >= K) break;
    out[static_cast<size_t>(row) * K + c] = o[c];
  }
}

}  // namespace tc

struct tc_rensemble {
  tc_ctx* ctx = nullptr;
  uint32_t T = 0, K = 1, F = 0;
  tc::RPNode* nodes = nullptr;
  uint32_t* root = nullptr;
  float* weight = nullptr;
  uint32_t* tree_class = nullptr;
  float* base = nullptr;
};

namespace tc {
namespace {

void rensemble_free(tc_rensemble* e) {
  if (!e) return;
  dfree(e->nodes);
  dfree(e->root);
  dfree(e->weight);
  dfree(e->tree_class);
  dfree(e->base);
  if (e->ctx) unref(e->ctx);
  delete e;
}

void rpredict_dev(tc_rensemble* e, const float* X, uint32_t n, uint32_t F, int32_t transform, float* out) {
  TC_REQUIRE(F == e->F, "tc_rpredict: feature count differs from the ensemble's");
  if (n == 0) return;
  const size_t lds = 64ull * (F + 1) * sizeof(float);
  if (lds <= 64 * 1024)
    hipLaunchKernelGGL(k_rpredict<true>, dim3(ceil_div(n, 64)), dim3(64), lds, e->ctx->stream,
                       e->nodes, e->root, e->weight, e->tree_class, e->base, e->T, e->K, X, n, F,
                       transform, out);
  else
    hipLaunchKernelGGL(k_rpredict<false>, dim3(ceil_div(n, 64)), dim3(64), 0, e->ctx->stream,
                       e->nodes, e->root, e->weight, e->tree_class, e->base, e->T, e->K, X, n, F,
                       transform, out);
  TC_HIP(hipGetLastError());
}

}  // namespace
}  // namespace tc

extern "C" {

int32_t tc_rtree_build(tc_dataset* ds, const float* g, const float* h, const uint32_t* rows,
                       uint32_t n_rows, const tc_rtree_params* p, tc_rmodel* out) {
  using namespace tc;
  return guard([&] {
    TC_REQUIRE(ds && g && h && out, "tc_rtree_build: NULL argument");
    check_rtree_params(p);
    for (uint32_t i = 0; i < ds->N; ++i)
      TC_REQUIRE(h[i] >= 0.f && std::isfinite(h[i]) && std::isfinite(g[i]),
                 "tc_rtree_build: h must be finite and >= 0, g finite");
    if (rows)
      for (uint32_t i = 0; i < n_rows; ++i) TC_REQUIRE(rows[i] < ds->N, "tc_rtree_build: row out of range");
    RSpec s;
    s.ds = ds;
    s.loss = TC_LOSS_LOGLOSS;  // no update runs; the packed h + count cells
    s.K = 1;
    s.n_iter = 1;
    s.tree = *p;
    RTrainer tr(ds->ctx, s);
    tr.set_gradients(g, h, rows, rows ? n_rows : ds->N);
    tr.enqueue_iteration(false);
    tr.download(out, 1);
  });
}

int32_t tc_rtrain(tc_dataset* ds, const float* y, const tc_rtrain_params* p, tc_rmodel* out) {
  return tc_rtrain_valid(ds, y, p, nullptr, nullptr, 0, out, nullptr);
}

int32_t tc_rtrain_valid(tc_dataset* ds, const float* y, const tc_rtrain_params* p,
                        tc_dataset* valid, const float* y_valid, uint32_t early_stopping_rounds,
                        tc_rmodel* out, tc_train_history* history) {
  using namespace tc;
  return guard([&] {
    TC_REQUIRE(ds && y && p && out, "tc_rtrain: NULL argument");
    TC_REQUIRE(!valid || y_valid, "tc_rtrain_valid: validation set without labels");
    TC_REQUIRE(p->loss == TC_LOSS_MSE || p->loss == TC_LOSS_LOGLOSS || p->loss == TC_LOSS_SOFTMAX,
               "tc_rtrain: unsupported loss");
    TC_REQUIRE(p->learning_rate == p->learning_rate, "learning_rate is NaN");
    RSpec s;
    s.ds = ds;
    s.y = y;
    s.loss = p->loss;
    s.K = p->loss == TC_LOSS_SOFTMAX ? p->n_classes : 1u;
    if (p->loss == TC_LOSS_SOFTMAX) {
      TC_REQUIRE(p->n_classes >= 2 && p->n_classes <= static_cast<uint32_t>(kMaxClasses),
                 "n_classes must be in [2, 16]");
      for (uint32_t i = 0; i < ds->N; ++i)
        TC_REQUIRE(y[i] >= 0.f && y[i] < static_cast<float>(p->n_classes) && y[i] == std::floor(y[i]),
                   "labels must be class indices 0..K-1");
    }
    s.n_iter = p->n_iterations;
    s.lr = p->learning_rate;
    s.tree = p->tree;
    s.subsample = p->subsample;
    s.colsample = p->colsample_bytree;
    // booster.cpp:198-219 (MVS, else GOSS); the multiclass loop takes GOSS or
    // the plain sample only (:1326-1354)
    s.use_mvs = p->use_mvs != 0 && p->loss != TC_LOSS_SOFTMAX;
    s.use_goss = p->use_goss != 0 && !s.use_mvs;
    s.goss_top = p->goss_top_rate;
    s.goss_other = p->goss_other_rate;
    s.mvs_ratio = p->mvs_subsample;
    if (s.use_goss)
      TC_REQUIRE(p->goss_top_rate >= 0.f && p->goss_other_rate >= 0.f &&
                     p->goss_top_rate + p->goss_other_rate <= 1.f,
                 "goss rates must be >= 0 with top + other <= 1");
    if (s.use_mvs) TC_REQUIRE(p->mvs_subsample > 0.f && p->mvs_subsample <= 1.f,
                              "mvs_subsample must be in (0, 1]");
    s.seed = p->seed;
    s.base = init_prediction_host(p->loss, y, ds->N);
    s.valid = valid;
    s.y_valid = y_valid;
    s.es_rounds = early_stopping_rounds;
    if (valid && p->loss == TC_LOSS_SOFTMAX)
      for (uint32_t i = 0; i < valid->N; ++i)
        TC_REQUIRE(y_valid[i] >= 0.f && y_valid[i] < static_cast<float>(p->n_classes) &&
                       y_valid[i] == std::floor(y_valid[i]),
                   "validation labels must be class indices 0..K-1");
    if (p->n_iterations == 0) {
      out->base = s.base;
      if (history) *history = tc_train_history{history->train_loss, history->valid_loss, 0, 0, 0, 0.f};
      return;
    }
    RTrainer tr(ds->ctx, s);
    tr.init_gradients();
    tr.run(p->use_graph != 0, p->time_hist != 0);
    tr.download(out, p->n_iterations);
    if (history) tr.history(history);
  });
}

int32_t tc_rensemble_create(tc_ctx* ctx, uint32_t n_trees, uint32_t n_classes,
                            const uint32_t* tree_off, const uint16_t* feature, const uint8_t* bin,
                            const uint8_t* flags, const uint32_t* left, const uint32_t* right,
                            const float* value, const float* weights, const uint32_t* tree_class,
                            const float* base, const float* edges, const uint32_t* n_edges,
                            uint32_t n_features, int32_t rule, tc_rensemble** out) {
  using namespace tc;
  return guard([&] {
    TC_REQUIRE(ctx && out && tree_off && weights && base && n_edges, "tc_rensemble_create: NULL argument");
    TC_REQUIRE(n_classes >= 1 && n_classes <= static_cast<uint32_t>(kMaxClasses), "n_classes must be in [1, 16]");
    TC_REQUIRE(rule == 0 || rule == 1, "rule must be 0 or 1");
    TC_REQUIRE(n_features >= 1 && n_features <= 65535, "n_features must be in [1, 65535]");
    const uint32_t total = tree_off[n_trees];
    TC_REQUIRE(total == 0 || (feature && bin && flags && left && right && value),
               "tc_rensemble_create: NULL node array");
    std::vector<RPNode> nodes(std::max(1u, total));
    std::vector<uint32_t> root(std::max(1u, n_trees)), cls(std::max(1u, n_trees), 0u);
    for (uint32_t t = 0; t < n_trees; ++t) {
      const uint32_t a = tree_off[t], b = tree_off[t + 1];
      TC_REQUIRE(a <= b && b <= total, "tree_off must be non-decreasing");
      root[t] = a == b ? kRNone : a;
      if (tree_class) {
        TC_REQUIRE(tree_class[t] < n_classes, "tree_class out of range");
        cls[t] = tree_class[t];
      }
      for (uint32_t k = a; k < b; ++k) {
        RPNode& x = nodes[k];
        if (flags[k] & 1u) {
          x.a = 3u << 16;
          x.v = value[k];
          x.left = x.right = k;
          continue;
        }
        const uint32_t f = feature[k];
        TC_REQUIRE(f < n_features, "split feature out of range");
        TC_REQUIRE(left[k] < b - a && right[k] < b - a, "child id out of range");
        const uint32_t ne = n_edges[f], thr = bin[k];
        uint32_t kind = 2u;
        float t_ = 0.f;
        if (thr < ne) {
          t_ = edges[static_cast<size_t>(f) * 255 + thr];
          kind = (rule == 1 && ne > 16) ? 1u : 0u;
        }
        x.a = f | (kind << 16);
        x.v = t_;
        x.left = a + left[k];
        x.right = a + right[k];
      }
    }
    // every walk must end at a leaf within max depth: children are later ids
    for (uint32_t t = 0; t < n_trees; ++t)
      for (uint32_t k = tree_off[t]; k < tree_off[t + 1]; ++k)
        if ((nodes[k].a >> 16) != 3u)
          TC_REQUIRE(nodes[k].left > k && nodes[k].right > k, "child ids must follow their parent's");
    auto* e = new tc_rensemble();
    try {
      e->ctx = ref(ctx);
      e->T = n_trees;
      e->K = n_classes;
      e->F = n_features;
      hipStream_t s = ctx->stream;
      e->nodes = dmalloc<RPNode>(nodes.size());
      e->root = dmalloc<uint32_t>(root.size());
      e->weight = dmalloc<float>(std::max(1u, n_trees));
      e->tree_class = dmalloc<uint32_t>(cls.size());
      e->base = dmalloc<float>(n_classes);
      TC_HIP(hipMemcpyAsync(e->nodes, nodes.data(), sizeof(RPNode) * nodes.size(), hipMemcpyHostToDevice, s));
      TC_HIP(hipMemcpyAsync(e->root, root.data(), 4ull * root.size(), hipMemcpyHostToDevice, s));
      if (n_trees) TC_HIP(hipMemcpyAsync(e->weight, weights, 4ull * n_trees, hipMemcpyHostToDevice, s));
      TC_HIP(hipMemcpyAsync(e->tree_class, cls.data(), 4ull * cls.size(), hipMemcpyHostToDevice, s));
      TC_HIP(hipMemcpyAsync(e->base, base, 4ull * n_classes, hipMemcpyHostToDevice, s));
      TC_HIP(hipStreamSynchronize(s));
    } catch (...) {
      rensemble_free(e);
      throw;
    }
    *out = e;
  });
}

int32_t tc_last_train_hist_rows(tc_ctx* ctx, uint64_t* rows, uint32_t* launches) {
  using namespace tc;
  return guard([&] {
    TC_REQUIRE(ctx, "NULL context");
    if (rows) *rows = ctx->last_hist_rows;
    if (launches) *launches = ctx->last_hist_launches;
  });
}

int32_t tc_rensemble_destroy(tc_rensemble* ens) {
  using namespace tc;
  return guard([&] { rensemble_free(ens); });
}

int32_t tc_rpredict_dev(tc_rensemble* ens, const float* X_dev, uint32_t n, uint32_t F,
                        int32_t transform, float* out_dev) {
  using namespace tc;
  return guard([&] {
    TC_REQUIRE(ens && (n == 0 || (X_dev && out_dev)), "tc_rpredict_dev: NULL argument");
    rpredict_dev(ens, X_dev, n, F, transform, out_dev);
  });
}

int32_t tc_rpredict(tc_rensemble* ens, const float* X, uint32_t n, uint32_t F, int32_t transform,
                    float* out) {
  using namespace tc;
  return guard([&] {
    TC_REQUIRE(ens && (n == 0 || (X && out)), "tc_rpredict: NULL argument");
    TC_REQUIRE(F == ens->F, "tc_rpredict: feature count differs from the ensemble's");
    if (n == 0) return;
    hipStream_t s = ens->ctx->stream;
    float* dx = dmalloc<float>(static_cast<size_t>(n) * F);
    float* dout = nullptr;
    try {
      dout = dmalloc<float>(static_cast<size_t>(n) * ens->K);
      TC_HIP(hipMemcpyAsync(dx, X, 4ull * n * F, hipMemcpyHostToDevice, s));
      rpredict_dev(ens, dx, n, F, transform, dout);
      TC_HIP(hipMemcpyAsync(out, dout, 4ull * n * ens->K, hipMemcpyDeviceToHost, s));
      TC_HIP(hipStreamSynchronize(s));
    } catch (...) {
      dfree(dx);
      dfree(dout);
      throw;
    }
    dfree(dx);
    dfree(dout);
  });
}

}  // extern "C"
