/* Test infrastructure only -- the CPU oracle.  See tc_oracle.h.
 *
 * Built with -std=gnu11 -O3 -mavx2 -mfma: GNU mode keeps GCC's default
 * floating-point contraction (-ffp-contract=fast), the same setting the
 * reference's -std=c++20 build gets, so the *_f32 functions contract the same
 * expressions into FMAs as the reference does.  Tests pin that claim against
 * the compiled reference bit for bit. */
#include "tc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- A2 / A3 */

static int cmp_float(const void* a, const void* b) {
    const float x = *(const float*)a, y = *(const float*)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}

/* upper_bound: number of edges <= v (src/core/dataset.cpp:336-344) */
static uint8_t find_bin(float v, const float* e, uint32_t ne) {
    uint32_t lo = 0, hi = ne;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v < e[mid]) hi = mid;
        else lo = mid + 1;
    }
    return (uint8_t)lo;
}

int tco_quantize(const float* X, uint32_t n, uint32_t F, uint32_t max_bins, int wrap,
                 uint8_t* bins, float* edges, uint32_t* n_edges) {
    /* compute_quantile_bins, src/core/dataset.cpp:278-334 */
    const uint8_t nb = (uint8_t)max_bins; /* BinIndex cast, :177 */
    float* vals = (float*)malloc(sizeof(float) * (n ? n : 1));
    for (uint32_t f = 0; f < F; ++f) {
        uint32_t U = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const float v = X[(size_t)i * F + f];
            if (!isnan(v)) vals[U++] = v;
        }
        float* e = edges + (size_t)f * 255;
        uint8_t* col = bins + (size_t)f * n;
        if (U == 0) { /* all NaN: one edge 0.0, every row in the NaN bin (:288-295) */
            e[0] = 0.0f;
            n_edges[f] = 1;
            for (uint32_t i = 0; i < n; ++i) col[i] = nb;
            continue;
        }
        qsort(vals, U, sizeof(float), cmp_float);
        /* std::unique with operator== (-0.0 == +0.0 collapse to one) */
        uint32_t u = 1;
        for (uint32_t i = 1; i < U; ++i)
            if (!(vals[i] == vals[u - 1])) vals[u++] = vals[i];
        /* actual_bins = min(n_bins, BinIndex(U)) -- the uint8 wrap (:308) */
        const uint32_t A = wrap ? ((uint32_t)(uint8_t)u < nb ? (uint32_t)(uint8_t)u : nb)
                                : (u < nb ? u : nb);
        uint32_t ne = 0;
        for (uint32_t b = 1; b < A; ++b) { /* :313-318 */
            const size_t idx = (size_t)b * u / A;
            if (ne == 0 || vals[idx] > e[ne - 1]) e[ne++] = vals[idx];
        }
        n_edges[f] = ne;
        for (uint32_t i = 0; i < n; ++i) {
            const float v = X[(size_t)i * F + f];
            col[i] = isnan(v) ? nb : find_bin(v, e, ne);
        }
    }
    free(vals);
    return 0;
}

void tco_apply_bins(const float* X, uint32_t n, uint32_t F, const float* edges,
                    const uint32_t* n_edges, uint8_t* bins) {
    for (uint32_t f = 0; f < F; ++f)
        for (uint32_t i = 0; i < n; ++i) {
            const float v = X[(size_t)i * F + f];
            bins[(size_t)f * n + i] = isnan(v) ? 255 : find_bin(v, edges + (size_t)f * 255, n_edges[f]);
        }
}

/* ---------------------------------------------------------------- A5 - A8 */

float tco_init_prediction(int loss, const float* y, uint32_t n) {
    float s = 0.0f; /* sequential float sum, loss.cpp:218-221 / :365-368 */
    for (uint32_t i = 0; i < n; ++i) s += y[i];
    if (loss == TCO_LOSS_MSE) return s / (float)n;
    float p = s / (float)n;
    const float eps = 1e-7f; /* clip, loss.hpp:103-105 */
    p = fmaxf(eps, fminf(1.0f - eps, p));
    return logf(p / (1.0f - p));
}

/* The reference's AVX2 exp: round-to-nearest range reduction t = x*log2(e),
 * degree-5 Taylor polynomial of 2^f in Horner form with fused multiply-adds,
 * scaled by 2^round(t) through the exponent bits (loss.cpp:34-69). */
static float ref_exp_poly(float x) {
    x = fmaxf(x, -88.0f);
    x = fminf(x, 88.0f);
    const float t = x * 1.4426950408889634f;
    const float ti = rintf(t);
    const float tf = t - ti;
    float poly = fmaf(0.0013333558146428443f, tf, 0.009618129107628477f);
    poly = fmaf(poly, tf, 0.05550410866482046f);
    poly = fmaf(poly, tf, 0.2402265069591007f);
    poly = fmaf(poly, tf, 0.6931471805599453f);
    poly = fmaf(poly, tf, 1.0f);
    union { uint32_t u; float f; } pw;
    pw.u = (uint32_t)((int32_t)ti + 127) << 23;
    return poly * pw.f;
}

void tco_gradients_f32(int loss, const float* y, const float* p, float* g, float* h, uint32_t n) {
    if (loss == TCO_LOSS_MSE) { /* loss.cpp:355-361 */
        for (uint32_t i = 0; i < n; ++i) {
            g[i] = 2.0f * (p[i] - y[i]);
            h[i] = 2.0f;
        }
        return;
    }
    const uint32_t nv = (n / 8) * 8; /* AVX2 rows, loss.cpp:130-155 */
    for (uint32_t i = 0; i < nv; ++i) {
        float x = fmaxf(p[i], -15.0f);
        x = fminf(x, 15.0f);
        const float e = ref_exp_poly(0.0f - x);
        const float s = 1.0f / (1.0f + e);
        g[i] = s - y[i];
        const float hh = s * (1.0f - s);
        h[i] = hh > 1e-6f ? hh : 1e-6f;
    }
    for (uint32_t i = nv; i < n; ++i) { /* scalar tail, :157-163 */
        const float pr = fmaxf(-15.0f, fminf(15.0f, p[i]));
        const float s = 1.0f / (1.0f + expf(-pr));
        g[i] = s - y[i];
        const float hh = s * (1.0f - s);
        h[i] = hh > 1e-6f ? hh : 1e-6f;
    }
}

void tco_gradients_f64(int loss, const float* y, const float* p, double* g, double* h, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        if (loss == TCO_LOSS_MSE) {
            g[i] = 2.0 * ((double)p[i] - (double)y[i]);
            h[i] = 2.0;
        } else {
            double x = (double)p[i];
            x = x < -15.0 ? -15.0 : (x > 15.0 ? 15.0 : x);
            const double s = 1.0 / (1.0 + exp(-x));
            g[i] = s - (double)y[i];
            const double hh = s * (1.0 - s);
            h[i] = hh > 1e-6 ? hh : 1e-6;
        }
    }
}

void tco_transform_f32(const float* raw, float* out, uint32_t n, uint32_t K) {
    /* probabilities from raw scores in the reference's float order.  K == 1:
     * LogLoss::transform_prediction, the sigmoid 1 / (1 + exp(-x)) (loss.cpp:
     * 213-215, loss.hpp:99-101); K > 1: CrossEntropyLoss::transform_to_proba
     * (loss.cpp:320-331), the softmax of loss.cpp:273-285 per row: max, exp of
     * the shifted scores, sequential sum, divide.  raw / out row-major n x K. */
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = raw + (size_t)i * K;
        float* o = out + (size_t)i * K;
        if (K == 1) {
            o[0] = 1.0f / (1.0f + expf(-r[0]));
            continue;
        }
        float m = r[0];
        for (uint32_t c = 1; c < K; ++c) m = r[c] > m ? r[c] : m;
        float e[64];
        float s = 0.0f;
        for (uint32_t c = 0; c < K; ++c) {
            e[c] = expf(r[c] - m);
            s += e[c];
        }
        for (uint32_t c = 0; c < K; ++c) o[c] = e[c] / s;
    }
}

void tco_gradients_mc_f32(const float* y, const float* p, float* g, float* h, uint32_t n, uint32_t K) {
    /* compute_multiclass_gradients + softmax in the reference's float order,
     * loss.cpp:273-318: max shift, exp, sequential sum, divide; p and g/h
     * row-major n x K */
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = p + (size_t)i * K;
        float m = r[0];
        for (uint32_t c = 1; c < K; ++c) m = r[c] > m ? r[c] : m;
        float e[64];
        float s = 0.0f;
        for (uint32_t c = 0; c < K; ++c) {
            e[c] = expf(r[c] - m);
            s += e[c];
        }
        const uint32_t lab = (uint32_t)y[i];
        for (uint32_t c = 0; c < K; ++c) {
            const float pc = e[c] / s;
            g[(size_t)i * K + c] = pc - (c == lab ? 1.0f : 0.0f);
            const float hh = pc * (1.0f - pc);
            h[(size_t)i * K + c] = hh > 1e-6f ? hh : 1e-6f;
        }
    }
}

void tco_gradients_mc_f64(const float* y, const float* p, double* g, double* h, uint32_t n, uint32_t K) {
    /* compute_multiclass_gradients + softmax, loss.cpp:273-318 */
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = p + (size_t)i * K;
        double m = r[0];
        for (uint32_t c = 1; c < K; ++c) m = r[c] > m ? r[c] : m;
        double s = 0.0;
        for (uint32_t c = 0; c < K; ++c) s += exp((double)r[c] - m);
        const uint32_t lab = (uint32_t)y[i];
        for (uint32_t c = 0; c < K; ++c) {
            const double pc = exp((double)r[c] - m) / s;
            g[(size_t)i * K + c] = pc - (c == lab ? 1.0 : 0.0);
            const double hh = pc * (1.0 - pc);
            h[(size_t)i * K + c] = hh > 1e-6 ? hh : 1e-6;
        }
    }
}

/* ---------------------------------------------------------------- A9 */

void tco_hist_f32(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                  const uint32_t* rows, uint32_t n_rows, float* out_gh, uint32_t* out_count) {
    memset(out_gh, 0, sizeof(float) * (size_t)F * 512);
    if (out_count) memset(out_count, 0, sizeof(uint32_t) * (size_t)F * 256);
    const uint32_t nr = rows ? n_rows : n;
    for (uint32_t f = 0; f < F; ++f) {
        const uint8_t* col = bins + (size_t)f * n;
        float* o = out_gh + (size_t)f * 512;
        for (uint32_t i = 0; i < nr; ++i) {
            const uint32_t r = rows ? rows[i] : i;
            o[2 * col[r]] += g[r];
            o[2 * col[r] + 1] += h[r];
            if (out_count) out_count[(size_t)f * 256 + col[r]] += 1;
        }
    }
}

void tco_hist_f64(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                  const uint32_t* rows, uint32_t n_rows, double* out_gh, double* abs_g,
                  uint32_t* out_count) {
    memset(out_gh, 0, sizeof(double) * (size_t)F * 512);
    if (abs_g) memset(abs_g, 0, sizeof(double) * (size_t)F * 256);
    if (out_count) memset(out_count, 0, sizeof(uint32_t) * (size_t)F * 256);
    const uint32_t nr = rows ? n_rows : n;
    for (uint32_t f = 0; f < F; ++f) {
        const uint8_t* col = bins + (size_t)f * n;
        double* o = out_gh + (size_t)f * 512;
        for (uint32_t i = 0; i < nr; ++i) {
            const uint32_t r = rows ? rows[i] : i;
            const uint32_t b = col[r];
            o[2 * b] += (double)g[r];
            o[2 * b + 1] += (double)h[r];
            if (abs_g) abs_g[(size_t)f * 256 + b] += fabs((double)g[r]);
            if (out_count) out_count[(size_t)f * 256 + b] += 1;
        }
    }
}

/* ---------------------------------------------------------------- A10 */

#define SPLIT_BODY(REAL)                                                                     \
    const uint32_t mb = max_bins & 0xFFu;                                                    \
    const REAL eps = (REAL)1e-10f;                                                           \
    REAL* pg = (REAL*)calloc(n_nodes, sizeof(REAL));                                         \
    REAL* L = (REAL*)calloc((size_t)n_nodes * 2, sizeof(REAL));                              \
    for (uint32_t nd = 0; nd < n_nodes; ++nd)                                                \
        pg[nd] = node_h[nd] > eps ? (node_g[nd] * node_g[nd]) / (node_h[nd] + lambda) : 0;   \
    REAL best = (REAL)-1e30f, sec = (REAL)-1e30f;                                            \
    int32_t bf = 0, bb = 0;                                                                  \
    for (uint32_t f = 0; f < F; ++f) {                                                       \
        memset(L, 0, sizeof(REAL) * n_nodes * 2);                                            \
        for (uint32_t b = 0; b + 1 < mb; ++b) {                                              \
            REAL total = 0;                                                                  \
            int any = 0;                                                                     \
            for (uint32_t nd = 0; nd < n_nodes; ++nd) {                                      \
                const REAL* hb = hist + ((size_t)nd * F + f) * 512 + 2 * b;                  \
                L[2 * nd] += hb[0];                                                          \
                L[2 * nd + 1] += hb[1];                                                      \
                if (!is_root && node_count[nd] == 0) continue;                               \
                const REAL rg = node_g[nd] - L[2 * nd], rh = node_h[nd] - L[2 * nd + 1];     \
                if (L[2 * nd + 1] < eps || rh < eps) continue;                               \
                const REAL lgn = (L[2 * nd] * L[2 * nd]) / (L[2 * nd + 1] + lambda);         \
                const REAL rgn = (rg * rg) / (rh + lambda);                                  \
                total += (REAL)0.5f * (lgn + rgn - pg[nd]);                                  \
                any = 1;                                                                     \
            }                                                                                \
            if (is_root && !any) continue;                                                   \
            if (total > best) { sec = best; best = total; bf = (int32_t)f; bb = (int32_t)b; } \
            else if (total > sec) sec = total;                                               \
        }                                                                                    \
    }                                                                                        \
    free(pg); free(L);                                                                       \
    *feature = bf; *bin = bb; *gain = best;

void tco_split_f32(const float* hist, const float* node_g, const float* node_h,
                   const uint32_t* node_count, uint32_t n_nodes, uint32_t F, uint32_t max_bins,
                   float lambda, int is_root, int32_t* feature, int32_t* bin, float* gain) {
    SPLIT_BODY(float)
    (void)sec;
}

void tco_split_f64(const double* hist, const double* node_g, const double* node_h,
                   const uint32_t* node_count, uint32_t n_nodes, uint32_t F, uint32_t max_bins,
                   double lambda, int is_root, int32_t* feature, int32_t* bin, double* gain,
                   double* second) {
    SPLIT_BODY(double)
    if (second) *second = sec;
}

/* ---------------------------------------------------------------- A9 - A14 */

#define REAL float
#define SUF(x) x##_f32
#include "tc_oracle_tree.inc"
#undef REAL
#undef SUF
#define REAL double
#define SUF(x) x##_f64
#include "tc_oracle_tree.inc"
#undef REAL
#undef SUF

int tco_tree_build_f32(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                       const uint32_t* root, uint32_t n_root,
                       const float* edges, const uint32_t* n_edges, const tco_tree_params* p,
                       uint16_t* feature, uint8_t* bin, float* fthr, float* gain,
                       float* leaves, uint32_t* leaf_idx) {
    return tree_build_f32(bins, n, F, 1, g, h, root, n_root, edges, n_edges, p, feature, bin,
                          fthr, gain, NULL, leaves, leaf_idx);
}

int tco_tree_build_f64(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                       const uint32_t* root, uint32_t n_root,
                       const float* edges, const uint32_t* n_edges, const tco_tree_params* p,
                       uint16_t* feature, uint8_t* bin, float* fthr, double* gain, double* gap,
                       double* leaves, uint32_t* leaf_idx) {
    double* gd = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* hd = (double*)malloc(sizeof(double) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) { gd[i] = g[i]; hd[i] = h[i]; }
    const int d = tree_build_f64(bins, n, F, 1, gd, hd, root, n_root, edges, n_edges, p, feature,
                                 bin, fthr, gain, gap, leaves, leaf_idx);
    free(gd); free(hd);
    return d;
}

int tco_tree_build_mc_f64(const uint8_t* bins, uint32_t n, uint32_t F, uint32_t K,
                          const double* g, const double* h, const float* edges,
                          const uint32_t* n_edges, const tco_tree_params* p,
                          uint16_t* feature, uint8_t* bin, float* fthr, double* gain,
                          double* gap, double* leaves, uint32_t* leaf_idx) {
    return tree_build_f64(bins, n, F, K, g, h, NULL, n, edges, n_edges, p, feature, bin, fthr,
                          gain, gap, leaves, leaf_idx);
}

/* ---------------------------------------------------------------- A15 - A16 */

void tco_leaf_index_binned(const uint8_t* bins, uint32_t n, uint32_t depth,
                           const uint16_t* feature, const uint8_t* bin, uint32_t* leaf_idx) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t li = 0;
        for (uint32_t d = 0; d < depth; ++d) {
            const uint8_t v = bins[(size_t)feature[d] * n + i];
            li = (li << 1) | ((v != 255 && v > bin[d]) ? 1u : 0u);
        }
        leaf_idx[i] = li;
    }
}

static uint32_t float_leaf(const float* row, const uint16_t* feat, const float* thr, uint32_t depth) {
    /* fast_float_ensemble.hpp:122-137: NaN goes left, v >= threshold goes right */
    uint32_t li = 0;
    for (uint32_t d = 0; d < depth; ++d) {
        const float v = row[feat[d]];
        li = (li << 1) | ((!isnan(v) && v >= thr[d]) ? 1u : 0u);
    }
    return li;
}

/* Sum of one row's tree values in the reference's float order.  lv[t] is
 * tree t's leaf value.  Rows of a full 8-row block (grouped != 0) take the
 * AVX2 block order (fast_float_ensemble.hpp:388-541 for raw floats,
 * fast_ensemble.hpp:256-426 for bins): groups of four trees, each
 * w0*l0 + w1*l1 + w2*l2 + w3*l3 as the reference compiler contracts it (w1*l1
 * rounded, the other three products fused onto it in order), the groups added
 * with a plain add, the T mod 4 leftover trees fused in one by one (the
 * compiler forwards the stored product into the vector add).  Other rows take
 * the scalar order sum += w_t * l_t, one fused multiply-add per tree in tree
 * order: the rows past the last full 8-row block (fast_float_ensemble.hpp:
 * 279-287, :543-560; fast_ensemble.hpp:428-446), every row of a binned call
 * under 16 rows (symmetric_tree.cpp:936-962) and every row of an ensemble
 * whose trees all have depth 0 (the fast ensembles are then empty and the
 * fallbacks add tree by tree, symmetric_tree.cpp:1144-1148, :1010-1012). */
static float ensemble_sum(const float* lv, const float* weight, uint32_t T, int grouped) {
    float sum = 0.0f;
    if (!grouped) {
        for (uint32_t t = 0; t < T; ++t) sum = fmaf(weight[t], lv[t], sum);
        return sum;
    }
    const uint32_t TB = T / 4;
    for (uint32_t tb = 0; tb < TB; ++tb) {
        const uint32_t t0 = tb * 4;
        float grp = weight[t0 + 1] * lv[t0 + 1];
        grp = fmaf(weight[t0], lv[t0], grp);
        grp = fmaf(weight[t0 + 2], lv[t0 + 2], grp);
        grp = fmaf(weight[t0 + 3], lv[t0 + 3], grp);
        sum = sum + grp;
    }
    for (uint32_t t = TB * 4; t < T; ++t) sum = fmaf(weight[t], lv[t], sum);
    return sum;
}

static int any_split(const uint16_t* depth, uint32_t T) {
    for (uint32_t t = 0; t < T; ++t)
        if (depth[t] > 0) return 1;
    return 0;
}

void tco_predict_f32(const uint16_t* feature, const float* fthr, const uint16_t* depth,
                     const float* leaves, const float* weight, uint32_t T, uint32_t D,
                     float base, const float* X, uint32_t n, uint32_t F, float* out,
                     uint32_t* leaf_idx) {
    const uint32_t LP = 1u << D;
    const uint32_t nblk = (n / 8) * 8;
    const int split = any_split(depth, T);
    float* lv = (float*)malloc(sizeof(float) * (T ? T : 1));
    for (uint32_t i = 0; i < n; ++i) {
        const float* row = X + (size_t)i * F;
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t l = float_leaf(row, feature + (size_t)t * D, fthr + (size_t)t * D, depth[t]);
            if (leaf_idx) leaf_idx[(size_t)i * T + t] = l;
            lv[t] = leaves[(size_t)t * LP + l];
        }
        out[i] = ensemble_sum(lv, weight, T, split && i < nblk) + base; /* booster.cpp:918-924 */
    }
    free(lv);
}

void tco_predict_binned_f32(const uint16_t* feature, const uint8_t* bin, const uint16_t* depth,
                            const float* leaves, const float* weight, uint32_t T, uint32_t D,
                            float base, const uint8_t* bins, uint32_t n, uint32_t F, float* out) {
    const uint32_t LP = 1u << D;
    const uint32_t nblk = n >= 16 ? (n / 8) * 8 : 0;
    const int split = any_split(depth, T);
    float* lv = (float*)malloc(sizeof(float) * (T ? T : 1));
    (void)F;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t t = 0; t < T; ++t) {
            /* go right = bin != 255 && bin > threshold (fast_ensemble.hpp:293-301) */
            uint32_t l = 0;
            for (uint32_t d = 0; d < depth[t]; ++d) {
                const uint8_t v = bins[(size_t)feature[(size_t)t * D + d] * n + i];
                l = (l << 1) | ((v != 255 && v > bin[(size_t)t * D + d]) ? 1u : 0u);
            }
            lv[t] = leaves[(size_t)t * LP + l];
        }
        out[i] = ensemble_sum(lv, weight, T, split && i < nblk) + base; /* booster.cpp:599-604 */
    }
    free(lv);
}

void tco_predict_f64(const uint16_t* feature, const float* fthr, const uint16_t* depth,
                     const float* leaves, const float* weight, uint32_t T, uint32_t D,
                     double base, const float* X, uint32_t n, uint32_t F, double* out,
                     double* abs_sum) {
    const uint32_t LP = 1u << D;
    for (uint32_t i = 0; i < n; ++i) {
        const float* row = X + (size_t)i * F;
        double s = 0.0, a = fabs(base);
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t l = float_leaf(row, feature + (size_t)t * D, fthr + (size_t)t * D, depth[t]);
            const double v = (double)weight[t] * (double)leaves[(size_t)t * LP + l];
            s += v;
            a += fabs(v);
        }
        out[i] = s + base;
        if (abs_sum) abs_sum[i] = a;
    }
}

/* K-channel ensemble (A18): out[i][c] = base + sum_t w_t * leaf_t[idx_t][c],
 * leaves [T][2^D][K], FP64; abs_sum[i][c] = |base| + sum_t |w_t * leaf|. */
void tco_predict_mc_f64(const uint16_t* feature, const float* fthr, const uint16_t* depth,
                        const float* leaves, const float* weight, uint32_t T, uint32_t D,
                        uint32_t K, double base, const float* X, uint32_t n, uint32_t F,
                        double* out, double* abs_sum) {
    const uint32_t LP = 1u << D;
    for (uint32_t i = 0; i < n; ++i) {
        const float* row = X + (size_t)i * F;
        for (uint32_t c = 0; c < K; ++c) {
            out[(size_t)i * K + c] = base;
            if (abs_sum) abs_sum[(size_t)i * K + c] = fabs(base);
        }
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t l = float_leaf(row, feature + (size_t)t * D, fthr + (size_t)t * D, depth[t]);
            for (uint32_t c = 0; c < K; ++c) {
                const double v = (double)weight[t] * (double)leaves[((size_t)t * LP + l) * K + c];
                out[(size_t)i * K + c] += v;
                if (abs_sum) abs_sum[(size_t)i * K + c] += fabs(v);
            }
        }
    }
}

/* ---------------------------------------------------------------- A5 - A15 */

void tco_fma_update(float* preds, const float* vals, float lr, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) preds[i] = fmaf(lr, vals[i], preds[i]);
}

float tco_fit_f32(const uint8_t* bins, uint32_t n, uint32_t F, const float* edges,
                  const uint32_t* n_edges, const float* y, int loss, uint32_t n_trees,
                  float lr, const tco_tree_params* p, uint16_t* feature, uint8_t* bin,
                  float* fthr, float* gain, uint16_t* depth, float* leaves, float* train_preds) {
    const uint32_t D = p->max_depth, LP = 1u << D;
    const float base = tco_init_prediction(loss, y, n);
    float* g = (float*)malloc(sizeof(float) * (n ? n : 1));
    float* h = (float*)malloc(sizeof(float) * (n ? n : 1));
    uint32_t* li = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) train_preds[i] = base;
    for (uint32_t t = 0; t < n_trees; ++t) {
        tco_gradients_f32(loss, y, train_preds, g, h, n); /* booster.cpp:505-519 */
        const int d = tco_tree_build_f32(bins, n, F, g, h, NULL, n, edges, n_edges, p,
                                         feature + (size_t)t * D, bin + (size_t)t * D,
                                         fthr + (size_t)t * D, gain + (size_t)t * D,
                                         leaves + (size_t)t * LP, NULL);
        depth[t] = (uint16_t)d;
        /* binned predict + preds += lr * tree (symmetric_tree.cpp:813-840,
         * booster.cpp:458-463) */
        tco_leaf_index_binned(bins, n, (uint32_t)d, feature + (size_t)t * D, bin + (size_t)t * D, li);
        for (uint32_t i = 0; i < n; ++i) train_preds[i] += lr * leaves[(size_t)t * LP + li[i]];
    }
    free(g); free(h); free(li);
    return base;
}

/* ---------------------------------------------------------------- A19 */

/* Loss::compute_loss, the mean per-row loss after each tree (booster.cpp:252
 * for the training rows, :283 for the validation rows), float accumulation in
 * row order (the reference's OpenMP reduction on one thread):
 * MSELoss (loss.cpp:336-346): (p - y)^2; LogLoss (:110-121):
 * -y log q - (1 - y) log(1 - q), q = clip(1 / (1 + exp(-p)), 1e-7)
 * (loss.hpp:99-105). */
float tco_compute_loss_f32(int loss, const float* y, const float* p, uint32_t n) {
    float s = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        if (loss == 0) {
            const float d = p[i] - y[i];
            s += d * d;
        } else {
            float q = 1.0f / (1.0f + expf(-p[i]));
            q = q < 1e-7f ? 1e-7f : (q > 1.0f - 1e-7f ? 1.0f - 1e-7f : q);
            s += -y[i] * logf(q) - (1 - y[i]) * logf(1 - q);
        }
    }
    return s / (float)n;
}

double tco_compute_loss_f64(int loss, const float* y, const float* p, uint32_t n) {
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        if (loss == 0) {
            const double d = (double)p[i] - (double)y[i];
            s += d * d;
        } else {
            double q = 1.0 / (1.0 + exp(-(double)p[i]));
            q = q < 1e-7 ? 1e-7 : (q > 1.0 - 1e-7 ? 1.0 - 1e-7 : q);
            s += -(double)y[i] * log(q) - (1.0 - (double)y[i]) * log(1.0 - q);
        }
    }
    return s / (double)n;
}

/* CrossEntropyLoss::compute_loss (loss.cpp:232-244): p row-major n x K,
 * softmax with the max shift, -log(max(q[label], 1e-7)), float, row order. */
float tco_compute_loss_mc_f32(const float* y, const float* p, uint32_t n, uint32_t K) {
    float s = 0.0f;
    float e[64];
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = p + (size_t)i * K;
        float m = r[0];
        for (uint32_t c = 1; c < K; ++c) m = r[c] > m ? r[c] : m;
        float sum = 0.0f;
        for (uint32_t c = 0; c < K; ++c) { e[c] = expf(r[c] - m); sum += e[c]; }
        const uint32_t lab = (uint32_t)y[i];
        const float q = e[lab] / sum;
        s += -logf(q > 1e-7f ? q : 1e-7f);
    }
    return s / (float)n;
}

/* Early stopping over a sequence of per-tree validation metrics
 * (booster.cpp:115-121, :326-352): improved = metric < best - 1e-10 (or
 * > best + 1e-10 when higher is better), best starting at +-1e30; after an
 * improvement the count resets, else it grows; the loop breaks after tree
 * iter when iter >= max(50, n_estimators / 10) and the count has reached
 * `rounds`.  Returns the trees kept (iter + 1 at the break, else n_iter);
 * best_iter / best get the last improvement. */
uint32_t tco_early_stop(const float* metric, uint32_t n_iter, uint32_t n_estimators,
                        uint32_t rounds, int higher_is_better, uint32_t* best_iter, float* best) {
    float b = higher_is_better ? -1e30f : 1e30f;
    uint32_t bi = 0, count = 0;
    const uint32_t min_it = n_estimators / 10 > 50u ? n_estimators / 10 : 50u;
    uint32_t kept = n_iter;
    for (uint32_t it = 0; it < n_iter; ++it) {
        const float v = metric[it];
        const int improved = higher_is_better ? (v > b + 1e-10f) : (v < b - 1e-10f);
        if (improved) { b = v; bi = it; count = 0; } else { ++count; }
        if (it >= min_it && count >= rounds) { kept = it + 1; break; }
    }
    if (best_iter) *best_iter = bi;
    if (best) *best = b;
    return kept;
}

/* ------------------------------------------------ regular depthwise trees */

#include "tc_oracle_regular.inc"
