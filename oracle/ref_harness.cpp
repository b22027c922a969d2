// Test infrastructure only: a C-ABI harness over the reference's own
// hot-path translation units (compiled in place from the read-only reference
// tree by oracle/Makefile, output in oracle/_ref/).  Used by tests/ to pin the
// oracle and write golden vectors, and by bench.py's cpu_baseline leg.  Nothing
// in the product links or loads this.
//
// The boosting loop below restates src/boosting/booster.cpp:92-100, :173-174,
// :198-229, :438-465, :505-519 and :551-583 (binary/regression, symmetric
// branch, row sampling) because booster.cpp itself does not compile here (it pulls in
// Eigen through tree.hpp).  Everything else is the reference's code:
//   Dataset::from_dense / compute_bins      src/core/dataset.cpp:50-94, :176-344
//   Dataset::compute_target_statistics      src/core/dataset.cpp:679-721
//   Loss::create / compute_gradients        include/turbocat/loss.hpp:511, src/boosting/loss.cpp
//   CPUHistogramBuilder::build              src/tree/histogram.cpp:62-205
//   SymmetricTree::build / predict_batch    src/tree/symmetric_tree.cpp:34-368, :813-840
//   SymmetricEnsemble::predict_batch_raw_fast  src/tree/symmetric_tree.cpp:1133-1149

#include "turbocat/config.hpp"
#include "turbocat/dataset.hpp"
#include "turbocat/histogram.hpp"
#include "turbocat/loss.hpp"
#include "turbocat/symmetric_tree.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <omp.h>
#include <memory>
#include <numeric>
#include <vector>

using namespace turbocat;

namespace {

struct RefData {
    std::unique_ptr<Dataset> ds;
    Config cfg;
};

double now_s() {
    return std::chrono::duration<double>(
        std::chrono::steady_clock::now().time_since_epoch()).count();
}

Config make_config(uint32_t max_bins, uint32_t max_depth, float lambda_l2,
                   float min_child_weight, int loss_id) {
    Config c;
    c.verbosity = 0;
    c.tree.max_bins = static_cast<uint16_t>(max_bins);
    c.tree.max_depth = static_cast<uint16_t>(max_depth);
    c.tree.lambda_l2 = lambda_l2;
    c.tree.min_child_weight = min_child_weight;
    c.tree.use_symmetric = true;
    c.boosting.subsample = 1.0f;
    c.boosting.colsample_bytree = 1.0f;
    if (loss_id == 0) {
        c.task = TaskType::Regression;
        c.loss.loss_type = LossType::MSE;
    } else {
        c.task = TaskType::BinaryClassification;
        c.loss.loss_type = LossType::LogLoss;
    }
    return c;
}

}  // namespace

extern "C" {

// Bin X (row-major n x F) exactly as the reference fit does.  bins_out is
// column-major F x n; edges_out holds 255 slots per feature.
void* tcref_dataset_create(const float* X, const float* y, uint32_t n, uint32_t F,
                           uint32_t max_bins, double* seconds) {
    auto* r = new RefData();
    r->cfg = make_config(max_bins, 6, 1.0f, 1.0f, 0);
    r->ds = std::make_unique<Dataset>();
    double t0 = now_s();
    r->ds->from_dense(X, n, static_cast<FeatureIndex>(F), y, nullptr, {});
    r->ds->compute_bins(r->cfg);
    if (seconds) *seconds = now_s() - t0;
    return r;
}

void tcref_dataset_free(void* h) { delete static_cast<RefData*>(h); }

void tcref_dataset_bins(void* h, uint8_t* bins_out, float* edges_out, uint32_t* n_edges_out) {
    auto* r = static_cast<RefData*>(h);
    const Dataset& ds = *r->ds;
    const uint32_t n = ds.n_samples(), F = ds.n_features();
    for (uint32_t f = 0; f < F; ++f) {
        std::memcpy(bins_out + static_cast<size_t>(f) * n, ds.binned().column(f), n);
        const auto& e = ds.bin_edges()[f];
        n_edges_out[f] = static_cast<uint32_t>(e.size());
        for (size_t k = 0; k < e.size() && k < 255; ++k) edges_out[f * 255 + k] = e[k];
    }
}

void tcref_gradients(int loss_id, const float* y, const float* p, float* g, float* h, uint32_t n) {
    LossConfig lc;
    lc.loss_type = loss_id == 0 ? LossType::MSE : LossType::LogLoss;
    auto loss = Loss::create(lc, loss_id == 0 ? TaskType::Regression : TaskType::BinaryClassification, 2);
    loss->compute_gradients(y, p, g, h, n);
}

// CrossEntropyLoss::compute_multiclass_gradients (loss.cpp:287-318):
// p, g, h row-major n x K.
void tcref_gradients_mc(const float* y, const float* p, float* g, float* h, uint32_t n,
                        uint32_t K) {
    CrossEntropyLoss loss(K);
    loss.compute_multiclass_gradients(y, p, g, h, n);
}

// Loss::compute_loss (loss.cpp:110-121, :232-244, :336-346) on one OpenMP
// thread, so the float reduction runs in row order.
float tcref_compute_loss(int loss_id, const float* y, const float* p, uint32_t n) {
    LossConfig lc;
    lc.loss_type = loss_id == 0 ? LossType::MSE : LossType::LogLoss;
    auto loss = Loss::create(lc, loss_id == 0 ? TaskType::Regression : TaskType::BinaryClassification, 2);
    const int threads = omp_get_max_threads();
    omp_set_num_threads(1);
    const float v = loss->compute_loss(y, p, n);
    omp_set_num_threads(threads);
    return v;
}

float tcref_compute_loss_mc(const float* y, const float* p, uint32_t n, uint32_t K) {
    CrossEntropyLoss loss(K);
    return loss.compute_loss(y, p, n);
}

// Probabilities from raw scores: LogLoss::transform_prediction (loss.cpp:213-215)
// for K == 1, CrossEntropyLoss::transform_to_proba (loss.cpp:320-331) for K > 1
// (raw and out row-major n x K).
void tcref_transform(const float* raw, float* out, uint32_t n, uint32_t K) {
    if (K == 1) {
        LossConfig lc;
        lc.loss_type = LossType::LogLoss;
        auto loss = Loss::create(lc, TaskType::BinaryClassification, 2);
        for (uint32_t i = 0; i < n; ++i) out[i] = loss->transform_prediction(raw[i]);
    } else {
        CrossEntropyLoss loss(K);
        loss.transform_to_proba(raw, out, n);
    }
}

float tcref_init_prediction(int loss_id, const float* y, uint32_t n) {
    LossConfig lc;
    lc.loss_type = loss_id == 0 ? LossType::MSE : LossType::LogLoss;
    auto loss = Loss::create(lc, loss_id == 0 ? TaskType::Regression : TaskType::BinaryClassification, 2);
    return loss->init_prediction(y, n);
}

// Histogram over a row list (rows == NULL means all rows in order).
// out_gh: F x 255 x {grad, hess}; out_count: F x 255.
void tcref_hist(void* h, const float* g, const float* hs, const uint32_t* rows, uint32_t n_rows,
                float* out_gh, uint32_t* out_count) {
    auto* r = static_cast<RefData*>(h);
    Dataset& ds = *r->ds;
    const uint32_t n = ds.n_samples(), F = ds.n_features();
    std::memcpy(ds.gradients(), g, n * sizeof(float));
    std::memcpy(ds.hessians(), hs, n * sizeof(float));
    std::vector<Index> idx;
    if (rows) idx.assign(rows, rows + n_rows);
    else { idx.resize(n); std::iota(idx.begin(), idx.end(), 0u); }
    Histogram hist(static_cast<FeatureIndex>(F), 255);
    CPUHistogramBuilder b;
    b.build(ds, idx, {}, hist);
    for (uint32_t f = 0; f < F; ++f) {
        const GradientPair* bp = hist.bins(static_cast<FeatureIndex>(f));
        for (uint32_t k = 0; k < 255; ++k) {
            out_gh[(f * 255 + k) * 2 + 0] = bp[k].grad;
            out_gh[(f * 255 + k) * 2 + 1] = bp[k].hess;
            out_count[f * 255 + k] = bp[k].count;
        }
    }
}

// One symmetric tree on given gradients (teacher forcing).
// split_* arrays hold max_depth entries; leaves 2^max_depth; leaf_idx n.
int tcref_tree_build(void* h, const float* g, const float* hs, uint32_t max_depth,
                     float lambda_l2, float min_child_weight,
                     uint16_t* split_feature, uint8_t* split_bin, float* split_fthr,
                     float* split_gain, float* leaves, uint32_t* leaf_idx) {
    auto* r = static_cast<RefData*>(h);
    Dataset& ds = *r->ds;
    const uint32_t n = ds.n_samples();
    std::memcpy(ds.gradients(), g, n * sizeof(float));
    std::memcpy(ds.hessians(), hs, n * sizeof(float));
    Config c = make_config(r->cfg.tree.max_bins, max_depth, lambda_l2, min_child_weight, 0);
    std::vector<Index> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    CPUHistogramBuilder b;
    SymmetricTree t(c.tree);
    t.build(ds, idx, b);
    const int depth = t.depth();
    for (int d = 0; d < depth; ++d) {
        split_feature[d] = t.splits()[d].feature;
        split_bin[d] = t.splits()[d].threshold;
        split_fthr[d] = t.splits()[d].float_threshold;
        split_gain[d] = t.splits()[d].gain;
    }
    for (uint32_t k = 0; k < (1u << depth); ++k) leaves[k] = t.leaf_values()[k];
    if (leaf_idx) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t li = 0;
            for (int d = 0; d < depth; ++d) {
                uint8_t bin = ds.binned().column(t.splits()[d].feature)[i];
                li = (li << 1) | ((bin != 255 && bin > t.splits()[d].threshold) ? 1u : 0u);
            }
            leaf_idx[i] = li;
        }
    }
    return depth;
}

// Dataset::random_subsample (src/core/dataset.cpp:552-562) on the handle's
// row count; returns the number of rows written to out.
uint32_t tcref_random_subsample(void* h, float ratio, uint64_t seed, uint32_t* out) {
    auto* r = static_cast<RefData*>(h);
    auto idx = r->ds->random_subsample(ratio, seed);
    std::memcpy(out, idx.data(), idx.size() * sizeof(uint32_t));
    return static_cast<uint32_t>(idx.size());
}

// Dataset::goss_subsample (src/core/dataset.cpp:494-550) on gradients g[n]
// of the handle's rows; returns the number of rows written to out.
uint32_t tcref_goss_subsample(void* h, const float* g, float top_rate, float other_rate,
                              uint64_t seed, uint32_t* out) {
    auto* r = static_cast<RefData*>(h);
    const uint32_t n = r->ds->n_samples();
    std::vector<float> ones(n, 1.0f);
    r->ds->set_gradients_copy(g, ones.data(), n);
    auto s = r->ds->goss_subsample(top_rate, other_rate, seed);
    std::memcpy(out, s.row_indices.data(), s.row_indices.size() * sizeof(uint32_t));
    return static_cast<uint32_t>(s.row_indices.size());
}

// Dataset::mvs_subsample (src/core/dataset.cpp:576-673) on gradients g[n] of
// the handle's rows, on one OpenMP thread (its float score sum is an OpenMP
// reduction: one thread adds in row order); returns the number of rows
// written to out.
static std::vector<Index> mvs_one_thread(const Dataset& ds, float ratio, uint64_t seed) {
    const int threads = omp_get_max_threads();
    omp_set_num_threads(1);
    auto s = ds.mvs_subsample(ratio, seed);
    omp_set_num_threads(threads);
    return s.row_indices;
}

uint32_t tcref_mvs_subsample(void* h, const float* g, float ratio, uint64_t seed, uint32_t* out) {
    auto* r = static_cast<RefData*>(h);
    const uint32_t n = r->ds->n_samples();
    std::vector<float> ones(n, 1.0f);
    r->ds->set_gradients_copy(g, ones.data(), n);
    auto rows = mvs_one_thread(*r->ds, ratio, seed);
    std::memcpy(out, rows.data(), rows.size() * sizeof(uint32_t));
    return static_cast<uint32_t>(rows.size());
}

// Full fit on the dataset (symmetric trees), mirroring
// booster.cpp's loop, row sampling included.  Model arrays: feature/bin/fthr/gain [T*D], depth [T],
// leaves [T*2^D].  train_preds [n] gets the final training predictions.
// seconds[0] = target statistics + init, seconds[1] = boosting loop,
// seconds[2] = time inside CPUHistogramBuilder-driven SymmetricTree::build.
static uint64_t xorshift64s(uint64_t& s) {  // Booster::next_random, booster.cpp:1219-1225
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s * 0x2545F4914F6CDD1DULL;
}

static float fit_impl(void* h, int loss_id, uint32_t n_trees, uint32_t max_depth, float lr,
                      float lambda_l2, float min_child_weight, float subsample, float colsample,
                      bool goss, float top_rate, float other_rate, float mvs_ratio, uint64_t seed,
                      uint16_t* feat, uint8_t* bin, float* fthr, float* gain, uint16_t* depths,
                      float* leaves, float* train_preds, double* seconds) {
    auto* r = static_cast<RefData*>(h);
    Dataset& ds = *r->ds;
    const uint32_t n = ds.n_samples();
    Config c = make_config(r->cfg.tree.max_bins, max_depth, lambda_l2, min_child_weight, loss_id);
    double t0 = now_s();
    auto loss = Loss::create(c.loss, c.task, c.n_classes);
    const float base = loss->init_prediction(ds.labels().data(), n);
    auto hb = HistogramBuilder::create(c.device);
    ds.compute_target_statistics(c);
    double t1 = now_s();
    std::vector<float> preds(n, base), buf(n);
    uint64_t rng = seed;
    std::vector<Index> all(n);
    std::iota(all.begin(), all.end(), 0u);
    const uint32_t L = 1u << max_depth;
    double tb = 0.0;
    for (uint32_t it = 0; it < n_trees; ++it) {
        ds.ensure_gradients_allocated(n);
        loss->compute_gradients(ds.labels().data(), preds.data(), ds.gradients(), ds.hessians(), n);
        // sample selection and feature draw (booster.cpp:198-229)
        std::vector<Index> rows = all;
        if (mvs_ratio > 0.0f) rows = mvs_one_thread(ds, mvs_ratio, xorshift64s(rng));
        else if (goss) rows = ds.goss_subsample(top_rate, other_rate, xorshift64s(rng)).row_indices;
        else if (subsample < 1.0f) rows = ds.random_subsample(subsample, xorshift64s(rng));
        if (colsample < 1.0f) (void)ds.random_feature_subsample(colsample, xorshift64s(rng));
        SymmetricTree t(c.tree);
        double a = now_s();
        t.build(ds, rows, *hb);
        tb += now_s() - a;
        t.predict_batch(ds, buf.data());
        #pragma omp parallel for
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) preds[i] += lr * buf[i];
        depths[it] = t.depth();
        for (int d = 0; d < t.depth(); ++d) {
            feat[it * max_depth + d] = t.splits()[d].feature;
            bin[it * max_depth + d] = t.splits()[d].threshold;
            fthr[it * max_depth + d] = t.splits()[d].float_threshold;
            gain[it * max_depth + d] = t.splits()[d].gain;
        }
        for (uint32_t k = 0; k < (1u << t.depth()); ++k) leaves[it * L + k] = t.leaf_values()[k];
    }
    double t2 = now_s();
    if (train_preds) std::memcpy(train_preds, preds.data(), n * sizeof(float));
    if (seconds) { seconds[0] = t1 - t0; seconds[1] = t2 - t1; seconds[2] = tb; }
    return base;
}

float tcref_fit(void* h, int loss_id, uint32_t n_trees, uint32_t max_depth, float lr,
                float lambda_l2, float min_child_weight, float subsample, float colsample,
                uint64_t seed,
                uint16_t* feat, uint8_t* bin, float* fthr, float* gain, uint16_t* depths,
                float* leaves, float* train_preds, double* seconds) {
    return fit_impl(h, loss_id, n_trees, max_depth, lr, lambda_l2, min_child_weight, subsample,
                    colsample, false, 0.f, 0.f, 0.f, seed, feat, bin, fthr, gain, depths, leaves,
                    train_preds, seconds);
}

// The same fit with GOSS row sampling (booster.cpp:207-212).
float tcref_fit_goss(void* h, int loss_id, uint32_t n_trees, uint32_t max_depth, float lr,
                     float lambda_l2, float min_child_weight, float top_rate, float other_rate,
                     float colsample, uint64_t seed,
                     uint16_t* feat, uint8_t* bin, float* fthr, float* gain, uint16_t* depths,
                     float* leaves, float* train_preds) {
    return fit_impl(h, loss_id, n_trees, max_depth, lr, lambda_l2, min_child_weight, 1.0f,
                    colsample, true, top_rate, other_rate, 0.f, seed, feat, bin, fthr, gain, depths,
                    leaves, train_preds, nullptr);
}

// The same fit with MVS row sampling (booster.cpp:201-205).
float tcref_fit_mvs(void* h, int loss_id, uint32_t n_trees, uint32_t max_depth, float lr,
                    float lambda_l2, float min_child_weight, float ratio, float colsample,
                    uint64_t seed,
                    uint16_t* feat, uint8_t* bin, float* fthr, float* gain, uint16_t* depths,
                    float* leaves, float* train_preds) {
    return fit_impl(h, loss_id, n_trees, max_depth, lr, lambda_l2, min_child_weight, 1.0f,
                    colsample, false, 0.f, 0.f, ratio, seed, feat, bin, fthr, gain, depths,
                    leaves, train_preds, nullptr);
}

// predict_raw_nobinning_fast without the Booster: flat model -> raw scores
// (base added, booster.cpp:921-924).
static void make_ensemble(SymmetricEnsemble& ens, const uint16_t* feat, const float* fthr,
                          const uint8_t* bin, const uint16_t* depths, const float* leaves,
                          uint32_t n_trees, uint32_t max_depth, float lr) {
    const uint32_t L = 1u << max_depth;
    for (uint32_t t = 0; t < n_trees; ++t) {
        auto tree = std::make_unique<SymmetricTree>();
        std::vector<SymmetricSplit> sp(depths[t]);
        for (int d = 0; d < depths[t]; ++d) {
            sp[d].feature = feat[t * max_depth + d];
            sp[d].threshold = bin[t * max_depth + d];
            sp[d].float_threshold = fthr[t * max_depth + d];
        }
        std::vector<float> lv(leaves + t * L, leaves + t * L + (1u << depths[t]));
        tree->set_data(depths[t], std::move(sp), std::move(lv));
        ens.add_tree(std::move(tree), lr);
    }
}

// Booster::predict_raw_nobinning_fast (booster.cpp:899-925) over the
// ensemble: SymmetricEnsemble::predict_batch_raw_fast, then + base.
void tcref_predict(const uint16_t* feat, const float* fthr, const uint8_t* bin, const uint16_t* depths,
                   const float* leaves, uint32_t n_trees, uint32_t max_depth, float lr, float base,
                   const float* X, uint32_t n, uint32_t F, float* out) {
    SymmetricEnsemble ens;
    make_ensemble(ens, feat, fthr, bin, depths, leaves, n_trees, max_depth, lr);
    ens.predict_batch_raw_fast(X, n, static_cast<FeatureIndex>(F), out);
    for (uint32_t i = 0; i < n; ++i) out[i] += base;
}

// Booster::predict_raw (booster.cpp:586-605) over the handle's binned rows:
// SymmetricEnsemble::predict_batch (symmetric_tree.cpp:934-962), then + base.
void tcref_predict_binned(void* h, const uint16_t* feat, const float* fthr, const uint8_t* bin,
                          const uint16_t* depths, const float* leaves, uint32_t n_trees,
                          uint32_t max_depth, float lr, float base, float* out) {
    auto* r = static_cast<RefData*>(h);
    SymmetricEnsemble ens;
    make_ensemble(ens, feat, fthr, bin, depths, leaves, n_trees, max_depth, lr);
    ens.predict_batch(*r->ds, out);
    for (uint32_t i = 0; i < r->ds->n_samples(); ++i) out[i] += base;
}

}  // extern "C"
