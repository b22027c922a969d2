/* Test infrastructure only.  CPU oracle for the TurboCat symmetric-tree hot
 * path: a plain-C restatement of the reference algorithm, in two precisions.
 *
 *   *_f32  follow the reference's FP32 arithmetic and summation order so that
 *          they can be checked bit-for-bit against the reference itself
 *          (oracle/_ref/libtcref.so, built from the reference sources).
 *   *_f64  the same algorithm in double precision: the yardstick for the
 *          1e-5 relative tolerance on sums, leaf values and predictions.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  Every function cites the
 * reference file:line it restates.  Layouts:
 *   X       row-major n x F float32
 *   bins    column-major F x n uint8 (reference BinnedData, dataset.hpp:27-82)
 *   edges   255 float slots per feature, n_edges[f] of them valid
 *   hist    F x 256 x {grad, hess}; slot 255 holds NaN rows (see below)
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TCO_LOSS_MSE = 0, TCO_LOSS_LOGLOSS = 1 };

/* A2: quantile binning.  wrap != 0 reproduces the reference's uint8 cast of
 * the unique count (src/core/dataset.cpp:308); wrap == 0 is the "fixed" mode
 * A = min(max_bins, U).  Returns 0. */
int tco_quantize(const float* X, uint32_t n, uint32_t F, uint32_t max_bins, int wrap,
                 uint8_t* bins, float* edges, uint32_t* n_edges);

/* A3: bin new rows with existing edges (src/core/dataset.cpp:346-379). */
void tco_apply_bins(const float* X, uint32_t n, uint32_t F, const float* edges,
                    const uint32_t* n_edges, uint8_t* bins);

/* A5: base score (src/boosting/loss.cpp:217-226, :364-370). */
float tco_init_prediction(int loss, const float* y, uint32_t n);

/* A6/A7: gradients, FP32 as the reference AVX2 build computes them
 * (src/boosting/loss.cpp:123-163, :348-362). */
void tco_gradients_f32(int loss, const float* y, const float* p, float* g, float* h, uint32_t n);
/* A6/A7 in double precision with an exact exp (yardstick). */
void tco_gradients_f64(int loss, const float* y, const float* p, double* g, double* h, uint32_t n);
/* A8: softmax gradients, K channels, row-major n x K (src/boosting/loss.cpp:273-318). */
void tco_gradients_mc_f32(const float* y, const float* p, float* g, float* h, uint32_t n, uint32_t K);
void tco_gradients_mc_f64(const float* y, const float* p, double* g, double* h, uint32_t n, uint32_t K);
/* Probabilities from raw scores (row-major n x K): the sigmoid for K == 1
 * (loss.cpp:213-215), the softmax for K > 1 (loss.cpp:273-285, :320-331). */
void tco_transform_f32(const float* raw, float* out, uint32_t n, uint32_t K);
void tco_predict_mc_f64(const uint16_t* feature, const float* fthr, const uint16_t* depth,
                        const float* leaves, const float* weight, uint32_t T, uint32_t D,
                        uint32_t K, double base, const float* X, uint32_t n, uint32_t F,
                        double* out, double* abs_sum);

/* A9: histogram of a row list (rows == NULL: all rows in order).
 * f32: sequential row-order sums (src/tree/histogram.cpp:93-133, :307-378).
 * f64: double sums, plus abs_g = sum |g| per slot for the normalized
 * tolerance.  out_count may be NULL. */
void tco_hist_f32(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                  const uint32_t* rows, uint32_t n_rows, float* out_gh, uint32_t* out_count);
void tco_hist_f64(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                  const uint32_t* rows, uint32_t n_rows, double* out_gh, double* abs_g,
                  uint32_t* out_count);

/* A10: level split scan over n_nodes histograms (layout n_nodes x F x 256 x 2)
 * and node totals.  is_root selects the d == 0 rules.  Writes the argmax
 * (feature, bin, gain) with ties to the lowest (feature, bin), and the best
 * gain among candidates with a different (feature, bin) as *second.
 * src/tree/symmetric_tree.cpp:130-262. */
void tco_split_f32(const float* hist, const float* node_g, const float* node_h,
                   const uint32_t* node_count, uint32_t n_nodes, uint32_t F, uint32_t max_bins,
                   float lambda, int is_root, int32_t* feature, int32_t* bin, float* gain);
void tco_split_f64(const double* hist, const double* node_g, const double* node_h,
                   const uint32_t* node_count, uint32_t n_nodes, uint32_t F, uint32_t max_bins,
                   double lambda, int is_root, int32_t* feature, int32_t* bin, double* gain,
                   double* second);

typedef struct {
    uint32_t max_depth;
    uint32_t max_bins;         /* 255 in every reference preset */
    float lambda_l2;
    float lambda_l1;
    float max_delta_step;
    uint32_t min_samples_leaf;
    float min_child_weight;
    float min_split_gain;
} tco_tree_params;

/* A9-A14: one symmetric tree on given gradients (SymmetricTree::build,
 * src/tree/symmetric_tree.cpp:34-368; leaves :753-775).  Returns the depth.
 * root/n_root: the tree's row list in order (the row sample of
 * booster.cpp:198-219); NULL = every row.  leaf_idx covers every row.
 * Per level d < depth: feature[d], bin[d], fthr[d] (A11), gain[d]; for the
 * f64 variant also gap[d] = best - second best candidate gain at that level.
 * leaves: 2^depth values.  leaf_idx (optional): per-row MSB-first index. */
int tco_tree_build_f32(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                       const uint32_t* root, uint32_t n_root,
                       const float* edges, const uint32_t* n_edges, const tco_tree_params* p,
                       uint16_t* feature, uint8_t* bin, float* fthr, float* gain,
                       float* leaves, uint32_t* leaf_idx);
int tco_tree_build_f64(const uint8_t* bins, uint32_t n, uint32_t F, const float* g, const float* h,
                       const uint32_t* root, uint32_t n_root,
                       const float* edges, const uint32_t* n_edges, const tco_tree_params* p,
                       uint16_t* feature, uint8_t* bin, float* fthr, double* gain, double* gap,
                       double* leaves, uint32_t* leaf_idx);

/* A18 (builder-defined, no reference counterpart): K-channel symmetric tree,
 * gradients row-major n x K.  gain = sum over nodes and channels; leaves are
 * 2^depth x K.  K == 1 reduces to tco_tree_build_f64. */
int tco_tree_build_mc_f64(const uint8_t* bins, uint32_t n, uint32_t F, uint32_t K,
                          const double* g, const double* h, const float* edges,
                          const uint32_t* n_edges, const tco_tree_params* p,
                          uint16_t* feature, uint8_t* bin, float* fthr, double* gain,
                          double* gap, double* leaves, uint32_t* leaf_idx);

/* A15: binned per-tree leaf index (src/tree/symmetric_tree.cpp:777-791). */
void tco_leaf_index_binned(const uint8_t* bins, uint32_t n, uint32_t depth,
                           const uint16_t* feature, const uint8_t* bin, uint32_t* leaf_idx);

/* A16: raw-float ensemble prediction.  Model arrays: feature/fthr [T*D],
 * depth [T], leaves [T * 2^D], weight [T].  out = sum_t w_t * leaf + base.
 * f32: the reference's order for a call of n rows -- rows of the full 8-row
 * blocks sum groups of four trees w0*l0 + w1*l1 + w2*l2 + w3*l3, then add
 * them, remaining trees fused in one by one
 * (include/turbocat/fast_float_ensemble.hpp:388-541); the last n mod 8 rows
 * (and every row when no tree has a split) fuse w_t * l_t into the sum tree
 * by tree (:279-287, :543-560).  f64: double sum.
 * leaf_idx (optional): n x T indices. */
void tco_predict_f32(const uint16_t* feature, const float* fthr, const uint16_t* depth,
                     const float* leaves, const float* weight, uint32_t T, uint32_t D,
                     float base, const float* X, uint32_t n, uint32_t F, float* out,
                     uint32_t* leaf_idx);
/* A17: binned ensemble prediction over column-major bins (F x n),
 * SymmetricEnsemble::predict_batch -> FastEnsemble::predict_batch_column_major
 * (src/tree/symmetric_tree.cpp:934-962, include/turbocat/fast_ensemble.hpp:
 * 256-446) + base (booster.cpp:599-604).  Same summation orders as
 * tco_predict_f32, except that a call of fewer than 16 rows takes the
 * tree-by-tree order for every row (symmetric_tree.cpp:954-961). */
void tco_predict_binned_f32(const uint16_t* feature, const uint8_t* bin, const uint16_t* depth,
                            const float* leaves, const float* weight, uint32_t T, uint32_t D,
                            float base, const uint8_t* bins, uint32_t n, uint32_t F, float* out);
void tco_predict_f64(const uint16_t* feature, const float* fthr, const uint16_t* depth,
                     const float* leaves, const float* weight, uint32_t T, uint32_t D,
                     double base, const float* X, uint32_t n, uint32_t F, double* out,
                     double* abs_sum);

/* preds[i] = fma(lr, vals[i], preds[i]): the training-prediction update as
 * the reference compiler contracts it (booster.cpp:460-463). */
void tco_fma_update(float* preds, const float* vals, float lr, uint32_t n);

/* A5-A15: the whole boosting loop in FP32 reference order (booster.cpp:92-465),
 * bins already computed.  Returns the base score.  train_preds: n floats. */
float tco_fit_f32(const uint8_t* bins, uint32_t n, uint32_t F, const float* edges,
                  const uint32_t* n_edges, const float* y, int loss, uint32_t n_trees,
                  float lr, const tco_tree_params* p, uint16_t* feature, uint8_t* bin,
                  float* fthr, float* gain, uint16_t* depth, float* leaves, float* train_preds);

/* A19: per-tree losses and the early-stopping rule (booster.cpp:252-352). */
float tco_compute_loss_f32(int loss, const float* y, const float* p, uint32_t n);
double tco_compute_loss_f64(int loss, const float* y, const float* p, uint32_t n);
float tco_compute_loss_mc_f32(const float* y, const float* p, uint32_t n, uint32_t K);
uint32_t tco_early_stop(const float* metric, uint32_t n_iter, uint32_t n_estimators,
                        uint32_t rounds, int higher_is_better, uint32_t* best_iter, float* best);


/* Regular (non-oblivious) depthwise trees, Tree::build (src/tree/tree.cpp:
 * 162-217, :297-390) with SplitFinder (src/tree/histogram.cpp:569-703).
 * Parameters of TreeConfig (include/turbocat/config.hpp:26-59). */
typedef struct {
    uint32_t max_depth;
    uint32_t max_leaves;
    uint32_t max_bins;              /* histogram stride: candidates b < max_bins - 1 */
    uint32_t min_samples_leaf;
    float min_child_weight;
    float lambda_l2;
    float min_split_gain;
    float gamma;                    /* subtracted from every candidate's gain */
    float leaf_smooth;              /* 0: off */
    float max_delta_step;           /* 0: off */
    int32_t learn_missing_direction;
    int32_t grow_policy;            /* 0 depthwise (Tree::build), 1 lossguide (build_leafwise) */
} tco_rtree_params;

/* One regular tree on given gradients, FP64 sums, nodes in the reference's
 * numbering (add_node order: a split node's two children get consecutive
 * ids; depthwise: the smaller child's subtree is built first; lossguide: in
 * the order the priority queue expands the nodes).  flags: bit 0 is_leaf,
 * bit 1 default_left.  gap[i] = best - second best valid gain at split node
 * i (+inf if it had one candidate; 0 when the best or a better candidate has
 * a side whose hessian is min_child_weight within float rounding).  Lossguide also reports hgap: the
 * smallest gain difference between two queued candidates when one of them
 * was popped (+inf without such a pair), the margin of the expansion order.
 * Returns 0, or -1 if more than cap nodes. */
int tco_rtree_build_f64(const uint8_t* bins, uint32_t n, uint32_t F, const float* g,
                        const float* h, const uint32_t* root, uint32_t n_root,
                        const tco_rtree_params* p, uint32_t cap, uint16_t* feature,
                        uint8_t* bin, uint8_t* flags, uint32_t* left, uint32_t* right,
                        double* value, double* gain, double* gap, uint32_t* n_nodes,
                        uint32_t* n_leaves, uint16_t* depth, double* hgap);
/* build_leafwise's candidate queue alone (std::priority_queue by gain):
 * ops >= 0 push that item, < 0 pop; popped items in order; returns the pops. */
uint32_t tco_rqueue_order(const double* gains, const int32_t* ops, uint32_t n_ops,
                          uint32_t* popped);
/* Tree::predict(dataset, row) over column-major bins: leaf value per row. */
void tco_rtree_apply_binned(const uint16_t* feature, const uint8_t* bin, const uint8_t* flags,
                            const uint32_t* left, const uint32_t* right, const float* value,
                            uint32_t n_nodes, const uint8_t* bins, uint32_t n, float* out);
/* Regular-tree ensemble prediction from float rows (rule 0: apply_bins, the
 * multiclass path; rule 1: predict_raw_fast, the binary/regression path);
 * out n x K. */
void tco_rpredict_f32(const uint32_t* tree_off, const uint16_t* feature, const uint8_t* bin,
                      const uint8_t* flags, const uint32_t* left, const uint32_t* right,
                      const float* value, const float* weight, const uint32_t* tree_class,
                      uint32_t T, uint32_t K, float base, const float* X, uint32_t n, uint32_t F,
                      const float* edges, const uint32_t* n_edges, int rule, float* out);

#ifdef __cplusplus
}
#endif
#endif
