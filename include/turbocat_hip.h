/* turbocat_hip.h -- C-ABI of the MI355X (gfx950) TurboCat hot path:
 * quantization, gradients, histograms, oblivious split scoring, tree build,
 * boosting loop, ensemble prediction, continuous-batching serving and batched
 * multi-job training for symmetric trees; regular depthwise trees (build,
 * boosting incl. the reference's multiclass loop, prediction).
 *
 * Conventions
 *   - Every function returns int32_t: TC_OK (0) or a negative TC_E* code;
 *     tc_last_error() gives the message of the calling thread's last error.
 *   - Host pointers are borrowed for the duration of the call; functions whose
 *     name ends in _dev take device pointers (memory from tc_device_alloc or
 *     any hipMalloc) and run asynchronously on the context's stream.
 *   - Device memory is owned by handles and released by their *_destroy.
 *   - One host thread at a time per context (the reference Booster is not
 *     re-entrant either); the server runs its own thread and streams.
 *   - Layouts: X row-major n x F float32; bins column-major F x n uint8;
 *     edges 255 float slots per feature (n_edges[f] valid); histograms
 *     F x 256 x {grad, hess} float32 with slot 255 holding the NaN bin.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   Dataset::compute_bins / apply_bins     include/turbocat/dataset.hpp:134,137
 *   Loss::compute_gradients               include/turbocat/loss.hpp:57-63
 *   HistogramBuilder::build / create      include/turbocat/histogram.hpp:72-119
 *   Histogram::subtract_from              src/tree/histogram.cpp:39-42
 *   SymmetricTree::build (split scan)     include/turbocat/symmetric_tree.hpp:49-53
 *   Booster::train (symmetric branch)     src/boosting/booster.cpp:47-418
 *   SymmetricEnsemble::predict_batch_raw_fast  include/turbocat/symmetric_tree.hpp:177
 *   SymmetricEnsemble::predict_batch (binned)  include/turbocat/symmetric_tree.hpp:171
 *   Tree::build (Depthwise)               src/tree/tree.cpp:162-217
 *   Booster::train (regular branch), Booster::train_multiclass
 *                                         src/boosting/booster.cpp:438-493, :1263-1400
 *   TreeEnsemble::predict_batch_optimized / _multiclass_optimized
 *                                         src/tree/tree.cpp:1471-1709
 */
#ifndef TURBOCAT_HIP_H
#define TURBOCAT_HIP_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_OK 0
#define TC_EINVAL (-1)   /* bad argument (maps to std::invalid_argument) */
#define TC_ENOMEM (-2)   /* device or host allocation failed */
#define TC_EDEVICE (-3)  /* HIP runtime error or no usable gfx950 device */
#define TC_ESTATE (-4)   /* call not valid in the object's state (e.g. not fitted) */
#define TC_EBUSY (-5)    /* serving: admission refused, retry later */
#define TC_ETIMEOUT (-6) /* serving: wait timed out */

#define TC_LOSS_MSE 0
#define TC_LOSS_LOGLOSS 1
/* K-class softmax (CrossEntropyLoss, loss.cpp:269-330) with K-channel
 * symmetric trees: labels are class indices 0..K-1. */
#define TC_LOSS_SOFTMAX 2

typedef struct tc_ctx tc_ctx;
typedef struct tc_dataset tc_dataset;
typedef struct tc_ensemble tc_ensemble;
typedef struct tc_server tc_server;

/* Tree hyper-parameters (TreeConfig, include/turbocat/config.hpp:26-59). */
typedef struct {
  uint32_t max_depth;        /* 1..10 */
  uint32_t max_bins;         /* histogram stride / split candidates b < max_bins-1; 255 */
  float lambda_l2;
  float lambda_l1;
  float max_delta_step;
  uint32_t min_samples_leaf;
  float min_child_weight;
  float min_split_gain;
} tc_tree_params;

/* Boosting parameters (BoostingConfig, config.hpp:66-96; symmetric trees).
 * Row sampling follows Booster::train (booster.cpp:198-229): per tree, when
 * subsample < 1 an xorshift64* draw seeds Dataset::random_subsample
 * (dataset.cpp:552-562: std::mt19937_64 + std::shuffle of 0..N-1, first
 * uint32(N * subsample) kept, in shuffled order); when colsample_bytree < 1
 * one more draw is consumed (symmetric trees ignore the feature subset).
 * With use_goss (booster.cpp:207-212, single channel only) the draw seeds
 * Dataset::goss_subsample on the current gradients instead (tc_goss_subsample)
 * and subsample is not used.  use_mvs (booster.cpp:201-205, single channel
 * only) takes precedence over both: the draw seeds Dataset::mvs_subsample on
 * the current gradients (tc_mvs_subsample), a sample whose size varies from
 * tree to tree. */
typedef struct {
  int32_t loss;              /* TC_LOSS_MSE or TC_LOSS_LOGLOSS */
  uint32_t n_trees;
  float learning_rate;
  tc_tree_params tree;
  uint32_t hist_chunk;       /* rows per histogram work item; 0 = autotuned (tc_trainer_hist_chunks) */
  int32_t use_graph;         /* 1: replay one captured hipGraph per tree */
  float subsample;           /* in (0, 1): row sample per tree; 0 or >= 1: every row */
  float colsample_bytree;    /* in (0, 1): consumes an RNG draw per tree */
  uint64_t seed;             /* Config::seed (config.hpp:235) */
  uint32_t n_classes;        /* TC_LOSS_SOFTMAX: K in [2, 16], (2^max_depth) * K <= 8192;
                                otherwise 0 or 1 */
  uint32_t hist_deep_items;  /* with hist_chunk != 0: 0 = hist_chunk rows per work item at
                                every level, else hist_chunk rows at the root and each deeper
                                level cut into at most about this many items (<= 65536);
                                ignored when autotuned.  Never changes the model. */
  int32_t use_goss;          /* BoostingConfig::use_goss (config.hpp:83) */
  float goss_top_rate;       /* in [0, 1]: rows kept by largest |g| (config.hpp:84) */
  float goss_other_rate;     /* in [0, 1]: share of rows drawn from the rest (config.hpp:85) */
  int32_t use_mvs;           /* BoostingConfig::use_mvs (config.hpp:88) */
  float mvs_subsample;       /* in (0, 1]: MVS target share of rows (config.hpp:89) */
} tc_train_params;

/* Host-side model arrays filled by training (caller-allocated).
 * feature/bin/fthr/gain: n_trees * max_depth; depth: n_trees;
 * leaves: (n_trees << max_depth) * K, [tree][leaf][class];
 * train_preds (optional): K x n, class-major.  K = 1 except TC_LOSS_SOFTMAX. */
typedef struct {
  uint16_t* feature;
  uint8_t* bin;
  float* fthr;
  float* gain;
  uint16_t* depth;
  float* leaves;
  float* train_preds;
  float base;                /* out: base score (A5) */
} tc_model;

/* ---------------------------------------------------------------- context */
int32_t tc_ctx_create(int32_t device_id, tc_ctx** out);
int32_t tc_ctx_destroy(tc_ctx* ctx);
int32_t tc_ctx_synchronize(tc_ctx* ctx);
const char* tc_last_error(void);
/* Device memory plumbing for the _dev entry points. */
int32_t tc_device_alloc(tc_ctx* ctx, size_t bytes, void** out);
int32_t tc_device_free(tc_ctx* ctx, void* p);
int32_t tc_memcpy_h2d(tc_ctx* ctx, void* dst, const void* src, size_t bytes);
int32_t tc_memcpy_d2h(tc_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------- quantization */
/* A2: Dataset::compute_bins -> compute_quantile_bins (src/core/dataset.cpp:
 * 176-334).  Computes per-feature edges (wrap != 0 reproduces the reference's
 * uint8 wrap of the unique count, dataset.cpp:308) and bins every row on the
 * GPU.  Returns a device-resident dataset; edges/n_edges copies optional. */
int32_t tc_quantize_fit(tc_ctx* ctx, const float* X, uint32_t n, uint32_t F, uint32_t max_bins,
                        int32_t wrap, float* edges_out, uint32_t* n_edges_out, tc_dataset** out);
/* A3: Dataset::apply_bins (dataset.cpp:346-379): bin X with given edges. */
int32_t tc_dataset_create(tc_ctx* ctx, const float* X, uint32_t n, uint32_t F,
                          const float* edges, const uint32_t* n_edges, uint32_t max_bins,
                          tc_dataset** out);
int32_t tc_dataset_bins(tc_dataset* ds, uint8_t* bins_out /* F x n column-major */);
int32_t tc_dataset_shape(tc_dataset* ds, uint32_t* n, uint32_t* F);
int32_t tc_dataset_destroy(tc_dataset* ds);

/* ------------------------------------------------------------- gradients */
/* A6/A7: Loss::compute_gradients (loss.hpp:57-63; loss.cpp:123-211, :348-362). */
int32_t tc_gradients(tc_ctx* ctx, int32_t loss, const float* y, const float* preds, float* g,
                     float* h, uint32_t n);
/* A8: CrossEntropyLoss::compute_multiclass_gradients (loss.cpp:287-318):
 * softmax of each row's K raw scores; preds, g, h row-major n x K. */
int32_t tc_gradients_multi(tc_ctx* ctx, const float* y, const float* preds, float* g, float* h,
                           uint32_t n, uint32_t K);
/* Row sample of one tree: Dataset::random_subsample(ratio, seed)
 * (dataset.cpp:552-562).  out needs n slots; *n_out = uint32(n * ratio). */
int32_t tc_random_subsample(uint32_t n, float ratio, uint64_t seed, uint32_t* out,
                            uint32_t* n_out);
/* Row sample of one tree under GOSS: Dataset::goss_subsample(top_rate,
 * other_rate, seed) (dataset.cpp:494-550) on gradients g[n]: the
 * uint32(n * top_rate) rows of largest |g| (std::nth_element order), then
 * min(uint32(n * other_rate), n - top) rows of the rest by a partial
 * Fisher-Yates draw (std::mt19937_64).  out needs n slots. */
int32_t tc_goss_subsample(const float* g, uint32_t n, float top_rate, float other_rate,
                          uint64_t seed, uint32_t* out, uint32_t* n_out);
/* Row sample of one tree under MVS: Dataset::mvs_subsample(ratio, seed)
 * (dataset.cpp:576-673) on gradients g[n]: each row kept with probability
 * min(1, s * max(|g|, t) / S * n) by a std::mt19937_64 uniform float draw,
 * t a floor at the (1 - ratio) quantile of |g|, topped up by a shuffled draw
 * of the rest when fewer than 80% of uint32(n * ratio) rows were kept.  The
 * float sum S is taken in row order (the reference's OpenMP reduction on one
 * thread).  out needs n slots; *n_out varies with the gradients. */
int32_t tc_mvs_subsample(const float* g, uint32_t n, float ratio, uint64_t seed, uint32_t* out,
                         uint32_t* n_out);
/* A5: Loss::init_prediction (loss.cpp:217-226, :364-370). */
int32_t tc_init_prediction(int32_t loss, const float* y, uint32_t n, float* base);

/* ------------------------------------------------------------ histograms */
/* A9: HistogramBuilder::build(dataset, rows, {}, out) (histogram.hpp:78-83;
 * histogram.cpp:62-205).  rows == NULL: all rows.  hist_out: F x 256 x 2
 * (grad, hess); count_out (optional): F x 256 row counts, the
 * GradientPair::count the reference's SplitFinder reads (types.hpp:123-146,
 * histogram.cpp:705-758), exact. */
int32_t tc_hist_build(tc_dataset* ds, const float* g, const float* h, const uint32_t* rows,
                      uint32_t n_rows, float* hist_out, uint32_t* count_out);
/* A13: Histogram::subtract_from(parent, other) (histogram.cpp:39-42). */
int32_t tc_hist_subtract(tc_ctx* ctx, const float* parent, const float* child, float* out,
                         uint32_t F);

/* --------------------------------------------------------- split scoring */
/* A10: the level split scan inlined in SymmetricTree::build
 * (symmetric_tree.cpp:130-262) over n_nodes node histograms (n_nodes x F x
 * 256 x 2) and node totals.  is_root selects the d == 0 rules.  Ties go to
 * the lowest (feature, bin); feature = -1 when no candidate is valid. */
int32_t tc_split_scan(tc_ctx* ctx, const float* hist, const float* node_g, const float* node_h,
                      const uint32_t* node_count, uint32_t n_nodes, uint32_t F,
                      uint32_t max_bins, float lambda_l2, int32_t is_root, int32_t* feature,
                      int32_t* bin, float* gain);

/* ------------------------------------------------------------ tree build */
/* A9-A14: SymmetricTree::build on given gradients (symmetric_tree.hpp:49-53).
 * Outputs per level d < *depth: feature, bin, fthr, gain; leaves 2^depth;
 * leaf_idx (optional) per row, MSB-first. */
int32_t tc_tree_build(tc_dataset* ds, const float* g, const float* h, const tc_tree_params* p,
                      uint16_t* feature, uint8_t* bin, float* fthr, float* gain, float* leaves,
                      uint32_t* leaf_idx, int32_t* depth);

/* A18: K-channel symmetric tree on given gradients (g, h class-major K x n):
 * one shared (feature, bin) per level whose gain is the sum over nodes and
 * classes of the per-class gains; leaves 2^depth x K.  K = 1 is exactly
 * tc_tree_build.  No reference counterpart (the reference's multiclass
 * grows regular per-class trees, booster.cpp:1367-1395). */
int32_t tc_tree_build_multi(tc_dataset* ds, const float* g, const float* h, uint32_t K,
                            const tc_tree_params* p, uint16_t* feature, uint8_t* bin, float* fthr,
                            float* gain, float* leaves, uint32_t* leaf_idx, int32_t* depth);

/* ------------------------------------------------------------- boosting */
/* A5-A15: Booster::train, symmetric branch without sampling
 * (booster.cpp:47-418, build_tree :438-465). */
int32_t tc_train(tc_dataset* ds, const float* y, const tc_train_params* p, tc_model* model);
/* A19: validation set and early stopping (booster.cpp:252-371), the metric
 * being the loss.  After every tree the validation rows (binned with the
 * training edges, tc_dataset_create) get preds += lr * tree value and the mean
 * loss (Loss::compute_loss: MSE or LogLoss) is taken; the fit stops after the
 * tree `iter` at which iter >= max(50, n_trees / 10) and early_stopping_rounds
 * trees in a row did not improve the best loss by more than 1e-10.  Trees
 * after the stop stay zero in the model; the reference keeps every tree up to
 * and including the stopping one, as here.  valid may be NULL (then only the
 * training losses are recorded, if asked for).  history: outputs; its
 * train_loss / valid_loss arrays (n_trees floats, optional) receive the mean
 * losses after each tree. */
typedef struct {
  float* train_loss;          /* in: n_trees floats or NULL */
  float* valid_loss;          /* in: n_trees floats or NULL */
  uint32_t n_trees_built;     /* out: trees built (the stopping tree included) */
  int32_t stopped_early;      /* out: 1 when early stopping ended the fit */
  uint32_t best_iteration;    /* out: tree of the best validation loss */
  float best_valid_loss;      /* out */
} tc_train_history;
int32_t tc_train_valid(tc_dataset* ds, const float* y, const tc_train_params* p, tc_dataset* valid,
                       const float* y_valid, uint32_t early_stopping_rounds, tc_model* model,
                       tc_train_history* history);
/* Batched form: n_jobs independent fits in lockstep, one launch per kernel
 * for all jobs.  All jobs share max_depth and n_trees; y[j], models[j] per
 * job.  No reference counterpart (the reference fits one model per call). */
int32_t tc_train_batch(tc_ctx* ctx, uint32_t n_jobs, tc_dataset* const* ds,
                       const float* const* y, const tc_train_params* p, tc_model* models);
/* Timing hook for benchmarks: time (ms, device events) of the boosting loop
 * of the last tc_train / tc_train_batch on this context, and of its
 * histogram kernels. */
int32_t tc_last_train_timing(tc_ctx* ctx, double* loop_ms, double* hist_ms);
/* Rows fed to histogram launches by the last regular-tree fit on this
 * context (every tree of every class), and those launches' count: the
 * algorithmic bytes are rows x (F + 8).  No reference counterpart. */
int32_t tc_last_train_hist_rows(tc_ctx* ctx, uint64_t* rows, uint32_t* launches);

/* Stepwise form of tc_train / tc_train_batch for callers that drive the
 * boosting loop themselves (benchmarks, schedulers): create once, then each
 * tc_trainer_step enqueues n more trees for every job (asynchronously, one
 * captured hipGraph replay per tree); tc_trainer_sync waits.  hist_ms /
 * hist_launches of tc_trainer_step_timed report HIP-event time of every
 * histogram kernel launch of those trees (graph not used).  No reference
 * counterpart: the reference's Booster::train loop (booster.cpp:160-408)
 * runs to completion in one call. */
typedef struct tc_trainer tc_trainer;
int32_t tc_trainer_create(tc_ctx* ctx, uint32_t n_jobs, tc_dataset* const* ds,
                          const float* const* y, const tc_train_params* p, tc_trainer** out);
int32_t tc_trainer_step(tc_trainer* tr, uint32_t n_trees);
int32_t tc_trainer_step_timed(tc_trainer* tr, uint32_t n_trees, double* hist_ms,
                              uint32_t* hist_launches);
int32_t tc_trainer_sync(tc_trainer* tr);
/* Model of job j so far (trees beyond the ones built are zero). */
int32_t tc_trainer_model(tc_trainer* tr, uint32_t job, tc_model* model);
/* Rows fed to histogram kernels for job j since creation (for algorithmic
 * bytes: rows x (F + 8)). */
int32_t tc_trainer_hist_rows(tc_trainer* tr, uint32_t job, uint64_t* rows);
/* Histogram work-item shape of job j: rows per work item at the root, and
 * the number of work items each deeper level is cut into (the rows per item
 * of a level then follow that level's rows; 0 = hist_chunk rows at every
 * level).  With hist_chunk = 0 the shape is picked by timing candidate
 * shapes on the fit's first trees before the per-tree graph is captured
 * (every histogram is an exact integer sum, so the shape never changes the
 * model); otherwise root_rows = hist_chunk and deep_items = hist_deep_items.
 * No reference counterpart. */
int32_t tc_trainer_hist_chunks(tc_trainer* tr, uint32_t job, uint32_t* root_rows,
                               uint32_t* deep_items);
int32_t tc_trainer_destroy(tc_trainer* tr);

/* ------------------------------------------------------------- inference */
/* A16/A17 model upload: FastFloatEnsemble::from_symmetric_ensemble
 * (fast_float_ensemble.cpp:23-68).  Arrays as tc_model; weights per tree. */
int32_t tc_ensemble_create(tc_ctx* ctx, uint32_t n_trees, uint32_t max_depth,
                           const uint16_t* feature, const float* fthr, const uint8_t* bin,
                           const uint16_t* depth, const float* leaves, const float* weights,
                           float base, uint32_t n_features, tc_ensemble** out);
/* A18 inference: K-channel ensemble, leaves [tree][leaf][class].  tc_predict,
 * tc_predict_dev and tc_predict_binned then write n x K (row-major) scores;
 * transform 1 applies softmax (CrossEntropyLoss::transform_to_proba,
 * loss.cpp:320-330).  No reference counterpart. */
int32_t tc_ensemble_create_multi(tc_ctx* ctx, uint32_t n_trees, uint32_t max_depth,
                                 uint32_t n_classes, const uint16_t* feature, const float* fthr,
                                 const uint8_t* bin, const uint16_t* depth, const float* leaves,
                                 const float* weights, float base, uint32_t n_features,
                                 tc_ensemble** out);
int32_t tc_ensemble_destroy(tc_ensemble* ens);
/* A16: predict_batch_raw_fast + base (booster.cpp:899-925).  transform: 0 raw,
 * 1 sigmoid (LogLoss::transform_prediction, loss.cpp:213-215).  leaf_idx
 * (optional) n x T. */
int32_t tc_predict(tc_ensemble* ens, const float* X, uint32_t n, uint32_t F, int32_t transform,
                   float* out, uint32_t* leaf_idx);
int32_t tc_predict_dev(tc_ensemble* ens, const float* X_dev, uint32_t n, uint32_t F,
                       int32_t transform, float* out_dev);
/* tc_predict without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc): the rows and the outputs are staged in a fixed
 * 64 MiB device scratch that the library holds on each GPU until the process
 * ends, so 4 * n * F (rounded up to 256) + 4 * n * K must be at most 64 MiB
 * (TC_EINVAL otherwise).  Every call on one GPU shares that scratch: calls
 * that overlap on one GPU (several threads, or several contexts on it) race.
 * Same results as tc_predict (transform 0 raw, 1 sigmoid / softmax). */
int32_t tc_predict_noalloc(tc_ensemble* ens, const float* X, uint32_t n, uint32_t F,
                           int32_t transform, float* out);
/* A17: binned prediction (SymmetricEnsemble::predict_batch, symmetric_tree.cpp:
 * 934-962) over a device dataset's bins. */
int32_t tc_predict_binned(tc_ensemble* ens, tc_dataset* ds, float* out);

/* --------------------------------------------------------- regular trees */
/* Regular (non-oblivious) depthwise trees: TreeConfig (include/turbocat/
 * config.hpp:26-59) as Tree::build reads it (src/tree/tree.cpp:162-217,
 * :297-390; SplitFinder, src/tree/histogram.cpp:569-761).  Every node picks
 * its own (feature, bin); a node is a leaf at max_depth, under 2 *
 * min_samples_leaf rows, under 2 * min_child_weight hessian, without a
 * candidate whose sides both have >= min_samples_leaf rows and >=
 * min_child_weight hessian and whose gain 0.5 (GL^2/(HL+l) + GR^2/(HR+l) -
 * G^2/(H+l)) - gamma is >= min_split_gain, or when max_leaves leaves exist at
 * its visit in the reference's depth-first, smaller-child-first order. */
typedef struct {
  uint32_t max_depth;          /* 1..12 */
  uint32_t max_leaves;         /* TreeConfig::max_leaves (config.hpp:29), >= 1 */
  uint32_t max_bins;           /* candidates b < max_bins - 1; 255 */
  float lambda_l2;
  float min_child_weight;
  uint32_t min_samples_leaf;
  float min_split_gain;
  float gamma;                 /* TreeConfig::gamma (config.hpp:37) */
  float max_delta_step;        /* > 0: leaf values clamped (tree.cpp:675-678) */
  float leaf_smooth;           /* > 0: leaf value * count / (count + leaf_smooth) */
  int32_t learn_missing_direction;  /* TreeConfig::learn_missing_direction (config.hpp:52) */
  int32_t grow_policy;         /* TreeConfig::grow_policy (config.hpp:21-32): 0 depthwise
                                  (Tree::build), 1 lossguide (Tree::build_leafwise,
                                  tree.cpp:685-925; max_depth 1..11) */
} tc_rtree_params;

/* Boosting with regular trees: Booster::train's regular branch (booster.cpp:
 * 160-232, build_tree :438-493) for TC_LOSS_MSE / TC_LOSS_LOGLOSS, and
 * Booster::train_multiclass (booster.cpp:1263-1400) for TC_LOSS_SOFTMAX: per
 * iteration the softmax gradients of every class, then one regular tree per
 * class on that class's gradients over the iteration's shared row sample.
 * Sampling as tc_train_params (the row sample's draw, then the colsample
 * draw): MVS, else GOSS, else the plain row sample (booster.cpp:198-219); a
 * softmax fit takes GOSS on each row's largest |g| over the classes, else
 * the plain row sample, and no MVS (booster.cpp:1326-1354). */
typedef struct {
  int32_t loss;                /* TC_LOSS_MSE, TC_LOSS_LOGLOSS or TC_LOSS_SOFTMAX */
  uint32_t n_iterations;       /* BoostingConfig::n_estimators; trees = n_iterations * K */
  float learning_rate;
  tc_rtree_params tree;
  float subsample;             /* in (0, 1): row sample per iteration; 0 or >= 1: every row */
  float colsample_bytree;      /* in (0, 1): consumes an RNG draw per iteration (trees use every feature) */
  uint64_t seed;
  uint32_t n_classes;          /* TC_LOSS_SOFTMAX: K in [3, 16]; otherwise 0 or 1 */
  int32_t use_graph;           /* 1: one captured hipGraph replay per iteration */
  int32_t time_hist;           /* 1: plain launches with HIP events around every histogram
                                  launch (tc_last_train_timing, tc_last_train_hist_rows) */
  int32_t use_goss;            /* GOSS row sampling (BoostingConfig, config.hpp:83-85) */
  float goss_top_rate, goss_other_rate;
  int32_t use_mvs;             /* MVS row sampling, before GOSS (config.hpp:88-89; binary and
                                  regression fits only, as in the reference) */
  float mvs_subsample;
} tc_rtrain_params;

/* Regular trees in the reference's numbering (Tree::nodes in add_node order;
 * a split node's children are consecutive ids, tree.cpp:357-363), tree after
 * tree in slots of node_cap = 2^(max_depth+1) - 1 nodes: node k of tree t at
 * t * node_cap + k.  flags: bit 0 is_leaf, bit 1 default_left (as Tree::save,
 * tree.cpp:1768).  Trees are iteration-major, class-minor (the order of
 * TreeEnsemble::add_tree_for_class, booster.cpp:1386). */
typedef struct {
  uint16_t* feature;
  uint8_t* bin;
  uint8_t* flags;
  uint32_t* left;
  uint32_t* right;
  float* value;
  float* gain;
  uint32_t* n_nodes;           /* per tree */
  uint32_t* n_leaves;          /* per tree */
  uint16_t* depth;             /* per tree */
  float* train_preds;          /* optional: K x n raw training predictions, class-major */
  float base;                  /* out: base score (0 for softmax) */
} tc_rmodel;

/* One regular tree on given gradients (Tree::build, tree.cpp:162-217) over the
 * rows `rows` (NULL: all rows; the root's row sample otherwise); h >= 0.
 * out: one tree (node_cap slots). */
int32_t tc_rtree_build(tc_dataset* ds, const float* g, const float* h, const uint32_t* rows,
                       uint32_t n_rows, const tc_rtree_params* p, tc_rmodel* out);
/* Boosting with regular trees (see tc_rtrain_params). */
int32_t tc_rtrain(tc_dataset* ds, const float* y, const tc_rtrain_params* p, tc_rmodel* out);
/* With a validation set and early stopping: after every iteration the new
 * trees are applied to the validation rows (binned with the training edges,
 * tc_dataset_create; NaN by each node's default_left), their mean loss taken
 * (MSE, LogLoss, or softmax cross-entropy -log max(p_label, 1e-7)), and the
 * fit stops after the iteration `iter` at which iter >= max(50, n_iterations
 * / 10) and early_stopping_rounds iterations in a row did not improve the
 * best loss by more than 1e-10 (booster.cpp:306-371, :1400-1430).  Trees of
 * iterations after the stop have no nodes; history->n_trees_built counts
 * iterations.  history may be NULL; its train_loss is zero-filled. */
int32_t tc_rtrain_valid(tc_dataset* ds, const float* y, const tc_rtrain_params* p,
                        tc_dataset* valid, const float* y_valid, uint32_t early_stopping_rounds,
                        tc_rmodel* out, tc_train_history* history);

/* Regular-tree ensemble prediction: TreeEnsemble::predict_batch_optimized /
 * predict_batch_multiclass_optimized (tree.cpp:1471-1709) over the bins of
 * rule 0 (Dataset::apply_bins: upper_bound, the multiclass path of
 * bindings.cpp:311-327) or rule 1 (Booster::predict_raw_fast, booster.cpp:
 * 680-871, the binary/regression path: up to 16 edges the count of edges <= v,
 * more the count of edges < v); a row walks every tree (bin != 255 && bin >
 * split bin: right, so NaN goes left) and adds weight * value into its
 * class's sum in tree order, then + base[class].  Trees: nodes
 * tree_off[t] .. tree_off[t+1]-1 with tree-local child ids; tree_class NULL:
 * every tree adds to class 0.  edges: 255 slots per feature. */
typedef struct tc_rensemble tc_rensemble;
int32_t tc_rensemble_create(tc_ctx* ctx, uint32_t n_trees, uint32_t n_classes,
                            const uint32_t* tree_off, const uint16_t* feature, const uint8_t* bin,
                            const uint8_t* flags, const uint32_t* left, const uint32_t* right,
                            const float* value, const float* weights, const uint32_t* tree_class,
                            const float* base, const float* edges, const uint32_t* n_edges,
                            uint32_t n_features, int32_t rule, tc_rensemble** out);
int32_t tc_rensemble_destroy(tc_rensemble* ens);
/* out: n x K; transform 0 raw, 1 sigmoid (K = 1, loss.cpp:213-215) or softmax
 * (K > 1, CrossEntropyLoss::transform_to_proba, loss.cpp:320-330). */
int32_t tc_rpredict(tc_rensemble* ens, const float* X, uint32_t n, uint32_t F, int32_t transform,
                    float* out);
int32_t tc_rpredict_dev(tc_rensemble* ens, const float* X_dev, uint32_t n, uint32_t F,
                        int32_t transform, float* out_dev);

/* --------------------------------------------------------------- serving */
/* Continuous-batching server over a resident ensemble (single-output or
 * K-channel).  No reference counterpart.  A serving kernel stays resident
 * while batches keep coming and takes them from a ring of descriptors in
 * uncached pinned host memory: no launch and no copy call per batch.  It ends
 * after 0.5 ms without a batch or 10 ms of running (HIP may put other
 * streams' work on its hardware queue, which runs in order), and the server
 * launches it again when a batch waits.  While a server runs, the library's
 * own frees are deferred to its stop (hipFree waits for an idle device); a
 * caller's device-wide wait (hipDeviceSynchronize, hipFree) may wait up to
 * one time slice.  Requests are admitted while their rows (4F + 4K bytes each) fit
 * memory_budget (a request larger than the whole budget is admitted alone);
 * admitted requests are packed, whole or in pieces, into batches of up to
 * max_batch_rows rows; requests that arrive while batches are on the device
 * join later batches; up to n_streams batches are on the device at once.
 * Each result is bitwise what tc_predict gives for the request alone.  The
 * host copies request rows into the server's pinned staging (uncached, read
 * by the kernel in place) and the outputs back. */
typedef struct {
  uint32_t max_batch_rows;   /* rows per batch, <= 2^24 */
  uint64_t memory_budget;    /* bytes of request data admitted at once */
  uint32_t n_streams;        /* batches on the device at once, 1..8 */
  int32_t transform;         /* 0 raw, 1 sigmoid (softmax for K > 1) */
  uint64_t fail_batch;       /* fault injection for tests: batch fail_batch - 1 is failed by the
                                host as if the device had reported an error (0: none) */
} tc_serve_params;
typedef struct {
  uint64_t batches, rows;
  uint64_t peak_admitted_bytes;    /* most request bytes admitted at once */
  uint64_t max_batches_in_flight;
  uint64_t cancelled, failed;      /* requests */
  uint64_t kernel_launches;        /* serving kernel launches (one per busy period or time slice) */
} tc_serve_counters_t;
int32_t tc_serve_start(tc_ensemble* ens, const tc_serve_params* p, tc_server** out);
/* Submit n rows; X and out must stay valid until tc_serve_wait returns for
 * ticket (out: n x K floats). */
int32_t tc_serve_submit(tc_server* s, const float* X, uint32_t n, float* out, uint64_t* ticket);
/* TC_OK, TC_ETIMEOUT (the request stays live), TC_ESTATE (cancelled) or
 * TC_EDEVICE (the server failed; every unfinished request fails). */
int32_t tc_serve_wait(tc_server* s, uint64_t ticket, int64_t timeout_us);
/* Cancels a request none of whose rows is on the device yet (TC_EBUSY
 * otherwise); its wait then returns TC_ESTATE. */
int32_t tc_serve_cancel(tc_server* s, uint64_t ticket);
int32_t tc_serve_stats(tc_server* s, uint64_t* batches, uint64_t* rows);
int32_t tc_serve_counters(tc_server* s, tc_serve_counters_t* out);
/* Latencies in microseconds of the requests finished so far, from
 * tc_serve_submit to the result in the caller's buffer, in completion order
 * (the first 2^20): copies up to cap of them into us; *count = how many. */
int32_t tc_serve_latencies(tc_server* s, float* us, uint64_t cap, uint64_t* count);
/* Drains what was submitted (unless the server failed), then stops. */
int32_t tc_serve_stop(tc_server* s);
/* Pinned host memory: tc_predict transfers rows from it to the device
 * without a staging copy. */
int32_t tc_host_alloc(size_t bytes, void** out);
int32_t tc_host_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
