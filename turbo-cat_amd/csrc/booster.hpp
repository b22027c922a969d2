// Host C++ booster over the C-ABI: the MI355X counterpart of the reference's
// Booster (src/boosting/booster.cpp:47-493, :599-945, :1049-1203, multiclass
// :1263-1482) for symmetric trees and for regular depthwise trees (mode
// "small", and every multiclass fit: the reference's train_multiclass always
// grows regular per-class trees).  Training, binning and prediction run on
// the GPU through include/turbocat_hip.h; this class owns the model on the
// host, keeps it resident on the device for prediction, and reads/writes the
// reference's TCAT v3 model files.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "turbocat_hip.h"

namespace tcamd {

// Task codes as stored in TCAT files (TaskType, include/turbocat/types.hpp:242-247).
enum Task : uint8_t { kBinary = 0, kMulticlass = 1, kRegression = 2 };

struct Params {
  // boosting (BoostingConfig, config.hpp:66-96)
  uint32_t n_estimators = 500;
  float learning_rate = 0.1f;
  float subsample = 1.0f;
  float colsample_bytree = 1.0f;
  bool use_goss = false;                // GOSS row sampling (config.hpp:83-85)
  float goss_top_rate = 0.2f;
  float goss_other_rate = 0.1f;
  bool use_mvs = false;                 // MVS row sampling, before GOSS (config.hpp:88-89)
  float mvs_subsample = 0.5f;
  uint32_t early_stopping_rounds = 50;  // with a validation set (booster.cpp:344-346)
  // tree (TreeConfig, config.hpp:26-59)
  uint32_t max_depth = 6;
  uint32_t max_bins = 255;
  float lambda_l2 = 1.0f;
  float lambda_l1 = 0.0f;
  float max_delta_step = 0.0f;
  uint32_t min_samples_leaf = 1;
  float min_child_weight = 1.0f;
  float min_split_gain = 0.0f;
  // regular trees (TreeConfig::max_leaves, gamma, leaf_smooth,
  // learn_missing_direction; config.hpp:29-52)
  bool use_symmetric = true;  // false: regular depthwise trees (mode "small")
  uint32_t max_leaves = 64;
  float gamma = 0.0f;
  float leaf_smooth = 0.0f;
  bool learn_missing = true;
  bool lossguide = false;     // regular trees grown leaf-wise (GrowPolicy::Lossguide)
  uint32_t n_classes = 2;     // > 2: softmax over regular per-class trees (train_multiclass)
  // run
  int32_t loss = TC_LOSS_MSE;
  Task task = kRegression;
  uint64_t seed = 42;
  int32_t verbosity = 0;
  int32_t device = 0;
  bool wrap_bins = true;      // reproduce the reference quantizer's uint8 wrap
  bool use_graph = true;
};

// Translate a nonzero C-ABI code into the reference's exception types.
void check(int32_t rc);

class Booster {
 public:
  explicit Booster(const Params& p);
  ~Booster();
  Booster(const Booster&) = delete;
  Booster& operator=(const Booster&) = delete;

  // Quantize X (row-major n x F), train n_estimators symmetric trees.  With a
  // validation set (Xv: nv x F, binned with the training edges) the fit early-
  // stops on the validation loss (booster.cpp:255-371) and keeps the trees built.
  void fit(const float* X, const float* y, uint32_t n, uint32_t F, const float* Xv = nullptr,
           const float* yv = nullptr, uint32_t nv = 0);
  // Raw scores + base (predict_raw_nobinning_fast, booster.cpp:899-925; regular
  // trees: predict_raw_fast, :680-871).  Multiclass: n x K.
  void predict_raw(const float* X, uint32_t n, uint32_t F, float* out);
  // sigmoid(raw) (predict_proba_nobinning_fast, booster.cpp:927-945);
  // multiclass: softmax over the classes of apply_bins + the multiclass
  // ensemble (predict_proba_multiclass, booster.cpp:1506-1536), n x K.
  void predict_proba(const float* X, uint32_t n, uint32_t F, float* out);
  // Per-feature sum of split gains, 256 entries (symmetric_tree.cpp:892, :1176).
  std::vector<float> feature_importance() const;

  void save(const std::string& path) const;   // TCAT v3 (booster.cpp:1049-1109)
  void load(const std::string& path);         // TCAT v1-3 symmetric (booster.cpp:1111-1203)

  bool fitted() const { return fitted_; }
  bool regular() const { return regular_; }
  uint32_t n_classes() const { return K_ > 1 ? K_ : 2u; }
  uint32_t outputs() const { return K_; }
  uint32_t n_trees() const { return T_; }
  uint32_t n_features() const { return F_; }
  float base_prediction() const { return base_; }
  const Params& params() const { return p_; }
  Params& params() { return p_; }
  double last_fit_seconds() const { return fit_seconds_; }

  // model (tree-major, D = max_depth slots per tree, 2^D leaves per tree)
  uint32_t depth_slots() const { return D_; }
  const std::vector<uint16_t>& split_feature() const { return feat_; }
  const std::vector<uint8_t>& split_bin() const { return bin_; }
  const std::vector<float>& split_threshold() const { return fthr_; }
  const std::vector<float>& split_gain() const { return gain_; }
  const std::vector<uint16_t>& tree_depth() const { return depth_; }
  const std::vector<float>& leaves() const { return leaves_; }
  const std::vector<float>& tree_weights() const { return weights_; }
  const std::vector<std::vector<float>>& bin_edges() const { return edges_; }

  // regular model: tree t's nodes are [tree_off[t], tree_off[t+1]) with
  // tree-local child ids (the reference's Tree::nodes)
  const std::vector<uint32_t>& tree_off() const { return roff_; }
  const std::vector<uint16_t>& node_feature() const { return rfeat_; }
  const std::vector<uint8_t>& node_bin() const { return rbin_; }
  const std::vector<uint8_t>& node_flags() const { return rflags_; }
  const std::vector<uint32_t>& node_left() const { return rleft_; }
  const std::vector<uint32_t>& node_right() const { return rright_; }
  const std::vector<float>& node_value() const { return rvalue_; }
  const std::vector<float>& node_gain() const { return rgain_; }
  const std::vector<uint32_t>& tree_class() const { return rclass_; }

 private:
  void ensure_ensemble();
  void drop_ensemble();
  void fit_regular(const float* X, const float* y, uint32_t n, uint32_t F, const float* Xv,
                   const float* yv, uint32_t nv);
  tc_rensemble* regular_ensemble(int rule);

  Params p_;
  tc_ctx* ctx_ = nullptr;
  tc_ensemble* ens_ = nullptr;
  bool fitted_ = false;
  uint32_t T_ = 0, D_ = 0, F_ = 0;
  float base_ = 0.f;
  double fit_seconds_ = 0.0;
  std::vector<std::vector<float>> edges_;
  std::vector<uint16_t> feat_;
  std::vector<uint8_t> bin_;
  std::vector<float> fthr_, gain_, leaves_, weights_;
  std::vector<uint16_t> depth_;
  // regular trees
  bool regular_ = false;
  uint32_t K_ = 1;              // outputs: classes of a multiclass model, else 1
  std::vector<uint32_t> roff_, rleft_, rright_, rclass_;
  std::vector<uint16_t> rfeat_, rdepth_, rnleaves_;
  std::vector<uint8_t> rbin_, rflags_;
  std::vector<float> rvalue_, rgain_;
  tc_rensemble* rens_[2] = {nullptr, nullptr};  // by binning rule
};

// One process-wide context per device (the reference binds one Booster to
// one thread; contexts here are shared the same way).
tc_ctx* device_context(int32_t device);

}  // namespace tcamd
