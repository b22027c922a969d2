// pybind11 module `_turbocat`: the reference's binding surface
// (python/bindings.cpp:37-1225), backed by the MI355X booster.  Same class
// names, keyword arguments, defaults, return types and error behaviour
// (RuntimeError "Model not fitted. Call fit() first." before fit; ValueError
// for invalid arguments).  mode="auto"/"large" grows symmetric trees,
// mode="small" regular trees (bindings.cpp:96-103, :150-158), grown depth-wise
// or, with grow_policy "lossguide" / "leaf" / "leafwise" (bindings.cpp:88-93),
// leaf-wise; labels with more than two classes train the reference's
// multiclass loop (regular per-class trees, booster.cpp:1263-1400) whatever
// the mode, and predict through those trees.  Symmetric trees ignore
// grow_policy, as in the reference.  Options whose reference code path lies
// outside these (GradTree, ordered boosting, categorical features,
// interactions, early-stopping metrics other than the loss, other losses)
// raise ValueError naming the option instead of silently training something
// else.  use_goss samples each tree's rows by GOSS, symmetric and regular
// trees alike (booster.cpp:198-219, :1326-1354).  A validation set early-stops
// on its loss (booster.cpp:255-371, :1400-1430).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <cstdio>

#include "booster.hpp"

namespace py = pybind11;
using tcamd::Booster;
using tcamd::Params;

namespace {

using FArray = py::array_t<float, py::array::c_style | py::array::forcecast>;

struct Matrix {
  const float* data;
  uint32_t n, F;
};

Matrix as_matrix(const FArray& X) {
  if (X.ndim() != 2) throw std::runtime_error("X must be 2-dimensional");
  return Matrix{X.data(), static_cast<uint32_t>(X.shape(0)), static_cast<uint32_t>(X.shape(1))};
}

void reject(bool cond, const std::string& what) {
  if (cond) throw std::invalid_argument(what + " is not supported by the MI355X backend");
}

bool lossguide(const std::string& g) { return g == "lossguide" || g == "leaf" || g == "leafwise"; }

// The validation set of fit (an empty X_val means none; bindings.cpp:261-279).
Matrix validation(const FArray& X_val, const FArray& y_val, uint32_t F) {
  if (X_val.size() == 0) return Matrix{nullptr, 0, F};
  const Matrix v = as_matrix(X_val);
  if (v.F != F) throw std::invalid_argument("X_val has " + std::to_string(v.F) + " features, X has " +
                                            std::to_string(F));
  if (y_val.ndim() != 1 || static_cast<uint32_t>(y_val.shape(0)) != v.n)
    throw std::invalid_argument("X_val and y_val differ in rows");
  return v;
}

py::dict importance_dict(const std::vector<float>& imp) {
  double s = 0;
  for (float v : imp) s += v;
  py::array_t<float> g(imp.size()), gn(imp.size());
  for (size_t i = 0; i < imp.size(); ++i) {
    g.mutable_at(i) = imp[i];
    gn.mutable_at(i) = s > 0 ? static_cast<float>(imp[i] / s) : 0.f;
  }
  py::dict d;
  d["gain"] = g;
  d["gain_normalized"] = gn;
  return d;
}

class TurboCatClassifier {
 public:
  TurboCatClassifier(int n_estimators, float learning_rate, int max_depth, int max_bins,
                     float subsample, float colsample_bytree, float min_child_weight,
                     float lambda_l2, const std::string& loss, bool use_goss, float goss_top_rate,
                     float goss_other_rate, bool use_gradtree, bool use_symmetric,
                     bool use_ordered_boosting, const std::string& grow_policy,
                     const std::string& mode, int early_stopping_rounds,
                     const std::string& early_stopping_metric, const std::string& cat_encoding,
                     bool auto_interactions, int max_interactions,
                     const std::string& interaction_method, int n_threads, int seed, int verbosity)
      : mode_(mode), loss_(loss) {
    if (n_estimators < 0) throw std::invalid_argument("n_estimators must be >= 0");
    if (max_depth < 1) throw std::invalid_argument("max_depth must be >= 1");
    p_.n_estimators = static_cast<uint32_t>(n_estimators);
    p_.learning_rate = learning_rate;
    p_.max_depth = static_cast<uint32_t>(max_depth);
    p_.max_bins = static_cast<uint32_t>(max_bins);
    p_.subsample = subsample;
    p_.colsample_bytree = colsample_bytree;
    p_.min_child_weight = min_child_weight;
    p_.lambda_l2 = lambda_l2;
    p_.seed = static_cast<uint64_t>(seed);
    p_.verbosity = verbosity;
    p_.loss = TC_LOSS_LOGLOSS;
    p_.task = tcamd::kBinary;
    p_.use_goss = use_goss;
    p_.goss_top_rate = goss_top_rate;
    p_.goss_other_rate = goss_other_rate;
    use_gradtree_ = use_gradtree;
    use_ordered_ = use_ordered_boosting;
    auto_interactions_ = auto_interactions;
    p_.early_stopping_rounds = static_cast<uint32_t>(std::max(0, early_stopping_rounds));
    es_metric_ = early_stopping_metric;
    grow_policy_ = grow_policy;
    // mode "small": regular trees, "large": symmetric, "auto": symmetric at fit
    // (bindings.cpp:80, :99-103, :154-158)
    p_.use_symmetric = mode == "small" ? false : (mode == "large" || mode == "auto" ? true : use_symmetric);
    (void)cat_encoding; (void)max_interactions; (void)interaction_method; (void)n_threads;
  }

  void fit(FArray X, FArray y, FArray X_val, FArray y_val, std::vector<int> cat_features) {
    const Matrix m = as_matrix(X);
    if (y.ndim() != 1) throw std::runtime_error("y must be 1-dimensional");
    if (static_cast<uint32_t>(y.shape(0)) != m.n) throw std::invalid_argument("X and y differ in rows");
    reject(use_gradtree_, "use_gradtree");
    reject(use_ordered_, "use_ordered_boosting");
    reject(auto_interactions_, "auto_interactions");
    reject(!cat_features.empty(), "cat_features");
    const Matrix v = validation(X_val, y_val, m.F);
    // the loss is the metric; the reference's other metrics rank probabilities
    reject(v.n > 0 && (es_metric_ == "roc_auc" || es_metric_ == "auc" || es_metric_ == "pr_auc" ||
                       es_metric_ == "f1"),
           "early_stopping_metric='" + es_metric_ + "'");
    reject(!(loss_ == "auto" || loss_ == "logloss" || loss_ == "binary_crossentropy"),
           "loss='" + loss_ + "'");
    float max_label = 0.f;
    for (uint32_t i = 0; i < m.n; ++i) max_label = std::max(max_label, y.data()[i]);
    n_classes_ = static_cast<uint32_t>(max_label) + 1;  // bindings.cpp:161-167
    if (n_classes_ <= 2) n_classes_ = 2;
    reject(n_classes_ > 16, "more than 16 classes");
    Params p = p_;
    p.n_classes = n_classes_;
    if (n_classes_ > 2) {  // CrossEntropy over regular per-class trees (bindings.cpp:193-198)
      p.task = tcamd::kMulticlass;
      p.loss = TC_LOSS_SOFTMAX;
    }
    p.lossguide = lossguide(grow_policy_);  // read by regular trees only
    booster_ = std::make_unique<Booster>(p);
    if (p_.verbosity > 0)
      std::printf("[TurboCat/MI355X] fit: n_samples=%u, n_features=%u, n_estimators=%u\n", m.n, m.F,
                  p_.n_estimators);
    py::gil_scoped_release nogil;
    booster_->fit(m.data, y.data(), m.n, m.F, v.n ? v.data : nullptr, v.n ? y_val.data() : nullptr, v.n);
  }

  py::array_t<float> predict_proba(FArray X) { return predict_proba_nobinning_fast(X); }

  py::array_t<float> predict_proba_nobinning_fast(FArray X) {
    fitted();
    const Matrix m = as_matrix(X);
    if (booster_->outputs() > 1) {  // multiclass: n x K softmax (booster.cpp:1522-1536)
      const uint32_t K = booster_->outputs();
      py::array_t<float> out({static_cast<py::ssize_t>(m.n), static_cast<py::ssize_t>(K)});
      {
        py::gil_scoped_release nogil;
        booster_->predict_proba(m.data, m.n, m.F, out.mutable_data());
      }
      return out;
    }
    std::vector<float> p1(m.n);
    {
      py::gil_scoped_release nogil;
      booster_->predict_proba(m.data, m.n, m.F, p1.data());
    }
    py::array_t<float> out({static_cast<py::ssize_t>(m.n), static_cast<py::ssize_t>(2)});
    auto r = out.mutable_unchecked<2>();
    for (uint32_t i = 0; i < m.n; ++i) {
      r(i, 0) = 1.0f - p1[i];
      r(i, 1) = p1[i];
    }
    return out;
  }

  py::array_t<int> predict(FArray X) {
    fitted();
    auto proba = predict_proba(X);
    auto p = proba.unchecked<2>();
    py::array_t<int> out(proba.shape(0));
    auto r = out.mutable_unchecked<1>();
    if (proba.shape(1) > 2) {  // argmax, the first maximum wins (bindings.cpp:399-411)
      for (py::ssize_t i = 0; i < proba.shape(0); ++i) {
        int best = 0;
        for (py::ssize_t c = 1; c < proba.shape(1); ++c)
          if (p(i, c) > p(i, best)) best = static_cast<int>(c);
        r(i) = best;
      }
      return out;
    }
    for (py::ssize_t i = 0; i < proba.shape(0); ++i) r(i) = p(i, 1) >= 0.5f ? 1 : 0;
    return out;
  }

  py::dict feature_importance() {
    if (!booster_ || !booster_->fitted()) throw std::runtime_error("Model not fitted");
    return importance_dict(booster_->feature_importance());
  }

  void save(const std::string& path) {
    if (!booster_ || !booster_->fitted()) throw std::runtime_error("Model not fitted");
    booster_->save(path);
  }

  void load(const std::string& path) {
    auto b = std::make_unique<Booster>(p_);
    b->load(path);
    booster_ = std::move(b);
    n_classes_ = booster_->n_classes();
  }

  py::dict get_params() const {
    py::dict d;
    d["n_estimators"] = p_.n_estimators;
    d["learning_rate"] = p_.learning_rate;
    d["max_depth"] = p_.max_depth;
    d["max_bins"] = p_.max_bins;
    d["subsample"] = p_.subsample;
    d["colsample_bytree"] = p_.colsample_bytree;
    d["use_goss"] = p_.use_goss;
    d["min_child_weight"] = p_.min_child_weight;
    d["lambda_l2"] = p_.lambda_l2;
    d["seed"] = p_.seed;
    d["mode"] = mode_;
    return d;
  }

  float get_base_prediction() const { fitted(); return booster_->base_prediction(); }
  uint32_t get_n_features() const { fitted(); return booster_->n_features(); }
  uint32_t n_trees() const { return booster_ && booster_->fitted() ? booster_->n_trees() : 0; }
  uint32_t n_classes() const { return n_classes_; }

 private:
  void fitted() const {
    if (!booster_ || !booster_->fitted()) throw std::runtime_error("Model not fitted. Call fit() first.");
  }
  Params p_;
  std::string mode_, loss_, es_metric_ = "loss", grow_policy_ = "depthwise";
  bool use_gradtree_ = false, use_ordered_ = false, auto_interactions_ = false;
  uint32_t n_classes_ = 2;
  std::unique_ptr<Booster> booster_;
};

class TurboCatRegressor {
 public:
  TurboCatRegressor(int n_estimators, float learning_rate, int max_depth, const std::string& loss,
                    const std::string& mode, float subsample, float colsample_bytree,
                    bool use_goss, float goss_top_rate, float goss_other_rate,
                    int early_stopping_rounds, float lambda_l2, float min_child_weight,
                    int max_leaves, const std::string& grow_policy, int n_threads, int seed,
                    int verbosity)
      : mode_(mode), loss_(loss) {
    if (n_estimators < 0) throw std::invalid_argument("n_estimators must be >= 0");
    if (max_depth < 1) throw std::invalid_argument("max_depth must be >= 1");
    p_.n_estimators = static_cast<uint32_t>(n_estimators);
    p_.learning_rate = learning_rate;
    p_.max_depth = static_cast<uint32_t>(max_depth);
    p_.subsample = subsample;
    p_.colsample_bytree = colsample_bytree;
    p_.lambda_l2 = lambda_l2;
    p_.min_child_weight = min_child_weight;
    p_.seed = static_cast<uint64_t>(seed);
    p_.verbosity = verbosity;
    p_.loss = TC_LOSS_MSE;
    p_.task = tcamd::kRegression;
    p_.use_goss = use_goss;
    p_.goss_top_rate = goss_top_rate;
    p_.goss_other_rate = goss_other_rate;
    p_.early_stopping_rounds = static_cast<uint32_t>(std::max(0, early_stopping_rounds));
    if (max_leaves < 1) throw std::invalid_argument("max_leaves must be >= 1");
    p_.max_leaves = static_cast<uint32_t>(max_leaves);
    p_.use_symmetric = mode != "small";  // bindings.cpp:584-588, :615-620
    grow_policy_ = grow_policy;
    (void)n_threads;
  }

  void fit(FArray X, FArray y, FArray X_val, FArray y_val) {
    const Matrix m = as_matrix(X);
    if (y.ndim() != 1) throw std::runtime_error("y must be 1-dimensional");
    if (static_cast<uint32_t>(y.shape(0)) != m.n) throw std::invalid_argument("X and y differ in rows");
    const Matrix v = validation(X_val, y_val, m.F);
    reject(!(loss_ == "mse" || loss_ == "l2"), "loss='" + loss_ + "'");
    Params p = p_;
    p.lossguide = lossguide(grow_policy_);  // read by regular trees only
    booster_ = std::make_unique<Booster>(p);
    if (p_.verbosity > 0)
      std::printf("[TurboCat/MI355X] fit: n_samples=%u, n_features=%u, n_estimators=%u\n", m.n, m.F,
                  p_.n_estimators);
    py::gil_scoped_release nogil;
    booster_->fit(m.data, y.data(), m.n, m.F, v.n ? v.data : nullptr, v.n ? y_val.data() : nullptr, v.n);
  }

  py::array_t<float> predict_raw(FArray X, bool timing) {
    fitted();
    const Matrix m = as_matrix(X);
    py::array_t<float> out(m.n);
    const auto t0 = std::chrono::steady_clock::now();
    {
      py::gil_scoped_release nogil;
      booster_->predict_raw(m.data, m.n, m.F, out.mutable_data());
    }
    if (timing)
      std::printf("[TIMING] predict: %.3f ms for %u rows\n",
                  std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                  m.n);
    return out;
  }

  py::dict feature_importance() {
    if (!booster_ || !booster_->fitted()) throw std::runtime_error("Model not fitted");
    return importance_dict(booster_->feature_importance());
  }

  void save(const std::string& path) {
    if (!booster_ || !booster_->fitted()) throw std::runtime_error("Model not fitted");
    booster_->save(path);
  }

  void load(const std::string& path) {
    auto b = std::make_unique<Booster>(p_);
    b->load(path);
    booster_ = std::move(b);
  }

  py::dict tree_info() const {
    fitted();
    py::dict d;
    d["n_trees"] = booster_->n_trees();
    d["max_depth"] = booster_->depth_slots();
    d["base_prediction"] = booster_->base_prediction();
    d["n_features"] = booster_->n_features();
    return d;
  }

  std::string debug_info() const {
    if (!booster_ || !booster_->fitted()) return "TurboCatRegressor (not fitted), MI355X backend";
    return "TurboCatRegressor: " + std::to_string(booster_->n_trees()) +
           (booster_->regular() ? " regular trees, max depth " : " symmetric trees, depth ") +
           std::to_string(booster_->depth_slots()) + ", " + std::to_string(booster_->n_features()) +
           " features, base " + std::to_string(booster_->base_prediction()) + ", MI355X backend";
  }

  uint32_t get_n_features() const { fitted(); return booster_->n_features(); }
  uint32_t n_trees() const { return booster_ && booster_->fitted() ? booster_->n_trees() : 0; }
  float base_prediction() const { fitted(); return booster_->base_prediction(); }

 private:
  void fitted() const {
    if (!booster_ || !booster_->fitted()) throw std::runtime_error("Model not fitted. Call fit() first.");
  }
  Params p_;
  std::string mode_, loss_, es_metric_ = "loss", grow_policy_ = "depthwise";
  std::unique_ptr<Booster> booster_;
};

}  // namespace

PYBIND11_MODULE(_turbocat, m) {
  m.doc() = "TurboCat gradient boosting (symmetric and regular trees) on MI355X (gfx950)";
  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const std::invalid_argument& e) {
      PyErr_SetString(PyExc_ValueError, e.what());
    }
  });
  m.def("print_info", []() {
    py::print("TurboCat (MI355X / gfx950 backend): symmetric and regular trees, HIP kernels, C-ABI turbocat_hip.h");
  });

  py::class_<TurboCatClassifier>(m, "TurboCatClassifier")
      .def(py::init<int, float, int, int, float, float, float, float, const std::string&, bool, float,
                    float, bool, bool, bool, const std::string&, const std::string&, int,
                    const std::string&, const std::string&, bool, int, const std::string&, int, int,
                    int>(),
           py::arg("n_estimators") = 500, py::arg("learning_rate") = 0.2f, py::arg("max_depth") = 6,
           py::arg("max_bins") = 255, py::arg("subsample") = 1.0f,
           py::arg("colsample_bytree") = 1.0f, py::arg("min_child_weight") = 1.0f,
           py::arg("lambda_l2") = 1.0f, py::arg("loss") = "logloss", py::arg("use_goss") = false,
           py::arg("goss_top_rate") = 0.2f, py::arg("goss_other_rate") = 0.1f,
           py::arg("use_gradtree") = false, py::arg("use_symmetric") = false,
           py::arg("use_ordered_boosting") = false, py::arg("grow_policy") = "depthwise",
           py::arg("mode") = "auto", py::arg("early_stopping_rounds") = 50,
           py::arg("early_stopping_metric") = "loss", py::arg("cat_encoding") = "ordered",
           py::arg("auto_interactions") = false, py::arg("max_interactions") = 10,
           py::arg("interaction_method") = "split", py::arg("n_threads") = -1, py::arg("seed") = 42,
           py::arg("verbosity") = 1)
      .def("fit", &TurboCatClassifier::fit, py::arg("X"), py::arg("y"),
           py::arg("X_val") = FArray(), py::arg("y_val") = FArray(),
           py::arg("cat_features") = std::vector<int>{})
      .def("predict", &TurboCatClassifier::predict, py::arg("X"))
      .def("predict_proba", &TurboCatClassifier::predict_proba, py::arg("X"))
      .def("predict_proba_nobinning_fast", &TurboCatClassifier::predict_proba_nobinning_fast, py::arg("X"))
      .def("feature_importance", &TurboCatClassifier::feature_importance)
      .def("save", &TurboCatClassifier::save, py::arg("path"))
      .def("load", &TurboCatClassifier::load, py::arg("path"))
      .def("get_params", &TurboCatClassifier::get_params)
      .def("get_base_prediction", &TurboCatClassifier::get_base_prediction)
      .def("get_n_features", &TurboCatClassifier::get_n_features)
      .def_property_readonly("n_trees", &TurboCatClassifier::n_trees)
      .def_property_readonly("n_classes_", &TurboCatClassifier::n_classes);

  py::class_<TurboCatRegressor>(m, "TurboCatRegressor")
      .def(py::init<int, float, int, const std::string&, const std::string&, float, float, bool, float,
                    float, int, float, float, int, const std::string&, int, int, int>(),
           py::arg("n_estimators") = 500, py::arg("learning_rate") = 0.1f, py::arg("max_depth") = 6,
           py::arg("loss") = "mse", py::arg("mode") = "auto", py::arg("subsample") = 1.0f,
           py::arg("colsample_bytree") = 1.0f, py::arg("use_goss") = false,
           py::arg("goss_top_rate") = 0.2f, py::arg("goss_other_rate") = 0.1f,
           py::arg("early_stopping_rounds") = 50, py::arg("lambda_l2") = 3.0f,
           py::arg("min_child_weight") = 1.0f, py::arg("max_leaves") = 64,
           py::arg("grow_policy") = "depthwise", py::arg("n_threads") = -1, py::arg("seed") = 42,
           py::arg("verbosity") = 1)
      .def("fit", &TurboCatRegressor::fit, py::arg("X"), py::arg("y"), py::arg("X_val") = FArray(),
           py::arg("y_val") = FArray())
      .def("predict", &TurboCatRegressor::predict_raw, py::arg("X"), py::arg("timing") = false)
      .def("predict_fast", &TurboCatRegressor::predict_raw, py::arg("X"), py::arg("timing") = false)
      .def("predict_nobinning", &TurboCatRegressor::predict_raw, py::arg("X"), py::arg("timing") = false)
      .def("predict_nobinning_fast", &TurboCatRegressor::predict_raw, py::arg("X"),
           py::arg("timing") = false)
      .def("feature_importance", &TurboCatRegressor::feature_importance)
      .def("save", &TurboCatRegressor::save, py::arg("path"))
      .def("load", &TurboCatRegressor::load, py::arg("path"))
      .def("tree_info", &TurboCatRegressor::tree_info)
      .def("debug_info", &TurboCatRegressor::debug_info)
      .def("get_n_features", &TurboCatRegressor::get_n_features)
      .def_property_readonly("n_trees", &TurboCatRegressor::n_trees)
      .def_property_readonly("base_prediction", &TurboCatRegressor::base_prediction);
}
