// Host C++ booster over the C-ABI (see booster.hpp).
#include "booster.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>

namespace tcamd {

void check(int32_t rc) {
  if (rc == TC_OK) return;
  const std::string msg = tc_last_error();
  if (rc == TC_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

tc_ctx* device_context(int32_t device) {
  static std::mutex mu;
  static std::vector<tc_ctx*> ctxs;
  std::lock_guard<std::mutex> lk(mu);
  if (device < 0) throw std::invalid_argument("device must be >= 0");
  if (ctxs.size() <= static_cast<size_t>(device)) ctxs.resize(device + 1, nullptr);
  if (!ctxs[device]) check(tc_ctx_create(device, &ctxs[device]));
  return ctxs[device];
}

Booster::Booster(const Params& p) : p_(p) {}

Booster::~Booster() { drop_ensemble(); }

void Booster::drop_ensemble() {
  if (ens_) {
    tc_ensemble_destroy(ens_);
    ens_ = nullptr;
  }
}

void Booster::fit(const float* X, const float* y, uint32_t n, uint32_t F, const float* Xv,
                  const float* yv, uint32_t nv) {
  if (n == 0) throw std::invalid_argument("X has no rows");
  if (F == 0) throw std::invalid_argument("X has no features");
  if (p_.max_depth < 1 || p_.max_depth > 10)
    throw std::invalid_argument("max_depth must be in [1, 10] for symmetric trees");
  if (p_.max_bins < 2 || p_.max_bins > 255) throw std::invalid_argument("max_bins must be in [2, 255]");
  if (!(p_.subsample > 0.0f && p_.subsample <= 1.0f))
    throw std::invalid_argument("subsample must be in (0, 1]");
  if (p_.use_goss && !(p_.goss_top_rate >= 0.0f && p_.goss_top_rate <= 1.0f &&
                       p_.goss_other_rate >= 0.0f && p_.goss_other_rate <= 1.0f))
    throw std::invalid_argument("goss_top_rate and goss_other_rate must be in [0, 1]");
  if (p_.use_mvs && !(p_.mvs_subsample > 0.0f && p_.mvs_subsample <= 1.0f))
    throw std::invalid_argument("mvs_subsample must be in (0, 1]");
  const auto t0 = std::chrono::steady_clock::now();
  ctx_ = device_context(p_.device);
  drop_ensemble();
  fitted_ = false;
  std::vector<float> edges(static_cast<size_t>(F) * 255);
  std::vector<uint32_t> ne(F);
  tc_dataset* ds = nullptr;
  check(tc_quantize_fit(ctx_, X, n, F, p_.max_bins, p_.wrap_bins ? 1 : 0, edges.data(), ne.data(),
                        &ds));
  tc_train_params tp{};
  tp.loss = p_.loss;
  tp.n_trees = p_.n_estimators;
  tp.learning_rate = p_.learning_rate;
  tp.tree.max_depth = p_.max_depth;
  tp.tree.max_bins = p_.max_bins;
  tp.tree.lambda_l2 = p_.lambda_l2;
  tp.tree.lambda_l1 = p_.lambda_l1;
  tp.tree.max_delta_step = p_.max_delta_step;
  tp.tree.min_samples_leaf = p_.min_samples_leaf;
  tp.tree.min_child_weight = p_.min_child_weight;
  tp.tree.min_split_gain = p_.min_split_gain;
  tp.hist_chunk = 0;
  tp.use_graph = p_.use_graph ? 1 : 0;
  tp.subsample = p_.subsample;
  tp.colsample_bytree = p_.colsample_bytree;
  tp.use_goss = p_.use_goss ? 1 : 0;
  tp.goss_top_rate = p_.goss_top_rate;
  tp.goss_other_rate = p_.goss_other_rate;
  tp.use_mvs = p_.use_mvs ? 1 : 0;
  tp.mvs_subsample = p_.mvs_subsample;
  tp.seed = p_.seed;
  const uint32_t T = p_.n_estimators, D = p_.max_depth;
  feat_.assign(static_cast<size_t>(T) * D, 0);
  bin_.assign(static_cast<size_t>(T) * D, 0);
  fthr_.assign(static_cast<size_t>(T) * D, 0.f);
  gain_.assign(static_cast<size_t>(T) * D, 0.f);
  depth_.assign(T, 0);
  leaves_.assign(static_cast<size_t>(T) << D, 0.f);
  tc_model m{};
  m.feature = feat_.data();
  m.bin = bin_.data();
  m.fthr = fthr_.data();
  m.gain = gain_.data();
  m.depth = depth_.data();
  m.leaves = leaves_.data();
  uint32_t built = T;
  int32_t rc;
  if (Xv != nullptr && nv > 0) {
    // Dataset::apply_bins with the training edges (bindings.cpp:261-279)
    tc_dataset* vds = nullptr;
    rc = tc_dataset_create(ctx_, Xv, nv, F, edges.data(), ne.data(), p_.max_bins, &vds);
    if (rc == 0) {
      tc_train_history h{};
      rc = tc_train_valid(ds, y, &tp, vds, yv, p_.early_stopping_rounds, &m, &h);
      built = h.n_trees_built;
      tc_dataset_destroy(vds);
    }
  } else {
    rc = tc_train(ds, y, &tp, &m);
  }
  tc_dataset_destroy(ds);
  check(rc);
  if (built < T) {  // the trees after the stopping one were never built
    feat_.resize(static_cast<size_t>(built) * D);
    bin_.resize(static_cast<size_t>(built) * D);
    fthr_.resize(static_cast<size_t>(built) * D);
    gain_.resize(static_cast<size_t>(built) * D);
    depth_.resize(built);
    leaves_.resize(static_cast<size_t>(built) << D);
  }
  const uint32_t T_built = built;
  T_ = T_built;
  D_ = D;
  F_ = F;
  base_ = m.base;
  weights_.assign(T_built, p_.learning_rate);  // add_tree(tree, lr), booster.cpp:465
  edges_.assign(F, {});
  for (uint32_t f = 0; f < F; ++f)
    edges_[f].assign(edges.begin() + static_cast<size_t>(f) * 255,
                     edges.begin() + static_cast<size_t>(f) * 255 + ne[f]);
  fitted_ = true;
  fit_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void Booster::ensure_ensemble() {
  if (!fitted_) throw std::runtime_error("Model not fitted. Call fit() first.");
  if (ens_) return;
  ctx_ = device_context(p_.device);
  check(tc_ensemble_create(ctx_, T_, std::max(1u, D_), feat_.data(), fthr_.data(), bin_.data(),
                           depth_.data(), leaves_.data(), weights_.data(), base_, F_, &ens_));
}

void Booster::predict_raw(const float* X, uint32_t n, uint32_t F, float* out) {
  ensure_ensemble();
  if (F != F_)
    throw std::invalid_argument("X has " + std::to_string(F) + " features, the model " +
                                std::to_string(F_));
  check(tc_predict(ens_, X, n, F, 0, out, nullptr));
}

void Booster::predict_proba(const float* X, uint32_t n, uint32_t F, float* out) {
  ensure_ensemble();
  if (F != F_)
    throw std::invalid_argument("X has " + std::to_string(F) + " features, the model " +
                                std::to_string(F_));
  check(tc_predict(ens_, X, n, F, 1, out, nullptr));
}

std::vector<float> Booster::feature_importance() const {
  if (!fitted_) throw std::runtime_error("Model not fitted");
  std::vector<float> imp(256, 0.f);
  for (uint32_t t = 0; t < T_; ++t)
    for (uint32_t d = 0; d < depth_[t]; ++d) {
      const uint16_t f = feat_[static_cast<size_t>(t) * D_ + d];
      if (f < 256) imp[f] += gain_[static_cast<size_t>(t) * D_ + d];
    }
  return imp;
}

namespace {
template <typename T>
void put(std::ostream& o, const T& v) { o.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
template <typename T>
void get(std::istream& i, T& v) {
  i.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!i) throw std::runtime_error("Truncated TurboCat model file");
}
}  // namespace

void Booster::save(const std::string& path) const {
  if (!fitted_) throw std::runtime_error("Model not fitted");
  std::ofstream o(path, std::ios::binary);
  if (!o) throw std::runtime_error("Cannot open file for writing: " + path);
  o.write("TCAT", 4);
  put(o, static_cast<uint32_t>(3));
  put(o, base_);
  put(o, static_cast<uint8_t>(p_.task));
  put(o, static_cast<uint32_t>(2));                     // n_classes
  put(o, p_.learning_rate);
  put(o, static_cast<uint16_t>(p_.max_depth));
  put(o, p_.lambda_l2);
  put(o, static_cast<uint8_t>(1));                      // use_symmetric
  put(o, static_cast<uint64_t>(edges_.size()));
  for (const auto& e : edges_) {
    put(o, static_cast<uint64_t>(e.size()));
    if (!e.empty()) o.write(reinterpret_cast<const char*>(e.data()), 4 * e.size());
  }
  put(o, static_cast<uint64_t>(T_));
  for (uint32_t t = 0; t < T_; ++t) put(o, weights_[t]);
  for (uint32_t t = 0; t < T_; ++t) {
    const uint16_t d = depth_[t];
    put(o, d);
    for (uint32_t l = 0; l < d; ++l) {
      const size_t k = static_cast<size_t>(t) * D_ + l;
      put(o, feat_[k]);
      put(o, bin_[k]);
      put(o, fthr_[k]);
      put(o, gain_[k]);
    }
    put(o, static_cast<uint64_t>(1u << d));
    o.write(reinterpret_cast<const char*>(leaves_.data() + (static_cast<size_t>(t) << D_)),
            4ull << d);
  }
  if (!o) throw std::runtime_error("Write failed: " + path);
}

void Booster::load(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("Cannot open file: " + path);
  char magic[4];
  in.read(magic, 4);
  if (!in || std::strncmp(magic, "TCAT", 4) != 0) throw std::runtime_error("Invalid TurboCat model file");
  uint32_t version = 0;
  get(in, version);
  if (version < 1 || version > 3)
    throw std::runtime_error("Unsupported TurboCat model version: " + std::to_string(version));
  float base = 0.f;
  get(in, base);
  uint8_t task = 0, sym = 0;
  uint32_t n_classes = 2;
  float lr = p_.learning_rate, l2 = p_.lambda_l2;
  uint16_t max_depth = static_cast<uint16_t>(p_.max_depth);
  if (version >= 2) {
    get(in, task);
    get(in, n_classes);
    get(in, lr);
    get(in, max_depth);
    get(in, l2);
  }
  std::vector<std::vector<float>> edges;
  if (version >= 3) {
    get(in, sym);
    uint64_t nf = 0;
    get(in, nf);
    edges.resize(nf);
    for (auto& e : edges) {
      uint64_t ne = 0;
      get(in, ne);
      e.resize(ne);
      if (ne) in.read(reinterpret_cast<char*>(e.data()), 4 * ne);
    }
  }
  if (version < 3 || !sym)
    throw std::runtime_error("Only symmetric-tree TurboCat models (TCAT v3) run on this backend");
  if (task == kMulticlass && n_classes > 2)
    throw std::runtime_error("Multiclass TurboCat models are not supported by this backend");
  uint64_t nt = 0;
  get(in, nt);
  std::vector<float> w(nt);
  for (auto& x : w) get(in, x);
  struct Tree { uint16_t d; std::vector<uint16_t> f; std::vector<uint8_t> b; std::vector<float> th, g, lv; };
  std::vector<Tree> trees(nt);
  uint32_t D = 1;
  for (auto& tr : trees) {
    get(in, tr.d);
    tr.f.resize(tr.d); tr.b.resize(tr.d); tr.th.resize(tr.d); tr.g.resize(tr.d);
    for (uint32_t l = 0; l < tr.d; ++l) {
      get(in, tr.f[l]); get(in, tr.b[l]); get(in, tr.th[l]); get(in, tr.g[l]);
    }
    uint64_t nl = 0;
    get(in, nl);
    if (nl != (1ull << tr.d)) throw std::runtime_error("Corrupt TurboCat tree (leaf count)");
    tr.lv.resize(nl);
    in.read(reinterpret_cast<char*>(tr.lv.data()), 4 * nl);
    if (!in) throw std::runtime_error("Truncated TurboCat model file");
    D = std::max<uint32_t>(D, tr.d);
  }
  uint32_t F = static_cast<uint32_t>(edges.size());
  for (const auto& tr : trees)
    for (uint16_t f : tr.f) F = std::max<uint32_t>(F, f + 1u);
  drop_ensemble();
  T_ = static_cast<uint32_t>(nt);
  D_ = D;
  F_ = F;
  base_ = base;
  feat_.assign(static_cast<size_t>(T_) * D_, 0);
  bin_.assign(static_cast<size_t>(T_) * D_, 0);
  fthr_.assign(static_cast<size_t>(T_) * D_, 0.f);
  gain_.assign(static_cast<size_t>(T_) * D_, 0.f);
  depth_.assign(T_, 0);
  leaves_.assign(static_cast<size_t>(T_) << D_, 0.f);
  for (uint32_t t = 0; t < T_; ++t) {
    const Tree& tr = trees[t];
    depth_[t] = tr.d;
    for (uint32_t l = 0; l < tr.d; ++l) {
      const size_t k = static_cast<size_t>(t) * D_ + l;
      feat_[k] = tr.f[l]; bin_[k] = tr.b[l]; fthr_[k] = tr.th[l]; gain_[k] = tr.g[l];
    }
    std::memcpy(leaves_.data() + (static_cast<size_t>(t) << D_), tr.lv.data(), 4 * tr.lv.size());
  }
  weights_ = std::move(w);
  edges_ = std::move(edges);
  p_.task = static_cast<Task>(task);
  p_.loss = task == kRegression ? TC_LOSS_MSE : TC_LOSS_LOGLOSS;
  p_.learning_rate = lr;
  p_.max_depth = max_depth;
  p_.lambda_l2 = l2;
  fitted_ = true;
}

}  // namespace tcamd
