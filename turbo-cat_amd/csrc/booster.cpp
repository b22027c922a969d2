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
  for (auto& e : rens_) {
    if (e) tc_rensemble_destroy(e);
    e = nullptr;
  }
}

void Booster::fit(const float* X, const float* y, uint32_t n, uint32_t F, const float* Xv,
                  const float* yv, uint32_t nv) {
  if (n == 0) throw std::invalid_argument("X has no rows");
  if (F == 0) throw std::invalid_argument("X has no features");
  // the reference's multiclass loop always grows regular trees (booster.cpp:1367-1395)
  if (!p_.use_symmetric || p_.n_classes > 2) {
    fit_regular(X, y, n, F, Xv, yv, nv);
    return;
  }
  regular_ = false;
  K_ = 1;
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

// Regular trees: Booster::train's regular branch (booster.cpp:160-232,
// :467-493) or Booster::train_multiclass (:1263-1400), on the device
// (tc_rtrain_valid); the model is kept as the reference's Tree::nodes.
void Booster::fit_regular(const float* X, const float* y, uint32_t n, uint32_t F,
                          const float* Xv, const float* yv, uint32_t nv) {
  if (p_.max_depth < 1 || p_.max_depth > (p_.lossguide ? 11u : 12u))
    throw std::invalid_argument(p_.lossguide ? "max_depth must be in [1, 11] for lossguide trees"
                                             : "max_depth must be in [1, 12] for regular trees");
  if (p_.max_bins < 2 || p_.max_bins > 255) throw std::invalid_argument("max_bins must be in [2, 255]");
  if (!(p_.subsample > 0.0f && p_.subsample <= 1.0f))
    throw std::invalid_argument("subsample must be in (0, 1]");
  if (p_.max_leaves < 1) throw std::invalid_argument("max_leaves must be >= 1");
  const auto t0 = std::chrono::steady_clock::now();
  ctx_ = device_context(p_.device);
  drop_ensemble();
  fitted_ = false;
  const bool multi = p_.n_classes > 2;
  const uint32_t K = multi ? p_.n_classes : 1u;
  std::vector<float> edges(static_cast<size_t>(F) * 255);
  std::vector<uint32_t> ne(F);
  tc_dataset* ds = nullptr;
  check(tc_quantize_fit(ctx_, X, n, F, p_.max_bins, p_.wrap_bins ? 1 : 0, edges.data(), ne.data(),
                        &ds));
  tc_rtrain_params tp{};
  tp.loss = multi ? TC_LOSS_SOFTMAX : p_.loss;
  tp.n_iterations = p_.n_estimators;
  tp.learning_rate = p_.learning_rate;
  tp.tree.max_depth = p_.max_depth;
  tp.tree.max_leaves = p_.max_leaves;
  tp.tree.max_bins = p_.max_bins;
  tp.tree.lambda_l2 = p_.lambda_l2;
  tp.tree.min_child_weight = p_.min_child_weight;
  tp.tree.min_samples_leaf = p_.min_samples_leaf;
  tp.tree.min_split_gain = p_.min_split_gain;
  tp.tree.gamma = p_.gamma;
  tp.tree.max_delta_step = p_.max_delta_step;
  tp.tree.leaf_smooth = p_.leaf_smooth;
  tp.tree.learn_missing_direction = p_.learn_missing ? 1 : 0;
  tp.tree.grow_policy = p_.lossguide ? 1 : 0;
  tp.subsample = p_.subsample;
  tp.colsample_bytree = p_.colsample_bytree;
  tp.use_goss = p_.use_goss ? 1 : 0;
  tp.goss_top_rate = p_.goss_top_rate;
  tp.goss_other_rate = p_.goss_other_rate;
  tp.use_mvs = p_.use_mvs ? 1 : 0;
  tp.mvs_subsample = p_.mvs_subsample;
  tp.seed = p_.seed;
  tp.n_classes = multi ? K : 0u;
  tp.use_graph = p_.use_graph ? 1 : 0;
  const uint32_t iters = p_.n_estimators, T = iters * K;
  const size_t NC = (2u << p_.max_depth) - 1u, M = NC * T;
  std::vector<uint16_t> f16(M), dep(T);
  std::vector<uint8_t> b8(M), fl(M);
  std::vector<uint32_t> l32(M), r32(M), nn(T), nl(T);
  std::vector<float> v32(M), g32(M);
  tc_rmodel m{};
  m.feature = f16.data(); m.bin = b8.data(); m.flags = fl.data(); m.left = l32.data();
  m.right = r32.data(); m.value = v32.data(); m.gain = g32.data(); m.n_nodes = nn.data();
  m.n_leaves = nl.data(); m.depth = dep.data();
  uint32_t built = iters;
  int32_t rc;
  if (Xv != nullptr && nv > 0) {
    tc_dataset* vds = nullptr;  // Dataset::apply_bins with the training edges (bindings.cpp:261-279)
    rc = tc_dataset_create(ctx_, Xv, nv, F, edges.data(), ne.data(), p_.max_bins, &vds);
    if (rc == 0) {
      tc_train_history h{};
      rc = tc_rtrain_valid(ds, y, &tp, vds, yv, p_.early_stopping_rounds, &m, &h);
      built = h.n_trees_built;
      tc_dataset_destroy(vds);
    }
  } else {
    rc = tc_rtrain(ds, y, &tp, &m);
  }
  tc_dataset_destroy(ds);
  check(rc);
  const uint32_t Tb = built * K;  // the stopping iteration's trees are kept (booster.cpp:1386-1430)
  roff_.assign(1, 0u);
  rfeat_.clear(); rbin_.clear(); rflags_.clear(); rleft_.clear(); rright_.clear();
  rvalue_.clear(); rgain_.clear(); rclass_.clear(); rdepth_.clear(); rnleaves_.clear();
  for (uint32_t t = 0; t < Tb; ++t) {
    const size_t o = NC * t;
    for (uint32_t k = 0; k < nn[t]; ++k) {
      rfeat_.push_back(f16[o + k]); rbin_.push_back(b8[o + k]); rflags_.push_back(fl[o + k]);
      rleft_.push_back(l32[o + k]); rright_.push_back(r32[o + k]); rvalue_.push_back(v32[o + k]);
      rgain_.push_back(g32[o + k]);
    }
    roff_.push_back(static_cast<uint32_t>(rfeat_.size()));
    rclass_.push_back(t % K);  // add_tree_for_class order (booster.cpp:1386)
    rdepth_.push_back(dep[t]);
    rnleaves_.push_back(static_cast<uint16_t>(nl[t]));
  }
  regular_ = true;
  K_ = K;
  T_ = Tb;
  D_ = p_.max_depth;
  F_ = F;
  base_ = m.base;
  weights_.assign(Tb, p_.learning_rate);
  edges_.assign(F, {});
  for (uint32_t f = 0; f < F; ++f)
    edges_[f].assign(edges.begin() + static_cast<size_t>(f) * 255,
                     edges.begin() + static_cast<size_t>(f) * 255 + ne[f]);
  fitted_ = true;
  fit_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The device ensemble of the regular trees for a binning rule: 0 the
// multiclass path's Dataset::apply_bins, 1 predict_raw_fast's.
tc_rensemble* Booster::regular_ensemble(int rule) {
  if (!fitted_) throw std::runtime_error("Model not fitted. Call fit() first.");
  if (rens_[rule]) return rens_[rule];
  ctx_ = device_context(p_.device);
  std::vector<float> e(static_cast<size_t>(F_) * 255, 0.f);
  std::vector<uint32_t> ne(F_, 0);
  for (uint32_t f = 0; f < F_ && f < edges_.size(); ++f) {
    ne[f] = static_cast<uint32_t>(edges_[f].size());
    std::copy(edges_[f].begin(), edges_[f].end(), e.begin() + static_cast<size_t>(f) * 255);
  }
  // base_predictions_multiclass_ are zeros (booster.cpp:1276-1278)
  std::vector<float> base(K_, K_ > 1 ? 0.f : base_);
  const bool any = !rfeat_.empty();
  check(tc_rensemble_create(ctx_, T_, K_, roff_.data(), any ? rfeat_.data() : nullptr,
                            any ? rbin_.data() : nullptr, any ? rflags_.data() : nullptr,
                            any ? rleft_.data() : nullptr, any ? rright_.data() : nullptr,
                            any ? rvalue_.data() : nullptr, weights_.data(), rclass_.data(),
                            base.data(), e.data(), ne.data(), F_, rule, &rens_[rule]));
  return rens_[rule];
}

void Booster::ensure_ensemble() {
  if (!fitted_) throw std::runtime_error("Model not fitted. Call fit() first.");
  if (ens_) return;
  ctx_ = device_context(p_.device);
  check(tc_ensemble_create(ctx_, T_, std::max(1u, D_), feat_.data(), fthr_.data(), bin_.data(),
                           depth_.data(), leaves_.data(), weights_.data(), base_, F_, &ens_));
}

void Booster::predict_raw(const float* X, uint32_t n, uint32_t F, float* out) {
  if (!fitted_) throw std::runtime_error("Model not fitted. Call fit() first.");
  if (F != F_)
    throw std::invalid_argument("X has " + std::to_string(F) + " features, the model " +
                                std::to_string(F_));
  if (regular_) {
    check(tc_rpredict(regular_ensemble(K_ > 1 ? 0 : 1), X, n, F, 0, out));
    return;
  }
  ensure_ensemble();
  check(tc_predict(ens_, X, n, F, 0, out, nullptr));
}

void Booster::predict_proba(const float* X, uint32_t n, uint32_t F, float* out) {
  if (!fitted_) throw std::runtime_error("Model not fitted. Call fit() first.");
  if (F != F_)
    throw std::invalid_argument("X has " + std::to_string(F) + " features, the model " +
                                std::to_string(F_));
  if (regular_) {
    check(tc_rpredict(regular_ensemble(K_ > 1 ? 0 : 1), X, n, F, 1, out));
    return;
  }
  ensure_ensemble();
  check(tc_predict(ens_, X, n, F, 1, out, nullptr));
}

std::vector<float> Booster::feature_importance() const {
  if (!fitted_) throw std::runtime_error("Model not fitted");
  std::vector<float> imp(256, 0.f);
  if (regular_) {  // Tree::feature_importance summed and normalised (tree.cpp:1015-1025, :1414-1433)
    for (size_t k = 0; k < rfeat_.size(); ++k)
      if (!(rflags_[k] & 1u) && rgain_[k] > 0.f && rfeat_[k] < 256) imp[rfeat_[k]] += rgain_[k];
    float sum = 0.f;
    for (float v : imp) sum += v;
    if (sum > 0.f)
      for (float& v : imp) v /= sum;
    return imp;
  }
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
  put(o, static_cast<uint32_t>(K_ > 1 ? K_ : 2u));      // n_classes
  put(o, p_.learning_rate);
  put(o, static_cast<uint16_t>(p_.max_depth));
  put(o, p_.lambda_l2);
  put(o, static_cast<uint8_t>(regular_ ? 0 : 1));       // use_symmetric
  put(o, static_cast<uint64_t>(edges_.size()));
  for (const auto& e : edges_) {
    put(o, static_cast<uint64_t>(e.size()));
    if (!e.empty()) o.write(reinterpret_cast<const char*>(e.data()), 4 * e.size());
  }
  put(o, static_cast<uint64_t>(T_));
  for (uint32_t t = 0; t < T_; ++t) put(o, weights_[t]);
  if (regular_) {  // Tree::save (tree.cpp:1748-1774): per-class trees carry n_classes 1
    for (uint32_t t = 0; t < T_; ++t) {
      const uint32_t a = roff_[t], b = roff_[t + 1];
      put(o, rdepth_[t]);
      put(o, static_cast<uint32_t>(1));
      put(o, b - a);
      put(o, rnleaves_[t]);
      put(o, static_cast<uint16_t>(p_.max_depth));
      put(o, p_.lambda_l2);
      for (uint32_t k = a; k < b; ++k) {
        put(o, rfeat_[k]);
        put(o, rbin_[k]);
        put(o, rflags_[k]);
        put(o, static_cast<uint16_t>(rleft_[k]));   // TreeIndex (types.hpp:47)
        put(o, static_cast<uint16_t>(rright_[k]));
        put(o, rvalue_[k]);
        put(o, rgain_[k]);
      }
    }
    if (!o) throw std::runtime_error("Write failed: " + path);
    return;
  }
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
  if (version < 2)
    throw std::runtime_error("TurboCat model files before version 2 carry no task and are not supported");
  uint64_t nt = 0;
  get(in, nt);
  std::vector<float> w(nt);
  for (auto& x : w) get(in, x);
  if (version < 3 || !sym) {  // regular trees (booster.cpp:1180-1194, Tree::load tree.cpp:1795-1845)
    const uint32_t K = (task == kMulticlass && n_classes > 2) ? n_classes : 1u;
    std::vector<uint32_t> off(1, 0u), le, ri, cls;
    std::vector<uint16_t> fe, dp, nls;
    std::vector<uint8_t> bi, fl;
    std::vector<float> va, ga;
    for (uint64_t t = 0; t < nt; ++t) {
      uint16_t d = 0, nlv = 0, md = 0;
      uint32_t tk = 1, nn = 0;
      float lam = 0.f;
      get(in, d); get(in, tk); get(in, nn); get(in, nlv); get(in, md); get(in, lam);
      if (nn > 65535u) throw std::runtime_error("Corrupt TurboCat tree (node count)");
      for (uint32_t k = 0; k < nn; ++k) {
        uint16_t f = 0, l = 0, r = 0;
        uint8_t b = 0, flags = 0;
        float v = 0.f, g = 0.f;
        get(in, f); get(in, b); get(in, flags); get(in, l); get(in, r); get(in, v); get(in, g);
        if (!(flags & 1u) && (l >= nn || r >= nn || l <= k || r <= k))
          throw std::runtime_error("Corrupt TurboCat tree (child ids)");
        fe.push_back(f); bi.push_back(b); fl.push_back(flags); le.push_back(l); ri.push_back(r);
        va.push_back(v); ga.push_back(g);
      }
      if (tk > 1) {  // multiclass leaf values and node map (not written by per-class trees)
        uint32_t c = 0;
        get(in, c);
        std::vector<float> skip(c);
        if (c) in.read(reinterpret_cast<char*>(skip.data()), 4ull * c);
        get(in, c);
        std::vector<uint16_t> skip2(c);
        if (c) in.read(reinterpret_cast<char*>(skip2.data()), 2ull * c);
        if (!in) throw std::runtime_error("Truncated TurboCat model file");
      }
      off.push_back(static_cast<uint32_t>(fe.size()));
      // the class of a tree is not stored; trees are written iteration-major,
      // class-minor, so tree t belongs to class t mod K
      cls.push_back(static_cast<uint32_t>(t % K));
      dp.push_back(d);
      nls.push_back(nlv);
    }
    uint32_t F = static_cast<uint32_t>(edges.size());
    for (size_t k = 0; k < fe.size(); ++k)
      if (!(fl[k] & 1u)) F = std::max<uint32_t>(F, fe[k] + 1u);
    drop_ensemble();
    regular_ = true;
    K_ = K;
    T_ = static_cast<uint32_t>(nt);
    D_ = max_depth;
    F_ = F;
    base_ = base;
    roff_ = std::move(off); rfeat_ = std::move(fe); rbin_ = std::move(bi); rflags_ = std::move(fl);
    rleft_ = std::move(le); rright_ = std::move(ri); rvalue_ = std::move(va); rgain_ = std::move(ga);
    rclass_ = std::move(cls); rdepth_ = std::move(dp); rnleaves_ = std::move(nls);
    weights_ = std::move(w);
    edges_ = std::move(edges);
    edges_.resize(F);
    p_.task = static_cast<Task>(task);
    p_.loss = task == kRegression ? TC_LOSS_MSE : (K > 1 ? TC_LOSS_SOFTMAX : TC_LOSS_LOGLOSS);
    p_.n_classes = K > 1 ? K : 2u;
    p_.use_symmetric = false;
    p_.learning_rate = lr;
    p_.max_depth = max_depth;
    p_.lambda_l2 = l2;
    fitted_ = true;
    return;
  }
  if (task == kMulticlass && n_classes > 2)
    throw std::runtime_error("Multiclass symmetric-tree TurboCat models are not supported by this backend");
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
  regular_ = false;
  K_ = 1;
  p_.use_symmetric = true;
  p_.task = static_cast<Task>(task);
  p_.loss = task == kRegression ? TC_LOSS_MSE : TC_LOSS_LOGLOSS;
  p_.learning_rate = lr;
  p_.max_depth = max_depth;
  p_.lambda_l2 = l2;
  fitted_ = true;
}

}  // namespace tcamd
