"""K-channel symmetric multiclass (A8 softmax gradients, A18 trees and
inference) on the GPU against the CPU oracle.

The reference has no symmetric multiclass (its multiclass grows regular
per-class trees, booster.cpp:1367-1395), so A18 is pinned by the FP64
oracle (tco_tree_build_mc_f64, tco_predict_mc_f64) and by its K = 1
reduction to the single-channel builder; A8 by the FP32 oracle, which is
bit-exact with the reference's CrossEntropyLoss (test_oracle.py).

Bit-exact claims: per-level (feature, bin) and per-row leaf indices (outside
FP64 near-ties), the K = 1 tree equal to tc_tree_build, per-row leaf indices
of inference; softmax gradients, and probabilities from given raw scores
(glibc's expf restated on the device).  Float claims: leaf values and raw
scores within 1e-5 relative to the FP64 oracle, normalized by the sum of
|terms|."""
import numpy as np
import pytest

from conftest import multiclass_data, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
NEAR_TIE = 1e-5


def compare_mc_tree(gpu, bins, g, h, edges, ne, params, oracle):
    """GPU K-channel tree vs the FP64 oracle; g, h row-major (n, K).
    Returns (levels compared, near_tie_hit)."""
    ref = oracle.tree_mc_f64(bins, g, h, edges, ne, params)
    lam = params.lambda_l2
    for d in range(max(gpu["depth"], ref["depth"])):
        assert d < gpu["depth"] and d < ref["depth"], (
            f"depth differs at level {d}: gpu {gpu['depth']} ref {ref['depth']}")
        if (gpu["feature"][d], gpu["bin"][d]) != (ref["feature"][d], ref["bin"][d]):
            assert ref["gap"][d] <= NEAR_TIE * max(1.0, abs(ref["gain"][d])), (
                f"level {d}: gpu ({gpu['feature'][d]},{gpu['bin'][d]}) ref "
                f"({ref['feature'][d]},{ref['bin'][d]}) gap {ref['gap'][d]}")
            return d, True
        assert gpu["fthr"][d] == ref["fthr"][d]
        assert abs(gpu["gain"][d] - ref["gain"][d]) <= 1e-4 * max(1.0, abs(ref["gain"][d]))
    assert np.array_equal(gpu["leaf_idx"], ref["leaf_idx"])
    L, K = 1 << ref["depth"], g.shape[1]
    li = ref["leaf_idx"]
    for c in range(K):
        absg = np.bincount(li, weights=np.abs(g[:, c]), minlength=L)
        hs = np.bincount(li, weights=h[:, c], minlength=L)
        scale = absg / (hs + lam + 1e-30)
        assert rel_err(gpu["leaves"][:, c], ref["leaves"][:, c], scale).max() < TOL, c
    return ref["depth"], False


@pytest.mark.parametrize("K", [2, 3, 10, 16])
def test_softmax_gradients(ctx, oracle, K):
    rng = np.random.default_rng(K)
    n = 20011
    p = (rng.standard_normal((n, K)) * 4).astype(np.float32)
    p[::97] = 0.0                                        # uniform rows
    p[1::101, 0] = 30.0                                  # saturated rows
    y = rng.integers(0, K, n).astype(np.float32)
    g, h = ctx.gradients_multi(y, p)
    rg, rh = oracle.gradients_mc_f32(y, p)
    # the reference's float order with glibc's expf (restated on the device)
    assert np.array_equal(g, rg) and np.array_equal(h, rh)
    assert (h >= 1e-6).all()
    assert np.allclose(g.sum(axis=1), 0.0, atol=1e-5)   # sum_c p_c - 1 = 0


@pytest.mark.parametrize("n,F,K,depth,lam,seed", [
    (6000, 5, 3, 4, 1.0, 11),
    (30011, 12, 5, 6, 1.0, 12),
    (20000, 9, 10, 5, 0.0, 13),
])
def test_tree_multi_teacher_forced(ctx, oracle, n, F, K, depth, lam, seed):
    from turbocat.capi import tree_params
    from oracle import tree_params as otp
    X, y = multiclass_data(n, F, K, seed)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal((n, K)) * 0.5).astype(np.float32)
    g, h = oracle.gradients_mc_f32(y, p)                # (n, K) float32
    gpu = ds.tree_multi(g.T, h.T, tree_params(max_depth=depth, lambda_l2=lam))
    levels, tie = compare_mc_tree(gpu, bins, g.astype(np.float64), h.astype(np.float64),
                                  edges, ne, otp(max_depth=depth, lambda_l2=lam), oracle)
    assert levels >= 1 and not tie


def test_tree_multi_k1_is_single_channel(ctx):
    """K = 1 through the K-channel entry point is the single-channel tree."""
    from turbocat.capi import tree_params
    X, y = multiclass_data(15000, 7, 2, 21)
    ds, _, _ = ctx.quantize_fit(X)
    rng = np.random.default_rng(5)
    g = rng.standard_normal(15000).astype(np.float32)
    h = rng.uniform(0.1, 1.0, 15000).astype(np.float32)
    a = ds.tree(g, h, tree_params(max_depth=6))
    b = ds.tree_multi(g[None, :], h[None, :], tree_params(max_depth=6))
    assert a["depth"] == b["depth"]
    for k in ("feature", "bin", "fthr", "gain", "leaf_idx"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["leaves"], b["leaves"][:, 0])


def random_mc_model(T, D, K, F, seed):
    rng = np.random.default_rng(seed)
    depth = np.full(T, D, np.uint16)
    depth[::5] = rng.integers(0, D, len(depth[::5]))
    feat = rng.integers(0, F, T * D).astype(np.uint16)
    thr = rng.standard_normal(T * D).astype(np.float32)
    leaves = (rng.standard_normal((T << D) * K) * 0.1).astype(np.float32)
    for t in range(T):
        leaves[((t << D) + (1 << int(depth[t]))) * K:((t + 1) << D) * K] = 0
        feat[t * D + int(depth[t]):(t + 1) * D] = 0
        thr[t * D + int(depth[t]):(t + 1) * D] = 0
    return dict(T=T, D=D, K=K, feature=feat, fthr=thr, bin=np.zeros(T * D, np.uint8),
                depth=depth, leaves=leaves, weight=np.full(T, 0.05, np.float32), base=0.0)


@pytest.mark.parametrize("T,D,K,F,n", [(1, 1, 2, 3, 7), (37, 6, 4, 15, 3001), (200, 8, 10, 40, 2049)])
def test_predict_multi(ctx, oracle, T, D, K, F, n):
    m = random_mc_model(T, D, K, F, seed=T + K)
    X, _ = multiclass_data(n, F, K, seed=T)
    X[::41, F // 2] = np.nan
    ens = ctx.ensemble(m, F)
    out, li = ens.predict(X, want_leaf_idx=True)
    ref, ab = oracle.predict_mc_f64(m, X)
    assert out.shape == (n, K)
    assert rel_err(out, ref, ab).max() < TOL
    # leaf indices are those of the single-output oracle on the same splits
    m1 = dict(m, K=1, leaves=m["leaves"][::K].copy())
    _, rli = oracle.predict_f32(m1, X, want_leaf_idx=True)
    assert np.array_equal(li, rli)
    p = ens.predict(X, transform=1)
    e = np.exp(ref - ref.max(axis=1, keepdims=True))
    assert np.allclose(p, e / e.sum(axis=1, keepdims=True), rtol=1e-5, atol=1e-6)
    # the softmax of the device's own raw scores: the reference's float order
    assert np.array_equal(p, oracle.transform_f32(out, K))
    # alone and in a bigger batch: identical bits
    assert np.array_equal(ens.predict(X[5:6]), out[5:6])


def test_softmax_fit(ctx, oracle):
    """Boosting with K-channel trees: the first tree is the oracle's tree on
    the oracle's gradients at the uniform start; train predictions agree
    with predicting the model; the fit learns the labels."""
    from turbocat.capi import train_params, TC_LOSS_SOFTMAX
    from oracle import tree_params as otp
    n, F, K, T, D, lr = 30000, 10, 4, 25, 5, 0.3
    X, y = multiclass_data(n, F, K, 31, sigma=1.0)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    m = ds.train(y, train_params(TC_LOSS_SOFTMAX, n_trees=T, learning_rate=lr, max_depth=D,
                                 lambda_l2=1.0, n_classes=K))
    assert m["base"] == 0.0
    g0, h0 = oracle.gradients_mc_f64(y, np.zeros((n, K), np.float32))
    ref = oracle.tree_mc_f64(bins, g0, h0, edges, ne, otp(max_depth=D, lambda_l2=1.0))
    d0 = int(m["depth"][0])
    # this seed's first tree has no near-tie at any level (best - second best
    # is 1.7e-4 of the gain at its closest), so it is always compared
    assert ref["gap"][:ref["depth"]].min() > NEAR_TIE * max(1.0, abs(ref["gain"][0]))
    assert d0 == ref["depth"]
    assert np.array_equal(m["feature"][:d0], ref["feature"])
    assert np.array_equal(m["bin"][:d0], ref["bin"])
    lv = m["leaves"][:(1 << D) * K].reshape(1 << D, K)[:1 << d0]
    assert rel_err(lv, ref["leaves"], np.abs(ref["leaves"]) + 1e-3).max() < 1e-4
    ens = ctx.ensemble(m, F)
    raw = ens.predict(X)
    tp = m["train_preds"].T                              # (n, K)
    r64, ab = oracle.predict_mc_f64(m, X)
    assert rel_err(tp, r64, ab).max() < TOL
    assert rel_err(raw, r64, ab).max() < TOL
    acc = (np.argmax(raw, axis=1) == y).mean()
    assert acc > 0.7, acc
    # thresholds from the training edges: bin > t <=> v >= edges[t] (no NaN here),
    # and both forms sum in the same order
    assert np.array_equal(ens.predict_binned(ds), raw)
