"""Pinning the CPU oracle (tests run without a GPU).

Each oracle function is checked against something other than itself:
  * known-answer tests derived from the reference source (8-row toy, the
    quantizer's uint8 wrap, threshold semantics);
  * brute force on tiny inputs (exhaustive split search written from the
    gain formula, counting histograms);
  * the golden vectors in tests/golden/ written by the reference's own
    hot-path code (tests/golden/make_golden.py), bit for bit;
  * the compiled reference itself (oracle/_ref/libtcref.so) where it is
    present, bit for bit.
"""
import glob
import itertools
import os

import numpy as np
import pytest

from conftest import binary_data, regression_data

from oracle import LOSS_LOGLOSS, LOSS_MSE, RefLib, tree_params

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
GOLDEN = sorted(p for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz"))
                if not os.path.basename(p).startswith(("ce_", "loss_", "pred_")))
HAVE_REF = RefLib.available()


# ---------------------------------------------------------------- known answers

def test_kat_eight_rows_tree(oracle):
    X = np.array([-2, -1, 1, 2, -1.5, -0.5, 0.5, 1.5], np.float32).reshape(8, 1)
    y = np.array([0, 0, 1, 1, 0, 0, 1, 1], np.float32)
    bins, edges, ne = oracle.quantize(X)
    assert list(bins[0]) == [0, 2, 5, 7, 1, 3, 4, 6]
    assert list(edges[0, :ne[0]]) == [-1.5, -1, -0.5, 0.5, 1, 1.5, 2]
    base = oracle.init_prediction(LOSS_LOGLOSS, y)
    assert base == 0.0
    g, h = oracle.gradients_f32(LOSS_LOGLOSS, y, np.zeros(8, np.float32))
    assert list(g) == [0.5, 0.5, -0.5, -0.5, 0.5, 0.5, -0.5, -0.5] and (h == 0.25).all()
    for build in (oracle.tree_f32, oracle.tree_f64):
        t = build(bins, g, h, edges, ne, tree_params(3, lambda_l2=1.0, min_child_weight=0.01))
        assert t["depth"] == 3
        assert list(t["feature"]) == [0, 0, 0] and list(t["bin"]) == [3, 3, 3]
        assert list(t["gain"]) == [2.0, 0.0, 0.0]           # zero-gain splits applied
        assert list(t["fthr"]) == [0.5, 0.5, 0.5]
        assert list(t["leaves"]) == [-1, 0, 0, 0, 0, 0, 0, 1]
        assert list(t["leaf_idx"]) == [0, 0, 7, 7, 0, 0, 7, 7]


def test_kat_level0_gains(oracle):
    """Level-0 gains of the toy by bin: 0.1455, 0.5333, 1.1429, 2.0, ... (closed form
    0.5 (GL^2/(HL+1) + GR^2/(HR+1)))."""
    hist = np.zeros((1, 1, 256, 2), np.float64)
    order = [0, 0, 0, 0, 1, 1, 1, 1]                 # y by bin 0..7 for the toy
    for b, yy in enumerate(order):
        hist[0, 0, b] = [0.5 - yy, 0.25]
    for b, exp in enumerate([0.145455, 0.533333, 1.142857, 2.0, 1.142857, 0.533333, 0.145455]):
        h1 = hist.copy()
        h1[0, 0, b + 1:] = 0
        gl, hl = h1[0, 0, :, 0].sum(), h1[0, 0, :, 1].sum()
        gr, hr = -gl, 2.0 - hl
        assert abs(0.5 * (gl * gl / (hl + 1) + gr * gr / (hr + 1)) - exp) < 1e-6
    f, b, gain, second = oracle.split_f64(hist, [0.0], [2.0], [8], lam=1.0, is_root=True)
    assert (f, b) == (0, 3) and abs(gain - 2.0) < 1e-12 and abs(second - 1.142857142857) < 1e-9


@pytest.mark.parametrize("U,A", [(300, 44), (256, 0), (257, 1), (511, 255), (255, 255), (2, 2)])
def test_kat_quantizer_wrap(oracle, U, A):
    """A = min(255, U mod 256) bins; A - 1 edges at unique[b*U/A] (dataset.cpp:308-318)."""
    vals = np.arange(U, dtype=np.float32)
    X = np.concatenate([vals, vals[:17]]).reshape(-1, 1)       # duplicates do not count
    bins, edges, ne = oracle.quantize(X)
    assert ne[0] == max(A - 1, 0)
    if A > 1:
        assert list(edges[0, :ne[0]]) == [float(b * U // A) for b in range(1, A)]
    assert bins.max() == max(A - 1, 0)
    # the fixed mode keeps min(255, U) bins
    _, _, ne2 = oracle.quantize(X, wrap=False)
    assert ne2[0] == min(255, U) - 1


def test_kat_threshold_semantics(oracle):
    """bin > t  <=>  v >= edges[t] (upper_bound binning), NaN -> bin 255 goes left."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((3000, 1)).astype(np.float32)
    bins, edges, ne = oracle.quantize(X)
    for t in range(ne[0]):
        assert np.array_equal(bins[0] > t, X[:, 0] >= edges[0, t])
    Xn = np.array([[np.nan], [edges[0, 5]], [np.nextafter(edges[0, 5], -np.inf)]], np.float32)
    bn = oracle.apply_bins(Xn, edges, ne)
    assert list(bn[0]) == [255, 6, 5]


def test_kat_all_nan_and_constant(oracle):
    X = np.full((50, 2), np.nan, np.float32)
    X[:, 1] = 3.0
    bins, edges, ne = oracle.quantize(X)
    assert ne[0] == 1 and edges[0, 0] == 0.0 and (bins[0] == 255).all()
    assert ne[1] == 0 and (bins[1] == 0).all()


# ---------------------------------------------------------------- brute force

def brute_split(hist, G, H, C, lam, is_root, max_bins=255):
    """Exhaustive search written directly from the gain formula."""
    best, arg = -1e30, (0, 0)
    nn, F = hist.shape[:2]
    for f in range(F):
        for b in range(max_bins - 1):
            total, any_ = 0.0, False
            for n in range(nn):
                if not is_root and C[n] == 0:
                    continue
                gl, hl = hist[n, f, :b + 1, 0].sum(), hist[n, f, :b + 1, 1].sum()
                gr, hr = G[n] - gl, H[n] - hl
                if hl < 1e-10 or hr < 1e-10:
                    continue
                pg = G[n] ** 2 / (H[n] + lam) if H[n] > 1e-10 else 0.0
                total += 0.5 * (gl * gl / (hl + lam) + gr * gr / (hr + lam) - pg)
                any_ = True
            if is_root and not any_:
                continue
            if total > best:
                best, arg = total, (f, b)
    return arg, best


@pytest.mark.parametrize("seed", range(6))
def test_split_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    nn = [1, 2, 4][seed % 3]
    F, nb = 3, 9
    hist = np.zeros((nn, F, 256, 2))
    hist[..., :nb, 0] = rng.integers(-3, 4, (nn, F, nb))       # small integers: exact sums
    hist[..., :nb, 1] = rng.integers(0, 3, (nn, F, nb))
    for n in range(nn):                                        # same node totals on every feature
        hist[n, 1:] = hist[n, 0][None]
        hist[n, 1:, :nb] = hist[n, 0, rng.permutation(nb)][None]
    G = hist[:, 0, :, 0].sum(1)
    H = hist[:, 0, :, 1].sum(1)
    C = (H > 0).astype(np.uint32) * 5
    for lam in (0.0, 1.0):
        f, b, gain, _ = oracle.split_f64(hist, G, H, C, max_bins=12, lam=lam, is_root=nn == 1)
        (ef, eb), eg = brute_split(hist, G, H, C, lam, nn == 1, max_bins=12)
        assert (f, b) == (ef, eb) and abs(gain - eg) < 1e-9


def test_hist_brute_force_and_sibling_identity(oracle):
    X, _ = regression_data(700, 4, 3, low_card_col=1)
    bins, _, _ = oracle.quantize(X)
    rng = np.random.default_rng(1)
    g = rng.integers(-8, 8, 700).astype(np.float32)            # integers: exact in float32
    h = rng.integers(1, 4, 700).astype(np.float32)
    rows = np.sort(rng.choice(700, 250, replace=False)).astype(np.uint32)
    gh, cnt = oracle.hist_f32(bins, g, h, rows)
    exp = np.zeros((4, 256, 2), np.float32)
    for r in rows:
        for f in range(4):
            exp[f, bins[f, r]] += [g[r], h[r]]
    assert np.array_equal(gh, exp)
    assert (cnt.sum(1) == 250).all()
    full, _ = oracle.hist_f32(bins, g, h)
    rest = np.setdiff1d(np.arange(700), rows).astype(np.uint32)
    other, _ = oracle.hist_f32(bins, g, h, rest)
    assert np.array_equal(full - gh, other)                    # larger = parent - smaller


def test_tree_f32_vs_f64_and_k1_reduction(oracle):
    X, y = regression_data(3000, 6, 5)
    bins, edges, ne = oracle.quantize(X)
    g, h = oracle.gradients_f32(LOSS_MSE, y, np.zeros(3000, np.float32))
    p = tree_params(5, lambda_l2=1.0)
    a = oracle.tree_f32(bins, g, h, edges, ne, p)
    b = oracle.tree_f64(bins, g, h, edges, ne, p)
    c = oracle.tree_mc_f64(bins, g.astype(np.float64)[:, None], h.astype(np.float64)[:, None],
                           edges, ne, p)
    assert np.array_equal(a["feature"], b["feature"]) and np.array_equal(a["bin"], b["bin"])
    assert np.array_equal(b["feature"], c["feature"]) and np.array_equal(b["bin"], c["bin"])
    assert np.allclose(a["leaves"], b["leaves"], rtol=1e-5)
    assert np.array_equal(b["leaves"], c["leaves"][:, 0])
    assert np.array_equal(a["leaf_idx"], oracle.leaf_index_binned(bins, a["feature"], a["bin"]))


def test_predict_f32_vs_f64(oracle):
    X, y = regression_data(2000, 5, 8)
    bins, edges, ne = oracle.quantize(X)
    m, tp = oracle.fit_f32(bins, edges, ne, y, LOSS_MSE, 23, 0.1, tree_params(6, lambda_l2=0.0))
    p32 = oracle.predict_f32(m, X)
    p64, ab = oracle.predict_f64(m, X)
    assert (np.abs(p32 - p64) / np.maximum(ab, 1e-30)).max() < 1e-6
    # raw-float predict on the training rows == binned training predictions
    assert np.allclose(p32, tp, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_matches_golden(oracle, path):
    z = np.load(path, allow_pickle=False)
    n, F, seed, loss, T, D = (int(v) for v in z["meta"])
    kind = str(z["kind"])
    X, y = regression_data(n, F, seed, low_card_col=2) if kind == "reg" else binary_data(n, F, seed)
    bins, edges, ne = oracle.quantize(X)
    assert np.array_equal(bins, z["bins"]) and np.array_equal(ne, z["n_edges"])
    assert np.array_equal(edges, z["edges"])
    g, h = oracle.gradients_f32(loss, y, z["p"])
    assert np.array_equal(g, z["g"]) and np.array_equal(h, z["h"])
    lam, lr = float(z["lam"]), float(z["lr"])
    t = oracle.tree_f32(bins, g, h, edges, ne, tree_params(D, lambda_l2=lam))
    assert np.array_equal(t["feature"], z["t_feature"]) and np.array_equal(t["bin"], z["t_bin"])
    assert np.array_equal(t["gain"], z["t_gain"]) and np.array_equal(t["fthr"], z["t_fthr"])
    assert np.array_equal(t["leaves"], z["t_leaves"]) and np.array_equal(t["leaf_idx"], z["t_leaf_idx"])
    m, tp = oracle.fit_f32(bins, edges, ne, y, loss, T, lr, tree_params(D, lambda_l2=lam))
    for k in ("feature", "bin", "fthr", "gain", "depth", "leaves"):
        assert np.array_equal(m[k], z["m_" + k]), k
    assert np.float32(m["base"]) == z["m_base"]
    assert np.array_equal(tp, z["train_preds"])
    assert np.array_equal(oracle.predict_f32(m, z["Xp"]), z["pred"])


def test_ce_gradients_match_golden(oracle):
    """A8: the FP32 restatement equals CrossEntropyLoss::compute_multiclass_gradients
    (loss.cpp:287-318) bit for bit on the reference's own outputs."""
    z = np.load(os.path.join(GOLDEN_DIR, "ce_grad.npz"), allow_pickle=False)
    g, h = oracle.gradients_mc_f32(z["y"], z["p"])
    assert np.array_equal(g, z["g"]) and np.array_equal(h, z["h"])
    g64, h64 = oracle.gradients_mc_f64(z["y"], z["p"])
    assert np.abs(g64 - z["g"]).max() < 1e-6 and np.abs(h64 - z["h"]).max() < 1e-6


def golden_pred_model():
    z = np.load(os.path.join(GOLDEN_DIR, "pred_tail.npz"), allow_pickle=False)
    n, F, seed, T, D = (int(v) for v in z["meta"])
    m = dict(T=T, D=D, feature=z["m_feature"], bin=z["m_bin"], fthr=z["m_fthr"],
             depth=z["m_depth"], leaves=z["m_leaves"], weight=np.full(T, z["lr"], np.float32),
             base=float(z["m_base"]))
    return z, m


def test_predict_tail_rows_match_golden(oracle):
    """A16/A17 in calls of 517, 13 and 7 rows: the reference's raw-float and
    binned outputs (tests/golden/make_golden.py) bit for bit, including the
    rows past the last full 8-row block and NaN rows; the tail order differs
    from the block order, so a restatement that used the block order for every
    row fails here."""
    z, m = golden_pred_model()
    for r in (517, 13, 7):
        assert np.array_equal(oracle.predict_f32(m, z[f"X{r}"]), z[f"pred{r}"]), r
        assert np.array_equal(oracle.predict_binned_f32(m, z[f"bins{r}"]), z[f"pbin{r}"]), r
    # the orders do differ on these rows
    full = oracle.predict_f32(m, np.concatenate([z["X517"], z["X517"][:3]]))
    assert not np.array_equal(full[512:517], z["pred517"][512:])


def test_predict_binned_equals_float_on_training_bins(oracle):
    """bin > t <=> v >= edges[t] and NaN -> bin 255 (both left): over the bins
    of X the binned prediction is the float one, bit for bit, for every call
    of 16 rows or more (below 16 the binned path adds tree by tree)."""
    X, y = regression_data(3000, 6, 12)
    X[::7, 2] = np.nan
    bins, edges, ne = oracle.quantize(X)
    m, _ = oracle.fit_f32(bins, edges, ne, y, LOSS_MSE, 19, 0.1, tree_params(5, lambda_l2=0.0))
    for r in (3000, 517, 16, 21):
        b = oracle.apply_bins(X[:r], edges, ne)
        assert np.array_equal(oracle.predict_binned_f32(m, b), oracle.predict_f32(m, X[:r])), r
    b = oracle.apply_bins(X[:13], edges, ne)
    seq = oracle.predict_binned_f32(m, b)
    assert np.allclose(seq, oracle.predict_f32(m, X[:13]), rtol=1e-6)


def test_predict_mc_k1_reduction(oracle):
    """The K-channel FP64 prediction with K = 1 is the single-channel one, and
    with K > 1 each class is the single-channel prediction of its leaves."""
    rng = np.random.default_rng(9)
    T, D, K, F, n = 23, 5, 3, 6, 700
    X = rng.standard_normal((n, F)).astype(np.float32)
    m = dict(T=T, D=D, K=K, feature=rng.integers(0, F, T * D).astype(np.uint16),
             fthr=rng.standard_normal(T * D).astype(np.float32),
             depth=rng.integers(0, D + 1, T).astype(np.uint16),
             leaves=rng.standard_normal((T << D) * K).astype(np.float32),
             weight=np.full(T, 0.1, np.float32), base=0.25)
    out, _ = oracle.predict_mc_f64(m, X)
    for c in range(K):
        mc = dict(m, K=1, leaves=np.ascontiguousarray(m["leaves"][c::K]))
        ref, _ = oracle.predict_f64(mc, X)
        assert np.allclose(out[:, c], ref, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- compiled reference

@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
@pytest.mark.parametrize("n,F,seed,D", [(2053, 5, 1, 6), (9000, 17, 2, 7), (16, 4, 3, 3)])
def test_oracle_matches_reference(oracle, n, F, seed, D):
    def body():
        R = RefLib()
        X, y = regression_data(n, F, seed, low_card_col=min(3, F - 1))
        X[::31, 0] = X[0, 0]                                   # duplicates
        yb = (y > np.median(y)).astype(np.float32)
        for loss, yy, lam in ((LOSS_MSE, y, 0.0), (LOSS_LOGLOSS, yb, 1.0)):
            ds = R.dataset(X, yy)
            rb, re, rn = ds.bins()
            ob, oe, on = oracle.quantize(X)
            assert np.array_equal(rb, ob) and np.array_equal(rn, on) and np.array_equal(re, oe)
            rm, rtp, _ = ds.fit(loss, 9, D, 0.1, lam, 1.0)
            om, otp = oracle.fit_f32(ob, oe, on, yy, loss, 9, 0.1, tree_params(D, lambda_l2=lam))
            for k in ("feature", "bin", "fthr", "gain", "depth", "leaves"):
                assert np.array_equal(rm[k], om[k]), k
            assert np.array_equal(rtp, otp)
            Xp = np.random.default_rng(seed).standard_normal((64 * 8, F)).astype(np.float32)
            assert np.array_equal(R.predict(rm, Xp), oracle.predict_f32(om, Xp))
    RefLib.isolated(body)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
def test_predict_matches_reference_every_call_size(oracle):
    """Raw-float (predict_batch_raw_fast) and binned (SymmetricEnsemble::
    predict_batch) predictions of the compiled reference, bit for bit, for call
    sizes around the 8-row blocks, the 16-row binned fallback and the 256-row
    transpose threshold, and for an ensemble whose trees all have depth 0."""
    def body():
        R = RefLib()
        X, y = regression_data(2400, 6, 21)
        X[::11, 1] = np.nan
        ds = R.dataset(X, y)
        for T in (23, 8, 1):
            m, _, _ = ds.fit(LOSS_MSE, T, 5, 0.1, 0.0, 1.0)
            for r in (1, 7, 8, 13, 15, 16, 17, 255, 256, 263, 517):
                Xp = X[:r]
                assert np.array_equal(R.predict(m, Xp), oracle.predict_f32(m, Xp)), (T, r)
                dr = R.dataset(Xp, y[:r])
                bins, _, _ = dr.bins()
                assert np.array_equal(dr.predict_binned(m), oracle.predict_binned_f32(m, bins)), (T, r)
        flat = dict(T=5, D=3, feature=np.zeros(15, np.uint16), fthr=np.zeros(15, np.float32),
                    bin=np.zeros(15, np.uint8), depth=np.zeros(5, np.uint16),
                    leaves=np.random.default_rng(0).standard_normal(40).astype(np.float32),
                    weight=np.full(5, 0.1, np.float32), base=0.3)
        for r in (300, 17, 5):
            assert np.array_equal(R.predict(flat, X[:r]), oracle.predict_f32(flat, X[:r]))
    RefLib.isolated(body)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
@pytest.mark.parametrize("K", [2, 3, 10])
def test_ce_gradients_match_reference(oracle, K):
    rng = np.random.default_rng(K)
    p = (rng.standard_normal((3001, K)) * 3).astype(np.float32)
    y = rng.integers(0, K, 3001).astype(np.float32)
    rg, rh = RefLib.isolated(lambda: RefLib().gradients_mc(y, p))
    g, h = oracle.gradients_mc_f32(y, p)
    assert np.array_equal(g, rg) and np.array_equal(h, rh)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
@pytest.mark.parametrize("K", [1, 2, 3, 10])
def test_transform_matches_reference(oracle, K):
    """Probabilities from raw scores: the oracle's sigmoid / softmax against
    LogLoss::transform_prediction and CrossEntropyLoss::transform_to_proba of
    the compiled reference, bit for bit (scores past +-88, equal scores and
    saturated rows included)."""
    rng = np.random.default_rng(40 + K)
    raw = (rng.standard_normal((4001, K)) * 6).astype(np.float32)
    raw[::37] = 0.0
    raw[1::53, 0] = 95.0
    raw[2::59, -1] = -110.0
    ref = RefLib.isolated(lambda: RefLib().transform(raw, K))
    out = oracle.transform_f32(raw, K)
    assert np.array_equal(out, ref)


# --------------------------------------------------------------- A19: losses, early stopping
@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
@pytest.mark.parametrize("loss", [0, 1])
def test_compute_loss_matches_reference(oracle, loss):
    """Loss::compute_loss restated (mean per-row loss, float, row order) equals the
    compiled reference's on one OpenMP thread, bit for bit; the FP64 form is
    within float rounding."""
    rng = np.random.default_rng(31 + loss)
    n = 5003
    p = (rng.standard_normal(n) * 3).astype(np.float32)
    p[:5] = [0.0, 40.0, -40.0, 1e-8, -17.0]  # clip range of the sigmoid
    y = (rng.random(n) < 0.4).astype(np.float32) if loss else rng.standard_normal(n).astype(np.float32)
    a = oracle.compute_loss_f32(loss, y, p)
    b = RefLib.isolated(lambda: RefLib().compute_loss(loss, y, p))
    assert a == b
    assert abs(oracle.compute_loss_f64(loss, y, p) - a) <= 1e-5 * abs(a)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
def test_compute_loss_mc_matches_reference(oracle):
    rng = np.random.default_rng(5)
    for K in (2, 3, 10):
        n = 2001
        p = (rng.standard_normal((n, K)) * 2).astype(np.float32)
        y = rng.integers(0, K, n).astype(np.float32)
        assert oracle.compute_loss_mc_f32(y, p) == RefLib.isolated(lambda: RefLib().compute_loss_mc(y, p))


def test_compute_loss_matches_golden(oracle):
    """The compiled reference's Loss::compute_loss values (tests/golden/
    make_golden.py), for hosts without the reference tree."""
    z = np.load(os.path.join(GOLDEN_DIR, "loss_values.npz"), allow_pickle=False)
    assert np.float32(oracle.compute_loss_f32(0, z["yr"], z["p"])) == z["mse"]
    assert np.float32(oracle.compute_loss_f32(1, z["yb"], z["p"])) == z["logloss"]
    assert np.float32(oracle.compute_loss_mc_f32(z["yk"], z["pk"])) == z["ce"]


def test_compute_loss_closed_forms(oracle):
    """MSE of a constant offset d is d^2; LogLoss at p = 0 is ln 2 for any labels;
    a confident wrong prediction is capped by the 1e-7 clip at -ln(1e-7)."""
    y = np.linspace(-1, 1, 64).astype(np.float32)
    assert oracle.compute_loss_f32(0, y, y + np.float32(0.5)) == pytest.approx(0.25, rel=1e-6)
    yb = (np.arange(64) % 2).astype(np.float32)
    assert oracle.compute_loss_f32(1, yb, np.zeros(64, np.float32)) == pytest.approx(np.log(2), rel=1e-6)
    worst = oracle.compute_loss_f32(1, np.ones(4, np.float32), np.full(4, -60.0, np.float32))
    assert worst == pytest.approx(-np.log(np.float32(1e-7)), rel=1e-5)


def test_early_stop_rule_known_answers(oracle):
    """booster.cpp:326-352: any improvement beyond 1e-10 resets the count; the
    loop breaks after tree `iter` once iter >= max(50, n_estimators / 10) and
    `rounds` trees in a row failed to improve."""
    # strictly improving: never stops, best is the last tree
    m = np.linspace(1.0, 0.5, 80).astype(np.float32)
    assert oracle.early_stop(m, 80, 5) == (80, 79, pytest.approx(0.5))
    # improves until tree 9, flat afterwards: the count reaches 5 at tree 14, but the
    # break waits for tree 50 (min iterations)
    m = np.concatenate([np.linspace(1.0, 0.9, 10), np.full(70, 0.9)]).astype(np.float32)
    kept, best_it, best = oracle.early_stop(m, 80, 5)
    assert (kept, best_it) == (51, 9) and best == pytest.approx(0.9)
    # n_estimators = 1000 moves the minimum to tree 100
    m = np.concatenate([np.linspace(1.0, 0.9, 10), np.full(190, 0.9)]).astype(np.float32)
    assert oracle.early_stop(m, 1000, 5)[0] == 101
    # improving through tree 55, four flat trees, an improvement at 60 resets the
    # count, which then reaches 5 at tree 65
    m = (1.0 - 0.001 * np.minimum(np.arange(100), 55)).astype(np.float32)
    m[60:] = m[55] - np.float32(0.01)
    assert oracle.early_stop(m, 100, 5)[:2] == (66, 60)
    # an improvement smaller than 1e-10 does not count; higher-is-better flips it
    m = np.full(70, 0.5, np.float32)
    assert oracle.early_stop(m, 70, 3)[:2] == (51, 0)
    assert oracle.early_stop(-m, 70, 3, higher_is_better=True)[:2] == (51, 0)
