"""Row sampling (booster.cpp:198-229, dataset.cpp:494-562) without a GPU.

The library draws each tree's row sample on the host with the reference's
engine (xorshift64* seeding std::mt19937_64 + std::shuffle, or GOSS:
std::nth_element on |g| then a partial Fisher-Yates draw).  Pinned two ways
against the reference's own code (oracle/_ref/libtcref.so):
  * tc_random_subsample equals Dataset::random_subsample, and
    tc_goss_subsample equals Dataset::goss_subsample, for the same seed
    (GOSS also on gradients with only a few distinct |g|, where the
    library's placement of ties decides the rows);
  * a whole sampled fit: the oracle, fed the library's per-tree samples,
    reproduces the reference's sampled fit bit for bit -- which fails if the
    draw sequence (next_random) or the colsample draw were wrong.
"""
import numpy as np
import pytest

from conftest import binary_data, regression_data

from oracle import LOSS_LOGLOSS, LOSS_MSE, RefLib, tree_params

pytestmark = pytest.mark.skipif(not RefLib.available(), reason="reference harness not built")


@pytest.mark.parametrize("n,ratio,seed", [(50, 0.8, 7), (1000, 0.5, 123), (100003, 0.8, 2 ** 63 + 5),
                                          (10, 0.05, 1), (7, 1.0, 3)])
def test_random_subsample_matches_reference(n, ratio, seed):
    from turbocat import capi

    def body():
        X = np.zeros((n, 1), np.float32)
        ds = RefLib().dataset(X)
        return ds.random_subsample(ratio, seed)
    ref = RefLib.isolated(body)
    got = capi.random_subsample(n, ratio, seed)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("loss,subsample,colsample,seed", [(LOSS_MSE, 0.8, 0.8, 42),
                                                           (LOSS_LOGLOSS, 0.5, 1.0, 123)])
def test_sampled_fit_replays_reference(oracle, loss, subsample, colsample, seed):
    from turbocat import capi
    n, F, T, D, lr = 3000, 6, 10, 5, 0.1
    lam = 0.0 if loss == LOSS_MSE else 1.0
    X, y = regression_data(n, F, 77) if loss == LOSS_MSE else binary_data(n, F, 77)

    def body():
        ds = RefLib().dataset(X, y)
        return ds.fit(loss, T, D, lr, lam, 1.0, subsample=subsample, colsample=colsample, seed=seed)
    rm, rtp, _ = RefLib.isolated(body)
    bins, edges, ne = oracle.quantize(X)
    preds = np.full(n, np.float32(oracle.init_prediction(loss, y)), np.float32)
    state = seed
    for t in range(T):
        s, state = capi.next_random(state)
        rows = capi.random_subsample(n, subsample, s)
        if colsample < 1.0:
            _, state = capi.next_random(state)
        g, h = oracle.gradients_f32(loss, y, preds)
        tr = oracle.tree_f32(bins, g, h, edges, ne, tree_params(D, lambda_l2=lam), rows=rows)
        d = tr["depth"]
        assert d == rm["depth"][t]
        assert np.array_equal(tr["feature"], rm["feature"][t * D:t * D + d])
        assert np.array_equal(tr["bin"], rm["bin"][t * D:t * D + d])
        assert np.array_equal(tr["gain"], rm["gain"][t * D:t * D + d])
        assert np.array_equal(tr["leaves"], rm["leaves"][t << D:(t << D) + (1 << d)])
        li = oracle.leaf_index_binned(bins, tr["feature"], tr["bin"])
        preds = oracle.fma_update(preds, np.float32(lr), tr["leaves"][li])
    assert np.array_equal(preds, rtp)


@pytest.mark.parametrize("n,top,other,seed,kind", [
    (1000, 0.2, 0.1, 7, "normal"), (5003, 0.3, 0.25, 2 ** 63 + 9, "ties"),
    (20000, 0.2, 0.1, 123, "two"), (10, 0.05, 0.05, 1, "normal"), (64, 1.0, 0.5, 3, "ties"),
    (64, 0.0, 0.5, 11, "normal"), (300, 0.5, 1.0, 5, "ties")])
def test_goss_subsample_matches_reference(n, top, other, seed, kind):
    from turbocat import capi
    rng = np.random.default_rng(n + seed % 1000)
    if kind == "normal":
        g = rng.standard_normal(n).astype(np.float32)
    elif kind == "ties":  # few distinct |g|, both signs
        g = (rng.integers(-4, 5, n) * 0.25).astype(np.float32)
    else:  # LogLoss at the base score: two values of |p - y|
        y = (rng.random(n) < 0.3).astype(np.float32)
        g = (np.float32(0.3) - y).astype(np.float32)

    def body():
        ds = RefLib().dataset(np.zeros((n, 1), np.float32))
        return ds.goss_subsample(g, top, other, seed)
    ref = RefLib.isolated(body)
    got = capi.goss_subsample(g, top, other, seed)
    assert np.array_equal(got, ref)
    assert len(np.unique(got)) == len(got)


@pytest.mark.parametrize("loss,colsample,seed", [(LOSS_LOGLOSS, 1.0, 42), (LOSS_MSE, 0.8, 321)])
def test_goss_fit_replays_reference(oracle, loss, colsample, seed):
    """The reference's GOSS fit, replayed by the oracle on the library's
    per-tree GOSS rows (drawn from the oracle's own gradients)."""
    from turbocat import capi
    n, F, T, D, lr, top, other = 4000, 6, 8, 5, 0.1, 0.2, 0.1
    lam = 0.0 if loss == LOSS_MSE else 1.0
    X, y = regression_data(n, F, 78) if loss == LOSS_MSE else binary_data(n, F, 78)

    def body():
        ds = RefLib().dataset(X, y)
        return ds.fit_goss(loss, T, D, lr, lam, 1.0, top_rate=top, other_rate=other,
                           colsample=colsample, seed=seed)
    rm, rtp = RefLib.isolated(body)
    bins, edges, ne = oracle.quantize(X)
    preds = np.full(n, np.float32(oracle.init_prediction(loss, y)), np.float32)
    state = seed
    for t in range(T):
        g, h = oracle.gradients_f32(loss, y, preds)
        s, state = capi.next_random(state)
        rows = capi.goss_subsample(g, top, other, s)
        if colsample < 1.0:
            _, state = capi.next_random(state)
        tr = oracle.tree_f32(bins, g, h, edges, ne, tree_params(D, lambda_l2=lam), rows=rows)
        d = tr["depth"]
        assert d == rm["depth"][t]
        assert np.array_equal(tr["feature"], rm["feature"][t * D:t * D + d])
        assert np.array_equal(tr["bin"], rm["bin"][t * D:t * D + d])
        assert np.array_equal(tr["leaves"], rm["leaves"][t << D:(t << D) + (1 << d)])
        li = oracle.leaf_index_binned(bins, tr["feature"], tr["bin"])
        preds = oracle.fma_update(preds, np.float32(lr), tr["leaves"][li])
    assert np.array_equal(preds, rtp)


@pytest.mark.parametrize("n,ratio,seed,kind", [
    (1000, 0.5, 7, "normal"), (5003, 0.3, 2 ** 63 + 9, "ties"), (20000, 0.5, 123, "two"),
    (4000, 0.5, 17, "spiky"), (3000, 0.3, 5, "zeros"), (10, 0.5, 1, "normal"),
    (1, 0.5, 3, "normal"), (777, 1.0, 11, "normal"), (2048, 0.05, 13, "spiky"),
    (100003, 0.7, 99, "ties"), (1048583, 0.5, 31, "ties"), (600011, 0.2, 8, "normal")])
def test_mvs_subsample_matches_reference(n, ratio, seed, kind):
    """MVS rows equal the compiled reference's Dataset::mvs_subsample on one
    OpenMP thread: the probability pass, and ("spiky": a few huge |g| take
    most of the score sum, so under 80% of the target is kept) the shuffled
    top-up."""
    from turbocat import capi
    rng = np.random.default_rng(n + seed % 1000)
    if kind == "normal":
        g = rng.standard_normal(n).astype(np.float32)
    elif kind == "ties":
        g = (rng.integers(-4, 5, n) * 0.25).astype(np.float32)
    elif kind == "two":  # LogLoss at the base score: two values of |p - y|
        y = (rng.random(n) < 0.3).astype(np.float32)
        g = (np.float32(0.3) - y).astype(np.float32)
    elif kind == "spiky":
        g = (rng.standard_normal(n) * 1e-3).astype(np.float32)
        g[rng.choice(n, max(1, n // 100), replace=False)] = 1000.0
    else:  # every |g| below the 1e-10 floor
        g = np.zeros(n, np.float32)

    def body():
        ds = RefLib().dataset(np.zeros((n, 1), np.float32))
        return ds.mvs_subsample(g, ratio, seed)
    ref = RefLib.isolated(body)
    got = capi.mvs_subsample(g, ratio, seed)
    assert np.array_equal(got, ref)
    assert len(np.unique(got)) == len(got)
    if kind == "spiky":  # the top-up ran: exactly the target
        assert len(got) == max(1, int(np.float32(n) * np.float32(ratio)))


@pytest.mark.parametrize("loss,seed", [(LOSS_LOGLOSS, 42), (LOSS_MSE, 321)])
def test_mvs_fit_replays_reference(oracle, loss, seed):
    """The reference's MVS fit, replayed by the oracle on the library's
    per-tree MVS rows (drawn from the oracle's own gradients)."""
    from turbocat import capi
    n, F, T, D, lr, ratio = 4000, 6, 8, 5, 0.1, 0.5
    lam = 0.0 if loss == LOSS_MSE else 1.0
    X, y = regression_data(n, F, 79) if loss == LOSS_MSE else binary_data(n, F, 79)

    def body():
        ds = RefLib().dataset(X, y)
        return ds.fit_mvs(loss, T, D, lr, lam, 1.0, ratio=ratio, seed=seed)
    rm, rtp = RefLib.isolated(body)
    bins, edges, ne = oracle.quantize(X)
    preds = np.full(n, np.float32(oracle.init_prediction(loss, y)), np.float32)
    state = seed
    for t in range(T):
        g, h = oracle.gradients_f32(loss, y, preds)
        s, state = capi.next_random(state)
        rows = capi.mvs_subsample(g, ratio, s)
        tr = oracle.tree_f32(bins, g, h, edges, ne, tree_params(D, lambda_l2=lam), rows=rows)
        d = tr["depth"]
        assert d == rm["depth"][t]
        assert np.array_equal(tr["feature"], rm["feature"][t * D:t * D + d])
        assert np.array_equal(tr["bin"], rm["bin"][t * D:t * D + d])
        assert np.array_equal(tr["leaves"], rm["leaves"][t << D:(t << D) + (1 << d)])
        li = oracle.leaf_index_binned(bins, tr["feature"], tr["bin"])
        preds = oracle.fma_update(preds, np.float32(lr), tr["leaves"][li])
    assert np.array_equal(preds, rtp)
