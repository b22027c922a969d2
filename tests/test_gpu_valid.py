"""A19: validation set and early stopping on the device (booster.cpp:252-371).

The validation path must not change training (the first trees equal a plain
fit's, bit for bit); the per-tree mean losses are checked against the oracle's
FP64 Loss::compute_loss on predictions replayed from the model with the
oracle's binned leaf index and the reference's fused update; the stopping
point is the oracle's early-stopping rule applied to the device's own loss
sequence (exact), and to the oracle's (equal unless two losses tie within
float rounding)."""
import numpy as np
import pytest

from conftest import binary_data, regression_data

pytestmark = pytest.mark.gpu


def _replay_losses(oracle, loss, model, base, lr, bins, y):
    """Mean loss after each tree of `model` over rows binned as `bins`."""
    T, D = model["T"], model["D"]
    p = np.full(bins.shape[1], base, np.float32)
    out = []
    for t in range(T):
        f = model["feature"][t * D:(t + 1) * D][: model["depth"][t]]
        b = model["bin"][t * D:(t + 1) * D][: model["depth"][t]]
        li = oracle.leaf_index_binned(bins, f, b)
        vals = model["leaves"][(t << D) + li]
        p = oracle.fma_update(p, lr, vals)
        out.append(oracle.compute_loss_f64(loss, y, p))
    return np.array(out), p


def _data(loss, n, F, seed):
    return regression_data(n, F, seed) if loss == 0 else binary_data(n, F, seed)


@pytest.mark.parametrize("loss", [0, 1])
def test_validation_losses_match_oracle_and_leave_training_alone(ctx, oracle, loss):
    from turbocat import capi
    X, y = _data(loss, 30000, 12, 51)
    Xv, yv = _data(loss, 9001, 12, 52)
    ds, edges, ne = ctx.quantize_fit(X)
    vds = ctx.dataset(Xv, edges, ne)
    lam = 0.0 if loss == 0 else 1.0
    prm = capi.train_params(loss, n_trees=30, learning_rate=0.2, max_depth=5, lambda_l2=lam)
    plain = ds.train(y, prm)
    m, h = ds.train_valid(y, prm, vds, yv, early_stopping_rounds=1000)
    for k in ("feature", "bin", "fthr", "gain", "depth", "leaves", "train_preds"):
        assert np.array_equal(m[k], plain[k]), k
    assert h["n_trees_built"] == 30 and not h["stopped_early"]   # 30 < the 50-tree minimum
    tl, tp = _replay_losses(oracle, loss, m, m["base"], 0.2, ds.bins(), y)
    assert np.array_equal(tp, m["train_preds"])  # the replay is the device's arithmetic
    vl, _ = _replay_losses(oracle, loss, m, m["base"], 0.2, vds.bins(), yv)
    np.testing.assert_allclose(h["train_loss"], tl, rtol=2e-6)
    np.testing.assert_allclose(h["valid_loss"], vl, rtol=2e-6)
    best = int(np.argmin(h["valid_loss"]))
    assert h["best_iteration"] == oracle.early_stop(h["valid_loss"], 30, 1000)[1]
    assert h["best_valid_loss"] == h["valid_loss"][h["best_iteration"]]
    assert h["valid_loss"][best] == h["valid_loss"].min()


@pytest.mark.parametrize("loss,rounds", [(0, 5), (1, 8), (0, 0)])
def test_early_stopping_follows_the_rule(ctx, oracle, loss, rounds):
    """Validation labels unrelated to X: the validation loss stops improving
    early, so the fit ends at the first tree index >= 50 whose no-improvement
    count has reached `rounds`."""
    from turbocat import capi
    X, y = _data(loss, 20000, 10, 61)
    rng = np.random.default_rng(62)
    Xv = rng.standard_normal((5000, 10)).astype(np.float32)
    yv = (rng.random(5000) < 0.5).astype(np.float32) if loss else rng.standard_normal(5000).astype(np.float32)
    ds, edges, ne = ctx.quantize_fit(X)
    vds = ctx.dataset(Xv, edges, ne)
    T = 200
    prm = capi.train_params(loss, n_trees=T, learning_rate=0.3, max_depth=4,
                            lambda_l2=0.0 if loss == 0 else 1.0)
    m, h = ds.train_valid(y, prm, vds, yv, early_stopping_rounds=rounds)
    kept = h["n_trees_built"]
    assert h["stopped_early"] and 51 <= kept < T
    vl = h["valid_loss"][:kept]
    assert np.all(h["valid_loss"][kept:] == 0)
    ok, best_it, best = oracle.early_stop(vl, T, rounds)
    assert (ok, best_it) == (kept, h["best_iteration"]) and best == h["best_valid_loss"]
    # the oracle's own losses give the same stop unless two of them tie within rounding
    ref, _ = _replay_losses(oracle, loss, {k: (v[: kept * 4] if k in ("feature", "bin", "fthr", "gain")
                                                else v) for k, v in m.items()} | {"T": kept},
                            m["base"], 0.3, vds.bins(), yv)
    d = np.abs(np.diff(np.minimum.accumulate(ref)))
    if not np.any((d > 0) & (d < 1e-6 * np.abs(ref[1:]))):
        assert oracle.early_stop(ref.astype(np.float32), T, rounds)[:2] == (kept, best_it)
    # the trees it kept are a plain fit's first trees
    plain = ds.train(y, capi.train_params(loss, n_trees=kept, learning_rate=0.3, max_depth=4,
                                          lambda_l2=0.0 if loss == 0 else 1.0))
    D = 4
    for k in ("feature", "bin", "fthr", "gain"):
        assert np.array_equal(m[k][: kept * D], plain[k]), k
    assert np.array_equal(m["depth"][:kept], plain["depth"])
    assert np.array_equal(m["leaves"][: kept << D], plain["leaves"])
    assert np.all(m["depth"][kept:] == 0) and np.all(m["leaves"][kept << D:] == 0)


def test_dropin_fit_with_validation_stops_early():
    """TurboCatRegressor / TurboCatClassifier.fit(X, y, X_val, y_val) early-stop
    and keep the trees built; predictions use exactly those trees."""
    import turbocat
    X, y = regression_data(20000, 8, 71)
    rng = np.random.default_rng(72)
    Xv = rng.standard_normal((4000, 8)).astype(np.float32)
    yv = rng.standard_normal(4000).astype(np.float32)
    r = turbocat.TurboCatRegressor(n_estimators=300, learning_rate=0.3, max_depth=4,
                                   early_stopping_rounds=5, verbosity=0)
    r.fit(X, y, X_val=Xv, y_val=yv)
    assert 51 <= r.n_trees < 300
    full = turbocat.TurboCatRegressor(n_estimators=r.n_trees, learning_rate=0.3, max_depth=4,
                                      verbosity=0)
    full.fit(X, y)
    assert np.array_equal(r.predict(X[:1000]), full.predict(X[:1000]))
    Xb, yb = binary_data(20000, 8, 73)
    yvb = (rng.random(4000) < 0.5).astype(np.float32)
    c = turbocat.TurboCatClassifier(n_estimators=300, learning_rate=0.3, max_depth=4,
                                    early_stopping_rounds=5, verbosity=0)
    c.fit(Xb, yb, X_val=Xv, y_val=yvb)
    assert 51 <= c.n_trees < 300
    from turbocat import _turbocat
    raw = _turbocat.TurboCatClassifier(n_estimators=10, early_stopping_metric="roc_auc", verbosity=0)
    with pytest.raises(ValueError, match="early_stopping_metric"):
        raw.fit(Xb, yb, Xv, yvb, [])
