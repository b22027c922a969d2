"""Parity at the full sizes of the headline workloads.

Where the C oracle finishes in seconds it is compared directly; elsewhere
the checks are properties that hold at any size:

- training, 1,000,000 rows x 100 features, binary LogLoss, depth 6: the
  bins of a column sample (Gaussian columns and Uniform columns with 5%
  duplicated values) bit-exact against the oracle's quantizer; the first
  tree teacher-forced against the FP64 oracle tree; a 12-tree fit identical
  under three work-item shapes (every histogram is an exact integer sum),
  its first four trees replayed by the oracle and its training predictions equal
  to the oracle's fused updates with the GPU's leaves;
- batched training, 64 regression jobs of 10,000 to 2,000,000 rows
  (log-uniform) and 16 to 256 features, 200 trees each: every job's model
  from the batched launches identical to the job trained alone, and the
  largest job's first tree teacher-forced against the FP64 oracle;
- multiclass, 4,000,000 rows x 256 features, 10-class softmax, depth 8: a
  1,000-tree fit identical under two work-item shapes, predictions of all rows equal
  to predictions of two row halves and to predictions from the training
  bins, class probabilities summing to one.

The data are the same seeded synthetic workloads the benchmark times."""
import numpy as np
import pytest

import bench
from test_gpu_tree import compare_tree, fma_update

pytestmark = pytest.mark.gpu


def fit(ctx, ds, y, loss, n_trees, depth, lr, shape, n_classes=0):
    from turbocat import capi
    prm = capi.train_params(loss, n_trees=n_trees, learning_rate=lr, max_depth=depth,
                            lambda_l2=1.0, hist_chunk=shape[0], hist_deep_items=shape[1],
                            n_classes=n_classes)
    tr = capi.Trainer(ctx, [ds], [y], prm)
    tr.step(n_trees)
    tr.sync()
    m = tr.model()
    tr.close()
    return m


def same_model(a, b):
    for k in ("feature", "bin", "fthr", "gain", "depth", "leaves", "train_preds"):
        assert np.array_equal(a[k], b[k]), k
    assert a["base"] == b["base"]


@pytest.fixture(scope="module")
def train_set(ctx):
    n, F = 1_000_000, 100
    print(f"\n[fullsize] training set {n} x {F}", flush=True)
    X, y = bench.binary_workload(n, F, seed=1001)
    ds, edges, ne = ctx.quantize_fit(X)
    yield X, y, ds, edges, ne
    ds.close()


def test_full_size_bins_of_a_column_sample(train_set, oracle):
    X, y, ds, edges, ne = train_set
    cols = [0, 1, 41, 79, 80, 81, 90, 99]  # Gaussian 0-79, Uniform with duplicates 80-99
    ob, oe, on = oracle.quantize(np.ascontiguousarray(X[:, cols]))
    assert np.array_equal(ne[cols], on)
    for i, f in enumerate(cols):
        assert np.array_equal(edges[f, :ne[f]], oe[i, :on[i]]), f
    bins = ds.bins()
    assert np.array_equal(bins[cols], ob)


def test_full_size_first_tree_teacher_forced(train_set, oracle):
    from turbocat.capi import tree_params
    from oracle import tree_params as otp
    X, y, ds, edges, ne = train_set
    base = oracle.init_prediction(1, y)
    g, h = oracle.gradients_f32(1, y, np.full(len(y), np.float32(base), np.float32))
    gpu = ds.tree(g, h, tree_params(max_depth=6, lambda_l2=1.0))
    levels, tie = compare_tree(gpu, ds.bins(), g, h, edges, ne, otp(max_depth=6, lambda_l2=1.0),
                               oracle)
    assert levels == 6 and not tie


def test_full_size_fit_is_shape_free_and_replays(ctx, train_set, oracle):
    from oracle import tree_params as otp
    X, y, ds, edges, ne = train_set
    T, D, lr = 12, 6, np.float32(0.1)
    models = [fit(ctx, ds, y, 1, T, D, 0.1, s) for s in ((0, 0), (13699, 256), (4096, 0))]
    for m in models[1:]:
        same_model(m, models[0])
    m = models[0]
    assert m["base"] == oracle.init_prediction(1, y)
    bins = ds.bins()
    preds = np.full(len(y), np.float32(m["base"]), np.float32)
    ties = 0
    for t in range(T):
        d = int(m["depth"][t])
        gpu = dict(depth=d, feature=m["feature"][t * D:t * D + d], bin=m["bin"][t * D:t * D + d],
                   fthr=m["fthr"][t * D:t * D + d], gain=m["gain"][t * D:t * D + d],
                   leaves=m["leaves"][t << D:(t << D) + (1 << d)])
        gpu["leaf_idx"] = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
        if t < 4:  # the oracle's FP64 tree on the same gradients
            g, h = oracle.gradients_f32(1, y, preds)
            _, tie = compare_tree(gpu, bins, g, h, edges, ne, otp(max_depth=D, lambda_l2=1.0),
                                  oracle)
            ties += tie
        preds = fma_update(oracle, preds, lr, gpu["leaves"][gpu["leaf_idx"]])
    assert ties == 0
    assert np.array_equal(m["train_preds"], preds)


def test_full_size_multiclass_fit_and_predict(ctx):
    from turbocat import capi
    n, F, K = 4_000_000, 256, 10
    print(f"\n[fullsize] multiclass set {n} x {F}, {K} classes", flush=True)
    X, y = bench.multiclass_workload(n, F, K, seed=1004)
    ds, edges, ne = ctx.quantize_fit(X)
    try:
        # the workload's 1,000 trees, under the autotuned and a fixed work-item shape
        a = fit(ctx, ds, y, capi.TC_LOSS_SOFTMAX, 1000, 8, 0.05, (0, 0), n_classes=K)
        print("[fullsize] multiclass: 1000 trees fitted (autotuned shape)", flush=True)
        b = fit(ctx, ds, y, capi.TC_LOSS_SOFTMAX, 1000, 8, 0.05, (4096, 512), n_classes=K)
        same_model(a, b)
        assert (a["depth"] > 0).all()
        ens = ctx.ensemble(a, F)
        try:
            raw = ens.predict(X)
            half = n // 2 + 7
            assert np.array_equal(np.concatenate([ens.predict(X[:half]), ens.predict(X[half:])]),
                                  raw)
            assert np.array_equal(ens.predict_binned(ds), raw)
            p = ens.predict(X[:200_000], transform=1)
            assert np.abs(p.astype(np.float64).sum(axis=1) - 1.0).max() < 1e-5
        finally:
            ens.close()
    finally:
        ds.close()


def test_full_size_batched_jobs_equal_each_job_alone(ctx, oracle):
    from turbocat import capi
    from oracle import tree_params as otp
    rng = np.random.default_rng(2002)
    J, T, D = 64, 200, 6  # the batched workload's jobs and trees
    rows = np.floor(np.exp(rng.uniform(np.log(1e4), np.log(2e6), J))).astype(int)
    feats = rng.integers(16, 257, J)
    big = int(np.argmax(rows))
    print(f"\n[fullsize] {J} jobs, {rows.sum()} rows", flush=True)
    dss, ys = [], []
    try:
        for j in range(J):
            X, y = bench.regression_workload(int(rows[j]), int(feats[j]), seed=5000 + j)
            ds, edges, ne = ctx.quantize_fit(X)
            if j == big:
                big_edges, big_ne = edges, ne
            dss.append(ds)
            ys.append(y)
            del X
        prm = capi.train_params(capi.TC_LOSS_MSE, n_trees=T, learning_rate=0.1, max_depth=D,
                                lambda_l2=0.0)
        tr = capi.Trainer(ctx, dss, ys, prm)
        tr.step(T)
        tr.sync()
        batched = [tr.model(j) for j in range(J)]
        tr.close()
        for j in range(J):
            tj = capi.Trainer(ctx, [dss[j]], [ys[j]], prm)
            tj.step(T)
            tj.sync()
            same_model(tj.model(), batched[j])
            tj.close()
        m, y, bins = batched[big], ys[big], dss[big].bins()
        assert m["base"] == oracle.init_prediction(0, y)
        g, h = oracle.gradients_f32(0, y, np.full(len(y), np.float32(m["base"]), np.float32))
        d = int(m["depth"][0])
        gpu = dict(depth=d, feature=m["feature"][:d], bin=m["bin"][:d], fthr=m["fthr"][:d],
                   gain=m["gain"][:d], leaves=m["leaves"][:1 << d])
        gpu["leaf_idx"] = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
        _, tie = compare_tree(gpu, bins, g, h, big_edges, big_ne,
                              otp(max_depth=D, lambda_l2=0.0), oracle)
        assert not tie
    finally:
        for ds in dss:
            ds.close()
