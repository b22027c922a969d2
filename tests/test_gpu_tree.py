"""GPU symmetric-tree build (A9-A14) and boosting loop (A5-A15) against the
CPU oracle.

Bit-exact claims: chosen (feature, bin) per level, tree depth, per-row leaf
indices, and -- given the GPU's own trees -- the training predictions.
Float claims: leaf values within 1e-5 of the FP64 oracle, normalized by
sum |g| / (H + lambda) of the leaf.  Near-tie rule: where the FP64 gap
between the best and second-best candidate of a level is below
NEAR_TIE * |best gain|, either choice is accepted and the rest of that tree
is not compared (the two builds then legitimately diverge); such events are
counted and must stay rare."""
import numpy as np
import pytest

from conftest import binary_data, regression_data, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
NEAR_TIE = 1e-5


def compare_tree(gpu, bins, g, h, edges, ne, params, oracle):
    """Compare one GPU tree with the FP64 oracle on the same gradients.
    Returns (levels compared, near_tie_hit)."""
    ref = oracle.tree_f64(bins, g, h, edges, ne, params)
    lam = params.lambda_l2
    for d in range(max(gpu["depth"], ref["depth"])):
        if d >= gpu["depth"] or d >= ref["depth"]:
            assert False, f"depth differs at level {d}: gpu {gpu['depth']} ref {ref['depth']}"
        same = (gpu["feature"][d] == ref["feature"][d]) and (gpu["bin"][d] == ref["bin"][d])
        if not same:
            assert ref["gap"][d] <= NEAR_TIE * max(1.0, abs(ref["gain"][d])), (
                f"level {d}: gpu ({gpu['feature'][d]},{gpu['bin'][d]}) ref "
                f"({ref['feature'][d]},{ref['bin'][d]}) gap {ref['gap'][d]} gain {ref['gain'][d]}")
            return d, True
        assert gpu["fthr"][d] == ref["fthr"][d]
        assert abs(gpu["gain"][d] - ref["gain"][d]) <= 1e-4 * max(1.0, abs(ref["gain"][d]))
    assert gpu["depth"] == ref["depth"]
    assert np.array_equal(gpu["leaf_idx"], ref["leaf_idx"])
    L = 1 << ref["depth"]
    absg = np.bincount(ref["leaf_idx"], weights=np.abs(g.astype(np.float64)), minlength=L)
    hs = np.bincount(ref["leaf_idx"], weights=h.astype(np.float64), minlength=L)
    scale = absg / (hs + lam + 1e-30)
    assert rel_err(gpu["leaves"], ref["leaves"], scale).max() < TOL
    return ref["depth"], False


def test_kat_eight_rows(ctx, oracle):
    """Known-answer test: the reference's 8-row toy (tests/test_turbocat.py:51-57)
    with one depth-3 tree, lr 1, lambda 1 -- splits (0,3) at every level with
    gains 2, 0, 0 (zero-gain splits applied), leaves [-1, 0 x 6, +1]."""
    from turbocat.capi import train_params, TC_LOSS_LOGLOSS
    X = np.array([-2, -1, 1, 2, -1.5, -0.5, 0.5, 1.5], np.float32).reshape(8, 1)
    y = np.array([0, 0, 1, 1, 0, 0, 1, 1], np.float32)
    ds, edges, ne = ctx.quantize_fit(X)
    assert list(ds.bins()[0]) == [0, 2, 5, 7, 1, 3, 4, 6]
    m = ds.train(y, train_params(TC_LOSS_LOGLOSS, n_trees=1, learning_rate=1.0, max_depth=3,
                                 lambda_l2=1.0, min_child_weight=0.01))
    assert m["base"] == 0.0
    assert list(m["depth"]) == [3]
    assert list(m["feature"]) == [0, 0, 0] and list(m["bin"]) == [3, 3, 3]
    assert list(m["gain"]) == [2.0, 0.0, 0.0]
    assert list(m["fthr"]) == [0.5, 0.5, 0.5]
    assert list(m["leaves"]) == [-1, 0, 0, 0, 0, 0, 0, 1]
    ens = ctx.ensemble(m, 1)
    p = ens.predict(X, transform=1)
    assert np.allclose(p, np.where(X[:, 0] >= 0.5, 0.7310586, 0.26894143), atol=1e-6)


@pytest.mark.parametrize("n,F,depth,lam,seed", [
    (5000, 6, 4, 1.0, 1),
    (50000, 20, 6, 0.0, 2),
    (30011, 33, 6, 1.0, 3),
    (20000, 10, 8, 1.0, 4),
])
def test_tree_teacher_forced(ctx, oracle, n, F, depth, lam, seed):
    from turbocat.capi import tree_params
    X, y = regression_data(n, F, seed, low_card_col=3)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(seed + 100)
    g = (rng.standard_normal(n) + 0.3 * X[:, 0] - np.sin(X[:, 1])).astype(np.float32)
    h = (rng.random(n) * 0.2 + 0.05).astype(np.float32)
    prm = tree_params(max_depth=depth, lambda_l2=lam)
    from oracle import tree_params as otp
    oprm = otp(max_depth=depth, lambda_l2=lam)
    gpu = ds.tree(g, h, prm)
    lv, tie = compare_tree(gpu, bins, g, h, edges, ne, oprm, oracle)
    assert not tie


@pytest.mark.parametrize("n,F,depth,seed", [
    (3000, 1, 10, 5),      # one feature, the deepest tree: 512 pairs at the last level
    (2500, 300, 3, 6),     # 19 tiles, a 384-byte row stride
    (17, 5, 4, 7),         # fewer rows than any work item
    (70000, 129, 7, 8),    # 9 tiles, the last with one feature
])
def test_tree_extreme_shapes(ctx, oracle, n, F, depth, seed):
    """Teacher-forced trees at the edges of the histogram layout (tile and
    stride rounding, tiny nodes and empty children, maximum depth) against
    the oracle; a split may differ only at an FP64 near-tie."""
    from turbocat.capi import tree_params
    from oracle import tree_params as otp
    X, y = regression_data(n, F, seed)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(seed + 200)
    g = (rng.standard_normal(n) + 0.5 * X[:, 0]).astype(np.float32)
    h = (rng.random(n) * 0.3 + 0.05).astype(np.float32)
    gpu = ds.tree(g, h, tree_params(max_depth=depth, lambda_l2=1.0, min_child_weight=0.0))
    compare_tree(gpu, bins, g, h, edges, ne, otp(max_depth=depth, lambda_l2=1.0,
                                                  min_child_weight=0.0), oracle)


def test_tree_stops_early(ctx, oracle):
    """min_child_weight large enough that the level stop test fires."""
    from turbocat.capi import tree_params
    from oracle import tree_params as otp
    X, y = regression_data(3000, 5, 9)
    ds, edges, ne = ctx.quantize_fit(X)
    g = (X[:, 0] + 0.1).astype(np.float32)
    h = np.full(3000, 1.0, np.float32)
    # on this data the oracle stops at depth 3, 1, 1 and 0 (root) respectively
    for mcw, want in ((600.0, 3), (800.0, 1), (1400.0, 1), (2000.0, 0)):
        gpu = ds.tree(g, h, tree_params(max_depth=6, min_child_weight=mcw))
        ref = oracle.tree_f64(ds.bins(), g, h, edges, ne, otp(max_depth=6, min_child_weight=mcw))
        assert gpu["depth"] == ref["depth"] == want
        assert np.array_equal(gpu["leaf_idx"], ref["leaf_idx"])


def fma_update(oracle, preds, lr, leaf):
    """preds += lr * leaf as one fused multiply-add per row (the reference
    compiler's contraction of booster.cpp:460-463), computed by the oracle."""
    return oracle.fma_update(preds, lr, leaf)


@pytest.mark.parametrize("loss,lam", [(0, 0.0), (1, 1.0)])
def test_fit_teacher_forced(ctx, oracle, loss, lam):
    from turbocat.capi import train_params
    n, F, T = 20000, 12, 25
    X, y = (regression_data(n, F, 31) if loss == 0 else binary_data(n, F, 31))
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    m = ds.train(y, train_params(loss, n_trees=T, learning_rate=0.1, max_depth=6, lambda_l2=lam))
    assert m["base"] == oracle.init_prediction(loss, y)
    from oracle import tree_params as otp
    preds = np.full(n, np.float32(m["base"]), np.float32)
    ties = 0
    D = 6
    for t in range(T):
        g, h = oracle.gradients_f32(loss, y, preds)
        d = int(m["depth"][t])
        gpu = dict(depth=d, feature=m["feature"][t * D:t * D + d], bin=m["bin"][t * D:t * D + d],
                   fthr=m["fthr"][t * D:t * D + d], gain=m["gain"][t * D:t * D + d],
                   leaves=m["leaves"][t << D:(t << D) + (1 << d)])
        gpu["leaf_idx"] = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
        _, tie = compare_tree(gpu, bins, g, h, edges, ne, otp(max_depth=D, lambda_l2=lam), oracle)
        ties += tie
        preds = fma_update(oracle, preds, np.float32(0.1), gpu["leaves"][gpu["leaf_idx"]])
    assert ties <= 1
    # the GPU's training predictions are the same fused updates of the same leaves
    assert np.array_equal(m["train_preds"], preds)


def test_fit_batched_matches_each_job(ctx, oracle):
    """Three jobs of different shapes and losses trained in one batched call:
    every job's trees pass the teacher-forced check against the oracle."""
    from turbocat.capi import train_params
    from oracle import tree_params as otp
    shapes = [(9000, 7, 0, 0.0, 41), (25000, 19, 1, 1.0, 42), (4000, 40, 0, 1.0, 43)]
    dss, ys, prms, data = [], [], [], []
    for n, F, loss, lam, seed in shapes:
        X, y = regression_data(n, F, seed) if loss == 0 else binary_data(n, F, seed)
        ds, edges, ne = ctx.quantize_fit(X)
        dss.append(ds); ys.append(y)
        prms.append(train_params(loss, n_trees=8, learning_rate=0.2, max_depth=5, lambda_l2=lam))
        data.append((ds.bins(), edges, ne, loss, lam))
    models = ctx.train_batch(dss, ys, prms)
    D = 5
    for (bins, edges, ne, loss, lam), y, m in zip(data, ys, models):
        n = bins.shape[1]
        preds = np.full(n, np.float32(m["base"]), np.float32)
        ties = 0
        for t in range(8):
            g, h = oracle.gradients_f32(loss, y, preds)
            d = int(m["depth"][t])
            gpu = dict(depth=d, feature=m["feature"][t * D:t * D + d], bin=m["bin"][t * D:t * D + d],
                       fthr=m["fthr"][t * D:t * D + d], gain=m["gain"][t * D:t * D + d],
                       leaves=m["leaves"][t << D:(t << D) + (1 << d)])
            gpu["leaf_idx"] = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
            _, tie = compare_tree(gpu, bins, g, h, edges, ne, otp(max_depth=D, lambda_l2=lam), oracle)
            ties += tie
            preds = fma_update(oracle, preds, np.float32(0.2), gpu["leaves"][gpu["leaf_idx"]])
        assert ties <= 1
        assert np.array_equal(m["train_preds"], preds)


def test_sampled_fit_graph_equals_stepwise(ctx):
    """A plain row sample is shuffled ahead on host threads (a ring of future
    trees) while the trainer captures and replays its per-tree graph: the
    graph fit, the fit without a graph and a fit stepped in two calls give the
    same trees, bit for bit (the draws never disturb a capture)."""
    from turbocat import capi
    X, y = binary_data(200000, 12, 8)
    ds, _, _ = ctx.quantize_fit(X)
    kw = dict(n_trees=40, learning_rate=0.1, max_depth=6, lambda_l2=1.0, subsample=0.8,
              colsample_bytree=0.8, seed=5)
    a = ds.train(y, capi.train_params(capi.TC_LOSS_LOGLOSS, use_graph=1, **kw))
    b = ds.train(y, capi.train_params(capi.TC_LOSS_LOGLOSS, use_graph=0, **kw))
    tr = capi.Trainer(ctx, [ds], [y], capi.train_params(capi.TC_LOSS_LOGLOSS, use_graph=1, **kw))
    tr.step(3)
    tr.step(37)
    tr.sync()
    c = tr.model()
    tr.close()
    for k in ("feature", "bin", "fthr", "gain", "depth", "leaves"):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k


@pytest.mark.parametrize("loss,subsample,colsample", [(0, 0.8, 0.8), (1, 0.5, 1.0)])
def test_fit_sampled_teacher_forced(ctx, oracle, loss, subsample, colsample):
    """Row sampling on the device: each tree is grown on the reference's row
    sample (same draws, same shuffle) and every row's prediction is updated."""
    from turbocat import capi
    from oracle import tree_params as otp
    n, F, T, D, seed = 15000, 9, 12, 6, 99
    lam = 0.0 if loss == 0 else 1.0
    X, y = regression_data(n, F, 5) if loss == 0 else binary_data(n, F, 5)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    m = ds.train(y, capi.train_params(loss, n_trees=T, learning_rate=0.1, max_depth=D,
                                      lambda_l2=lam, subsample=subsample,
                                      colsample_bytree=colsample, seed=seed))
    preds = np.full(n, np.float32(m["base"]), np.float32)
    state, ties = seed, 0
    for t in range(T):
        s, state = capi.next_random(state)
        rows = capi.random_subsample(n, subsample, s)
        if colsample < 1.0:
            _, state = capi.next_random(state)
        g, h = oracle.gradients_f32(loss, y, preds)
        d = int(m["depth"][t])
        gpu = dict(depth=d, feature=m["feature"][t * D:t * D + d], bin=m["bin"][t * D:t * D + d],
                   fthr=m["fthr"][t * D:t * D + d], gain=m["gain"][t * D:t * D + d],
                   leaves=m["leaves"][t << D:(t << D) + (1 << d)])
        ref = oracle.tree_f64(bins, g, h, edges, ne, otp(max_depth=D, lambda_l2=lam), rows=rows)
        same = d == ref["depth"] and np.array_equal(gpu["feature"], ref["feature"]) and \
            np.array_equal(gpu["bin"], ref["bin"])
        if not same:
            k = next(i for i in range(min(d, ref["depth"]))
                     if (gpu["feature"][i], gpu["bin"][i]) != (ref["feature"][i], ref["bin"][i]))
            assert ref["gap"][k] <= NEAR_TIE * max(1.0, abs(ref["gain"][k]))
            ties += 1
        else:
            li = ref["leaf_idx"][rows]
            L = 1 << d
            absg = np.bincount(li, weights=np.abs(g[rows].astype(np.float64)), minlength=L)
            hs = np.bincount(li, weights=h[rows].astype(np.float64), minlength=L)
            assert rel_err(gpu["leaves"], ref["leaves"], absg / (hs + lam + 1e-30)).max() < TOL
        li_all = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
        preds = fma_update(oracle, preds, np.float32(0.1), gpu["leaves"][li_all])
    assert ties <= 1
    assert np.array_equal(m["train_preds"], preds)


@pytest.mark.parametrize("loss,colsample,top,other", [(1, 1.0, 0.2, 0.1), (0, 0.8, 0.3, 0.2)])
def test_fit_goss_teacher_forced(ctx, oracle, loss, colsample, top, other):
    """GOSS on the device's gradients: each tree is grown on the rows the
    reference's Dataset::goss_subsample draws from the same gradients, and
    every row's prediction is updated."""
    from turbocat import capi
    from oracle import tree_params as otp
    n, F, T, D, seed = 15000, 9, 10, 6, 77
    lam = 0.0 if loss == 0 else 1.0
    X, y = regression_data(n, F, 6) if loss == 0 else binary_data(n, F, 6)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    m = ds.train(y, capi.train_params(loss, n_trees=T, learning_rate=0.1, max_depth=D,
                                      lambda_l2=lam, colsample_bytree=colsample, seed=seed,
                                      use_goss=True, goss_top_rate=top, goss_other_rate=other))
    preds = np.full(n, np.float32(m["base"]), np.float32)
    state, ties = seed, 0
    for t in range(T):
        g, h = oracle.gradients_f32(loss, y, preds)
        s, state = capi.next_random(state)
        rows = capi.goss_subsample(g, top, other, s)
        if colsample < 1.0:
            _, state = capi.next_random(state)
        d = int(m["depth"][t])
        gpu = dict(depth=d, feature=m["feature"][t * D:t * D + d], bin=m["bin"][t * D:t * D + d],
                   fthr=m["fthr"][t * D:t * D + d], gain=m["gain"][t * D:t * D + d],
                   leaves=m["leaves"][t << D:(t << D) + (1 << d)])
        ref = oracle.tree_f64(bins, g, h, edges, ne, otp(max_depth=D, lambda_l2=lam), rows=rows)
        same = d == ref["depth"] and np.array_equal(gpu["feature"], ref["feature"]) and \
            np.array_equal(gpu["bin"], ref["bin"])
        if not same:
            k = next(i for i in range(min(d, ref["depth"]))
                     if (gpu["feature"][i], gpu["bin"][i]) != (ref["feature"][i], ref["bin"][i]))
            assert ref["gap"][k] <= NEAR_TIE * max(1.0, abs(ref["gain"][k]))
            ties += 1
        else:
            li = ref["leaf_idx"][rows]
            L = 1 << d
            absg = np.bincount(li, weights=np.abs(g[rows].astype(np.float64)), minlength=L)
            hs = np.bincount(li, weights=h[rows].astype(np.float64), minlength=L)
            assert rel_err(gpu["leaves"], ref["leaves"], absg / (hs + lam + 1e-30)).max() < TOL
        li_all = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
        preds = fma_update(oracle, preds, np.float32(0.1), gpu["leaves"][li_all])
    assert ties <= 1
    assert np.array_equal(m["train_preds"], preds)


@pytest.mark.parametrize("loss,ratio,colsample,n", [(1, 0.5, 1.0, 15000), (0, 0.3, 0.8, 15000),
                                                    (1, 0.4, 0.8, 15005)])  # LogLoss tail rows
def test_fit_mvs_teacher_forced(ctx, oracle, loss, ratio, colsample, n):
    """MVS on the device's gradients: each tree is grown on the rows the
    reference's Dataset::mvs_subsample draws from the same gradients (a
    sample of its own size each tree), and every row's prediction is
    updated."""
    from turbocat import capi
    from oracle import tree_params as otp
    F, T, D, seed = 9, 10, 6, 91
    lam = 0.0 if loss == 0 else 1.0
    X, y = regression_data(n, F, 8) if loss == 0 else binary_data(n, F, 8)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    m = ds.train(y, capi.train_params(loss, n_trees=T, learning_rate=0.1, max_depth=D,
                                      lambda_l2=lam, colsample_bytree=colsample, seed=seed,
                                      subsample=0.9, use_mvs=True, mvs_subsample=ratio))
    preds = np.full(n, np.float32(m["base"]), np.float32)
    state, ties, sizes = seed, 0, set()
    for t in range(T):
        g, h = oracle.gradients_f32(loss, y, preds)
        s, state = capi.next_random(state)
        rows = capi.mvs_subsample(g, ratio, s)
        sizes.add(len(rows))
        if colsample < 1.0:
            _, state = capi.next_random(state)
        d = int(m["depth"][t])
        gpu = dict(depth=d, feature=m["feature"][t * D:t * D + d], bin=m["bin"][t * D:t * D + d],
                   fthr=m["fthr"][t * D:t * D + d], gain=m["gain"][t * D:t * D + d],
                   leaves=m["leaves"][t << D:(t << D) + (1 << d)])
        ref = oracle.tree_f64(bins, g, h, edges, ne, otp(max_depth=D, lambda_l2=lam), rows=rows)
        same = d == ref["depth"] and np.array_equal(gpu["feature"], ref["feature"]) and \
            np.array_equal(gpu["bin"], ref["bin"])
        if not same:
            k = next(i for i in range(min(d, ref["depth"]))
                     if (gpu["feature"][i], gpu["bin"][i]) != (ref["feature"][i], ref["bin"][i]))
            assert ref["gap"][k] <= NEAR_TIE * max(1.0, abs(ref["gain"][k]))
            ties += 1
        else:
            li = ref["leaf_idx"][rows]
            L = 1 << d
            absg = np.bincount(li, weights=np.abs(g[rows].astype(np.float64)), minlength=L)
            hs = np.bincount(li, weights=h[rows].astype(np.float64), minlength=L)
            assert rel_err(gpu["leaves"], ref["leaves"], absg / (hs + lam + 1e-30)).max() < TOL
        li_all = oracle.leaf_index_binned(bins, gpu["feature"], gpu["bin"])
        preds = fma_update(oracle, preds, np.float32(0.1), gpu["leaves"][li_all])
    assert ties <= 1
    assert len(sizes) > 1  # the sample size followed the gradients
    assert np.array_equal(m["train_preds"], preds)


@pytest.mark.parametrize("loss", [0, 1])
def test_work_item_shape_never_changes_the_model(ctx, loss):
    """Histograms are exact integer sums with a fixed-point scale that does not
    depend on the work-item size, so an autotuned fit (shapes timed on its own
    first trees) and fits with fixed work-item sizes give identical models."""
    from turbocat import capi
    X, y = regression_data(60000, 12, 41) if loss == 0 else binary_data(60000, 12, 41)
    ds, _, _ = ctx.quantize_fit(X)
    lam = 0.0 if loss == 0 else 1.0
    models = []
    for hc, di in ((0, 0), (1024, 0), (4096, 0), (32768, 0), (2048, 64), (16384, 4096)):
        prm = capi.train_params(loss, n_trees=16, learning_rate=0.1, max_depth=6, lambda_l2=lam,
                                hist_chunk=hc, hist_deep_items=di)
        tr = capi.Trainer(ctx, [ds], [y], prm)
        tr.step(16)
        tr.sync()
        root, deep_items = tr.hist_chunks()
        if hc:
            assert (root, deep_items) == (hc, di)
        else:
            assert 1024 <= root <= 16384 and deep_items in (128, 256, 512, 1024)
        models.append(tr.model())
        tr.close()
    for m in models[1:]:
        for k in ("feature", "bin", "fthr", "gain", "depth", "leaves", "train_preds"):
            assert np.array_equal(m[k], models[0][k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("loss,K", [(0, 1), (1, 1), (2, 3)])
def test_atomic_histogram_mode_gives_the_same_model(ctx, loss, K, monkeypatch):
    """The two histogram flushes -- stored cells with partial slots summed by
    the split scan (default), and int64 atomics into zeroed node histograms
    (TURBOCAT_HIST_ATOMIC=1, the fallback when the slots would not fit) --
    produce bit-identical models: every histogram is an exact integer sum."""
    from turbocat import capi
    if K > 1:
        rng = np.random.default_rng(7)
        X = rng.standard_normal((40000, 20)).astype(np.float32)
        y = (np.argmax(X[:, :K] + 0.3 * rng.standard_normal((40000, K)), axis=1)).astype(np.float32)
    else:
        X, y = regression_data(40000, 20, 9) if loss == 0 else binary_data(40000, 20, 9)
    ds, _, _ = ctx.quantize_fit(X)
    models = []
    for mode in ("0", "1"):
        monkeypatch.setenv("TURBOCAT_HIST_ATOMIC", mode)
        prm = capi.train_params(loss, n_trees=12, learning_rate=0.1, max_depth=6,
                                lambda_l2=1.0, hist_chunk=2048, hist_deep_items=64,
                                n_classes=K if K > 1 else 0)
        tr = capi.Trainer(ctx, [ds], [y], prm)
        tr.step(12)
        tr.sync()
        models.append(tr.model())
        tr.close()
    for k in ("feature", "bin", "fthr", "gain", "depth", "leaves", "train_preds"):
        assert np.array_equal(models[0][k], models[1][k]), k
