"""GPU regular (non-oblivious) depthwise trees -- tc_rtree_build, tc_rtrain and
the regular-tree ensemble prediction -- against the CPU oracle
(tco_rtree_build_f64, tco_rtree_apply_binned, tco_rpredict_f32).

Bit-exact claims: every node's (feature, bin), is_leaf and default_left
flags and child ids in the reference's numbering (add_node order, the
max_leaves cut in depth-first smaller-child-first order), tree depth and
leaf count; predictions of an ensemble of given trees (both binning rules,
NaN, values equal to an edge); training predictions given the GPU's own
trees.  Float claims: leaf values and gains within 1e-5 of the FP64 oracle
(relative, on the scale of the sums they come from).  Near-tie rule: where
the FP64 gap between a node's best and second-best candidate is below
NEAR_TIE * max(1, |gain|), either choice is accepted and the rest of the tree
(whose numbering then differs) is not compared; such events are counted and
must stay rare.

Lossguide (Tree::build_leafwise): the same claims, in the expansion order's
numbering, on cases whose candidate-queue margin (the oracle's hgap: the
smallest gain difference at a pop) is above the float gains' rounding; with
room for every leaf the leaf-wise tree equals the depthwise one (compared
from the root, independent of numbering); boosting fits are compared tree by
tree the same way."""
import numpy as np
import pytest

from conftest import binary_data, multiclass_data, regression_data

pytestmark = pytest.mark.gpu

NEAR_TIE = 1e-5


def oparams(p):
    """The oracle's parameter struct for a capi RTreeParams."""
    from oracle import rtree_params
    return rtree_params(max_depth=p.max_depth, max_leaves=p.max_leaves, max_bins=p.max_bins,
                        min_samples_leaf=p.min_samples_leaf, min_child_weight=p.min_child_weight,
                        lambda_l2=p.lambda_l2, min_split_gain=p.min_split_gain, gamma=p.gamma,
                        leaf_smooth=p.leaf_smooth, max_delta_step=p.max_delta_step,
                        learn_missing_direction=bool(p.learn_missing_direction),
                        grow_policy=p.grow_policy)


def term_scales(ref, bins, g, h, lam, rows=None):
    """Per node of the oracle's tree, the largest of the gain's terms
    G^2 / (H + lambda) over the node and its two children (rows routed as the
    training partition routes them: NaN and bin <= split bin to the left).
    A float gain carries rounding of about 1e-7 of that, however small the
    gain itself."""
    n = bins.shape[1]
    g64, h64 = g.astype(np.float64), h.astype(np.float64)
    sc = np.zeros(ref["n_nodes"])
    if ref["n_nodes"] == 0:
        return sc
    root = np.zeros(n, bool)
    root[np.arange(n) if rows is None else np.asarray(rows, np.int64)] = True
    term = lambda m: g64[m].sum() ** 2 / (h64[m].sum() + lam)  # noqa: E731
    stack = [(0, root)]
    while stack:
        i, m = stack.pop()
        if int(ref["flags"][i]) & 1:
            sc[i] = term(m)
            continue
        col = bins[int(ref["feature"][i])]
        left = m & ((col == 255) | (col <= int(ref["bin"][i])))
        right = m & ~left
        sc[i] = max(term(m), term(left), term(right))
        stack += [(int(ref["left"][i]), left), (int(ref["right"][i]), right)]
    return sc


def compare_rtree(gpu, ref, g, h, p, bins=None, rows=None):
    """The two trees walked together from the root, left subtree first (the
    order the reference's depthwise recursion numbers and caps nodes in), so
    the first difference met is the first one the build made; then, with no
    near-tie met, the numbering itself.  Returns (nodes compared, near-tie
    hit)."""
    vscale = float(np.abs(g.astype(np.float64)).sum()) + 1.0
    sc = term_scales(ref, bins, g, h, p.lambda_l2, rows) if bins is not None else \
        np.zeros(ref["n_nodes"])
    stack, n = ([(0, 0)] if min(gpu["n_nodes"], ref["n_nodes"]) else []), 0
    while stack:
        i, j = stack.pop()
        n += 1
        gl, rl = int(gpu["flags"][i]) & 1, int(ref["flags"][j]) & 1
        eps = NEAR_TIE * max(1.0, abs(ref["gain"][j])) + 1e-6 * sc[j]
        if not rl and ref["gap"][j] <= eps:
            if gl or (gpu["feature"][i], gpu["bin"][i]) != (ref["feature"][j], ref["bin"][j]):
                return n, True
        if gl != rl:
            # a leaf on one side only: the decision must be borderline (the
            # best gain at min_split_gain)
            gain = float(ref["gain"][j]) if not rl else float(gpu["gain"][i])
            assert abs(gain - p.min_split_gain) <= 1e-4 * max(1.0, abs(gain)) + 1e-6 * sc[j], (
                f"node {j}: leaf on one side only (gpu leaf {gl}, ref leaf {rl}, gain {gain})")
            return n, True
        assert int(gpu["flags"][i]) == int(ref["flags"][j]), f"node {j}: flags"
        if rl:
            v_ref = float(ref["value"][j])
            assert abs(float(gpu["value"][i]) - v_ref) <= 1e-5 * max(1.0, abs(v_ref)) + \
                1e-6 * vscale / (p.lambda_l2 + 1e-3), (f"node {j}: leaf {gpu['value'][i]} vs {v_ref}")
            continue
        assert (int(gpu["feature"][i]), int(gpu["bin"][i])) == \
            (int(ref["feature"][j]), int(ref["bin"][j])), (
                f"node {j}: gpu ({gpu['feature'][i]},{gpu['bin'][i]}) ref "
                f"({ref['feature'][j]},{ref['bin'][j]}) gap {ref['gap'][j]}")
        assert abs(float(gpu["gain"][i]) - float(ref["gain"][j])) <= \
            1e-4 * max(1.0, abs(float(ref["gain"][j]))) + 1e-6 * sc[j], f"node {j}: gain"
        stack.append((int(gpu["right"][i]), int(ref["right"][j])))
        stack.append((int(gpu["left"][i]), int(ref["left"][j])))
    assert gpu["n_nodes"] == ref["n_nodes"], (gpu["n_nodes"], ref["n_nodes"])
    m = int(ref["n_nodes"])
    for key in ("left", "right", "flags"):
        assert np.array_equal(np.asarray(gpu[key])[:m], np.asarray(ref[key])[:m]), key
    assert gpu["n_leaves"] == ref["n_leaves"] and gpu["depth"] == ref["depth"]
    return n, False


def data_with_nan(n, F, seed, nan_frac=0.0):
    X, y = regression_data(n, F, seed, low_card_col=1)
    if nan_frac:
        rng = np.random.default_rng(seed + 99)
        X[rng.random((n, F)) < nan_frac] = np.nan
    return X, y


PARAMS = [
    dict(max_depth=6),
    dict(max_depth=8, max_leaves=20),                       # the max_leaves cut binds
    dict(max_depth=5, min_samples_leaf=40, min_child_weight=3.0),
    dict(max_depth=4, lambda_l2=0.0, min_split_gain=0.5, gamma=0.2),
    dict(max_depth=5, leaf_smooth=5.0, max_delta_step=0.3),
    dict(max_depth=3, learn_missing_direction=False, max_bins=16),
    dict(max_depth=1),
    dict(max_depth=10, max_leaves=64, min_child_weight=0.5),
]


@pytest.mark.parametrize("pi", range(len(PARAMS)))
@pytest.mark.parametrize("loss", ["logloss", "mse"])
def test_rtree_teacher_forced(ctx, oracle, pi, loss):
    """One tree on the oracle's gradients: the GPU's nodes are the FP64
    oracle's, in its numbering (NaN in 5% of the values: NaN rows go left at
    the partition and the learned default_left is recorded)."""
    from oracle import LOSS_LOGLOSS, LOSS_MSE
    from turbocat.capi import rtree_params
    X, yr = data_with_nan(6000, 9, 300 + pi, nan_frac=0.05)
    lid = LOSS_LOGLOSS if loss == "logloss" else LOSS_MSE
    y = (yr > 0).astype(np.float32) if loss == "logloss" else yr
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    p0 = np.full(len(y), oracle.init_prediction(lid, y), np.float32)
    g, h = oracle.gradients_f32(lid, y, p0)
    p = rtree_params(**PARAMS[pi])
    gpu = ds.rtree(g, h, p)
    ref = oracle.rtree_f64(bins, g, h, oparams(p))
    n, tie = compare_rtree(gpu, ref, g, h, p, bins)
    assert n >= 1
    if PARAMS[pi].get("max_leaves", 64) < (1 << PARAMS[pi]["max_depth"]) and not tie:
        assert gpu["n_leaves"] >= min(PARAMS[pi]["max_leaves"], 2)


def test_rtree_row_sample_and_empty_root(ctx, oracle):
    """The root's rows as a sample (the trainer's subsample), and no rows at
    all: a tree without nodes (tree.cpp:172-174)."""
    from oracle import LOSS_LOGLOSS
    from turbocat.capi import rtree_params
    X, yr = data_with_nan(5000, 6, 77, nan_frac=0.02)
    y = (yr > 0.3).astype(np.float32)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    g, h = oracle.gradients_f32(LOSS_LOGLOSS, y, np.zeros(len(y), np.float32))
    rows = np.random.default_rng(5).permutation(len(y))[:3100].astype(np.uint32)
    p = rtree_params(max_depth=7, max_leaves=30)
    _, tie = compare_rtree(ds.rtree(g, h, p, rows), oracle.rtree_f64(bins, g, h, oparams(p), rows),
                           g, h, p, bins, rows)
    empty = ds.rtree(g, h, p, np.zeros(0, np.uint32))
    assert empty["n_nodes"] == 0 and empty["n_leaves"] == 0


def test_rtree_kat_eight_rows(ctx, oracle):
    """The 8-row toy (tests/test_turbocat.py:55-56) with mode='small': the root
    splits (0, 3) with gain 2, both children are pure and become leaves -1 and
    +1 (every candidate gain there is negative), default_left stays 1."""
    from turbocat.capi import TC_LOSS_LOGLOSS, rtrain_params
    X = np.array([-2, -1, 1, 2, -1.5, -0.5, 0.5, 1.5], np.float32).reshape(8, 1)
    y = np.array([0, 0, 1, 1, 0, 0, 1, 1], np.float32)
    ds, _, _ = ctx.quantize_fit(X)
    m = ds.rtrain(y, rtrain_params(TC_LOSS_LOGLOSS, n_iterations=1, learning_rate=1.0, max_depth=3,
                                   lambda_l2=1.0, min_child_weight=0.01))
    t = m["trees"][0]
    assert m["base"] == 0.0
    assert t["n_nodes"] == 3 and t["n_leaves"] == 2 and t["depth"] == 1
    assert (int(t["feature"][0]), int(t["bin"][0]), float(t["gain"][0])) == (0, 3, 2.0)
    assert list(t["flags"]) == [2, 3, 3] and int(t["left"][0]) == 1 and int(t["right"][0]) == 2
    assert list(t["value"][1:]) == [-1.0, 1.0]
    assert list(m["train_preds"][0]) == [-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0]


def replay_binary(oracle, bins, y, lid, m, p, lr):
    """Teacher-forced replay of a binary / regression fit: per tree the
    oracle's gradients on the replayed predictions, its FP64 tree compared
    with the GPU's, then the GPU's tree applied (Tree::predict over bins, one
    fused multiply-add).  Returns (near ties, final predictions)."""
    preds = np.full(len(y), np.float32(m["base"]), np.float32)
    ties = 0
    for t, tree in enumerate(m["trees"]):
        g, h = oracle.gradients_f32(lid, y, preds)
        ref = oracle.rtree_f64(bins, g, h, oparams(p))
        _, tie = compare_rtree(tree, ref, g, h, p, bins)
        ties += tie
        preds = oracle.fma_update(preds, np.float32(lr), oracle.rtree_apply_binned(tree, bins))
    return ties, preds


@pytest.mark.parametrize("loss,subsample", [("logloss", 1.0), ("mse", 1.0), ("logloss", 0.7)])
def test_rtrain_matches_oracle(ctx, oracle, loss, subsample):
    """Booster::train's regular branch: 8 trees, every tree teacher-forced
    against the oracle on the replayed predictions; the GPU's training
    predictions equal the replay bit for bit (n % 8 == 0: every row takes the
    reference's vectorised gradient).  With subsample the trees grow on the
    reference's row samples (same engine and draws), which the oracle is
    given."""
    from oracle import LOSS_LOGLOSS, LOSS_MSE
    from turbocat import capi
    X, yr = data_with_nan(8000, 8, 11, nan_frac=0.01)
    lid = LOSS_LOGLOSS if loss == "logloss" else LOSS_MSE
    y = (yr > 0).astype(np.float32) if loss == "logloss" else yr
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    p = capi.rtree_params(max_depth=6, max_leaves=24, lambda_l2=1.0)
    lr = 0.3
    tp = capi.rtrain_params(lid, n_iterations=8, learning_rate=lr, subsample=subsample, seed=9)
    tp.tree = p
    m = ds.rtrain(y, tp)
    assert m["base"] == np.float32(oracle.init_prediction(lid, y))
    if subsample >= 1.0:
        ties, preds = replay_binary(oracle, bins, y, lid, m, p, lr)
        assert ties <= 1
        assert np.array_equal(m["train_preds"][0], preds)
        return
    # sampled: the oracle grows each tree on the same rows
    state = 9
    preds = np.full(len(y), np.float32(m["base"]), np.float32)
    ties = 0
    for tree in m["trees"]:
        seed, state = capi.next_random(state)
        rows = capi.random_subsample(len(y), subsample, seed)
        g, h = oracle.gradients_f32(lid, y, preds)
        _, tie = compare_rtree(tree, oracle.rtree_f64(bins, g, h, oparams(p), rows), g, h, p, bins,
                               rows)
        ties += tie
        preds = oracle.fma_update(preds, np.float32(lr), oracle.rtree_apply_binned(tree, bins))
    assert ties <= 1
    assert np.array_equal(m["train_preds"][0], preds)


def test_rtrain_multiclass_matches_oracle(ctx, oracle):
    """Booster::train_multiclass (booster.cpp:1263-1400): per iteration the
    softmax gradients, then one regular tree per class on that class's
    gradients.  Each tree is compared with the oracle's on the oracle's
    gradients of the replayed predictions (the device's softmax is the
    reference's float order with glibc's expf); the replayed predictions
    equal the GPU's bit for bit."""
    from turbocat import capi
    K = 3
    X, y = multiclass_data(6000, 10, K, 21)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    p = capi.rtree_params(max_depth=5, max_leaves=20, lambda_l2=1.0)
    lr = 0.2
    tp = capi.rtrain_params(capi.TC_LOSS_SOFTMAX, n_iterations=4, learning_rate=lr, n_classes=K)
    tp.tree = p
    m = ds.rtrain(y, tp)
    assert m["base"] == 0.0 and len(m["trees"]) == 4 * K
    preds = np.zeros((len(y), K), np.float32)
    ties = 0
    for it in range(4):
        g, h = oracle.gradients_mc_f32(y, preds)
        nxt = preds.copy()
        for k in range(K):
            tree = m["trees"][it * K + k]
            gk = np.ascontiguousarray(g[:, k]); hk = np.ascontiguousarray(h[:, k])
            _, tie = compare_rtree(tree, oracle.rtree_f64(bins, gk, hk, oparams(p)), gk, hk, p, bins)
            ties += tie
            nxt[:, k] = oracle.fma_update(preds[:, k].copy(), np.float32(lr),
                                          oracle.rtree_apply_binned(tree, bins))
        preds = nxt
    assert ties <= 1
    assert np.array_equal(m["train_preds"].T, preds)


@pytest.mark.parametrize("n", [8000, 7997])  # 7997: LogLoss's scalar tail rows select too
@pytest.mark.parametrize("kind", ["goss", "mvs"])
def test_rtrain_goss_mvs_matches_oracle(ctx, oracle, kind, n):
    """Booster::train's regular branch with GOSS or MVS (booster.cpp:198-219):
    each tree grows on the rows the reference's selection draws from the
    iteration's gradients (the device's gradients are the oracle's bit for
    bit, so capi.goss_subsample / mvs_subsample on the oracle's give the
    same rows), then the colsample draw; every tree teacher-forced against
    the oracle on those rows, and the replay equals the GPU's training
    predictions."""
    from oracle import LOSS_LOGLOSS
    from turbocat import capi
    X, yr = data_with_nan(n, 8, 13, nan_frac=0.01)
    y = (yr > 0).astype(np.float32)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    p = capi.rtree_params(max_depth=6, max_leaves=24, lambda_l2=1.0)
    lr, top, other, ratio = 0.3, 0.2, 0.2, 0.5
    tp = capi.rtrain_params(capi.TC_LOSS_LOGLOSS, n_iterations=8, learning_rate=lr, seed=9,
                            colsample_bytree=0.8, use_goss=kind == "goss",
                            goss_top_rate=top, goss_other_rate=other, use_mvs=kind == "mvs",
                            mvs_subsample=ratio)
    tp.tree = p
    m = ds.rtrain(y, tp)
    state = 9
    preds = np.full(len(y), np.float32(m["base"]), np.float32)
    ties = 0
    for tree in m["trees"]:
        g, h = oracle.gradients_f32(LOSS_LOGLOSS, y, preds)
        seed, state = capi.next_random(state)
        rows = capi.goss_subsample(g, top, other, seed) if kind == "goss" else \
            capi.mvs_subsample(g, ratio, seed)
        _, state = capi.next_random(state)  # colsample: drawn, unused
        _, tie = compare_rtree(tree, oracle.rtree_f64(bins, g, h, oparams(p), rows), g, h, p, bins,
                               rows)
        ties += tie
        preds = oracle.fma_update(preds, np.float32(lr), oracle.rtree_apply_binned(tree, bins))
    assert ties <= 1
    assert np.array_equal(m["train_preds"][0], preds)


def test_rtrain_multiclass_goss_matches_oracle(ctx, oracle):
    """The multiclass loop with GOSS (booster.cpp:1326-1347): one sample per
    iteration, selected on each row's largest |g| over the classes, shared by
    the K class trees; each tree against the oracle's on those rows."""
    from turbocat import capi
    K = 3
    X, y = multiclass_data(6000, 10, K, 23)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    p = capi.rtree_params(max_depth=5, max_leaves=20, lambda_l2=1.0)
    lr, top, other = 0.2, 0.3, 0.2
    tp = capi.rtrain_params(capi.TC_LOSS_SOFTMAX, n_iterations=4, learning_rate=lr, n_classes=K,
                            seed=17, use_goss=True, goss_top_rate=top, goss_other_rate=other,
                            use_mvs=True)  # MVS is not a multiclass option: GOSS is used
    tp.tree = p
    m = ds.rtrain(y, tp)
    state = 17
    preds = np.zeros((len(y), K), np.float32)
    ties = 0
    for it in range(4):
        g, h = oracle.gradients_mc_f32(y, preds)
        gsel = np.abs(g).max(axis=1).astype(np.float32)
        seed, state = capi.next_random(state)
        rows = capi.goss_subsample(gsel, top, other, seed)
        nxt = preds.copy()
        for k in range(K):
            tree = m["trees"][it * K + k]
            gk = np.ascontiguousarray(g[:, k]); hk = np.ascontiguousarray(h[:, k])
            ref = oracle.rtree_f64(bins, gk, hk, oparams(p), rows)
            _, tie = compare_rtree(tree, ref, gk, hk, p, bins, rows)
            ties += tie
            nxt[:, k] = oracle.fma_update(preds[:, k].copy(), np.float32(lr),
                                          oracle.rtree_apply_binned(tree, bins))
        preds = nxt
    assert ties <= 1
    assert np.array_equal(m["train_preds"].T, preds)


def test_dropin_regular_goss(ctx):
    """use_goss with mode="small" and with more than two classes trains
    (the reference samples regular trees' rows the same way)."""
    from turbocat import TurboCatClassifier
    X, y = multiclass_data(4000, 8, 3, 29)
    clf = TurboCatClassifier(n_estimators=15, mode="small", use_goss=True, subsample=1.0,
                             colsample_bytree=1.0, verbosity=0)
    clf.fit(X, (y == 1).astype(np.float32))
    assert clf.predict_proba(X).shape == (4000, 2)
    mc = TurboCatClassifier(n_estimators=10, use_goss=True, subsample=1.0, colsample_bytree=1.0,
                            verbosity=0)
    mc.fit(X, y)
    assert (mc.predict(X) == y).mean() > 0.6


@pytest.mark.parametrize("kind", ["binary", "multiclass"])
def test_rpredict_matches_oracle(ctx, oracle, kind):
    """tc_rpredict of trained regular trees equals the oracle's
    tco_rpredict_f32 bit for bit: rule 1 (predict_raw_fast, binary) and rule
    0 (apply_bins, multiclass with the trees' classes), with NaN values,
    values equal to an edge and a row count that is not a multiple of 64."""
    from oracle import LOSS_LOGLOSS
    from turbocat import capi
    if kind == "binary":
        X, yr = data_with_nan(4000, 7, 3, nan_frac=0.03)
        y = (yr > 0).astype(np.float32)
        K, loss, rule = 1, LOSS_LOGLOSS, 1
    else:
        X, y = multiclass_data(4000, 7, 4, 8)
        K, loss, rule = 4, capi.TC_LOSS_SOFTMAX, 0
    ds, edges, ne = ctx.quantize_fit(X)
    tp = capi.rtrain_params(loss, n_iterations=12, learning_rate=0.2, n_classes=K if K > 1 else 0)
    tp.tree = capi.rtree_params(max_depth=6, max_leaves=40)
    m = ds.rtrain(y, tp)
    rng = np.random.default_rng(4)
    Xp = rng.standard_normal((1003, X.shape[1])).astype(np.float32)
    Xp[rng.random(Xp.shape) < 0.03] = np.nan
    for f in range(X.shape[1]):  # values that equal an edge
        if ne[f]:
            Xp[f * 7:(f + 1) * 7, f] = edges[f, rng.integers(0, ne[f], 7)]
    ens = capi.REnsemble(ctx, m["trees"], m["weight"], edges, ne, X.shape[1], rule=rule, K=K,
                         tree_class=m["tree_class"] if K > 1 else None, base=m["base"])
    ref = oracle.rpredict_f32(m["trees"], m["weight"], Xp, edges, ne, rule, K=K,
                              tree_class=m["tree_class"] if K > 1 else None, base=m["base"])
    out = ens.predict(Xp)
    assert np.array_equal(out.reshape(-1, K), ref)
    prob = ens.predict(Xp, transform=1).reshape(-1, K)
    assert np.array_equal(prob, oracle.transform_f32(ref, K).reshape(-1, K))
    if K == 1:
        assert np.allclose(prob[:, 0], 1 / (1 + np.exp(-ref[:, 0].astype(np.float64))), atol=1e-6)
    else:
        e = np.exp(ref - ref.max(1, keepdims=True))
        assert np.allclose(prob, e / e.sum(1, keepdims=True), atol=1e-6)
    if rule == 0:
        # apply_bins bins as training did (upper_bound) and the data has no
        # NaN, so the training rows through prediction reproduce the training
        # update's sums (the order of the additions differs: within 1e-5).
        # Not so under rule 1: with more than 16 edges a value equal to an edge
        # -- every training value is one -- falls one bin lower there, and NaN
        # always goes left, as in the reference.
        tr = ens.predict(X).reshape(-1, K)
        assert np.abs(tr.T - m["train_preds"]).max() <= 1e-5 * max(1.0, float(np.abs(tr).max()))


def test_rtrain_graph_equals_stepwise(ctx):
    """The per-iteration graph replay and plain launches build identical
    models."""
    from turbocat import capi
    X, y = binary_data(5000, 12, 4)
    ds, _, _ = ctx.quantize_fit(X)
    a = []
    for use_graph in (0, 1):
        tp = capi.rtrain_params(capi.TC_LOSS_LOGLOSS, n_iterations=5, learning_rate=0.3,
                                use_graph=use_graph)
        tp.tree = capi.rtree_params(max_depth=6, max_leaves=20)
        a.append(ds.rtrain(y, tp))
    for k in ("feature", "bin", "flags", "left", "right", "value", "n_nodes"):
        assert np.array_equal(a[0]["raw"][k], a[1]["raw"][k]), k
    assert np.array_equal(a[0]["train_preds"], a[1]["train_preds"])


@pytest.mark.parametrize("kind", ["regression", "multiclass"])
def test_rtrain_validation_and_early_stopping(ctx, oracle, kind):
    """Validation rows after every iteration (booster.cpp:306-371,
    :1400-1430): the GPU's per-iteration validation losses equal the oracle's
    replay -- each tree applied to the validation bins (apply_bins, NaN by
    default_left), one fused multiply-add, Loss::compute_loss -- within 2e-6
    relative; the iterations built follow the reference's stopping rule on
    the GPU's own losses; the trees are those of a fit without validation."""
    from oracle import LOSS_MSE
    from turbocat import capi
    rounds = 4
    if kind == "regression":
        # NaN-free: with NaN the reference's child sums leave out the NaN rows
        # that the partition sends left, and MSE at this learning rate diverges
        # (the oracle's own replay does too)
        X, y = data_with_nan(6000, 8, 31)
        Xv, yv = data_with_nan(1500, 8, 32)
        loss, K = capi.TC_LOSS_MSE, 1
    else:
        X, y = multiclass_data(5000, 8, 3, 41)
        Xv, yv = multiclass_data(1200, 8, 3, 42)
        loss, K = capi.TC_LOSS_SOFTMAX, 3
    ds, edges, ne = ctx.quantize_fit(X)
    vds = ctx.dataset(Xv, edges, ne)
    vb = vds.bins()
    tp = capi.rtrain_params(loss, n_iterations=120, learning_rate=0.5, n_classes=K if K > 1 else 0)
    tp.tree = capi.rtree_params(max_depth=6, max_leaves=40, min_child_weight=0.5)
    m, h = ds.rtrain_valid(y, tp, vds, yv, early_stopping_rounds=rounds)
    built = h["n_built"]
    assert 1 <= built <= 120
    vl = h["valid_loss"]
    kept, best_it, _ = oracle.early_stop(vl, 120, rounds)
    assert (kept, best_it) == (built, h["best_iteration"])
    if built < 120:
        assert h["stopped"]
        assert all(tr["n_nodes"] == 0 for tr in m["trees"][built * K:])
    # replay of the validation predictions and losses
    vp = np.full((len(yv), K), np.float32(m["base"]), np.float32)
    for it in range(built):
        for k in range(K):
            tr = m["trees"][it * K + k]
            vp[:, k] = oracle.fma_update(vp[:, k].copy(), np.float32(tp.learning_rate),
                                         oracle.rtree_apply_binned(tr, vb))
        ref = oracle.compute_loss_f32(LOSS_MSE, yv, vp[:, 0]) if K == 1 else \
            oracle.compute_loss_mc_f32(yv, vp)
        assert abs(vl[it] - ref) <= 2e-6 * abs(ref), (it, vl[it], ref)
    # validation does not change the trees
    plain = ds.rtrain(y, tp)
    for a, b in zip(m["trees"][:built * K], plain["trees"][:built * K]):
        for key in ("feature", "bin", "flags", "left", "right", "value"):
            assert np.array_equal(a[key], b[key]), key


def test_dropin_mode_small_and_multiclass(tmp_path):
    """The drop-in classes: mode='small' grows regular trees and predicts
    through predict_raw_fast's binning (rule 1); labels with three classes
    train the reference's multiclass loop and predict through apply_bins
    (rule 0) and softmax; both equal the C-ABI path bit for bit, survive a
    TCAT save/load, and fit the training data."""
    import turbocat
    from turbocat import capi
    ctx = capi.Context(0)
    # continuous features: with more than 16 edges predict_raw_fast bins a value
    # equal to an edge one lower than training did (std::lower_bound), so on a
    # low-cardinality feature, every value of which is an edge, the reference's
    # mode="small" predictions of its own training rows are off (reproduced
    # bit for bit in test_rpredict_matches_oracle)
    X, yr = regression_data(3000, 6, 51)
    y = (yr > 0).astype(np.float32)
    clf = turbocat.TurboCatClassifier(mode="small", n_estimators=20, learning_rate=0.3,
                                      max_depth=5, subsample=1.0, colsample_bytree=1.0,
                                      verbosity=0).fit(X, y)
    assert clf.n_trees == 20
    p = clf.predict_proba(X)
    assert p.shape == (3000, 2) and p.dtype == np.float32 and np.allclose(p.sum(1), 1.0)
    assert (clf.predict(X) == y).mean() > 0.8
    ds, edges, ne = ctx.quantize_fit(X)
    tp = capi.rtrain_params(capi.TC_LOSS_LOGLOSS, n_iterations=20, learning_rate=0.3, seed=42)
    tp.tree = capi.rtree_params(max_depth=5, max_leaves=64, lambda_l2=1.0)
    m = ds.rtrain(y, tp)
    ens = capi.REnsemble(ctx, m["trees"], m["weight"], edges, ne, 6, rule=1, base=m["base"])
    assert np.array_equal(p[:, 1], ens.predict(X, transform=1))
    path = tmp_path / "small.tcat"
    clf.save(str(path))
    again = turbocat.TurboCatClassifier().load(str(path))
    assert np.array_equal(again.predict_proba(X), p)

    from conftest import multiclass_data
    Xm, ym = multiclass_data(3000, 6, 3, 52)
    mc = turbocat.TurboCatClassifier(n_estimators=15, learning_rate=0.3, max_depth=5,
                                     subsample=1.0, colsample_bytree=1.0, verbosity=0).fit(Xm, ym)
    assert mc.n_classes_ == 3 and mc.n_trees == 45
    pm = mc.predict_proba(Xm)
    assert pm.shape == (3000, 3) and np.allclose(pm.sum(1), 1.0, atol=1e-6)
    assert (mc.predict(Xm) == ym).mean() > 0.7
    dsm, em, nem = ctx.quantize_fit(Xm)
    tpm = capi.rtrain_params(capi.TC_LOSS_SOFTMAX, n_iterations=15, learning_rate=0.3, n_classes=3)
    tpm.tree = capi.rtree_params(max_depth=5, max_leaves=64, lambda_l2=1.0)
    mm = dsm.rtrain(ym, tpm)
    ensm = capi.REnsemble(ctx, mm["trees"], mm["weight"], em, nem, 6, rule=0, K=3,
                          tree_class=mm["tree_class"], base=0.0)
    assert np.array_equal(pm, ensm.predict(Xm, transform=1))
    path = tmp_path / "multi.tcat"
    mc.save(str(path))
    again = turbocat.TurboCatClassifier().load(str(path))
    assert again.n_classes_ == 3 and np.array_equal(again.predict_proba(Xm), pm)
    reg = turbocat.TurboCatRegressor(mode="small", n_estimators=30, subsample=1.0,
                                     colsample_bytree=1.0, verbosity=0).fit(X, yr)
    assert reg.n_trees == 30
    assert np.corrcoef(reg.predict(X), yr)[0, 1] > 0.8


# ------------------------------------------------------------------ lossguide
LW_CASES = [  # (tree params, loss, data seed): every choice with a margin above rounding
    (dict(max_depth=6, max_leaves=16), "mse", 680),
    (dict(max_depth=4, max_leaves=64), "logloss", 437),
    (dict(max_depth=4, max_leaves=64), "mse", 402),
    (dict(max_depth=5, max_leaves=6, min_child_weight=3.0, lambda_l2=0.0, min_split_gain=0.5),
     "logloss", 410),
    (dict(max_depth=5, max_leaves=6, min_child_weight=3.0, lambda_l2=0.0, min_split_gain=0.5),
     "mse", 403),
    (dict(max_depth=8, max_leaves=10, min_samples_leaf=50), "logloss", 507),
    (dict(max_depth=8, max_leaves=10, min_samples_leaf=50), "mse", 500),
    (dict(max_depth=10, max_leaves=8, min_child_weight=5.0), "logloss", 556),
    (dict(max_depth=10, max_leaves=8, min_child_weight=5.0), "mse", 521),
    (dict(max_depth=7, max_leaves=6), "mse", 500),
]


def lw_data(ctx, oracle, loss, seed):
    from oracle import LOSS_LOGLOSS, LOSS_MSE
    X, yr = data_with_nan(6000, 9, seed, nan_frac=0.05)
    lid = LOSS_LOGLOSS if loss == "logloss" else LOSS_MSE
    y = (yr > 0).astype(np.float32) if loss == "logloss" else yr
    ds, edges, ne = ctx.quantize_fit(X)
    p0 = np.full(len(y), oracle.init_prediction(lid, y), np.float32)
    g, h = oracle.gradients_f32(lid, y, p0)
    return ds, ds.bins(), g, h


def lossguide_margin(ref, bins, g, h, p):
    """The oracle tree's smallest margin, in units of the comparison's
    rounding allowance: every split's gap to its runner-up (0 at a hessian
    border), its distance from min_split_gain, and the expansion order's gap
    (hgap).  Above 1, a float implementation makes the same choices and
    expands the nodes in the same order."""
    with np.errstate(all="ignore"):
        sc = np.nan_to_num(term_scales(ref, bins, g, h, p.lambda_l2))
    m = ref["hgap"] / (NEAR_TIE * max(1.0, float(np.abs(ref["gain"]).max())) + 1e-6 * float(sc.max()))
    for i in range(ref["n_nodes"]):
        if int(ref["flags"][i]) & 1:
            continue
        eps = NEAR_TIE * max(1.0, abs(ref["gain"][i])) + 1e-6 * sc[i]
        m = min(m, ref["gap"][i] / eps, abs(ref["gain"][i] - p.min_split_gain) / (10 * eps))
    return m


@pytest.mark.parametrize("ci", range(len(LW_CASES)))
def test_rtree_lossguide_teacher_forced(ctx, oracle, ci):
    """One leaf-wise tree on the oracle's gradients: the GPU's nodes are the
    FP64 oracle's build_leafwise, numbered in expansion order, on cases whose
    every choice has a margin (a near-tie anywhere may reorder every later
    expansion, and the numbering with it)."""
    from turbocat.capi import rtree_params
    kw, loss, seed = LW_CASES[ci]
    ds, bins, g, h = lw_data(ctx, oracle, loss, seed)
    p = rtree_params(grow_policy=1, **kw)
    gpu = ds.rtree(g, h, p)
    ref = oracle.rtree_f64(bins, g, h, oparams(p))
    assert lossguide_margin(ref, bins, g, h, p) > 1.0
    n, tie = compare_rtree(gpu, ref, g, h, p, bins)
    assert not tie and n == ref["n_nodes"]


def walk_compare(gpu, ref, p, vscale, i=0, j=0):
    """Two trees from the root, independent of numbering: the same splits
    (outside per-node FP64 near-ties) and leaf values within 1e-5.
    Returns True when a near-tie stopped the walk."""
    gl, rl = int(gpu["flags"][i]) & 1, int(ref["flags"][j]) & 1
    if not rl and ref["gap"][j] <= NEAR_TIE * max(1.0, abs(ref["gain"][j])):
        if gl or (gpu["feature"][i], gpu["bin"][i]) != (ref["feature"][j], ref["bin"][j]):
            return True
    assert gl == rl and int(gpu["flags"][i]) == int(ref["flags"][j]), (i, j)
    if rl:
        v = float(ref["value"][j])
        assert abs(float(gpu["value"][i]) - v) <= 1e-5 * max(1.0, abs(v)) + \
            1e-6 * vscale / (p.lambda_l2 + 1e-3), (i, j)
        return False
    assert (int(gpu["feature"][i]), int(gpu["bin"][i])) == (int(ref["feature"][j]), int(ref["bin"][j]))
    a = walk_compare(gpu, ref, p, vscale, int(gpu["left"][i]), int(ref["left"][j]))
    b = walk_compare(gpu, ref, p, vscale, int(gpu["right"][i]), int(ref["right"][j]))
    return a or b


@pytest.mark.parametrize("D,seed", [(3, 410), (6, 411), (9, 412)])
def test_rtree_lossguide_without_a_cut_is_the_depthwise_tree(ctx, oracle, D, seed):
    """With room for every leaf (max_leaves >= 2^max_depth) a leaf-wise tree
    grows every splittable node, as the depthwise tree does: the two GPU
    trees agree from the root down (numbered differently), and the leaf-wise
    tree matches the oracle's the same way."""
    from turbocat.capi import rtree_params
    ds, bins, g, h = lw_data(ctx, oracle, "logloss", seed)
    kw = dict(max_depth=D, max_leaves=1 << D, min_samples_leaf=10)
    a = ds.rtree(g, h, rtree_params(**kw))
    p = rtree_params(grow_policy=1, **kw)
    b = ds.rtree(g, h, p)
    assert (a["n_nodes"], a["n_leaves"], a["depth"]) == (b["n_nodes"], b["n_leaves"], b["depth"])

    def canon(t, i=0):
        if int(t["flags"][i]) & 1:
            return ("leaf", int(t["flags"][i]), float(t["value"][i]))
        return (int(t["feature"][i]), int(t["bin"][i]), int(t["flags"][i]),
                canon(t, int(t["left"][i])), canon(t, int(t["right"][i])))
    assert canon(a) == canon(b)
    ref = oracle.rtree_f64(bins, g, h, oparams(p))
    vscale = float(np.abs(g.astype(np.float64)).sum()) + 1.0
    assert not walk_compare(b, ref, p, vscale)


def test_rtrain_lossguide_matches_oracle(ctx, oracle):
    """Boosting with leaf-wise trees (Booster::train's regular branch with
    GrowPolicy::Lossguide): each tree against the oracle's on the replayed
    predictions, compared from the root; the replay equals the GPU's
    training predictions bit for bit."""
    from turbocat import capi
    from oracle import LOSS_LOGLOSS
    X, yr = regression_data(8000, 8, 31)
    y = (yr > 0).astype(np.float32)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    p = capi.rtree_params(max_depth=5, max_leaves=32, lambda_l2=1.0, grow_policy=1)
    tp = capi.rtrain_params(capi.TC_LOSS_LOGLOSS, n_iterations=6, learning_rate=0.3)
    tp.tree = p
    m = ds.rtrain(y, tp)
    preds = np.full(len(y), np.float32(m["base"]), np.float32)
    ties = 0
    for t in m["trees"]:
        g, h = oracle.gradients_f32(LOSS_LOGLOSS, y, preds)
        ref = oracle.rtree_f64(bins, g, h, oparams(p))
        vscale = float(np.abs(g.astype(np.float64)).sum()) + 1.0
        ties += walk_compare(t, ref, p, vscale)
        preds = oracle.fma_update(preds, np.float32(0.3), oracle.rtree_apply_binned(t, bins))
    assert ties <= 1
    assert np.array_equal(m["train_preds"][0], preds)


def test_rtree_lossguide_parameter_checks(ctx):
    from turbocat import capi
    X, y = regression_data(500, 4, 7)
    ds, _, _ = ctx.quantize_fit(X)
    g = np.ones(500, np.float32)
    with pytest.raises(capi.TurboCatError, match="grow_policy"):
        ds.rtree(g, g, capi.rtree_params(grow_policy=2))
    with pytest.raises(capi.TurboCatError, match="lossguide"):
        ds.rtree(g, g, capi.rtree_params(max_depth=12, grow_policy=1))


def test_dropin_lossguide(ctx):
    """grow_policy='lossguide' (also 'leaf', 'leafwise') grows leaf-wise
    regular trees through the drop-in classes; saving and loading keeps the
    predictions."""
    import os
    import tempfile
    from turbocat import TurboCatRegressor
    X, y = regression_data(5000, 6, 52)
    reg = TurboCatRegressor(n_estimators=30, mode="small", grow_policy="lossguide", max_depth=8,
                            max_leaves=24, subsample=1.0, colsample_bytree=1.0, verbosity=0)
    reg.fit(X, y)
    pr = reg.predict(X)
    assert 1.0 - np.mean((pr - y) ** 2) / np.var(y) > 0.8
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "m.tcat")
        reg.save(path)
        r2 = TurboCatRegressor(mode="small", verbosity=0).load(path)
        assert np.array_equal(r2.predict(X), pr)
    # the classifier's grow_policy is a keyword of the pybind class (the
    # reference's Python wrapper does not forward it), as in the reference
    from turbocat import _turbocat
    yc = (y > np.median(y)).astype(np.float32)
    Xe, ye = np.zeros((0, X.shape[1]), np.float32), np.zeros(0, np.float32)
    for gp in ("leaf", "leafwise"):
        clf = _turbocat.TurboCatClassifier(n_estimators=20, mode="small", grow_policy=gp,
                                           max_depth=6, subsample=1.0, colsample_bytree=1.0,
                                           verbosity=0)
        clf.fit(X, yc, Xe, ye, [])
        assert (np.asarray(clf.predict(X)) == yc).mean() > 0.85
