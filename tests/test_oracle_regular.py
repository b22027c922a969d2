"""Pinning the regular-tree oracle (tco_rtree_build_f64, tco_rtree_apply_binned,
tco_rpredict_f32) on the CPU.

The reference's regular trees (src/tree/tree.cpp) cannot be compiled here:
tree.hpp includes Eigen, which the image lacks.  The oracle is pinned instead
by:

* an exhaustive builder below, written from the reference's rules in exact
  rational arithmetic: Tree::build's Depthwise recursion (tree.cpp:162-217,
  :297-390) -- stop at max_depth, at max_leaves leaves made so far, under
  2*min_samples_leaf rows or 2*min_child_weight hessian; SplitFinder
  (histogram.cpp:569-703) -- every (feature, bin < max_bins - 1), both sides
  with count >= min_samples_leaf and hess >= min_child_weight, gain
  0.5*(GL^2/(HL+l) + GR^2/(HR+l) - G^2/(H+l)) - gamma, strict > so ties go to
  the lowest (f, b), gain >= min_split_gain; NaN rows to the left, then
  default_left learned from the side sizes; children numbered in add_node
  order and the smaller child's subtree built first; a child's node stats
  being the split's prefix (NaN rows excluded) or parent minus it; make_leaf
  (tree.cpp:655-681) with smoothing and clamp;
* known-answer trees derived by hand (the 8-row toy);
* the compiled reference itself where the two tree kinds coincide: a depth-1
  regular tree is the symmetric tree's root split (oracle/_ref);
* planted mistakes that these comparisons must notice.

Lossguide (Tree::build_leafwise, tree.cpp:685-925) is pinned the same way: an
exact leaf-wise builder below (the candidate of largest gain expanded first;
children queued left first when they may split, one more leaf would fit and
their split is valid; depth-limited candidates and the queue's remainder
become leaves), the invariant that with max_leaves >= 2^max_depth it grows
the depthwise tree (numbered differently), and the candidate queue's pop
order against std::priority_queue itself (compiled here with g++), ties
included.
"""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from conftest import regression_data
from oracle import LOSS_LOGLOSS, LOSS_MSE, RefLib, rtree_params

HAVE_REF = RefLib.available()


def exact_tree(bins, g, h, p, rows=None, mistake=None):
    """Regular depthwise tree by exhaustive search in exact arithmetic.
    Returns node dicts in the reference's numbering."""
    F, n = bins.shape
    lam = Fraction(float(p.lambda_l2))
    gamma = Fraction(float(p.gamma))
    msg = Fraction(float(p.min_split_gain))
    mcw = Fraction(float(p.min_child_weight))
    gF = [Fraction(float(v)) for v in g]
    hF = [Fraction(float(v)) for v in h]
    nodes, state = [], dict(leaves=0, depth=0)

    def add():
        nodes.append(dict(feature=0, bin=0, leaf=True, dl=True, left=0, right=0, value=0,
                          gain=0))
        return len(nodes) - 1

    def leaf(i, G, H, C):
        v = -G / (H + lam)
        if p.leaf_smooth > 0:
            v = v * (Fraction(C) / (Fraction(C) + Fraction(float(p.leaf_smooth))))
        if p.max_delta_step > 0:
            m = Fraction(float(p.max_delta_step))
            v = max(-m, min(m, v))
        nodes[i]["value"] = v
        state["leaves"] += 1

    def build(i, rs, G, H, C, d):
        state["depth"] = max(state["depth"], d)
        ml = p.max_leaves if mistake != "no_max_leaves" else 1 << 30
        if (d >= p.max_depth or state["leaves"] >= ml or len(rs) < 2 * p.min_samples_leaf
                or H < 2 * mcw):
            leaf(i, G, H, C)
            return
        best = None
        pg = G * G / (H + lam)
        for f in range(F):
            col = bins[f]
            top = int(max((col[r] for r in rs if col[r] != 255), default=0))
            for b in range(min(p.max_bins - 1, top + 1)):
                L = [r for r in rs if col[r] <= b]
                GL = sum((gF[r] for r in L), Fraction(0))
                HL = sum((hF[r] for r in L), Fraction(0))
                CL = len(L)
                GR, HR, CR = G - GL, H - HL, C - CL
                if CL < p.min_samples_leaf or HL < mcw or CR < p.min_samples_leaf or HR < mcw:
                    continue
                gn = Fraction(1, 2) * (GL * GL / (HL + lam) + GR * GR / (HR + lam) - pg) - gamma
                if gn >= msg and (best is None or gn > best[0]):
                    best = (gn, f, b, GL, HL, CL)
        if best is None:
            leaf(i, G, H, C)
            return
        gn, f, b, GL, HL, CL = best
        nd = nodes[i]
        nd.update(feature=f, bin=b, leaf=False, gain=gn)
        col = bins[f]
        lrows = [r for r in rs if (col[r] == 255 and nd["dl"]) or (col[r] != 255 and col[r] <= b)]
        rrows = [r for r in rs if r not in set(lrows)]
        if p.learn_missing_direction and 0 < len(lrows) < len(rs):
            nd["dl"] = len(lrows) >= len(rrows)
        lc, rc = add(), add()
        nd["left"], nd["right"] = lc, rc
        left_first = len(lrows) <= len(rrows)
        if mistake == "larger_first":
            left_first = not left_first
        jobs = [(lc, lrows, GL, HL, CL), (rc, rrows, G - GL, H - HL, C - CL)]
        if not left_first:
            jobs.reverse()
        for j in jobs:
            build(j[0], j[1], j[2], j[3], j[4], d + 1)

    rs = list(range(n)) if rows is None else [int(r) for r in rows]
    if rs:
        G = sum((gF[r] for r in rs), Fraction(0))
        H = sum((hF[r] for r in rs), Fraction(0))
        build(add(), rs, G, H, len(rs), 0)
    return nodes, state["leaves"], state["depth"]


def assert_same(tree, exact):
    nodes, nl, depth = exact
    assert tree["n_nodes"] == len(nodes)
    assert tree["n_leaves"] == nl and tree["depth"] == depth
    for i, nd in enumerate(nodes):
        flags = (1 if nd["leaf"] else 0) | (2 if nd["dl"] else 0)
        assert int(tree["flags"][i]) == flags, f"node {i} flags"
        if nd["leaf"]:
            assert abs(tree["value"][i] - float(nd["value"])) <= 1e-9 * max(1.0, abs(float(nd["value"])))
        else:
            assert (int(tree["feature"][i]), int(tree["bin"][i])) == (nd["feature"], nd["bin"]), f"node {i}"
            assert (int(tree["left"][i]), int(tree["right"][i])) == (nd["left"], nd["right"])
            assert abs(tree["gain"][i] - float(nd["gain"])) <= 1e-9 * max(1.0, abs(float(nd["gain"])))


def exact_leafwise(bins, g, h, p, rows=None, mistake=None):
    """Regular lossguide tree in exact arithmetic (build_leafwise,
    tree.cpp:685-925).  Returns (nodes, leaves, depth, tie): tie is True when
    two queued candidates had equal exact gains at a pop (the library queue
    then decides; pinned separately)."""
    F, n = bins.shape
    lam = Fraction(float(p.lambda_l2))
    gamma = Fraction(float(p.gamma))
    msg = Fraction(float(p.min_split_gain))
    mcw = Fraction(float(p.min_child_weight))
    gF = [Fraction(float(v)) for v in g]
    hF = [Fraction(float(v)) for v in h]
    nodes, st = [], dict(leaves=0, depth=0, tie=False)

    def add():
        nodes.append(dict(feature=0, bin=0, leaf=True, dl=True, left=0, right=0, value=0,
                          gain=0))
        return len(nodes) - 1

    def leaf(i, G, H, C):
        v = -G / (H + lam)
        if p.leaf_smooth > 0:
            v = v * (Fraction(C) / (Fraction(C) + Fraction(float(p.leaf_smooth))))
        if p.max_delta_step > 0:
            m = Fraction(float(p.max_delta_step))
            v = max(-m, min(m, v))
        nodes[i]["value"] = v
        st["leaves"] += 1

    def find(rs, G, H, C):
        best = None
        pg = G * G / (H + lam)
        for f in range(F):
            col = bins[f]
            top = int(max((col[r] for r in rs if col[r] != 255), default=0))
            for b in range(min(p.max_bins - 1, top + 1)):
                L = [r for r in rs if col[r] <= b]
                GL = sum((gF[r] for r in L), Fraction(0))
                HL = sum((hF[r] for r in L), Fraction(0))
                CL = len(L)
                GR, HR, CR = G - GL, H - HL, C - CL
                if CL < p.min_samples_leaf or HL < mcw or CR < p.min_samples_leaf or HR < mcw:
                    continue
                gn = Fraction(1, 2) * (GL * GL / (HL + lam) + GR * GR / (HR + lam) - pg) - gamma
                if gn >= msg and (best is None or gn > best[0]):
                    best = (gn, f, b, GL, HL, CL)
        return best

    rs = list(range(n)) if rows is None else [int(r) for r in rows]
    if not rs:
        return nodes, 0, 0, False
    G = sum((gF[r] for r in rs), Fraction(0))
    H = sum((hF[r] for r in rs), Fraction(0))
    root = add()
    best = find(rs, G, H, len(rs))
    if best is None:
        leaf(root, G, H, len(rs))
        return nodes, st["leaves"], 0, False
    queue = [dict(node=root, rows=rs, depth=0, G=G, H=H, C=len(rs), best=best)]
    max_leaves = p.max_leaves
    while queue and st["leaves"] < max_leaves:
        gains = sorted((c["best"][0] for c in queue), reverse=True)
        if len(gains) > 1 and gains[0] == gains[1]:
            st["tie"] = True
        k = max(range(len(queue)), key=lambda i: queue[i]["best"][0])
        if mistake == "min_first":
            k = min(range(len(queue)), key=lambda i: queue[i]["best"][0])
        c = queue.pop(k)
        if c["depth"] >= p.max_depth and mistake != "no_depth_check":
            leaf(c["node"], c["G"], c["H"], c["C"])
            continue
        gn, f, b, GL, HL, CL = c["best"]
        nd = nodes[c["node"]]
        nd.update(feature=f, bin=b, leaf=False, gain=gn)
        col = bins[f]
        lrows = [r for r in c["rows"] if (col[r] == 255 and nd["dl"]) or (col[r] != 255 and col[r] <= b)]
        lset = set(lrows)
        rrows = [r for r in c["rows"] if r not in lset]
        if p.learn_missing_direction and 0 < len(lrows) < len(c["rows"]):
            nd["dl"] = len(lrows) >= len(rrows)
        lc, rc = add(), add()
        nd["left"], nd["right"] = lc, rc
        st["depth"] = max(st["depth"], c["depth"] + 1)
        gate = (lambda: st["leaves"] + 1 < max_leaves) if mistake != "gate_le" else \
            (lambda: st["leaves"] + 1 <= max_leaves)
        for cid, rws, G2, H2, C2 in ((lc, lrows, GL, HL, CL),
                                     (rc, rrows, c["G"] - GL, c["H"] - HL, c["C"] - CL)):
            can = len(rws) >= 2 * p.min_samples_leaf and H2 >= 2 * mcw
            if can and gate():
                b2 = find(rws, G2, H2, C2)
                if b2 is not None:
                    queue.append(dict(node=cid, rows=rws, depth=c["depth"] + 1, G=G2, H=H2, C=C2,
                                      best=b2))
                    continue
            leaf(cid, G2, H2, C2)
    for c in queue:
        leaf(c["node"], c["G"], c["H"], c["C"])
    return nodes, st["leaves"], st["depth"], st["tie"]


def small_case(seed, n=36, F=3, nan=True):
    rng = np.random.default_rng(seed)
    bins = rng.integers(0, 6, (F, n)).astype(np.uint8)
    if nan:
        bins[rng.integers(0, F, 3), rng.integers(0, n, 3)] = 255
    g = rng.standard_normal(n).astype(np.float32)
    h = rng.uniform(0.05, 1.0, n).astype(np.float32)
    return bins, g, h


PARAMS = [
    dict(max_depth=3),
    dict(max_depth=4, max_leaves=5),                        # DFS cut-off binds
    dict(max_depth=3, min_samples_leaf=4, min_child_weight=0.8),
    dict(max_depth=3, lambda_l2=0.0, min_split_gain=0.05, gamma=0.02),
    dict(max_depth=3, leaf_smooth=2.0, max_delta_step=0.7),
    dict(max_depth=2, learn_missing_direction=False, max_bins=4),
]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("pi", range(len(PARAMS)))
def test_rtree_matches_exact_builder(oracle, seed, pi):
    bins, g, h = small_case(seed)
    p = rtree_params(**{**dict(min_child_weight=0.1), **PARAMS[pi]})
    rows = None
    if seed % 3 == 2:
        rows = np.random.default_rng(seed + 50).permutation(bins.shape[1])[:28]
    assert_same(oracle.rtree_f64(bins, g, h, p, rows), exact_tree(bins, g, h, p, rows))


@pytest.mark.parametrize("mistake", ["larger_first", "no_max_leaves"])
def test_rtree_comparison_notices_planted_mistakes(oracle, mistake):
    """Each planted mistake changes some tree of these cases, so the
    comparison above would fail on it."""
    p = rtree_params(max_depth=4, max_leaves=5, min_child_weight=0.1)
    noticed = False
    for seed in range(6):
        bins, g, h = small_case(seed)
        try:
            assert_same(oracle.rtree_f64(bins, g, h, p), exact_tree(bins, g, h, p, mistake=mistake))
        except AssertionError:
            noticed = True
    assert noticed


def test_rtree_kat_eight_rows(oracle):
    """The 8-row toy (tests/test_turbocat.py:55-56) with mode='small': the
    root takes (0, 3) with gain 2 as in the symmetric tree; both children are
    pure (every candidate gain is negative), so they become leaves -1 and +1;
    4 rows each side, so default_left stays 1."""
    X = np.array([-2, -1, 1, 2, -1.5, -0.5, 0.5, 1.5], np.float32).reshape(8, 1)
    y = np.array([0, 0, 1, 1, 0, 0, 1, 1], np.float32)
    bins, _, _ = oracle.quantize(X)
    g, h = oracle.gradients_f32(LOSS_LOGLOSS, y, np.zeros(8, np.float32))
    t = oracle.rtree_f64(bins, g, h, rtree_params(max_depth=3, min_child_weight=0.01))
    assert t["n_nodes"] == 3 and t["n_leaves"] == 2 and t["depth"] == 1
    assert (t["feature"][0], t["bin"][0], t["gain"][0]) == (0, 3, 2.0)
    assert list(t["flags"]) == [2, 3, 3] and list(t["left"][:1]) == [1] and list(t["right"][:1]) == [2]
    assert list(t["value"][1:]) == [-1.0, 1.0]
    # its training update walks to the same leaves
    v = oracle.rtree_apply_binned(t, bins)
    assert np.array_equal(v, np.where(y == 0, -1.0, 1.0).astype(np.float32))


def test_rtree_empty_root(oracle):
    bins, g, h = small_case(1)
    t = oracle.rtree_f64(bins, g, h, rtree_params(), rows=np.zeros(0, np.uint32))
    assert t["n_nodes"] == 0 and t["n_leaves"] == 0


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libtcref.so not built (no reference tree)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rtree_stump_matches_reference_symmetric_root(oracle, seed):
    """A depth-1 regular tree with min_samples_leaf 1 and min_child_weight 0
    has the candidates of the symmetric tree's root (non-empty sides), the
    same gain and the same leaf values, so it must reproduce the compiled
    reference's depth-1 SymmetricTree (up to FP32 near-ties)."""
    def body():
        X, y = regression_data(3000, 6, seed)
        R = RefLib()
        for loss, yy, lam in ((LOSS_MSE, y, 0.0), (LOSS_LOGLOSS, (y > 0).astype(np.float32), 1.0)):
            ds = R.dataset(X, yy)
            rb, _, _ = ds.bins()
            p0 = np.full(len(yy), oracle.init_prediction(loss, yy), np.float32)
            g, h = oracle.gradients_f32(loss, yy, p0)
            ref = ds.tree(g, h, max_depth=1, lambda_l2=lam, min_child_weight=0.0)
            t = oracle.rtree_f64(rb, g, h, rtree_params(max_depth=1, lambda_l2=lam,
                                                        min_child_weight=0.0))
            assert ref["depth"] == 1 and t["n_nodes"] == 3
            if t["gap"][0] > 1e-5 * abs(t["gain"][0]):
                assert (t["feature"][0], t["bin"][0]) == (ref["feature"][0], ref["bin"][0])
                # leaves: -G / (H + lambda) over each side's rows -- the
                # reference's from float sums in row order (bit for bit), the
                # oracle's from FP64 sums
                col = rb[t["feature"][0]]
                left = (col <= t["bin"][0]) | (col == 255)
                for k, side in enumerate((left, ~left)):
                    gs, hs = np.float32(0), np.float32(0)
                    for a, b in zip(g[side], h[side]):
                        gs, hs = np.float32(gs + a), np.float32(hs + b)
                    assert ref["leaves"][k] == np.float32(-gs / np.float32(hs + np.float32(lam)))
                    v64 = -g[side].astype(np.float64).sum() / (h[side].astype(np.float64).sum() + lam)
                    assert abs(t["value"][1 + k] - v64) <= 1e-12 * abs(v64)
    RefLib.isolated(body)


def _hand_tree():
    """Root: feature 1, bin 1 (edge 0.5); left child a leaf; right child
    splits feature 0 at bin 0."""
    return dict(n_nodes=5, feature=np.array([1, 0, 0, 0, 0], np.uint16),
                bin=np.array([1, 0, 0, 0, 0], np.uint8),
                flags=np.array([2, 3, 2, 3, 3], np.uint8),
                left=np.array([1, 0, 3, 0, 0], np.uint32),
                right=np.array([2, 0, 4, 0, 0], np.uint32),
                value=np.array([0, 0.25, 0, -1.5, 3.0], np.float32))


def test_rpredict_threshold_rules(oracle):
    """Edge-exact values: with up to 16 edges both prediction paths count the
    edges <= v (v equal to an edge goes right); with more than 16 edges
    predict_raw_fast takes std::lower_bound (v equal to the edge stays left),
    the multiclass path keeps upper_bound.  NaN goes left in both."""
    t = _hand_tree()
    few = np.full((2, 255), 0, np.float32)
    few[0, :3] = [-1.0, 0.0, 1.0]
    few[1, :3] = [0.0, 0.5, 2.0]
    many = np.zeros((2, 255), np.float32)
    many[0, :20] = np.arange(20) - 0.0
    many[1, :20] = np.linspace(-1, 1, 20)
    many[1, 1] = 0.5 - 1e-3
    many[1, 2] = 0.5                 # bin 1's edge is edges[1]
    many[1, 3:20] = np.linspace(0.6, 2, 17)
    X = np.array([[5.0, 0.5], [5.0, 0.4], [0.0, 2.5], [-3.0, 2.5], [np.nan, 3.0],
                  [0.0, np.nan]], np.float32)
    w = np.array([1.0], np.float32)
    # few edges: row 0 v1 = 0.5 = edges[1] -> bin 2 > 1 -> right; then v0 = 5 -> bin 3 > 0 -> right
    out = oracle.rpredict_f32([t], w, X, few, [3, 3], rule=1)[:, 0]
    assert list(out) == [3.0, 0.25, 3.0, -1.5, -1.5, 0.25]
    # many edges, rule 1: v1 = 0.5 equals edges[2] -> lower_bound 2, bin 2 > 1 -> right (edge of
    # bin 1 is edges[1] = 0.499); v1 = 0.4 -> bin 1 -> left
    out1 = oracle.rpredict_f32([t], w, X, many, [20, 20], rule=1)[:, 0]
    out0 = oracle.rpredict_f32([t], w, X, many, [20, 20], rule=0)[:, 0]
    # row 2: v0 = 0.0 equals edges[0]: lower_bound 0 (left, -1.5), upper_bound 1 (right, 3.0)
    assert out1[2] == -1.5 and out0[2] == 3.0
    assert out1[4] == -1.5 and out0[4] == -1.5     # NaN feature 0 -> left
    assert out1[5] == 0.25 and out0[5] == 0.25     # NaN feature 1 -> left


def test_rpredict_sum_is_one_fma_per_tree(oracle):
    """out = fma(w_T, v_T, ... fma(w_1, v_1, 0)) + base, per class, with the
    trees' values chosen so that separate products and adds round
    differently."""
    def stump(v):
        return dict(n_nodes=1, feature=np.zeros(1, np.uint16), bin=np.zeros(1, np.uint8),
                    flags=np.array([3], np.uint8), left=np.zeros(1, np.uint32),
                    right=np.zeros(1, np.uint32), value=np.array([v], np.float32))
    rng = np.random.default_rng(5)
    X = np.zeros((1, 1), np.float32)
    e = np.zeros((1, 255), np.float32)
    for _ in range(100):
        vals = rng.standard_normal(5).astype(np.float32)
        w = np.full(5, 0.1, np.float32)
        fused, sep = np.float32(0), np.float32(0)
        for t in range(5):
            fused = np.float32(np.float64(w[t]) * np.float64(vals[t]) + np.float64(fused))
            sep = np.float32(sep + np.float32(w[t] * vals[t]))
        if fused != sep:
            break
    assert fused != sep
    out = oracle.rpredict_f32([stump(v) for v in vals], w, X, e, [0], rule=1, base=0.5)[0, 0]
    assert out == np.float32(fused + np.float32(0.5))
    # K classes: trees go to their own class's sum
    out3 = oracle.rpredict_f32([stump(v) for v in vals[:3]], w[:3], X, e, [0], rule=0, K=3,
                               tree_class=np.array([2, 0, 2]))[0]
    c0 = np.float32(np.float64(w[1]) * np.float64(vals[1]))
    c2 = np.float32(np.float64(w[0]) * np.float64(vals[0]))
    c2 = np.float32(np.float64(w[2]) * np.float64(vals[2]) + np.float64(c2))
    assert list(out3) == [c0, 0.0, c2]


# ------------------------------------------------------------------ lossguide
LW_PARAMS = [
    dict(max_depth=3, max_leaves=64),                       # no cut: every splittable node
    dict(max_depth=4, max_leaves=2),                        # the loop stops at 2 leaves made
    dict(max_depth=4, max_leaves=4),
    dict(max_depth=2, max_leaves=5),                        # depth-limited pops
    dict(max_depth=4, max_leaves=6, min_samples_leaf=4, min_child_weight=0.8),
    dict(max_depth=3, max_leaves=5, lambda_l2=0.0, min_split_gain=0.05, gamma=0.02),
    dict(max_depth=3, max_leaves=7, leaf_smooth=2.0, max_delta_step=0.7),
    dict(max_depth=3, max_leaves=3, learn_missing_direction=False, max_bins=4),
]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("pi", range(len(LW_PARAMS)))
def test_lossguide_matches_exact_builder(oracle, seed, pi):
    bins, g, h = small_case(seed)
    p = rtree_params(**{**dict(min_child_weight=0.1, grow_policy=1), **LW_PARAMS[pi]})
    rows = None
    if seed % 3 == 2:
        rows = np.random.default_rng(seed + 50).permutation(bins.shape[1])[:28]
    nodes, nl, depth, tie = exact_leafwise(bins, g, h, p, rows)
    assert not tie  # no equal exact gains at a pop on these cases (the queue order is pinned below)
    assert_same(oracle.rtree_f64(bins, g, h, p, rows), (nodes, nl, depth))


@pytest.mark.parametrize("mistake", ["min_first", "no_depth_check", "gate_le"])
def test_lossguide_comparison_notices_planted_mistakes(oracle, mistake):
    hits = 0
    for seed in range(6):
        bins, g, h = small_case(seed)
        for kw in LW_PARAMS[1:5]:
            p = rtree_params(**{**dict(min_child_weight=0.1, grow_policy=1), **kw})
            wrong = exact_leafwise(bins, g, h, p, mistake=mistake)[:3]
            try:
                assert_same(oracle.rtree_f64(bins, g, h, p), wrong)
            except AssertionError:
                hits += 1
    assert hits > 0


def canon_tree(t, i=0):
    """A tree as nested tuples from its root: the same tree in any numbering
    gives the same value."""
    if t["flags"][i] & 1:
        return ("leaf", int(t["flags"][i]), round(float(t["value"][i]), 12))
    return (int(t["feature"][i]), int(t["bin"][i]), int(t["flags"][i]),
            canon_tree(t, int(t["left"][i])), canon_tree(t, int(t["right"][i])))


@pytest.mark.parametrize("D,msl,mcw", [(1, 1, 1.0), (3, 5, 1.0), (5, 5, 0.5), (6, 20, 3.0)])
def test_lossguide_without_a_cut_grows_the_depthwise_tree(oracle, D, msl, mcw):
    """Each node's split depends on its own rows only, so with room for every
    leaf (max_leaves >= 2^max_depth) the leaf-wise tree is the depthwise one,
    numbered in expansion order instead of depth first."""
    rng = np.random.default_rng(D)
    n, F = 4000, 6
    bins = rng.integers(0, 40, (F, n)).astype(np.uint8)
    bins[rng.integers(0, F, 50), rng.integers(0, n, 50)] = 255
    g = rng.standard_normal(n).astype(np.float32)
    h = rng.uniform(0.1, 0.6, n).astype(np.float32)
    kw = dict(max_depth=D, max_leaves=1 << D, min_samples_leaf=msl, min_child_weight=mcw)
    a = oracle.rtree_f64(bins, g, h, rtree_params(**kw))
    b = oracle.rtree_f64(bins, g, h, rtree_params(grow_policy=1, **kw))
    assert (a["n_nodes"], a["n_leaves"], a["depth"]) == (b["n_nodes"], b["n_leaves"], b["depth"])
    assert canon_tree(a) == canon_tree(b)


QUEUE_PROGRAM = r"""
#include <cstdio>
#include <queue>
#include <vector>
struct Item {
  double gain;
  unsigned id;
  bool operator<(const Item& o) const { return gain < o.gain; }
};
int main() {
  unsigned n_gains, n_ops;
  if (scanf("%u %u", &n_gains, &n_ops) != 2) return 1;
  std::vector<double> g(n_gains);
  for (auto& x : g) if (scanf("%lf", &x) != 1) return 1;
  std::priority_queue<Item> q;
  for (unsigned i = 0; i < n_ops; ++i) {
    int op;
    if (scanf("%d", &op) != 1) return 1;
    if (op >= 0) q.push(Item{g[op], static_cast<unsigned>(op)});
    else if (!q.empty()) { printf("%u\n", q.top().id); q.pop(); }
  }
  return 0;
}
"""


def test_lossguide_queue_order_matches_std_priority_queue(oracle, tmp_path):
    """The candidate queue's pop order, with many equal gains, against
    std::priority_queue<LeafCandidate>-style items compiled here."""
    src, exe = tmp_path / "q.cpp", tmp_path / "q"
    src.write_text(QUEUE_PROGRAM)
    subprocess.run(["g++", "-O1", "-std=c++17", str(src), "-o", str(exe)], check=True)
    for seed in range(20):
        rng = np.random.default_rng(seed)
        m = 300
        gains = rng.integers(0, 4 if seed % 2 else 40, m).astype(np.float64) * 0.5
        ops, pushed = [], 0
        while pushed < m:
            if rng.random() < 0.6:
                ops.append(pushed)
                pushed += 1
            else:
                ops.append(-1)
        ops += [-1] * m
        inp = f"{m} {len(ops)}\n" + " ".join(repr(float(x)) for x in gains) + "\n" + \
            " ".join(str(o) for o in ops) + "\n"
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True)
        ref = np.array([int(x) for x in out.stdout.split()], np.uint32)
        assert np.array_equal(oracle.rqueue_order(gains, ops), ref), seed
