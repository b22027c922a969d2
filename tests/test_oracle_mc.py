"""Pinning the K-channel oracle (tco_tree_build_mc_f64) for K > 1, on the CPU.

The K-channel symmetric tree has no reference counterpart; its definition is
the reference's single-channel rule (SymmetricTree::build,
src/tree/symmetric_tree.cpp:97-368) with the level gain summed over nodes and
classes:

    gain(f, b) = sum_n sum_c 0.5 * (GL^2/(HL+l) + GR^2/(HR+l) - G^2/(H+l)),

the parent term 0 when H <= 1e-10, a (node, class) unit skipped when HL or HR
is below 1e-10, a node skipped below the root when it has no rows, a root
candidate valid only when some unit is; ties to the lowest (feature, bin);
leaves v_c = -G_c / (H_c + l) per class.  The builder below restates that
definition directly, in exact rational arithmetic over integer gradients, and
every tree of the FP64 oracle must equal it.  A plausible mistake -- one
class's stride used for another, a parent term shared across classes, G
summed over classes before the gain, a dropped unit skip -- changes the
argmax or the leaves on these inputs (the planted-mistake test checks that
the comparison notices one).
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import tree_params

EPS = 1e-10


def exact_tree(bins, g, h, K, D, lam, max_bins=255, min_child_weight=1.0, min_samples_leaf=1,
               mistake=None):
    """K-channel symmetric tree by exhaustive search, exact arithmetic.
    bins F x n (no NaN), g and h n x K integers.  Returns the splits, the
    exact level gains and second-best gaps, the leaves and the leaf index."""
    F, n = bins.shape
    lam = Fraction(lam)
    gF = [[Fraction(int(v)) for v in row] for row in g]
    hF = [[Fraction(int(v)) for v in row] for row in h]
    node = np.zeros(n, np.int64)               # node of each row at the current level
    feats, bins_, gains, gaps = [], [], [], []
    for d in range(D):
        nn = 1 << d
        members = [np.flatnonzero(node == k) for k in range(nn)]
        G = [[sum((gF[i][c] for i in m), Fraction(0)) for c in range(K)] for m in members]
        H = [[sum((hF[i][c] for i in m), Fraction(0)) for c in range(K)] for m in members]
        # level stop: some node with enough rows and enough hessian (classes summed)
        if not any(len(members[k]) >= 2 * min_samples_leaf and
                   sum(H[k]) >= Fraction(2 * min_child_weight) for k in range(nn)):
            break
        cands = []
        for f in range(F):
            for b in range(max_bins - 1):
                total, valid = Fraction(0), False
                for k in range(nn):
                    if d > 0 and len(members[k]) == 0:
                        continue
                    left = members[k][bins[f, members[k]] <= b]
                    for c in range(K):
                        cc = (c + 1) % K if mistake == "class_stride" else c
                        gl = sum((gF[i][cc] for i in left), Fraction(0))
                        hl = sum((hF[i][c] for i in left), Fraction(0))
                        gr, hr = G[k][c] - gl, H[k][c] - hl
                        if hl < EPS or hr < EPS:
                            continue
                        if mistake == "summed_classes":
                            Gs, Hs = sum(G[k]), sum(H[k])
                            par = Gs * Gs / (Hs + lam) if Hs > EPS else Fraction(0)
                        else:
                            par = G[k][c] ** 2 / (H[k][c] + lam) if H[k][c] > EPS else Fraction(0)
                        total += Fraction(1, 2) * (gl * gl / (hl + lam) + gr * gr / (hr + lam) - par)
                        valid = True
                if d == 0 and not valid:
                    continue
                cands.append((total, f, b))
        if not cands:
            break
        best = max(c[0] for c in cands)
        f, b = min((c[1], c[2]) for c in cands if c[0] == best)   # ties: lowest (f, b)
        others = [c[0] for c in cands if (c[1], c[2]) != (f, b)]
        if best < 0:                                              # min_split_gain = 0
            break
        feats.append(f)
        bins_.append(b)
        gains.append(best)
        gaps.append(best - max(others) if others else None)
        node = 2 * node + (bins[f] > b).astype(np.int64)
    depth = len(feats)
    leaves = np.zeros((1 << depth, K))
    for k in range(1 << depth):
        m = np.flatnonzero(node == k)
        for c in range(K):
            Gk = sum((gF[i][c] for i in m), Fraction(0))
            Hk = sum((hF[i][c] for i in m), Fraction(0))
            leaves[k, c] = float(-Gk / (Hk + lam)) if Hk >= EPS else 0.0
    return dict(depth=depth, feature=feats, bin=bins_, gain=gains, gap=gaps, leaves=leaves,
                leaf_idx=node.astype(np.uint32))


def tiny_case(seed, K, n=48, F=3, levels=6):
    """Integer gradients with distinct per-class values, few bins per feature."""
    rng = np.random.default_rng(seed)
    bins = rng.integers(0, levels, (F, n)).astype(np.uint8)
    g = rng.integers(-4, 5, (n, K)).astype(np.float64)
    h = rng.integers(1, 4, (n, K)).astype(np.float64)
    edges = np.zeros((F, 255), np.float32)
    edges[:, :levels - 1] = np.arange(1, levels, dtype=np.float32)
    ne = np.full(F, levels - 1, np.uint32)
    return bins, g, h, edges, ne


def compare(oracle_tree, ex, lam):
    """Whole-tree equality, up to the first level whose exact gap between the
    best and second-best candidate is a tie within float64 rounding (then the
    oracle may take either side).  Returns the levels compared."""
    d = 0
    for d in range(ex["depth"]):
        gap = ex["gap"][d]
        near = gap is not None and float(gap) <= 1e-12 * max(1.0, abs(float(ex["gain"][d])))
        if near and (oracle_tree["feature"][d], oracle_tree["bin"][d]) != (ex["feature"][d], ex["bin"][d]):
            return d
        assert (int(oracle_tree["feature"][d]), int(oracle_tree["bin"][d])) == \
            (ex["feature"][d], ex["bin"][d]), f"level {d}"
        assert abs(oracle_tree["gain"][d] - float(ex["gain"][d])) <= 1e-12 * max(1.0, abs(float(ex["gain"][d])))
    assert oracle_tree["depth"] == ex["depth"]
    np.testing.assert_allclose(oracle_tree["leaves"], ex["leaves"], rtol=1e-13, atol=1e-15)
    assert np.array_equal(oracle_tree["leaf_idx"], ex["leaf_idx"])
    return ex["depth"]


@pytest.mark.parametrize("K", [2, 3])
@pytest.mark.parametrize("seed", range(8))
def test_tree_mc_equals_exact_builder(oracle, K, seed):
    bins, g, h, edges, ne = tiny_case(100 * K + seed, K)
    lam = [0.0, 1.0][seed % 2]
    D = 3
    ot = oracle.tree_mc_f64(bins, g, h, edges, ne, tree_params(D, lambda_l2=lam, max_bins=8))
    ex = exact_tree(bins, g, h, K, D, lam, max_bins=8)
    assert ex["depth"] >= 2, "case too small to exercise deeper levels"
    assert compare(ot, ex, lam) == ex["depth"], "a near-tie cut the comparison short"


@pytest.mark.parametrize("K", [2, 3])
def test_tree_mc_first_level_and_leaves_many_seeds(oracle, K):
    """Level-0 argmax and the one-split leaves over many small cases, at the
    root's own rules (unit skips and the validity of a candidate)."""
    compared = 0
    for seed in range(40):
        bins, g, h, edges, ne = tiny_case(7000 + 31 * K + seed, K, n=24, F=2, levels=5)
        h[seed % 24] = 0.0                       # a row with zero hessian in every class
        ot = oracle.tree_mc_f64(bins, g, h, edges, ne,
                                tree_params(1, lambda_l2=1.0, max_bins=6, min_child_weight=0.5))
        ex = exact_tree(bins, g, h, K, 1, 1.0, max_bins=6, min_child_weight=0.5)
        compared += compare(ot, ex, 1.0)
    assert compared >= 35


@pytest.mark.parametrize("mistake", ["class_stride", "summed_classes"])
def test_exact_builder_detects_planted_mistake(oracle, mistake):
    """The comparison is sensitive: a K-channel builder with a planted mistake
    differs from the oracle on some of the cases above."""
    differs = 0
    for seed in range(8):
        bins, g, h, edges, ne = tiny_case(200 + seed, 3)
        ot = oracle.tree_mc_f64(bins, g, h, edges, ne, tree_params(3, lambda_l2=1.0, max_bins=8))
        ex = exact_tree(bins, g, h, 3, 3, 1.0, max_bins=8, mistake=mistake)
        try:
            compare(ot, ex, 1.0)
        except AssertionError:
            differs += 1
    assert differs >= 4


def test_tree_mc_classes_are_not_interchangeable(oracle):
    """Permuting the classes permutes the leaf columns and nothing else; a
    permutation that also moves one class's hessians onto another's changes
    the tree (the channels are read at their own stride)."""
    bins, g, h, edges, ne = tiny_case(77, 3)
    p = tree_params(3, lambda_l2=1.0, max_bins=8)
    a = oracle.tree_mc_f64(bins, g, h, edges, ne, p)
    perm = [2, 0, 1]
    b = oracle.tree_mc_f64(bins, g[:, perm], h[:, perm], edges, ne, p)
    assert np.array_equal(a["feature"], b["feature"]) and np.array_equal(a["bin"], b["bin"])
    assert np.array_equal(a["leaves"][:, perm], b["leaves"])
    c = oracle.tree_mc_f64(bins, g, h[:, perm], edges, ne, p)
    ex = exact_tree(bins, g, h[:, perm], 3, 3, 1.0, max_bins=8)
    compare(c, ex, 1.0)
    assert not np.allclose(c["leaves"], a["leaves"])
