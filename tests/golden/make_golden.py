"""Write the golden vectors under tests/golden/ from the reference's own hot
path, compiled from its sources by oracle/Makefile (oracle/_ref/libtcref.so).

Run in a container that has the reference tree:
    make -C oracle && python tests/golden/make_golden.py

Each fixture is data only: seeded inputs (regenerated from the seed, so only
the seed and shape are stored), the reference's outputs, as .npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import binary_data, regression_data  # noqa: E402
from oracle import LOSS_LOGLOSS, LOSS_MSE, RefLib  # noqa: E402

CASES = [
    # name, kind, n, F, seed, loss, n_trees, depth, lr, lambda
    ("reg_small", "reg", 4000, 8, 101, LOSS_MSE, 12, 6, 0.1, 0.0),
    ("bin_small", "bin", 6000, 10, 102, LOSS_LOGLOSS, 12, 6, 0.1, 1.0),
    ("reg_deep", "reg", 3000, 5, 103, LOSS_MSE, 6, 8, 0.2, 3.0),
]


def make(name, kind, n, F, seed, loss, T, D, lr, lam):
    X, y = regression_data(n, F, seed, low_card_col=2) if kind == "reg" else binary_data(n, F, seed)
    R = RefLib()
    ds = R.dataset(X, y)
    bins, edges, ne = ds.bins()
    rng = np.random.default_rng(seed + 7)
    p = (rng.standard_normal(n) * 2).astype(np.float32)
    g, h = R.gradients(loss, y, p)
    t1 = ds.tree(g, h, max_depth=D, lambda_l2=lam, min_child_weight=1.0)
    m, tp, _ = ds.fit(loss, T, D, lr, lam, 1.0)
    Xp = rng.standard_normal((512, F)).astype(np.float32)
    pred = R.predict(m, Xp)
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        meta=np.array([n, F, seed, loss, T, D], np.int64), kind=np.array(kind),
        lr=np.float32(lr), lam=np.float32(lam),
        bins=bins, edges=edges, n_edges=ne, p=p, g=g, h=h,
        t_feature=t1["feature"], t_bin=t1["bin"], t_fthr=t1["fthr"], t_gain=t1["gain"],
        t_leaves=t1["leaves"], t_leaf_idx=t1["leaf_idx"],
        m_feature=m["feature"], m_bin=m["bin"], m_fthr=m["fthr"], m_gain=m["gain"],
        m_depth=m["depth"], m_leaves=m["leaves"], m_base=np.float32(m["base"]),
        train_preds=tp, Xp=Xp, pred=pred)


def make_ce(name, n, K, seed):
    """CrossEntropyLoss::compute_multiclass_gradients on seeded raw scores
    (uniform, saturated and random rows)."""
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal((n, K)) * 4).astype(np.float32)
    p[::13] = 0.0
    p[1::17, 0] = 40.0
    y = rng.integers(0, K, n).astype(np.float32)
    g, h = RefLib().gradients_mc(y, p)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), y=y, p=p, g=g, h=h)


def make_loss(name, n, seed):
    """Loss::compute_loss of MSE, LogLoss (saturated scores included) and
    CrossEntropyLoss (K = 3) on seeded predictions, one OpenMP thread."""
    rng = np.random.default_rng(seed)
    R = RefLib()
    p = (rng.standard_normal(n) * 4).astype(np.float32)
    p[:4] = [0.0, 45.0, -45.0, 1e-9]
    yr = rng.standard_normal(n).astype(np.float32)
    yb = (rng.random(n) < 0.3).astype(np.float32)
    pk = (rng.standard_normal((n, 3)) * 3).astype(np.float32)
    yk = rng.integers(0, 3, n).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), p=p, yr=yr, yb=yb, pk=pk, yk=yk,
                        mse=np.float32(R.compute_loss(0, yr, p)),
                        logloss=np.float32(R.compute_loss(1, yb, p)),
                        ce=np.float32(R.compute_loss_mc(yk, pk)))


def make_pred(name, n, F, seed, T, D, lr):
    """Raw-float and binned predictions of a reference-trained model in calls
    of several row counts: full 8-row blocks, a 5-row tail, a call under 16
    rows (the binned path's tree-by-tree fallback) and under 8 rows."""
    X, y = regression_data(n, F, seed, low_card_col=1)
    R = RefLib()
    ds = R.dataset(X, y)
    m, _, _ = ds.fit(LOSS_MSE, T, D, lr, 0.0, 1.0)
    rng = np.random.default_rng(seed + 1)
    out = dict(meta=np.array([n, F, seed, T, D], np.int64), lr=np.float32(lr),
               m_feature=m["feature"], m_bin=m["bin"], m_fthr=m["fthr"], m_depth=m["depth"],
               m_leaves=m["leaves"], m_base=np.float32(m["base"]))
    for r in (517, 13, 7):
        Xp = rng.standard_normal((r, F)).astype(np.float32)
        Xp[::5, 0] = np.nan  # NaN rows go left
        out[f"X{r}"] = Xp
        out[f"pred{r}"] = R.predict(m, Xp)
        dr = R.dataset(Xp, np.zeros(r, np.float32))
        bins, _, _ = dr.bins()
        out[f"bins{r}"] = bins
        out[f"pbin{r}"] = dr.predict_binned(m)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


if __name__ == "__main__":
    for c in CASES:
        RefLib.isolated(make, *c)  # fresh thread: see RefLib.isolated
        print("wrote", c[0])
    RefLib.isolated(make_ce, "ce_grad", 517, 7, 104)
    print("wrote ce_grad")
    RefLib.isolated(make_loss, "loss_values", 1501, 105)
    print("wrote loss_values")
    RefLib.isolated(make_pred, "pred_tail", 3000, 6, 106, 23, 5, 0.1)
    print("wrote pred_tail")
