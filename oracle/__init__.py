"""Test infrastructure only: ctypes front-end to the CPU checkers.

* ``Oracle``  -- oracle/_ref/liboracle.so, the plain-C restatement of the
  reference hot path (tc_oracle.c), FP32 reference order and FP64 yardstick.
* ``RefLib``  -- oracle/_ref/libtcref.so, the reference's own hot-path sources
  compiled by oracle/Makefile (present only where the reference tree was
  available at build time).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, and only to check or to time the CPU side; the product
never routes through it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
LIB_ORACLE = os.path.join(REF_DIR, "liboracle.so")
LIB_REF = os.path.join(REF_DIR, "libtcref.so")

LOSS_MSE, LOSS_LOGLOSS = 0, 1


def build(reference_root: str = "/root/reference") -> None:
    """Compile the checkers (make -C oracle); libtcref only if the reference exists."""
    subprocess.run(["make", "-s", "-C", HERE, f"REF={reference_root}", "oracle"], check=True)
    if os.path.exists(os.path.join(reference_root, "src", "tree", "symmetric_tree.cpp")):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, f"REF={reference_root}", "ref"], check=True)


def _p(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


class TreeParams(ctypes.Structure):
    _fields_ = [
        ("max_depth", c_uint32),
        ("max_bins", c_uint32),
        ("lambda_l2", c_float),
        ("lambda_l1", c_float),
        ("max_delta_step", c_float),
        ("min_samples_leaf", c_uint32),
        ("min_child_weight", c_float),
        ("min_split_gain", c_float),
    ]


def tree_params(max_depth=6, max_bins=255, lambda_l2=1.0, lambda_l1=0.0, max_delta_step=0.0,
                min_samples_leaf=1, min_child_weight=1.0, min_split_gain=0.0) -> TreeParams:
    return TreeParams(max_depth, max_bins, lambda_l2, lambda_l1, max_delta_step,
                      min_samples_leaf, min_child_weight, min_split_gain)


class RTreeParams(ctypes.Structure):
    """TreeConfig fields of the regular trees (config.hpp:21-59)."""
    _fields_ = [
        ("max_depth", c_uint32),
        ("max_leaves", c_uint32),
        ("max_bins", c_uint32),
        ("min_samples_leaf", c_uint32),
        ("min_child_weight", c_float),
        ("lambda_l2", c_float),
        ("min_split_gain", c_float),
        ("gamma", c_float),
        ("leaf_smooth", c_float),
        ("max_delta_step", c_float),
        ("learn_missing_direction", c_int32),
        ("grow_policy", c_int32),
    ]


def rtree_params(max_depth=6, max_leaves=64, max_bins=255, min_samples_leaf=1,
                 min_child_weight=1.0, lambda_l2=1.0, min_split_gain=0.0, gamma=0.0,
                 leaf_smooth=0.0, max_delta_step=0.0, learn_missing_direction=True,
                 grow_policy=0) -> RTreeParams:
    """grow_policy: 0 depthwise, 1 lossguide (TreeConfig::grow_policy)."""
    return RTreeParams(max_depth, max_leaves, max_bins, min_samples_leaf, min_child_weight,
                       lambda_l2, min_split_gain, gamma, leaf_smooth, max_delta_step,
                       int(learn_missing_direction), int(grow_policy))


class Oracle:
    def __init__(self, path: str = LIB_ORACLE):
        if not os.path.exists(path):
            build()
        self.L = L = ctypes.CDLL(path)
        V = c_void_p
        L.tco_quantize.argtypes = [V, c_uint32, c_uint32, c_uint32, c_int, V, V, V]
        L.tco_apply_bins.argtypes = [V, c_uint32, c_uint32, V, V, V]
        L.tco_init_prediction.argtypes = [c_int, V, c_uint32]
        L.tco_init_prediction.restype = c_float
        for fn in ("tco_gradients_f32", "tco_gradients_f64"):
            getattr(L, fn).argtypes = [c_int, V, V, V, V, c_uint32]
        L.tco_gradients_mc_f64.argtypes = [V, V, V, V, c_uint32, c_uint32]
        L.tco_gradients_mc_f32.argtypes = [V, V, V, V, c_uint32, c_uint32]
        L.tco_transform_f32.argtypes = [V, V, c_uint32, c_uint32]
        L.tco_predict_mc_f64.argtypes = [V, V, V, V, V, c_uint32, c_uint32, c_uint32, c_double, V,
                                         c_uint32, c_uint32, V, V]
        L.tco_hist_f32.argtypes = [V, c_uint32, c_uint32, V, V, V, c_uint32, V, V]
        L.tco_hist_f64.argtypes = [V, c_uint32, c_uint32, V, V, V, c_uint32, V, V, V]
        L.tco_split_f32.argtypes = [V, V, V, V, c_uint32, c_uint32, c_uint32, c_float, c_int, V, V, V]
        L.tco_split_f64.argtypes = [V, V, V, V, c_uint32, c_uint32, c_uint32, c_double, c_int, V, V, V, V]
        L.tco_tree_build_f32.argtypes = [V, c_uint32, c_uint32, V, V, V, c_uint32, V, V,
                                         POINTER(TreeParams), V, V, V, V, V, V]
        L.tco_tree_build_f64.argtypes = [V, c_uint32, c_uint32, V, V, V, c_uint32, V, V,
                                         POINTER(TreeParams), V, V, V, V, V, V, V]
        L.tco_tree_build_mc_f64.argtypes = [V, c_uint32, c_uint32, c_uint32, V, V, V, V,
                                            POINTER(TreeParams), V, V, V, V, V, V, V]
        L.tco_leaf_index_binned.argtypes = [V, c_uint32, c_uint32, V, V, V]
        L.tco_predict_f32.argtypes = [V, V, V, V, V, c_uint32, c_uint32, c_float, V, c_uint32,
                                      c_uint32, V, V]
        L.tco_predict_f64.argtypes = [V, V, V, V, V, c_uint32, c_uint32, c_double, V, c_uint32,
                                      c_uint32, V, V]
        L.tco_predict_binned_f32.argtypes = [V, V, V, V, V, c_uint32, c_uint32, c_float, V,
                                             c_uint32, c_uint32, V]
        L.tco_fit_f32.argtypes = [V, c_uint32, c_uint32, V, V, V, c_int, c_uint32, c_float,
                                  POINTER(TreeParams), V, V, V, V, V, V, V]
        L.tco_fit_f32.restype = c_float
        L.tco_fma_update.argtypes = [V, V, c_float, c_uint32]
        L.tco_compute_loss_f32.argtypes = [c_int, V, V, c_uint32]
        L.tco_compute_loss_f32.restype = c_float
        L.tco_compute_loss_f64.argtypes = [c_int, V, V, c_uint32]
        L.tco_compute_loss_f64.restype = c_double
        L.tco_compute_loss_mc_f32.argtypes = [V, V, c_uint32, c_uint32]
        L.tco_compute_loss_mc_f32.restype = c_float
        L.tco_early_stop.argtypes = [V, c_uint32, c_uint32, c_uint32, c_int, V, V]
        L.tco_early_stop.restype = c_uint32
        for fn in ("tco_tree_build_f32", "tco_tree_build_f64", "tco_tree_build_mc_f64"):
            getattr(L, fn).restype = c_int
        L.tco_rtree_build_f64.argtypes = [V, c_uint32, c_uint32, V, V, V, c_uint32,
                                          POINTER(RTreeParams), c_uint32, V, V, V, V, V, V, V,
                                          V, V, V, V, V]
        L.tco_rtree_build_f64.restype = c_int
        L.tco_rqueue_order.argtypes = [V, V, c_uint32, V]
        L.tco_rqueue_order.restype = c_uint32
        L.tco_rtree_apply_binned.argtypes = [V, V, V, V, V, V, c_uint32, V, c_uint32, V]
        L.tco_rpredict_f32.argtypes = [V, V, V, V, V, V, V, V, V, c_uint32, c_uint32, c_float, V,
                                       c_uint32, c_uint32, V, V, c_int, V]

    # ---------------------------------------------------------------- A19
    def compute_loss_f32(self, loss, y, p):
        """Loss::compute_loss (mean per-row loss, float, row order)."""
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        return float(self.L.tco_compute_loss_f32(loss, _p(y), _p(p), len(y)))

    def compute_loss_f64(self, loss, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        return float(self.L.tco_compute_loss_f64(loss, _p(y), _p(p), len(y)))

    def compute_loss_mc_f32(self, y, p):
        """CrossEntropyLoss::compute_loss; p row-major n x K."""
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        return float(self.L.tco_compute_loss_mc_f32(_p(y), _p(p), p.shape[0], p.shape[1]))

    def early_stop(self, metric, n_estimators, rounds, higher_is_better=False):
        """Trees kept, best iteration and best metric under the early-stopping rule."""
        m = np.ascontiguousarray(metric, np.float32)
        bi = np.zeros(1, np.uint32); b = np.zeros(1, np.float32)
        kept = self.L.tco_early_stop(_p(m), len(m), n_estimators, rounds, int(higher_is_better),
                                     _p(bi), _p(b))
        return int(kept), int(bi[0]), float(b[0])

    # ---------------------------------------------------------------- A2/A3
    def quantize(self, X, max_bins=255, wrap=True):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        bins = np.zeros((F, n), np.uint8)
        edges = np.zeros((F, 255), np.float32)
        ne = np.zeros(F, np.uint32)
        self.L.tco_quantize(_p(X), n, F, max_bins, int(wrap), _p(bins), _p(edges), _p(ne))
        return bins, edges, ne

    def apply_bins(self, X, edges, n_edges):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        bins = np.zeros((F, n), np.uint8)
        self.L.tco_apply_bins(_p(X), n, F, _p(np.ascontiguousarray(edges, np.float32)),
                              _p(np.ascontiguousarray(n_edges, np.uint32)), _p(bins))
        return bins

    # ---------------------------------------------------------------- A5-A8
    def init_prediction(self, loss, y):
        y = np.ascontiguousarray(y, np.float32)
        return float(self.L.tco_init_prediction(loss, _p(y), len(y)))

    def gradients_f32(self, loss, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        g = np.zeros_like(y); h = np.zeros_like(y)
        self.L.tco_gradients_f32(loss, _p(y), _p(p), _p(g), _p(h), len(y))
        return g, h

    def gradients_f64(self, loss, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        g = np.zeros(len(y)); h = np.zeros(len(y))
        self.L.tco_gradients_f64(loss, _p(y), _p(p), _p(g), _p(h), len(y))
        return g, h

    def gradients_mc_f32(self, y, p):
        """A8 in the reference's float order; p, g, h row-major n x K."""
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        n, K = p.shape
        g = np.zeros((n, K), np.float32); h = np.zeros((n, K), np.float32)
        self.L.tco_gradients_mc_f32(_p(y), _p(p), _p(g), _p(h), n, K)
        return g, h

    def transform_f32(self, raw, K=1):
        """Probabilities from raw scores in the reference's float order:
        sigmoid (K == 1) or per-row softmax (raw (n, K))."""
        raw = np.ascontiguousarray(raw, np.float32)
        n = raw.size // K
        out = np.zeros_like(raw)
        self.L.tco_transform_f32(_p(raw), _p(out), n, K)
        return out

    def gradients_mc_f64(self, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        n, K = p.shape
        g = np.zeros((n, K)); h = np.zeros((n, K))
        self.L.tco_gradients_mc_f64(_p(y), _p(p), _p(g), _p(h), n, K)
        return g, h

    # ---------------------------------------------------------------- A9
    def hist_f32(self, bins, g, h, rows=None):
        F, n = bins.shape
        gh = np.zeros((F, 256, 2), np.float32); cnt = np.zeros((F, 256), np.uint32)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        self.L.tco_hist_f32(_p(bins), n, F, _p(np.ascontiguousarray(g, np.float32)),
                            _p(np.ascontiguousarray(h, np.float32)), _p(r),
                            0 if r is None else len(r), _p(gh), _p(cnt))
        return gh, cnt

    def hist_f64(self, bins, g, h, rows=None):
        F, n = bins.shape
        gh = np.zeros((F, 256, 2)); ab = np.zeros((F, 256)); cnt = np.zeros((F, 256), np.uint32)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        self.L.tco_hist_f64(_p(bins), n, F, _p(np.ascontiguousarray(g, np.float32)),
                            _p(np.ascontiguousarray(h, np.float32)), _p(r),
                            0 if r is None else len(r), _p(gh), _p(ab), _p(cnt))
        return gh, ab, cnt

    # ---------------------------------------------------------------- A10
    def split_f32(self, hist, node_g, node_h, node_count, max_bins=255, lam=1.0, is_root=False):
        hist = np.ascontiguousarray(hist, np.float32)
        nn, F = hist.shape[0], hist.shape[1]
        f = c_int32(); b = c_int32(); gn = c_float()
        self.L.tco_split_f32(_p(hist), _p(np.ascontiguousarray(node_g, np.float32)),
                             _p(np.ascontiguousarray(node_h, np.float32)),
                             _p(np.ascontiguousarray(node_count, np.uint32)), nn, F, max_bins,
                             lam, int(is_root), ctypes.byref(f), ctypes.byref(b), ctypes.byref(gn))
        return f.value, b.value, gn.value

    def split_f64(self, hist, node_g, node_h, node_count, max_bins=255, lam=1.0, is_root=False):
        hist = np.ascontiguousarray(hist, np.float64)
        nn, F = hist.shape[0], hist.shape[1]
        f = c_int32(); b = c_int32(); gn = c_double(); sec = c_double()
        self.L.tco_split_f64(_p(hist), _p(np.ascontiguousarray(node_g, np.float64)),
                             _p(np.ascontiguousarray(node_h, np.float64)),
                             _p(np.ascontiguousarray(node_count, np.uint32)), nn, F, max_bins,
                             lam, int(is_root), ctypes.byref(f), ctypes.byref(b), ctypes.byref(gn),
                             ctypes.byref(sec))
        return f.value, b.value, gn.value, sec.value

    # ---------------------------------------------------------------- A9-A14
    def tree_f32(self, bins, g, h, edges, n_edges, params: TreeParams, rows=None):
        F, n = bins.shape
        D = params.max_depth
        feat = np.zeros(D, np.uint16); b = np.zeros(D, np.uint8); ft = np.zeros(D, np.float32)
        gn = np.zeros(D, np.float32); lv = np.zeros(1 << D, np.float32); li = np.zeros(n, np.uint32)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        d = self.L.tco_tree_build_f32(_p(bins), n, F, _p(np.ascontiguousarray(g, np.float32)),
                                      _p(np.ascontiguousarray(h, np.float32)), _p(r),
                                      n if r is None else len(r), _p(edges),
                                      _p(n_edges), ctypes.byref(params), _p(feat), _p(b), _p(ft),
                                      _p(gn), _p(lv), _p(li))
        return dict(depth=d, feature=feat[:d], bin=b[:d], fthr=ft[:d], gain=gn[:d],
                    leaves=lv[:1 << d], leaf_idx=li)

    def tree_f64(self, bins, g, h, edges, n_edges, params: TreeParams, rows=None):
        F, n = bins.shape
        D = params.max_depth
        feat = np.zeros(D, np.uint16); b = np.zeros(D, np.uint8); ft = np.zeros(D, np.float32)
        gn = np.zeros(D); gap = np.zeros(D); lv = np.zeros(1 << D); li = np.zeros(n, np.uint32)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        d = self.L.tco_tree_build_f64(_p(bins), n, F, _p(np.ascontiguousarray(g, np.float32)),
                                      _p(np.ascontiguousarray(h, np.float32)), _p(r),
                                      n if r is None else len(r), _p(edges),
                                      _p(n_edges), ctypes.byref(params), _p(feat), _p(b), _p(ft),
                                      _p(gn), _p(gap), _p(lv), _p(li))
        return dict(depth=d, feature=feat[:d], bin=b[:d], fthr=ft[:d], gain=gn[:d], gap=gap[:d],
                    leaves=lv[:1 << d], leaf_idx=li)

    def tree_mc_f64(self, bins, g, h, edges, n_edges, params: TreeParams):
        F, n = bins.shape
        K = g.shape[1]
        D = params.max_depth
        feat = np.zeros(D, np.uint16); b = np.zeros(D, np.uint8); ft = np.zeros(D, np.float32)
        gn = np.zeros(D); gap = np.zeros(D); lv = np.zeros((1 << D) * K); li = np.zeros(n, np.uint32)
        d = self.L.tco_tree_build_mc_f64(_p(bins), n, F, K, _p(np.ascontiguousarray(g, np.float64)),
                                         _p(np.ascontiguousarray(h, np.float64)), _p(edges),
                                         _p(n_edges), ctypes.byref(params), _p(feat), _p(b),
                                         _p(ft), _p(gn), _p(gap), _p(lv), _p(li))
        return dict(depth=d, feature=feat[:d], bin=b[:d], fthr=ft[:d], gain=gn[:d], gap=gap[:d],
                    leaves=lv[:(1 << d) * K].reshape(1 << d, K), leaf_idx=li)

    def leaf_index_binned(self, bins, feature, bin_):
        F, n = bins.shape
        li = np.zeros(n, np.uint32)
        self.L.tco_leaf_index_binned(_p(bins), n, len(feature),
                                     _p(np.ascontiguousarray(feature, np.uint16)),
                                     _p(np.ascontiguousarray(bin_, np.uint8)), _p(li))
        return li

    # ---------------------------------------------------------------- A16
    def predict_f32(self, model, X, want_leaf_idx=False):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        T, D = model["T"], model["D"]
        out = np.zeros(n, np.float32)
        li = np.zeros((n, T), np.uint32) if want_leaf_idx else None
        self.L.tco_predict_f32(_p(model["feature"]), _p(model["fthr"]), _p(model["depth"]),
                               _p(model["leaves"]), _p(model["weight"]), T, D, model["base"],
                               _p(X), n, F, _p(out), _p(li))
        return (out, li) if want_leaf_idx else out

    def predict_binned_f32(self, model, bins):
        """A17: binned prediction over column-major bins (F x n) in the
        reference's float order, + base."""
        bins = np.ascontiguousarray(bins, np.uint8)
        F, n = bins.shape
        out = np.zeros(n, np.float32)
        self.L.tco_predict_binned_f32(_p(model["feature"]), _p(model["bin"]), _p(model["depth"]),
                                      _p(model["leaves"]), _p(model["weight"]), model["T"],
                                      model["D"], model["base"], _p(bins), n, F, _p(out))
        return out

    def predict_f64(self, model, X):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        out = np.zeros(n); ab = np.zeros(n)
        self.L.tco_predict_f64(_p(model["feature"]), _p(model["fthr"]), _p(model["depth"]),
                               _p(model["leaves"]), _p(model["weight"]), model["T"], model["D"],
                               model["base"], _p(X), n, F, _p(out), _p(ab))
        return out, ab

    def predict_mc_f64(self, model, X):
        """K-channel ensemble: (n, K) raw scores and sum |terms|."""
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        K = model["K"]
        out = np.zeros((n, K)); ab = np.zeros((n, K))
        self.L.tco_predict_mc_f64(_p(model["feature"]), _p(model["fthr"]), _p(model["depth"]),
                                  _p(np.ascontiguousarray(model["leaves"], np.float32)),
                                  _p(model["weight"]), model["T"], model["D"], K, model["base"],
                                  _p(X), n, F, _p(out), _p(ab))
        return out, ab

    def fma_update(self, preds, lr, vals):
        out = np.array(preds, np.float32, copy=True)
        v = np.ascontiguousarray(vals, np.float32)
        self.L.tco_fma_update(_p(out), _p(v), lr, len(out))
        return out

    # ---------------------------------------------------------------- fit
    def fit_f32(self, bins, edges, n_edges, y, loss, n_trees, lr, params: TreeParams):
        F, n = bins.shape
        D = params.max_depth
        m = new_model(n_trees, D)
        tp = np.zeros(n, np.float32)
        base = self.L.tco_fit_f32(_p(bins), n, F, _p(edges), _p(n_edges),
                                  _p(np.ascontiguousarray(y, np.float32)), loss, n_trees, lr,
                                  ctypes.byref(params), _p(m["feature"]), _p(m["bin"]),
                                  _p(m["fthr"]), _p(m["gain"]), _p(m["depth"]), _p(m["leaves"]),
                                  _p(tp))
        m["base"] = float(base)
        m["weight"][:] = lr
        return m, tp

    # ------------------------------------------------------- regular trees
    def rtree_f64(self, bins, g, h, params: RTreeParams, rows=None):
        """One regular tree (Tree::build, tree.cpp:162-217: depthwise, or
        build_leafwise :685-925 when params.grow_policy is 1) in the
        reference's node numbering; FP64 sums.  Returns a dict of node arrays
        (flags bit 0 is_leaf, bit 1 default_left) plus n_leaves, depth and,
        for lossguide, hgap (the expansion order's smallest gain margin)."""
        F, n = bins.shape
        cap = (1 << (params.max_depth + 1)) + 1
        feat = np.zeros(cap, np.uint16); b = np.zeros(cap, np.uint8); fl = np.zeros(cap, np.uint8)
        le = np.zeros(cap, np.uint32); ri = np.zeros(cap, np.uint32); val = np.zeros(cap)
        gn = np.zeros(cap); gap = np.zeros(cap)
        nn = c_uint32(); nl = c_uint32(); dp = ctypes.c_uint16(); hg = ctypes.c_double()
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        rc = self.L.tco_rtree_build_f64(_p(np.ascontiguousarray(bins, np.uint8)), n, F,
                                        _p(np.ascontiguousarray(g, np.float32)),
                                        _p(np.ascontiguousarray(h, np.float32)), _p(r),
                                        n if r is None else len(r), ctypes.byref(params), cap,
                                        _p(feat), _p(b), _p(fl), _p(le), _p(ri), _p(val), _p(gn),
                                        _p(gap), ctypes.byref(nn), ctypes.byref(nl),
                                        ctypes.byref(dp), ctypes.byref(hg))
        assert rc == 0, "node capacity exceeded"
        k = nn.value
        return dict(n_nodes=k, n_leaves=nl.value, depth=dp.value, feature=feat[:k], bin=b[:k],
                    flags=fl[:k], left=le[:k], right=ri[:k], value=val[:k], gain=gn[:k],
                    gap=gap[:k], hgap=hg.value)

    def rqueue_order(self, gains, ops):
        """Items popped by build_leafwise's candidate queue for a sequence of
        pushes (item index >= 0) and pops (-1)."""
        gains = np.ascontiguousarray(gains, np.float64)
        ops = np.ascontiguousarray(ops, np.int32)
        out = np.zeros(max(1, len(ops)), np.uint32)
        k = self.L.tco_rqueue_order(_p(gains), _p(ops), len(ops), _p(out))
        return out[:k]

    def rtree_apply_binned(self, tree, bins):
        """Tree::predict(dataset, row) per row (NaN follows default_left)."""
        bins = np.ascontiguousarray(bins, np.uint8)
        F, n = bins.shape
        out = np.zeros(n, np.float32)
        self.L.tco_rtree_apply_binned(_p(np.ascontiguousarray(tree["feature"], np.uint16)),
                                      _p(np.ascontiguousarray(tree["bin"], np.uint8)),
                                      _p(np.ascontiguousarray(tree["flags"], np.uint8)),
                                      _p(np.ascontiguousarray(tree["left"], np.uint32)),
                                      _p(np.ascontiguousarray(tree["right"], np.uint32)),
                                      _p(np.ascontiguousarray(tree["value"], np.float32)),
                                      int(tree["n_nodes"]), _p(bins), n, _p(out))
        return out

    def rpredict_f32(self, trees, weights, X, edges, n_edges, rule, K=1, tree_class=None,
                     base=0.0):
        """Regular-tree ensemble prediction from floats (n x K): rule 1 is
        predict_raw_fast's binning (binary/regression), rule 0 apply_bins'
        (multiclass)."""
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        off = np.zeros(len(trees) + 1, np.uint32)
        for t, tr in enumerate(trees):
            off[t + 1] = off[t] + int(tr["n_nodes"])
        cat = lambda key, dt: np.ascontiguousarray(  # noqa: E731
            np.concatenate([np.asarray(tr[key])[:int(tr["n_nodes"])] for tr in trees] +
                           [np.zeros(1)]).astype(dt))
        out = np.zeros((n, K), np.float32)
        tc = None if tree_class is None else np.ascontiguousarray(tree_class, np.uint32)
        self.L.tco_rpredict_f32(_p(off), _p(cat("feature", np.uint16)), _p(cat("bin", np.uint8)),
                                _p(cat("flags", np.uint8)), _p(cat("left", np.uint32)),
                                _p(cat("right", np.uint32)), _p(cat("value", np.float32)),
                                _p(np.ascontiguousarray(weights, np.float32)), _p(tc), len(trees),
                                K, base, _p(X), n, F, _p(np.ascontiguousarray(edges, np.float32)),
                                _p(np.ascontiguousarray(n_edges, np.uint32)), rule, _p(out))
        return out


def new_model(T, D):
    return dict(T=T, D=D, feature=np.zeros(T * D, np.uint16), bin=np.zeros(T * D, np.uint8),
                fthr=np.zeros(T * D, np.float32), gain=np.zeros(T * D, np.float32),
                depth=np.zeros(T, np.uint16), leaves=np.zeros(T * (1 << D), np.float32),
                weight=np.zeros(T, np.float32), base=0.0)


class RefLib:
    """The reference's own hot path, compiled from its sources (oracle/Makefile)."""

    def __init__(self, path: str = LIB_REF):
        self.L = L = ctypes.CDLL(path)
        V = c_void_p
        L.tcref_dataset_create.restype = V
        L.tcref_dataset_create.argtypes = [V, V, c_uint32, c_uint32, c_uint32, V]
        L.tcref_dataset_free.argtypes = [V]
        L.tcref_dataset_bins.argtypes = [V, V, V, V]
        L.tcref_gradients.argtypes = [c_int, V, V, V, V, c_uint32]
        L.tcref_gradients_mc.argtypes = [V, V, V, V, c_uint32, c_uint32]
        L.tcref_transform.argtypes = [V, V, c_uint32, c_uint32]
        L.tcref_init_prediction.argtypes = [c_int, V, c_uint32]
        L.tcref_init_prediction.restype = c_float
        L.tcref_compute_loss.argtypes = [c_int, V, V, c_uint32]
        L.tcref_compute_loss.restype = c_float
        L.tcref_compute_loss_mc.argtypes = [V, V, c_uint32, c_uint32]
        L.tcref_compute_loss_mc.restype = c_float
        L.tcref_hist.argtypes = [V, V, V, V, c_uint32, V, V]
        L.tcref_tree_build.argtypes = [V, V, V, c_uint32, c_float, c_float, V, V, V, V, V, V]
        L.tcref_tree_build.restype = c_int
        L.tcref_fit.argtypes = [V, c_int, c_uint32, c_uint32, c_float, c_float, c_float,
                                c_float, c_float, ctypes.c_uint64, V, V, V, V, V, V, V, V]
        L.tcref_random_subsample.argtypes = [V, c_float, ctypes.c_uint64, V]
        L.tcref_random_subsample.restype = c_uint32
        L.tcref_goss_subsample.argtypes = [V, V, c_float, c_float, ctypes.c_uint64, V]
        L.tcref_goss_subsample.restype = c_uint32
        L.tcref_fit_goss.argtypes = [V, c_int, c_uint32, c_uint32, c_float, c_float, c_float,
                                     c_float, c_float, c_float, ctypes.c_uint64, V, V, V, V, V, V,
                                     V]
        L.tcref_fit_goss.restype = c_float
        L.tcref_mvs_subsample.argtypes = [V, V, c_float, ctypes.c_uint64, V]
        L.tcref_mvs_subsample.restype = c_uint32
        L.tcref_fit_mvs.argtypes = [V, c_int, c_uint32, c_uint32, c_float, c_float, c_float,
                                    c_float, c_float, ctypes.c_uint64, V, V, V, V, V, V, V]
        L.tcref_fit_mvs.restype = c_float
        L.tcref_fit.restype = c_float
        L.tcref_predict.argtypes = [V, V, V, V, V, c_uint32, c_uint32, c_float, c_float, V,
                                    c_uint32, c_uint32, V]
        L.tcref_predict_binned.argtypes = [V, V, V, V, V, V, c_uint32, c_uint32, c_float, c_float,
                                           V]

    @staticmethod
    def available(path: str = LIB_REF) -> bool:
        return os.path.exists(path)

    @staticmethod
    def isolated(fn, *args, **kw):
        """Run fn in a fresh OS thread.  SymmetricTree::build keeps static
        thread_local histogram pools that are regrown only when 2^max_depth
        grows, not when n_features grows (symmetric_tree.cpp:57-74), so a
        second dataset with more features on the same thread writes past the
        pools.  A fresh thread gets fresh pools."""
        import threading
        box = {}

        def run():
            try:
                box["r"] = fn(*args, **kw)
            except BaseException as e:  # noqa: BLE001
                box["e"] = e

        t = threading.Thread(target=run)
        t.start()
        t.join()
        if "e" in box:
            raise box["e"]
        return box["r"]

    def dataset(self, X, y=None, max_bins=255):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        y = np.zeros(n, np.float32) if y is None else np.ascontiguousarray(y, np.float32)
        sec = np.zeros(1)
        h = self.L.tcref_dataset_create(_p(X), _p(y), n, F, max_bins, _p(sec))
        return RefDataset(self, h, n, F, float(sec[0]), y)

    def gradients(self, loss, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        g = np.zeros_like(y); h = np.zeros_like(y)
        self.L.tcref_gradients(loss, _p(y), _p(p), _p(g), _p(h), len(y))
        return g, h

    def init_prediction(self, loss, y):
        y = np.ascontiguousarray(y, np.float32)
        return float(self.L.tcref_init_prediction(loss, _p(y), len(y)))

    def compute_loss(self, loss, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        return float(self.L.tcref_compute_loss(loss, _p(y), _p(p), len(y)))

    def compute_loss_mc(self, y, p):
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        return float(self.L.tcref_compute_loss_mc(_p(y), _p(p), p.shape[0], p.shape[1]))

    def gradients_mc(self, y, p):
        """CrossEntropyLoss::compute_multiclass_gradients; p row-major n x K."""
        y = np.ascontiguousarray(y, np.float32); p = np.ascontiguousarray(p, np.float32)
        n, K = p.shape
        g = np.zeros((n, K), np.float32); h = np.zeros((n, K), np.float32)
        self.L.tcref_gradients_mc(_p(y), _p(p), _p(g), _p(h), n, K)
        return g, h

    def transform(self, raw, K=1):
        """LogLoss::transform_prediction (K == 1) / CrossEntropyLoss::
        transform_to_proba (K > 1); raw row-major (n, K)."""
        raw = np.ascontiguousarray(raw, np.float32)
        out = np.zeros_like(raw)
        self.L.tcref_transform(_p(raw), _p(out), raw.size // K, K)
        return out

    def predict(self, model, X):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        out = np.zeros(n, np.float32)
        self.L.tcref_predict(_p(model["feature"]), _p(model["fthr"]), _p(model["bin"]),
                             _p(model["depth"]), _p(model["leaves"]), model["T"], model["D"],
                             float(model["weight"][0]) if model["T"] else 0.0, model["base"],
                             _p(X), n, F, _p(out))
        return out


class RefDataset:
    def __init__(self, lib, h, n, F, seconds, y):
        self.lib, self.h, self.n, self.F, self.seconds, self.y = lib, h, n, F, seconds, y

    def __del__(self):
        try:
            self.lib.L.tcref_dataset_free(self.h)
        except Exception:
            pass

    def bins(self):
        bins = np.zeros((self.F, self.n), np.uint8)
        edges = np.zeros((self.F, 255), np.float32)
        ne = np.zeros(self.F, np.uint32)
        self.lib.L.tcref_dataset_bins(self.h, _p(bins), _p(edges), _p(ne))
        return bins, edges, ne

    def hist(self, g, h, rows=None):
        gh = np.zeros((self.F, 255, 2), np.float32); cnt = np.zeros((self.F, 255), np.uint32)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        self.lib.L.tcref_hist(self.h, _p(np.ascontiguousarray(g, np.float32)),
                              _p(np.ascontiguousarray(h, np.float32)), _p(r),
                              0 if r is None else len(r), _p(gh), _p(cnt))
        return gh, cnt

    def tree(self, g, h, max_depth=6, lambda_l2=1.0, min_child_weight=1.0):
        D = max_depth
        feat = np.zeros(D, np.uint16); b = np.zeros(D, np.uint8); ft = np.zeros(D, np.float32)
        gn = np.zeros(D, np.float32); lv = np.zeros(1 << D, np.float32); li = np.zeros(self.n, np.uint32)
        d = self.lib.L.tcref_tree_build(self.h, _p(np.ascontiguousarray(g, np.float32)),
                                        _p(np.ascontiguousarray(h, np.float32)), D, lambda_l2,
                                        min_child_weight, _p(feat), _p(b), _p(ft), _p(gn), _p(lv),
                                        _p(li))
        return dict(depth=d, feature=feat[:d], bin=b[:d], fthr=ft[:d], gain=gn[:d],
                    leaves=lv[:1 << d], leaf_idx=li)

    def predict_binned(self, model):
        """Booster::predict_raw over this handle's binned rows (one weight,
        the learning rate, for every tree)."""
        out = np.zeros(self.n, np.float32)
        self.lib.L.tcref_predict_binned(self.h, _p(model["feature"]), _p(model["fthr"]),
                                        _p(model["bin"]), _p(model["depth"]), _p(model["leaves"]),
                                        model["T"], model["D"],
                                        float(model["weight"][0]) if model["T"] else 0.0,
                                        model["base"], _p(out))
        return out

    def random_subsample(self, ratio, seed):
        out = np.zeros(self.n, np.uint32)
        k = self.lib.L.tcref_random_subsample(self.h, ratio, seed, _p(out))
        return out[:k]

    def goss_subsample(self, g, top_rate, other_rate, seed):
        """Dataset::goss_subsample on gradients g of the handle's rows."""
        out = np.zeros(self.n, np.uint32)
        k = self.lib.L.tcref_goss_subsample(self.h, _p(np.ascontiguousarray(g, np.float32)),
                                            top_rate, other_rate, seed, _p(out))
        return out[:k]

    def mvs_subsample(self, g, ratio, seed):
        """Dataset::mvs_subsample on gradients g of the handle's rows (one
        OpenMP thread)."""
        out = np.zeros(self.n, np.uint32)
        k = self.lib.L.tcref_mvs_subsample(self.h, _p(np.ascontiguousarray(g, np.float32)),
                                           ratio, seed, _p(out))
        return out[:k]

    def fit_mvs(self, loss, n_trees, max_depth=6, lr=0.1, lambda_l2=1.0, min_child_weight=1.0,
                ratio=0.5, colsample=1.0, seed=42):
        """The reference's fit loop with MVS row sampling (booster.cpp:201-205)."""
        m = new_model(n_trees, max_depth)
        tp = np.zeros(self.n, np.float32)
        base = self.lib.L.tcref_fit_mvs(self.h, loss, n_trees, max_depth, lr, lambda_l2,
                                        min_child_weight, ratio, colsample, seed,
                                        _p(m["feature"]), _p(m["bin"]), _p(m["fthr"]),
                                        _p(m["gain"]), _p(m["depth"]), _p(m["leaves"]), _p(tp))
        m["base"] = float(base)
        m["weight"][:] = lr
        return m, tp

    def fit_goss(self, loss, n_trees, max_depth=6, lr=0.1, lambda_l2=1.0, min_child_weight=1.0,
                 top_rate=0.2, other_rate=0.1, colsample=1.0, seed=42):
        """The reference's fit loop with GOSS row sampling (booster.cpp:207-212)."""
        m = new_model(n_trees, max_depth)
        tp = np.zeros(self.n, np.float32)
        base = self.lib.L.tcref_fit_goss(self.h, loss, n_trees, max_depth, lr, lambda_l2,
                                         min_child_weight, top_rate, other_rate, colsample, seed,
                                         _p(m["feature"]), _p(m["bin"]), _p(m["fthr"]),
                                         _p(m["gain"]), _p(m["depth"]), _p(m["leaves"]), _p(tp))
        m["base"] = float(base)
        m["weight"][:] = lr
        return m, tp

    def fit(self, loss, n_trees, max_depth=6, lr=0.1, lambda_l2=1.0, min_child_weight=1.0,
            subsample=1.0, colsample=1.0, seed=42):
        m = new_model(n_trees, max_depth)
        tp = np.zeros(self.n, np.float32); sec = np.zeros(3)
        base = self.lib.L.tcref_fit(self.h, loss, n_trees, max_depth, lr, lambda_l2,
                                    min_child_weight, subsample, colsample, seed,
                                    _p(m["feature"]), _p(m["bin"]), _p(m["fthr"]),
                                    _p(m["gain"]), _p(m["depth"]), _p(m["leaves"]), _p(tp), _p(sec))
        m["base"] = float(base)
        m["weight"][:] = lr
        return m, tp, sec
