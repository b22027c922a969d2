"""ctypes front-end to the C-ABI in include/turbocat_hip.h.

This is the thin Python view of the drop-in boundary used by the parity
tests and the benchmark; the user-facing classes live in
``turbocat.TurboCatClassifier`` / ``TurboCatRegressor``.  Every call raises
``TurboCatError`` (a RuntimeError, like the reference's pybind-translated
std::runtime_error) when the library returns a nonzero code; the library
itself refuses to run on anything but a gfx950 device.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_float, c_int32, c_int64,
                    c_size_t, c_uint8, c_uint16, c_uint32, c_uint64, c_void_p)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libturbocat_hip.so")

TC_LOSS_MSE, TC_LOSS_LOGLOSS, TC_LOSS_SOFTMAX = 0, 1, 2
TC_EBUSY, TC_ETIMEOUT = -5, -6

# Every exported symbol of include/turbocat_hip.h (checked by the tests).
EXPORTS = [
    "tc_ctx_create", "tc_ctx_destroy", "tc_ctx_synchronize", "tc_last_error",
    "tc_device_alloc", "tc_device_free", "tc_memcpy_h2d", "tc_memcpy_d2h",
    "tc_quantize_fit", "tc_dataset_create", "tc_dataset_bins", "tc_dataset_shape",
    "tc_dataset_destroy", "tc_gradients", "tc_gradients_multi", "tc_random_subsample",
    "tc_goss_subsample",
    "tc_mvs_subsample",
    "tc_init_prediction", "tc_hist_build",
    "tc_hist_subtract", "tc_split_scan", "tc_tree_build", "tc_tree_build_multi", "tc_train",
    "tc_train_valid", "tc_train_batch",
    "tc_last_train_timing", "tc_trainer_create", "tc_trainer_step", "tc_trainer_step_timed",
    "tc_trainer_sync", "tc_trainer_model", "tc_trainer_hist_rows", "tc_trainer_hist_chunks",
    "tc_trainer_destroy",
    "tc_ensemble_create", "tc_ensemble_create_multi", "tc_ensemble_destroy", "tc_predict",
    "tc_predict_dev", "tc_predict_noalloc", "tc_predict_binned", "tc_serve_start", "tc_serve_submit",
    "tc_serve_wait", "tc_serve_cancel", "tc_serve_stats", "tc_serve_counters",
    "tc_serve_latencies", "tc_serve_stop",
    "tc_host_alloc", "tc_host_free",
    "tc_rtree_build", "tc_rtrain", "tc_rtrain_valid", "tc_rensemble_create", "tc_rensemble_destroy", "tc_rpredict",
    "tc_rpredict_dev", "tc_last_train_hist_rows",
]


class TurboCatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class TreeParams(Structure):
    _fields_ = [("max_depth", c_uint32), ("max_bins", c_uint32), ("lambda_l2", c_float),
                ("lambda_l1", c_float), ("max_delta_step", c_float),
                ("min_samples_leaf", c_uint32), ("min_child_weight", c_float),
                ("min_split_gain", c_float)]


class TrainParams(Structure):
    _fields_ = [("loss", c_int32), ("n_trees", c_uint32), ("learning_rate", c_float),
                ("tree", TreeParams), ("hist_chunk", c_uint32), ("use_graph", c_int32),
                ("subsample", c_float), ("colsample_bytree", c_float), ("seed", c_uint64),
                ("n_classes", c_uint32), ("hist_deep_items", c_uint32),
                ("use_goss", c_int32), ("goss_top_rate", c_float), ("goss_other_rate", c_float),
                ("use_mvs", c_int32), ("mvs_subsample", c_float)]


class Model(Structure):
    _fields_ = [("feature", c_void_p), ("bin", c_void_p), ("fthr", c_void_p), ("gain", c_void_p),
                ("depth", c_void_p), ("leaves", c_void_p), ("train_preds", c_void_p),
                ("base", c_float)]


class TrainHistory(Structure):
    _fields_ = [("train_loss", c_void_p), ("valid_loss", c_void_p), ("n_trees_built", c_uint32),
                ("stopped_early", c_int32), ("best_iteration", c_uint32),
                ("best_valid_loss", c_float)]


class RTreeParams(Structure):
    _fields_ = [("max_depth", c_uint32), ("max_leaves", c_uint32), ("max_bins", c_uint32),
                ("lambda_l2", c_float), ("min_child_weight", c_float),
                ("min_samples_leaf", c_uint32), ("min_split_gain", c_float), ("gamma", c_float),
                ("max_delta_step", c_float), ("leaf_smooth", c_float),
                ("learn_missing_direction", c_int32), ("grow_policy", c_int32)]


class RTrainParams(Structure):
    _fields_ = [("loss", c_int32), ("n_iterations", c_uint32), ("learning_rate", c_float),
                ("tree", RTreeParams), ("subsample", c_float), ("colsample_bytree", c_float),
                ("seed", c_uint64), ("n_classes", c_uint32), ("use_graph", c_int32),
                ("time_hist", c_int32), ("use_goss", c_int32), ("goss_top_rate", c_float),
                ("goss_other_rate", c_float), ("use_mvs", c_int32), ("mvs_subsample", c_float)]


class RModel(Structure):
    _fields_ = [("feature", c_void_p), ("bin", c_void_p), ("flags", c_void_p), ("left", c_void_p),
                ("right", c_void_p), ("value", c_void_p), ("gain", c_void_p),
                ("n_nodes", c_void_p), ("n_leaves", c_void_p), ("depth", c_void_p),
                ("train_preds", c_void_p), ("base", c_float)]


def rtree_params(max_depth=6, max_leaves=64, max_bins=255, lambda_l2=1.0, min_child_weight=1.0,
                 min_samples_leaf=1, min_split_gain=0.0, gamma=0.0, max_delta_step=0.0,
                 leaf_smooth=0.0, learn_missing_direction=True, grow_policy=0) -> RTreeParams:
    """TreeConfig of a regular tree (include/turbocat/config.hpp:21-59 defaults);
    grow_policy 0 depthwise, 1 lossguide."""
    return RTreeParams(max_depth, max_leaves, max_bins, lambda_l2, min_child_weight,
                       min_samples_leaf, min_split_gain, gamma, max_delta_step, leaf_smooth,
                       1 if learn_missing_direction else 0, int(grow_policy))


def rtrain_params(loss=TC_LOSS_MSE, n_iterations=100, learning_rate=0.1, subsample=1.0,
                  colsample_bytree=1.0, seed=42, n_classes=0, use_graph=1, time_hist=0,
                  use_goss=False, goss_top_rate=0.2, goss_other_rate=0.1, use_mvs=False,
                  mvs_subsample=0.5, **tree) -> RTrainParams:
    return RTrainParams(loss, n_iterations, learning_rate, rtree_params(**tree), subsample,
                        colsample_bytree, seed, n_classes, use_graph, time_hist,
                        1 if use_goss else 0, goss_top_rate, goss_other_rate, 1 if use_mvs else 0,
                        mvs_subsample)


def _new_rmodel(n_trees, max_depth, n=0, K=1):
    NC = (2 << max_depth) - 1
    a = dict(node_cap=NC, n_trees=n_trees, feature=np.zeros(n_trees * NC, np.uint16),
             bin=np.zeros(n_trees * NC, np.uint8), flags=np.zeros(n_trees * NC, np.uint8),
             left=np.zeros(n_trees * NC, np.uint32), right=np.zeros(n_trees * NC, np.uint32),
             value=np.zeros(n_trees * NC, np.float32), gain=np.zeros(n_trees * NC, np.float32),
             n_nodes=np.zeros(n_trees, np.uint32), n_leaves=np.zeros(n_trees, np.uint32),
             depth=np.zeros(n_trees, np.uint16),
             train_preds=np.zeros((K, n), np.float32) if n else None)
    m = RModel(*[_p(a[k]) for k in ("feature", "bin", "flags", "left", "right", "value", "gain",
                                    "n_nodes", "n_leaves", "depth", "train_preds")], 0.0)
    return m, a


def rtrees(a):
    """Per-tree node dicts (the reference's numbering) of a regular model."""
    NC, out = a["node_cap"], []
    for t in range(a["n_trees"]):
        k = int(a["n_nodes"][t])
        sl = slice(t * NC, t * NC + k)
        out.append(dict(feature=a["feature"][sl].copy(), bin=a["bin"][sl].copy(),
                        flags=a["flags"][sl].copy(), left=a["left"][sl].copy(),
                        right=a["right"][sl].copy(), value=a["value"][sl].copy(),
                        gain=a["gain"][sl].copy(), n_nodes=k, n_leaves=int(a["n_leaves"][t]),
                        depth=int(a["depth"][t])))
    return out


class ServeParams(Structure):
    _fields_ = [("max_batch_rows", c_uint32), ("memory_budget", c_uint64),
                ("n_streams", c_uint32), ("transform", c_int32), ("fail_batch", c_uint64)]


class ServeCounters(Structure):
    _fields_ = [("batches", c_uint64), ("rows", c_uint64), ("peak_admitted_bytes", c_uint64),
                ("max_batches_in_flight", c_uint64), ("cancelled", c_uint64),
                ("failed", c_uint64), ("kernel_launches", c_uint64)]


def tree_params(max_depth=6, max_bins=255, lambda_l2=1.0, lambda_l1=0.0, max_delta_step=0.0,
                min_samples_leaf=1, min_child_weight=1.0, min_split_gain=0.0) -> TreeParams:
    return TreeParams(max_depth, max_bins, lambda_l2, lambda_l1, max_delta_step,
                      min_samples_leaf, min_child_weight, min_split_gain)


def train_params(loss=TC_LOSS_MSE, n_trees=100, learning_rate=0.1, hist_chunk=0, use_graph=1,
                 subsample=1.0, colsample_bytree=1.0, seed=42, n_classes=0, hist_deep_items=0,
                 use_goss=False, goss_top_rate=0.2, goss_other_rate=0.1, use_mvs=False,
                 mvs_subsample=0.5, **tree) -> TrainParams:
    return TrainParams(loss, n_trees, learning_rate, tree_params(**tree), hist_chunk, use_graph,
                       subsample, colsample_bytree, seed, n_classes, hist_deep_items,
                       1 if use_goss else 0, goss_top_rate, goss_other_rate,
                       1 if use_mvs else 0, mvs_subsample)


def mvs_subsample(g, ratio, seed):
    """The row sample the trainer draws for one tree under MVS, on gradients g
    (tc_mvs_subsample); its size depends on the gradients."""
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros(max(len(g), 1), np.uint32)
    k = c_uint32()
    check(load().tc_mvs_subsample(_p(g), len(g), ratio, seed, _p(out), byref(k)))
    return out[:k.value]


def channels(p: TrainParams):
    """Gradient channels of a fit: K classes for softmax, else 1."""
    return p.n_classes if p.loss == TC_LOSS_SOFTMAX else 1


def random_subsample(n, ratio, seed):
    """The row sample the trainer draws for one tree (tc_random_subsample)."""
    out = np.zeros(max(n, 1), np.uint32)
    k = c_uint32()
    check(load().tc_random_subsample(n, ratio, seed, _p(out), byref(k)))
    return out[:k.value]


def goss_subsample(g, top_rate, other_rate, seed):
    """The row sample the trainer draws for one tree under GOSS, on gradients g
    (tc_goss_subsample)."""
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros(max(len(g), 1), np.uint32)
    k = c_uint32()
    check(load().tc_goss_subsample(_p(g), len(g), top_rate, other_rate, seed, _p(out), byref(k)))
    return out[:k.value]


def next_random(state):
    """xorshift64* step (Booster::next_random); returns (value, new_state)."""
    m = (1 << 64) - 1
    state ^= state >> 12
    state ^= (state << 25) & m
    state ^= state >> 27
    return (state * 0x2545F4914F6CDD1D) & m, state


_lib = None


def load(path: str = LIB_PATH):
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TurboCatError(-3, f"{path} not built; run python turbo-cat_amd/build.py")
    L = ctypes.CDLL(path)
    V, I, U = c_void_p, c_int32, c_uint32
    sig = {
        "tc_ctx_create": [I, V], "tc_ctx_destroy": [V], "tc_ctx_synchronize": [V],
        "tc_device_alloc": [V, c_size_t, V], "tc_device_free": [V, V],
        "tc_memcpy_h2d": [V, V, V, c_size_t], "tc_memcpy_d2h": [V, V, V, c_size_t],
        "tc_quantize_fit": [V, V, U, U, U, I, V, V, V],
        "tc_dataset_create": [V, V, U, U, V, V, U, V],
        "tc_dataset_bins": [V, V], "tc_dataset_shape": [V, V, V], "tc_dataset_destroy": [V],
        "tc_gradients": [V, I, V, V, V, V, U], "tc_init_prediction": [I, V, U, V],
        "tc_gradients_multi": [V, V, V, V, V, U, U],
        "tc_random_subsample": [U, c_float, c_uint64, V, V],
        "tc_goss_subsample": [V, U, c_float, c_float, c_uint64, V, V],
        "tc_mvs_subsample": [V, U, c_float, c_uint64, V, V],
        "tc_hist_build": [V, V, V, V, U, V, V], "tc_hist_subtract": [V, V, V, V, U],
        "tc_split_scan": [V, V, V, V, V, U, U, U, c_float, I, V, V, V],
        "tc_tree_build": [V, V, V, POINTER(TreeParams), V, V, V, V, V, V, V],
        "tc_tree_build_multi": [V, V, V, U, POINTER(TreeParams), V, V, V, V, V, V, V],
        "tc_train": [V, V, POINTER(TrainParams), POINTER(Model)],
        "tc_train_valid": [V, V, POINTER(TrainParams), V, V, U, POINTER(Model),
                           POINTER(TrainHistory)],
        "tc_train_batch": [V, U, V, V, V, V],
        "tc_last_train_timing": [V, V, V],
        "tc_trainer_create": [V, U, V, V, V, V], "tc_trainer_step": [V, U],
        "tc_trainer_step_timed": [V, U, V, V], "tc_trainer_sync": [V],
        "tc_trainer_model": [V, U, POINTER(Model)], "tc_trainer_hist_rows": [V, U, V],
        "tc_trainer_hist_chunks": [V, U, V, V],
        "tc_trainer_destroy": [V],
        "tc_ensemble_create": [V, U, U, V, V, V, V, V, V, c_float, U, V],
        "tc_ensemble_create_multi": [V, U, U, U, V, V, V, V, V, V, c_float, U, V],
        "tc_ensemble_destroy": [V], "tc_predict": [V, V, U, U, I, V, V],
        "tc_predict_dev": [V, V, U, U, I, V], "tc_predict_binned": [V, V, V],
        "tc_predict_noalloc": [V, V, U, U, I, V],
        "tc_serve_start": [V, POINTER(ServeParams), V], "tc_serve_submit": [V, V, U, V, V],
        "tc_serve_wait": [V, c_uint64, c_int64], "tc_serve_stats": [V, V, V],
        "tc_serve_cancel": [V, c_uint64], "tc_serve_counters": [V, POINTER(ServeCounters)],
        "tc_serve_latencies": [V, V, c_uint64, V],
        "tc_serve_stop": [V], "tc_host_alloc": [c_size_t, V], "tc_host_free": [V],
        "tc_rtree_build": [V, V, V, V, U, POINTER(RTreeParams), POINTER(RModel)],
        "tc_rtrain": [V, V, POINTER(RTrainParams), POINTER(RModel)],
        "tc_rtrain_valid": [V, V, POINTER(RTrainParams), V, V, U, POINTER(RModel),
                            POINTER(TrainHistory)],
        "tc_rensemble_create": [V, U, U, V, V, V, V, V, V, V, V, V, V, V, V, U, I, V],
        "tc_rensemble_destroy": [V], "tc_rpredict": [V, V, U, U, I, V],
        "tc_rpredict_dev": [V, V, U, U, I, V], "tc_last_train_hist_rows": [V, V, V],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = c_int32
    L.tc_last_error.restype = c_char_p
    L.tc_last_error.argtypes = []
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def check(rc):
    if rc != 0:
        raise TurboCatError(rc, load().tc_last_error().decode(errors="replace"))
    return rc


class Context:
    """A device context (one HIP stream on one GPU)."""

    def __init__(self, device: int = 0):
        self.L = load()
        self.h = c_void_p()
        check(self.L.tc_ctx_create(device, byref(self.h)))

    def close(self):
        if self.h:
            self.L.tc_ctx_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.L.tc_ctx_synchronize(self.h))

    # ------------------------------------------------------------- datasets
    def quantize_fit(self, X, max_bins=255, wrap=True):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        edges = np.zeros((F, 255), np.float32)
        ne = np.zeros(F, np.uint32)
        h = c_void_p()
        check(self.L.tc_quantize_fit(self.h, _p(X), n, F, max_bins, int(wrap), _p(edges), _p(ne),
                                     byref(h)))
        return Dataset(self, h, n, F), edges, ne

    def dataset(self, X, edges, n_edges, max_bins=255):
        X = np.ascontiguousarray(X, np.float32)
        n, F = X.shape
        h = c_void_p()
        check(self.L.tc_dataset_create(self.h, _p(X), n, F,
                                       _p(np.ascontiguousarray(edges, np.float32)),
                                       _p(np.ascontiguousarray(n_edges, np.uint32)), max_bins,
                                       byref(h)))
        return Dataset(self, h, n, F)

    # ------------------------------------------------------------ primitives
    def gradients(self, loss, y, preds):
        y = np.ascontiguousarray(y, np.float32)
        p = np.ascontiguousarray(preds, np.float32)
        g = np.zeros_like(y)
        h = np.zeros_like(y)
        check(self.L.tc_gradients(self.h, loss, _p(y), _p(p), _p(g), _p(h), len(y)))
        return g, h

    def gradients_multi(self, y, preds):
        """A8 softmax gradients; preds (n, K) row-major -> g, h (n, K)."""
        y = np.ascontiguousarray(y, np.float32)
        p = np.ascontiguousarray(preds, np.float32)
        n, K = p.shape
        g = np.zeros_like(p)
        h = np.zeros_like(p)
        check(self.L.tc_gradients_multi(self.h, _p(y), _p(p), _p(g), _p(h), n, K))
        return g, h

    def hist_subtract(self, parent, child):
        parent = np.ascontiguousarray(parent, np.float32)
        child = np.ascontiguousarray(child, np.float32)
        out = np.zeros_like(parent)
        check(self.L.tc_hist_subtract(self.h, _p(parent), _p(child), _p(out), parent.shape[0]))
        return out

    def split_scan(self, hist, node_g, node_h, node_count, max_bins=255, lambda_l2=1.0,
                   is_root=None):
        hist = np.ascontiguousarray(hist, np.float32)
        nn, F = hist.shape[0], hist.shape[1]
        if is_root is None:
            is_root = nn == 1
        f, b, g = c_int32(), c_int32(), c_float()
        check(self.L.tc_split_scan(self.h, _p(hist), _p(np.ascontiguousarray(node_g, np.float32)),
                                   _p(np.ascontiguousarray(node_h, np.float32)),
                                   _p(np.ascontiguousarray(node_count, np.uint32)), nn, F,
                                   max_bins, lambda_l2, int(is_root), byref(f), byref(b),
                                   byref(g)))
        return f.value, b.value, g.value

    # -------------------------------------------------------------- training
    def train_batch(self, datasets, ys, params):
        n = len(datasets)
        if isinstance(params, TrainParams):
            params = [params] * n
        P = (TrainParams * n)(*params)
        D = (c_void_p * n)(*[d.h for d in datasets])
        ys = [np.ascontiguousarray(y, np.float32) for y in ys]
        Y = (c_void_p * n)(*[y.ctypes.data for y in ys])
        models, arrs = [], []
        for d, p in zip(datasets, params):
            m, a = _new_model(p.n_trees, p.tree.max_depth, d.n, channels(p))
            models.append(m)
            arrs.append(a)
        M = (Model * n)(*models)
        check(self.L.tc_train_batch(self.h, n, D, Y, byref(P[0]), byref(M[0])))
        out = []
        for k in range(n):
            a = arrs[k]
            a["base"] = float(M[k].base)
            a["weight"] = np.full(params[k].n_trees, params[k].learning_rate, np.float32)
            out.append(a)
        return out

    def last_train_hist_rows(self):
        """(rows fed to histogram launches, launches) of the last regular-tree fit."""
        r, k = c_uint64(), c_uint32()
        check(self.L.tc_last_train_hist_rows(self.h, byref(r), byref(k)))
        return r.value, k.value

    def last_train_timing(self):
        a, b = c_double(), c_double()
        check(self.L.tc_last_train_timing(self.h, byref(a), byref(b)))
        return a.value, b.value

    def ensemble(self, model, n_features):
        return Ensemble(self, model, n_features)


def _new_model(T, D, n, K=1):
    """Host arrays of a model; with K channels leaves are [tree][leaf][class]
    and train_preds class-major (K, n)."""
    a = dict(T=T, D=D, K=K, feature=np.zeros(T * D, np.uint16), bin=np.zeros(T * D, np.uint8),
             fthr=np.zeros(T * D, np.float32), gain=np.zeros(T * D, np.float32),
             depth=np.zeros(T, np.uint16), leaves=np.zeros((T << D) * K, np.float32),
             train_preds=np.zeros(n * K, np.float32) if K == 1 else np.zeros((K, n), np.float32))
    m = Model(_p(a["feature"]), _p(a["bin"]), _p(a["fthr"]), _p(a["gain"]), _p(a["depth"]),
              _p(a["leaves"]), _p(a["train_preds"]), 0.0)
    return m, a


class Dataset:
    def __init__(self, ctx: Context, h, n, F):
        self.ctx, self.h, self.n, self.F = ctx, h, n, F

    def close(self):
        if self.h:
            self.ctx.L.tc_dataset_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bins(self):
        out = np.zeros((self.F, self.n), np.uint8)
        check(self.ctx.L.tc_dataset_bins(self.h, _p(out)))
        return out

    def hist(self, g, h, rows=None, counts=False):
        """F x 256 x (grad, hess) float32; with counts=True also the F x 256
        row counts (GradientPair::count)."""
        out = np.zeros((self.F, 256, 2), np.float32)
        cnt = np.zeros((self.F, 256), np.uint32) if counts else None
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        check(self.ctx.L.tc_hist_build(self.h, _p(np.ascontiguousarray(g, np.float32)),
                                       _p(np.ascontiguousarray(h, np.float32)), _p(r),
                                       0 if r is None else len(r), _p(out), _p(cnt)))
        return (out, cnt) if counts else out

    def tree(self, g, h, params: TreeParams, want_leaf_idx=True):
        D = params.max_depth
        feat = np.zeros(D, np.uint16); b = np.zeros(D, np.uint8); ft = np.zeros(D, np.float32)
        gn = np.zeros(D, np.float32); lv = np.zeros(1 << D, np.float32)
        li = np.zeros(self.n, np.uint32) if want_leaf_idx else None
        dep = c_int32()
        check(self.ctx.L.tc_tree_build(self.h, _p(np.ascontiguousarray(g, np.float32)),
                                       _p(np.ascontiguousarray(h, np.float32)), byref(params),
                                       _p(feat), _p(b), _p(ft), _p(gn), _p(lv), _p(li),
                                       byref(dep)))
        d = dep.value
        return dict(depth=d, feature=feat[:d], bin=b[:d], fthr=ft[:d], gain=gn[:d],
                    leaves=lv[:1 << d], leaf_idx=li)

    def tree_multi(self, g, h, params: TreeParams, want_leaf_idx=True):
        """A18 teacher-forced K-channel tree; g, h class-major (K, n)."""
        g = np.ascontiguousarray(g, np.float32); h = np.ascontiguousarray(h, np.float32)
        K = g.shape[0]
        D = params.max_depth
        feat = np.zeros(D, np.uint16); b = np.zeros(D, np.uint8); ft = np.zeros(D, np.float32)
        gn = np.zeros(D, np.float32); lv = np.zeros((1 << D) * K, np.float32)
        li = np.zeros(self.n, np.uint32) if want_leaf_idx else None
        dep = c_int32()
        check(self.ctx.L.tc_tree_build_multi(self.h, _p(g), _p(h), K, byref(params), _p(feat),
                                             _p(b), _p(ft), _p(gn), _p(lv), _p(li), byref(dep)))
        d = dep.value
        return dict(depth=d, feature=feat[:d], bin=b[:d], fthr=ft[:d], gain=gn[:d],
                    leaves=lv[:(1 << d) * K].reshape(1 << d, K), leaf_idx=li)

    def train(self, y, params: TrainParams):
        m, a = _new_model(params.n_trees, params.tree.max_depth, self.n, channels(params))
        y = np.ascontiguousarray(y, np.float32)
        check(self.ctx.L.tc_train(self.h, _p(y), byref(params), byref(m)))
        a["base"] = float(m.base)
        a["weight"] = np.full(params.n_trees, params.learning_rate, np.float32)
        return a

    def rtree(self, g, h, params: RTreeParams, rows=None):
        """One regular tree on given gradients (tc_rtree_build): node dict in
        the reference's numbering."""
        m, a = _new_rmodel(1, params.max_depth)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        check(self.ctx.L.tc_rtree_build(self.h, _p(np.ascontiguousarray(g, np.float32)),
                                        _p(np.ascontiguousarray(h, np.float32)), _p(r),
                                        0 if r is None else len(r), byref(params), byref(m)))
        return rtrees(a)[0]

    def rtrain(self, y, params: RTrainParams):
        """Boosting with regular trees (tc_rtrain).  Returns the model dict:
        trees (iteration-major, class-minor), weights, classes, base."""
        K = params.n_classes if params.loss == TC_LOSS_SOFTMAX else 1
        T = params.n_iterations * K
        m, a = _new_rmodel(T, params.tree.max_depth, self.n, K)
        y = np.ascontiguousarray(y, np.float32)
        check(self.ctx.L.tc_rtrain(self.h, _p(y), byref(params), byref(m)))
        return dict(trees=rtrees(a), weight=np.full(T, params.learning_rate, np.float32),
                    tree_class=np.tile(np.arange(K, dtype=np.uint32), params.n_iterations),
                    K=K, base=float(m.base), train_preds=a["train_preds"], raw=a)

    def rtrain_valid(self, y, params: RTrainParams, valid, y_valid, early_stopping_rounds=50):
        """tc_rtrain_valid: (model dict as rtrain, history dict)."""
        K = params.n_classes if params.loss == TC_LOSS_SOFTMAX else 1
        T = params.n_iterations * K
        m, a = _new_rmodel(T, params.tree.max_depth, self.n, K)
        vl = np.zeros(max(params.n_iterations, 1), np.float32)
        h = TrainHistory(None, _p(vl), 0, 0, 0, 0.0)
        check(self.ctx.L.tc_rtrain_valid(self.h, _p(np.ascontiguousarray(y, np.float32)),
                                         byref(params), valid.h,
                                         _p(np.ascontiguousarray(y_valid, np.float32)),
                                         early_stopping_rounds, byref(m), byref(h)))
        model = dict(trees=rtrees(a), weight=np.full(T, params.learning_rate, np.float32),
                     tree_class=np.tile(np.arange(K, dtype=np.uint32), params.n_iterations),
                     K=K, base=float(m.base), train_preds=a["train_preds"], raw=a)
        hist = dict(valid_loss=vl[:h.n_trees_built].copy(), n_built=h.n_trees_built,
                    stopped=bool(h.stopped_early), best_iteration=h.best_iteration,
                    best_valid_loss=h.best_valid_loss)
        return model, hist

    def train_valid(self, y, params: TrainParams, valid=None, y_valid=None,
                    early_stopping_rounds=50):
        """tc_train_valid: returns (model, history); history holds the mean
        training / validation loss after each tree, the trees built and the
        best iteration."""
        T = params.n_trees
        m, a = _new_model(T, params.tree.max_depth, self.n, channels(params))
        y = np.ascontiguousarray(y, np.float32)
        tl = np.zeros(max(1, T), np.float32)
        vl = np.zeros(max(1, T), np.float32)
        h = TrainHistory(tl.ctypes.data, vl.ctypes.data, 0, 0, 0, 0.0)
        yv = None if y_valid is None else np.ascontiguousarray(y_valid, np.float32)
        check(self.ctx.L.tc_train_valid(self.h, _p(y), byref(params),
                                        None if valid is None else valid.h,
                                        None if yv is None else _p(yv),
                                        early_stopping_rounds, byref(m), byref(h)))
        a["base"] = float(m.base)
        a["weight"] = np.full(T, params.learning_rate, np.float32)
        hist = {"train_loss": tl[:T].copy(), "valid_loss": vl[:T].copy(),
                "n_trees_built": int(h.n_trees_built), "stopped_early": bool(h.stopped_early),
                "best_iteration": int(h.best_iteration), "best_valid_loss": float(h.best_valid_loss)}
        return a, hist


class Trainer:
    """Stepwise boosting over one or more resident datasets (tc_trainer_*)."""

    def __init__(self, ctx: Context, datasets, ys, params):
        n = len(datasets)
        if isinstance(params, TrainParams):
            params = [params] * n
        self.ctx, self.params, self.datasets = ctx, list(params), list(datasets)
        self._ys = [np.ascontiguousarray(y, np.float32) for y in ys]
        P = (TrainParams * n)(*params)
        D = (c_void_p * n)(*[d.h for d in datasets])
        Y = (c_void_p * n)(*[y.ctypes.data for y in self._ys])
        self.h = c_void_p()
        check(ctx.L.tc_trainer_create(ctx.h, n, D, Y, byref(P[0]), byref(self.h)))

    def step(self, n_trees):
        check(self.ctx.L.tc_trainer_step(self.h, n_trees))

    def step_timed(self, n_trees):
        ms, k = c_double(), c_uint32()
        check(self.ctx.L.tc_trainer_step_timed(self.h, n_trees, byref(ms), byref(k)))
        return ms.value, k.value

    def sync(self):
        check(self.ctx.L.tc_trainer_sync(self.h))

    def hist_chunks(self, job=0):
        """Histogram work-item shape of job: (rows per item at the root, items
        each deeper level is cut into; 0 when a fixed hist_chunk is used at
        every level)."""
        r, d = c_uint32(), c_uint32()
        check(self.ctx.L.tc_trainer_hist_chunks(self.h, job, byref(r), byref(d)))
        return r.value, d.value

    def hist_rows(self, job=0):
        v = c_uint64()
        check(self.ctx.L.tc_trainer_hist_rows(self.h, job, byref(v)))
        return v.value

    def model(self, job=0):
        p = self.params[job]
        m, a = _new_model(p.n_trees, p.tree.max_depth, self.datasets[job].n, channels(p))
        check(self.ctx.L.tc_trainer_model(self.h, job, byref(m)))
        a["base"] = float(m.base)
        a["weight"] = np.full(p.n_trees, p.learning_rate, np.float32)
        return a

    def close(self):
        if self.h:
            self.ctx.L.tc_trainer_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    def __init__(self, ctx: Context, model: dict, n_features: int):
        self.ctx = ctx
        self.T, self.D, self.F = model["T"], model["D"], n_features
        self.K = model.get("K", 1)
        self.h = c_void_p()
        w = model.get("weight")
        if w is None:
            w = np.ones(self.T, np.float32)
        check(ctx.L.tc_ensemble_create_multi(ctx.h, self.T, self.D, self.K, _p(model["feature"]),
                                             _p(model["fthr"]), _p(model.get("bin")),
                                             _p(model["depth"]),
                                             _p(np.ascontiguousarray(model["leaves"], np.float32)),
                                             _p(np.ascontiguousarray(w, np.float32)),
                                             float(model["base"]), n_features, byref(self.h)))

    def close(self):
        if self.h:
            self.ctx.L.tc_ensemble_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def predict(self, X, transform=0, want_leaf_idx=False):
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        out = np.zeros(n, np.float32) if self.K == 1 else np.zeros((n, self.K), np.float32)
        li = np.zeros((n, self.T), np.uint32) if want_leaf_idx else None
        check(self.ctx.L.tc_predict(self.h, _p(X), n, self.F, transform, _p(out), _p(li)))
        return (out, li) if want_leaf_idx else out

    def predict_dev(self, x_dev, n, out_dev, transform=0):
        check(self.ctx.L.tc_predict_dev(self.h, x_dev, n, self.F, transform, out_dev))

    def predict_binned(self, ds: Dataset):
        out = np.zeros(ds.n, np.float32) if self.K == 1 else np.zeros((ds.n, self.K), np.float32)
        check(self.ctx.L.tc_predict_binned(self.h, ds.h, _p(out)))
        return out


class REnsemble:
    """Regular-tree ensemble on the device (tc_rensemble_create).  trees: node
    dicts; rule 0 (apply_bins) or 1 (predict_raw_fast)."""

    def __init__(self, ctx: Context, trees, weights, edges, n_edges, n_features, rule=1, K=1,
                 tree_class=None, base=0.0):
        self.ctx, self.F, self.K = ctx, n_features, K
        off = np.zeros(len(trees) + 1, np.uint32)
        for t, tr in enumerate(trees):
            off[t + 1] = off[t] + tr["n_nodes"]

        def cat(k, dt):
            if not trees:
                return np.zeros(1, dt)
            return np.ascontiguousarray(np.concatenate([tr[k][:tr["n_nodes"]] for tr in trees]), dt)
        self._keep = [off, cat("feature", np.uint16), cat("bin", np.uint8), cat("flags", np.uint8),
                      cat("left", np.uint32), cat("right", np.uint32), cat("value", np.float32)]
        w = np.ascontiguousarray(weights, np.float32)
        tc = None if tree_class is None else np.ascontiguousarray(tree_class, np.uint32)
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(base, np.float32), (K,)), np.float32)
        e = np.ascontiguousarray(edges, np.float32)
        ne = np.ascontiguousarray(n_edges, np.uint32)
        self._keep += [w, tc, b, e, ne]
        self.h = c_void_p()
        check(ctx.L.tc_rensemble_create(ctx.h, len(trees), K, *[_p(x) for x in self._keep[:7]],
                                        _p(w), _p(tc), _p(b), _p(e), _p(ne), n_features, rule,
                                        byref(self.h)))

    def predict(self, X, transform=0):
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        out = np.zeros(n, np.float32) if self.K == 1 else np.zeros((n, self.K), np.float32)
        check(self.ctx.L.tc_rpredict(self.h, _p(X), n, self.F, transform, _p(out)))
        return out

    def predict_dev(self, x_dev, n, out_dev, transform=0):
        check(self.ctx.L.tc_rpredict_dev(self.h, x_dev, n, self.F, transform, out_dev))

    def close(self):
        if self.h:
            self.ctx.L.tc_rensemble_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory (tc_host_alloc) viewed as a numpy array: rows in it
    reach the device without a staging copy."""

    def __init__(self, shape, dtype=np.float32):
        self.L = load()
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        self.ptr = c_void_p()
        check(self.L.tc_host_alloc(max(n, 1), byref(self.ptr)))
        buf = (ctypes.c_char * max(n, 1)).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            check(self.L.tc_host_free(self.ptr))
            self.ptr = c_void_p()


class Server:
    """Continuous-batching server over a resident ensemble."""

    def __init__(self, ens: Ensemble, max_batch_rows=65536, memory_budget=1 << 30, n_streams=2,
                 transform=0, fail_batch=0):
        self.ens = ens
        self.L = ens.ctx.L
        self.h = c_void_p()
        p = ServeParams(max_batch_rows, memory_budget, n_streams, transform, fail_batch)
        check(self.L.tc_serve_start(ens.h, byref(p), byref(self.h)))
        self._keep = {}

    def cancel(self, ticket):
        """True when cancelled; False when the request is already on the device."""
        rc = self.L.tc_serve_cancel(self.h, ticket)
        if rc == TC_EBUSY:
            return False
        check(rc)
        self._keep.pop(ticket, None)
        return True

    def latencies_us(self):
        """Submit-to-result latency of every request finished so far (us)."""
        n = c_uint64()
        check(self.L.tc_serve_latencies(self.h, None, 0, byref(n)))
        out = np.zeros(n.value, np.float32)
        check(self.L.tc_serve_latencies(self.h, _p(out), n.value, byref(n)))
        return out

    def counters(self):
        c = ServeCounters()
        check(self.L.tc_serve_counters(self.h, byref(c)))
        return {k: getattr(c, k) for k, _ in ServeCounters._fields_}

    def submit(self, X, out):
        t = c_uint64()
        check(self.L.tc_serve_submit(self.h, _p(X), X.shape[0], _p(out), byref(t)))
        self._keep[t.value] = (X, out)
        return t.value

    def wait(self, ticket, timeout_us=-1):
        rc = self.L.tc_serve_wait(self.h, ticket, timeout_us)
        if rc == TC_ETIMEOUT:
            return False
        check(rc)
        self._keep.pop(ticket, None)
        return True

    def stats(self):
        b, r = c_uint64(), c_uint64()
        check(self.L.tc_serve_stats(self.h, byref(b), byref(r)))
        return b.value, r.value

    def stop(self):
        if self.h:
            check(self.L.tc_serve_stop(self.h))
            self.h = c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
