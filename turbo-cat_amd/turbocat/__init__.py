"""TurboCat on MI355X: drop-in ``turbocat`` package.

Mirrors the reference's Python wrappers (python/turbocat/__init__.py:290-874):
the same class names, constructor keywords and defaults, numpy in / numpy out
(``predict`` -> int32 labels, ``predict_proba`` -> float32 (n, n_classes)),
and the same errors.  mode="auto"/"large" grows symmetric trees, mode="small"
regular depthwise trees; labels with more than two classes train the
reference's multiclass loop (one regular tree per class and iteration).
Quantization, training and prediction run on the GPU through the C-ABI in
include/turbocat_hip.h (``_turbocat`` is the pybind11 module over it; there is
no CPU fallback -- without a gfx950 device the calls raise).
"""
from __future__ import annotations

import numpy as np

__version__ = "0.4.2+mi355x"

from ._turbocat import TurboCatClassifier as _TurboCatClassifier  # noqa: E402
from ._turbocat import TurboCatRegressor as _TurboCatRegressor  # noqa: E402
from ._turbocat import print_info  # noqa: E402,F401


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class TurboCatClassifier:
    """Gradient-boosted trees for binary (LogLoss) and multiclass (softmax)
    classification."""

    def __init__(self, n_estimators=1000, learning_rate=0.1, max_depth=6, max_bins=255,
                 subsample=0.8, colsample_bytree=0.8, min_child_weight=1.0, lambda_l2=1.0,
                 loss="logloss", mode="auto", use_goss=False, goss_top_rate=0.2,
                 goss_other_rate=0.1, use_gradtree=False, use_symmetric=False,
                 use_ordered_boosting=False, early_stopping_rounds=50, n_jobs=-1,
                 n_threads=None, seed=42, verbosity=1):
        effective_threads = n_threads if n_threads is not None else n_jobs
        if use_symmetric and mode == "auto":
            mode = "large"
        self._model = _TurboCatClassifier(
            n_estimators=n_estimators, learning_rate=learning_rate, max_depth=max_depth,
            max_bins=max_bins, subsample=subsample, colsample_bytree=colsample_bytree,
            min_child_weight=min_child_weight, lambda_l2=lambda_l2, loss=loss, mode=mode,
            use_goss=use_goss, goss_top_rate=goss_top_rate, goss_other_rate=goss_other_rate,
            use_gradtree=use_gradtree, use_ordered_boosting=use_ordered_boosting,
            early_stopping_rounds=early_stopping_rounds, n_threads=effective_threads, seed=seed,
            verbosity=verbosity)
        self._params = dict(n_estimators=n_estimators, learning_rate=learning_rate,
                            max_depth=max_depth, max_bins=max_bins, subsample=subsample,
                            colsample_bytree=colsample_bytree, min_child_weight=min_child_weight,
                            lambda_l2=lambda_l2, loss=loss, mode=mode, use_goss=use_goss,
                            goss_top_rate=goss_top_rate, goss_other_rate=goss_other_rate,
                            use_gradtree=use_gradtree, use_ordered_boosting=use_ordered_boosting,
                            early_stopping_rounds=early_stopping_rounds, n_jobs=n_jobs, seed=seed,
                            verbosity=verbosity)

    def fit(self, X, y, X_val=None, y_val=None, cat_features=None):
        X = _f32(X)
        y = _f32(y)
        if X_val is not None:
            X_val, y_val = _f32(X_val), _f32(y_val)
        else:
            X_val = np.array([], dtype=np.float32).reshape(0, X.shape[1] if X.ndim > 1 else 0)
            y_val = np.array([], dtype=np.float32)
        self._model.fit(X, y, X_val, y_val, [] if cat_features is None else list(cat_features))
        return self

    def predict_proba(self, X):
        return np.asarray(self._model.predict_proba(_f32(X)), dtype=np.float32)

    def predict_proba_nobinning_fast(self, X):
        return np.asarray(self._model.predict_proba_nobinning_fast(_f32(X)), dtype=np.float32)

    def predict(self, X, threshold=0.5):
        proba = self.predict_proba(X)
        if proba.shape[1] == 2:
            return (proba[:, 1] >= threshold).astype(np.int32)
        return np.argmax(proba, axis=1).astype(np.int32)

    def feature_importance(self):
        return self._model.feature_importance()

    def save(self, path):
        self._model.save(path)

    def load(self, path):
        self._model.load(path)
        return self

    def get_params(self, deep=True):
        return self._params.copy()

    def set_params(self, **params):
        self._params.update(params)
        return self

    @property
    def n_trees(self):
        return self._model.n_trees

    @property
    def n_classes_(self):
        return self._model.n_classes_


class TurboCatRegressor:
    """Gradient-boosted trees for regression (MSE)."""

    def __init__(self, n_estimators=1000, learning_rate=0.1, max_depth=6, loss="mse",
                 mode="auto", subsample=0.8, colsample_bytree=0.8, use_goss=False,
                 goss_top_rate=0.2, goss_other_rate=0.1, early_stopping_rounds=50, lambda_l2=0.0,
                 min_child_weight=1.0, max_leaves=64, grow_policy="depthwise", n_jobs=-1,
                 n_threads=None, seed=42, verbosity=1):
        effective_threads = n_threads if n_threads is not None else n_jobs
        self._model = _TurboCatRegressor(
            n_estimators=n_estimators, learning_rate=learning_rate, max_depth=max_depth, loss=loss,
            mode=mode, subsample=subsample, colsample_bytree=colsample_bytree, use_goss=use_goss,
            goss_top_rate=goss_top_rate, goss_other_rate=goss_other_rate,
            early_stopping_rounds=early_stopping_rounds, lambda_l2=lambda_l2,
            min_child_weight=min_child_weight, max_leaves=max_leaves, grow_policy=grow_policy,
            n_threads=effective_threads, seed=seed, verbosity=verbosity)
        self._params = dict(n_estimators=n_estimators, learning_rate=learning_rate,
                            max_depth=max_depth, loss=loss, mode=mode, subsample=subsample,
                            colsample_bytree=colsample_bytree, use_goss=use_goss,
                            goss_top_rate=goss_top_rate, goss_other_rate=goss_other_rate,
                            early_stopping_rounds=early_stopping_rounds, lambda_l2=lambda_l2,
                            min_child_weight=min_child_weight, max_leaves=max_leaves,
                            grow_policy=grow_policy, n_jobs=n_jobs, seed=seed, verbosity=verbosity)

    def fit(self, X, y, X_val=None, y_val=None):
        X = _f32(X)
        y = _f32(y)
        if X_val is not None:
            X_val, y_val = _f32(X_val), _f32(y_val)
        else:
            X_val = np.array([], dtype=np.float32).reshape(0, X.shape[1] if X.ndim > 1 else 0)
            y_val = np.array([], dtype=np.float32)
        self._model.fit(X, y, X_val, y_val)
        return self

    def predict(self, X):
        return np.asarray(self._model.predict_nobinning_fast(_f32(X)), dtype=np.float32)

    def feature_importance(self):
        return self._model.feature_importance()

    def save(self, path):
        self._model.save(path)

    def load(self, path):
        self._model.load(path)
        return self

    def get_params(self, deep=True):
        return self._params.copy()

    def set_params(self, **params):
        self._params.update(params)
        return self

    @property
    def n_trees(self):
        return self._model.n_trees

    @property
    def base_prediction(self):
        return self._model.base_prediction


__all__ = ["TurboCatClassifier", "TurboCatRegressor", "print_info"]
