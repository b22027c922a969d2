import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "turbo-cat_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ctx():
    from turbocat.capi import Context
    c = Context(0)
    yield c
    c.close()


def regression_data(n, F, seed, low_card_col=None):
    """Synthetic regression set in the spirit of the 50K x 20 workload:
    y = x0 + 2 sin(x1) + 0.5 x2 x3 + N(0, 0.1)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, F)).astype(np.float32)
    if low_card_col is not None:
        X[:, low_card_col] = np.round(X[:, low_card_col] * 3)
    c = [X[:, i % F] for i in range(4)]  # narrow sets (F < 4) reuse columns
    y = (c[0] + 2 * np.sin(c[1]) + 0.5 * c[2] * c[3]
         + 0.1 * rng.standard_normal(n)).astype(np.float32)
    return X, y


def multiclass_data(n, F, K, seed, sigma=0.6):
    """Synthetic K-class set in the spirit of the 4M x 256 workload:
    X ~ N(0,1), W K x F ~ N(0, sigma), label = argmax(W x + Gumbel(0, 1))."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, F)).astype(np.float32)
    W = rng.normal(0.0, sigma, (K, F)).astype(np.float32)
    z = X @ W.T + rng.gumbel(0.0, 1.0, (n, K))
    return X, np.argmax(z, axis=1).astype(np.float32)


def binary_data(n, F, seed):
    """Synthetic binary set: label ~ Bernoulli(sigmoid(w.x)), w with 10 nonzeros."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, F)).astype(np.float32)
    w = np.zeros(F, np.float32)
    nz = rng.choice(F, min(10, F), replace=False)
    w[nz] = rng.standard_normal(len(nz))
    p = 1.0 / (1.0 + np.exp(-(X @ w)))
    y = (rng.random(n) < p).astype(np.float32)
    return X, y


def rel_err(x, ref, scale):
    """Normalized error |x - ref| / max(|ref|, scale): the 1e-5 tolerance
    measure used throughout (scale = sum of |terms|)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), np.asarray(scale, np.float64))
    den = np.where(den == 0, 1.0, den)
    return np.abs(x - ref) / den
