"""The allocation-free prediction entry point (tc_predict_noalloc), through
the C-ABI: the same outputs as tc_predict, bit for bit, on two calls in a row
(the second smaller, so stale scratch contents would show), for single-output
and K-channel ensembles; inputs beyond its fixed scratch are refused."""
import ctypes

import numpy as np
import pytest

from conftest import regression_data

pytestmark = pytest.mark.gpu


def _model(T, D, F, K, seed):
    rng = np.random.default_rng(seed)
    depth = np.full(T, D, np.uint16)
    depth[::5] = rng.integers(0, D, len(depth[::5]))
    return dict(T=T, D=D, K=K, feature=rng.integers(0, F, T * D).astype(np.uint16),
                fthr=rng.standard_normal(T * D).astype(np.float32),
                bin=np.zeros(T * D, np.uint8), depth=depth,
                leaves=(rng.standard_normal((T << D) * K) * 0.1).astype(np.float32),
                weight=np.full(T, 0.1, np.float32), base=0.25)


def _noalloc(ens, X, transform):
    X = np.ascontiguousarray(X, np.float32)
    n = X.shape[0]
    out = np.zeros(n, np.float32) if ens.K == 1 else np.zeros((n, ens.K), np.float32)
    rc = ens.ctx.L.tc_predict_noalloc(ens.h, X.ctypes.data_as(ctypes.c_void_p), n, ens.F,
                                      transform, out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


@pytest.mark.parametrize("K", [1, 3])
def test_noalloc_predict_equals_predict(ctx, K):
    F = 40
    ens = ctx.ensemble(_model(157, 7, F, K, seed=K), F)
    for n, seed, transform in ((20003, 1, 0), (517, 2, 1)):  # second call smaller
        X, _ = regression_data(n, F, seed)
        X[::13, 5] = np.nan
        rc, out = _noalloc(ens, X, transform)
        assert rc == 0, ctx.L.tc_last_error()
        assert np.array_equal(out, ens.predict(X, transform=transform))


def test_noalloc_predict_refuses_more_than_its_scratch(ctx):
    F = 200
    ens = ctx.ensemble(_model(8, 4, F, 1, seed=7), F)
    n = (64 << 20) // (4 * F + 4) + 64           # just past the 64 MiB scratch
    X = np.zeros((n, F), np.float32)
    rc, _ = _noalloc(ens, X, 0)
    assert rc == -1 and b"scratch" in ctx.L.tc_last_error()
    rc, out = _noalloc(ens, X[:1000], 0)          # and it still works afterwards
    assert rc == 0 and np.array_equal(out, ens.predict(X[:1000]))
