"""GPU parity of the hot-path primitives against the CPU oracle, through the
C-ABI: quantization (bit-exact), gradients, histograms, sibling subtraction
and the level split scan."""
import numpy as np
import pytest

from conftest import binary_data, regression_data, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative, normalized by the sum of |terms|


def test_quantize_bitexact(ctx, oracle):
    X, _ = regression_data(20011, 13, seed=3, low_card_col=5)
    X[::97, 2] = np.nan                      # NaN -> bin 255
    X[:, 7] = X[0, 7]                        # a constant column: 1 unique value
    X[:300, 8] = np.arange(300)              # 300 unique + the rest ...
    X[300:, 8] = 299.0                       # ... -> U = 300 -> wrap to 44 bins
    ds, edges, ne = ctx.quantize_fit(X)
    ob, oe, on = oracle.quantize(X)
    assert np.array_equal(ne, on)
    for f in range(X.shape[1]):
        assert np.array_equal(edges[f, :ne[f]], oe[f, :on[f]]), f
    assert ne[8] == 43                       # uint8 wrap: min(255, 300 % 256) - 1 edges
    assert np.array_equal(ds.bins(), ob)
    # A3: binning new rows with the training edges
    Xn, _ = regression_data(3001, 13, seed=4)
    ds2 = ctx.dataset(Xn, edges, ne)
    assert np.array_equal(ds2.bins(), oracle.apply_bins(Xn, oe, on))


def test_quantize_edge_cases(ctx, oracle):
    """Columns at the rule's corners, quantized on the device and by the
    oracle: no value at all, U = 255, 256, 257 and 511 under the uint8 wrap
    (A = 255, 0, 1, 255), +-inf, both signed zeros (equal under the reference's
    comparisons: edges compare equal, bins bit-exact), huge magnitudes, heavy
    duplicates, a single non-NaN value, and rows wide enough (F = 700) to
    stream through the pinned staging in several blocks."""
    rng = np.random.default_rng(21)
    n, F = 14003, 700
    X = rng.standard_normal((n, F)).astype(np.float32)
    X[:, 0] = np.nan
    for c, U in ((1, 255), (2, 256), (3, 257), (4, 511)):
        X[:, c] = rng.permutation(np.arange(n) % U).astype(np.float32)
    X[::3, 5] = np.inf
    X[1::3, 5] = -np.inf
    X[:, 6] = np.where(rng.random(n) < 0.5, -0.0, 0.0).astype(np.float32)
    X[::7, 6] = 1.5
    X[1::7, 6] = -1.0                                # U = 3: the zero is an edge
    X[:, 7] = (rng.standard_normal(n) * 1e37).astype(np.float32)
    X[:, 8] = np.round(rng.standard_normal(n) * 3).astype(np.float32)
    X[:, 9] = np.nan
    X[17, 9] = -2.5
    X[::11, 10] = np.nan
    for wrap in (True, False):
        for mb in (255, 16):
            ds, edges, ne = ctx.quantize_fit(X, max_bins=mb, wrap=wrap)
            ob, oe, on = oracle.quantize(X, max_bins=mb, wrap=wrap)
            assert np.array_equal(ne, on), (wrap, mb)
            for f in range(F):
                assert np.array_equal(edges[f, :ne[f]], oe[f, :on[f]]), (wrap, mb, f)
            assert np.array_equal(ds.bins(), ob), (wrap, mb)
            ds.close()
    assert on[0] == 1 and on[9] == 0                 # no value: one 0.0 edge; one value: none
    assert on[6] == 2 and oe[6, 0] == 0.0


def test_quantize_fixed_mode(ctx, oracle):
    X, _ = regression_data(5000, 4, seed=5)
    ds, edges, ne = ctx.quantize_fit(X, wrap=False)
    ob, oe, on = oracle.quantize(X, wrap=False)
    assert np.array_equal(ne, on) and (ne == 254).all()
    assert np.array_equal(ds.bins(), ob)


@pytest.mark.parametrize("loss", [0, 1])
def test_gradients(ctx, oracle, loss):
    rng = np.random.default_rng(11)
    n = 10007
    y = (rng.random(n) < 0.4).astype(np.float32) if loss else rng.standard_normal(n).astype(np.float32)
    p = (rng.standard_normal(n) * 4).astype(np.float32)
    p[:5] = [0.0, 15.5, -16.0, 1e-3, -88.0]
    g, h = ctx.gradients(loss, y, p)
    og, oh = oracle.gradients_f32(loss, y, p)
    # the reference's vectorized rows (same float operations) and its scalar
    # tail rows (glibc's expf, restated on the device): identical bits
    assert np.array_equal(g, og) and np.array_equal(h, oh)
    g64, h64 = oracle.gradients_f64(loss, y, p)
    # tolerance scale = sum of |terms|: g = p - y (2(p - y) for MSE), h = p(1 - p)
    if loss == 0:
        assert rel_err(g, g64, 2 * (np.abs(p) + np.abs(y))).max() < TOL
        assert (h == 2.0).all()
    else:
        prob = g64 + y
        assert rel_err(g, g64, np.abs(prob) + np.abs(y)).max() < TOL
        assert rel_err(h, h64, 1.0).max() < TOL


def test_gradients_scalar_tail(ctx, oracle):
    """The rows past the last full 8-row block take the reference's scalar
    sigmoid, 1 / (1 + exp(-x)) with glibc's expf: calls of 7 rows (all of them
    tail rows) over the clamped range [-15, 15] and its ends, bit for bit
    against the oracle (glibc on the host)."""
    rng = np.random.default_rng(12)
    p = np.concatenate([rng.uniform(-16, 16, 7 * 300), np.linspace(-15, 15, 7 * 60),
                        [-15.0, 15.0, -14.999999, 14.999999, 0.0, -0.0, 1e-30]]).astype(np.float32)
    p = p[: len(p) // 7 * 7]
    y = (rng.random(len(p)) < 0.5).astype(np.float32)
    for i in range(0, len(p), 7):
        g, h = ctx.gradients(1, y[i:i + 7], p[i:i + 7])
        og, oh = oracle.gradients_f32(1, y[i:i + 7], p[i:i + 7])
        assert np.array_equal(g, og) and np.array_equal(h, oh), p[i:i + 7]


@pytest.mark.parametrize("n,F,seed", [(1000, 5, 1), (50000, 20, 2), (70001, 37, 3)])
def test_hist_all_rows(ctx, oracle, n, F, seed):
    X, y = binary_data(n, F, seed)
    ds, edges, ne = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(n).astype(np.float32)
    h = (rng.random(n) * 0.25 + 1e-6).astype(np.float32)
    gh = ds.hist(g, h)
    o64, absg, cnt = oracle.hist_f64(bins, g, h)
    assert rel_err(gh[..., 0], o64[..., 0], absg).max() < TOL
    assert rel_err(gh[..., 1], o64[..., 1], o64[..., 1]).max() < TOL
    # empty bins stay exactly zero; populated bins carry every row
    assert np.array_equal(gh[..., 1] > 0, cnt > 0)
    # sum over bins equals node totals, per feature
    assert np.allclose(gh[..., 1].sum(1, dtype=np.float64), h.sum(dtype=np.float64), rtol=1e-5)


@pytest.mark.parametrize("with_rows", [False, True])
def test_hist_count_plane(ctx, oracle, with_rows):
    """tc_hist_build's per-bin row counts (GradientPair::count, the field the
    reference's SplitFinder reads) equal the oracle's counts exactly, NaN rows
    in slot 255 included."""
    n, F = 30011, 19
    X, _ = binary_data(n, F, 4)
    X[::9, 3] = np.nan
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(5)
    g = rng.standard_normal(n).astype(np.float32)
    h = (rng.random(n) * 0.25 + 1e-6).astype(np.float32)
    rows = np.sort(rng.choice(n, 7777, replace=False)).astype(np.uint32) if with_rows else None
    gh, cnt = ds.hist(g, h, rows, counts=True)
    o64, absg, ocnt = oracle.hist_f64(bins, g, h, rows)
    assert np.array_equal(cnt, ocnt)
    assert (cnt.sum(1) == (n if rows is None else len(rows))).all()
    assert cnt[3, 255] == (len(range(0, n, 9)) if rows is None else np.isnan(X[rows, 3]).sum())
    assert rel_err(gh[..., 0], o64[..., 0], absg).max() < TOL
    assert np.array_equal(ds.hist(g, h, rows), gh)


def test_hist_row_subset_and_subtraction(ctx, oracle):
    n, F = 40000, 11
    X, _ = regression_data(n, F, 7)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(8)
    g = rng.standard_normal(n).astype(np.float32)
    h = np.full(n, 2.0, np.float32)
    left = np.sort(rng.choice(n, 13000, replace=False)).astype(np.uint32)
    mask = np.ones(n, bool); mask[left] = False
    right = np.nonzero(mask)[0].astype(np.uint32)
    parent = ds.hist(g, h)
    hl = ds.hist(g, h, left)
    o64, absg, cnt = oracle.hist_f64(bins, g, h, left)
    assert rel_err(hl[..., 0], o64[..., 0], absg).max() < TOL
    assert np.array_equal(hl[..., 1], 2.0 * cnt)   # hess 2 per row: counts exactly
    hr = ctx.hist_subtract(parent, hl)
    r64, rabs, rcnt = oracle.hist_f64(bins, g, h, right)
    # larger = parent - smaller matches a direct build of the larger child
    assert np.abs(hr[..., 0] - r64[..., 0]).max() < 1e-4 * max(1.0, np.abs(g).sum() / F / 64)
    assert np.array_equal(hr[..., 1], 2.0 * rcnt)
    # empty row list: all zero
    assert not ds.hist(g, h, np.zeros(0, np.uint32)).any()


def test_split_scan_bitexact_root_and_levels(ctx, oracle):
    n, F = 30000, 9
    X, y = regression_data(n, F, 21, low_card_col=4)
    ds, _, _ = ctx.quantize_fit(X)
    bins = ds.bins()
    rng = np.random.default_rng(22)
    g = rng.standard_normal(n).astype(np.float32)
    h = (rng.random(n) + 0.5).astype(np.float32)
    # root: feed the same float histogram to both -> identical (f, b, gain)
    hist, _ = oracle.hist_f32(bins, g, h)
    G, H = np.float32(g.sum(dtype=np.float64)), np.float32(h.sum(dtype=np.float64))
    for lam in (0.0, 1.0, 3.0):
        got = ctx.split_scan(hist[None], [G], [H], [n], lambda_l2=lam)
        exp = oracle.split_f32(hist[None], [G], [H], [n], lam=lam, is_root=True)
        assert got[:2] == exp[:2] and np.float32(got[2]) == np.float32(exp[2]), (lam, got, exp)
    # a deeper level: 8 nodes including an empty one and a one-row node
    node = rng.integers(0, 8, n)
    node[node == 3] = 2
    node[0] = 3
    hs, Gs, Hs, Cs = [], [], [], []
    for k in range(8):
        rows = np.nonzero(node == k)[0].astype(np.uint32)
        hk, _ = oracle.hist_f32(bins, g, h, rows)
        hs.append(hk)
        Gs.append(np.float32(g[rows].sum(dtype=np.float64)))
        Hs.append(np.float32(h[rows].sum(dtype=np.float64)))
        Cs.append(len(rows))
    hs = np.stack(hs)
    for lam in (0.0, 1.0):
        got = ctx.split_scan(hs, Gs, Hs, Cs, lambda_l2=lam)
        exp = oracle.split_f32(hs, Gs, Hs, Cs, lam=lam, is_root=False)
        assert got[:2] == exp[:2] and np.float32(got[2]) == np.float32(exp[2]), (lam, got, exp)


def test_split_scan_ties_and_zero_gain(ctx, oracle):
    # all candidates invalid at a deeper level -> total 0.0 at (0, 0)
    F = 3
    hist = np.zeros((2, F, 256, 2), np.float32)
    hist[0, :, 10, :] = [1.0, 1.0]            # node 0: one bin only
    got = ctx.split_scan(hist, [1.0, 0.0], [1.0, 0.0], [5, 0], lambda_l2=1.0)
    assert got == (0, 0, 0.0)
    # exact tie between two features -> lowest feature wins
    hist = np.zeros((1, 2, 256, 2), np.float32)
    for f in range(2):
        hist[0, f, 3] = [-1.0, 1.0]
        hist[0, f, 9] = [1.0, 1.0]
    got = ctx.split_scan(hist, [0.0], [2.0], [2], lambda_l2=0.0)
    exp = oracle.split_f32(hist, [0.0], [2.0], [2], lam=0.0, is_root=True)
    assert got[:2] == (0, 3) == exp[:2]
