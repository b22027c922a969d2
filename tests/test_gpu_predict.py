"""GPU ensemble inference (A15-A17) and the continuous-batching server.

Bit-exact claims: per-row, per-tree leaf indices; raw scores equal the
oracle's FP32 restatement of the reference's summation order bit for bit;
every row's result is identical alone and inside a coalesced batch.  Float
claim: raw scores within 1e-5 of the FP64 sum, normalized by sum |w * leaf|."""
import threading

import numpy as np
import pytest

from conftest import regression_data, rel_err

pytestmark = pytest.mark.gpu


def random_model(T, D, F, seed, shallow_every=7):
    rng = np.random.default_rng(seed)
    depth = np.full(T, D, np.uint16)
    depth[::shallow_every] = rng.integers(0, D, len(depth[::shallow_every]))
    feat = rng.integers(0, F, T * D).astype(np.uint16)
    thr = rng.standard_normal(T * D).astype(np.float32)
    leaves = (rng.standard_normal(T << D) * 0.1).astype(np.float32)
    for t in range(T):   # unused slots stay zero, as the trainer leaves them
        leaves[(t << D) + (1 << int(depth[t])):(t + 1) << D] = 0
        feat[t * D + int(depth[t]):(t + 1) * D] = 0
        thr[t * D + int(depth[t]):(t + 1) * D] = 0
    return dict(T=T, D=D, feature=feat, fthr=thr, bin=np.zeros(T * D, np.uint8), depth=depth,
                leaves=leaves, weight=np.full(T, 0.1, np.float32), base=0.25)


@pytest.mark.parametrize("T,D,F,n", [(1, 1, 3, 5), (7, 4, 11, 1000), (100, 6, 20, 50000),
                                     (1003, 8, 200, 3001)])
def test_predict_bitexact(ctx, oracle, T, D, F, n):
    m = random_model(T, D, F, seed=T + n)
    X, _ = regression_data(n, F, seed=T)
    X[::53, F // 2] = np.nan
    ens = ctx.ensemble(m, F)
    out, li = ens.predict(X, want_leaf_idx=True)
    ref, rli = oracle.predict_f32(m, X, want_leaf_idx=True)
    assert np.array_equal(li, rli)
    assert np.array_equal(out, ref)
    r64, ab = oracle.predict_f64(m, X)
    assert rel_err(out, r64, ab).max() < 1e-5
    p = ens.predict(X, transform=1)
    assert np.allclose(p, 1.0 / (1.0 + np.exp(-r64)), rtol=1e-6, atol=1e-7)
    assert np.array_equal(p, oracle.transform_f32(ref))  # the reference's sigmoid, bit for bit


@pytest.mark.parametrize("n", [1, 7, 8, 13, 16, 23, 517, 3001])
def test_predict_tail_rows_bitexact(ctx, oracle, n):
    """The rows past a call's last full 8-row block take the reference's
    tree-by-tree order (fast_float_ensemble.hpp:543-560): bit for bit against
    the oracle, which is pinned to the compiled reference on these rows."""
    from turbocat.capi import train_params
    X, y = regression_data(6000, 7, 44)
    ds, _, _ = ctx.quantize_fit(X)
    m = ds.train(y, train_params(0, n_trees=23, learning_rate=0.1, max_depth=5, lambda_l2=0.0))
    m["weight"] = np.full(23, 0.1, np.float32)
    ens = ctx.ensemble(m, 7)
    Xp, _ = regression_data(n, 7, 45)
    Xp[::3, 2] = np.nan
    assert np.array_equal(ens.predict(Xp), oracle.predict_f32(m, Xp))
    mr = random_model(101, 6, 7, seed=n)
    assert np.array_equal(ctx.ensemble(mr, 7).predict(Xp), oracle.predict_f32(mr, Xp))


def test_predict_matches_reference_golden(ctx):
    """The reference's own raw-float and binned predictions of a reference-
    trained model (tests/golden/pred_tail.npz), in calls of 517, 13 and 7
    rows with NaN rows, reproduced bit for bit on the GPU."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "pred_tail.npz"),
                allow_pickle=False)
    n, F, seed, T, D = (int(v) for v in z["meta"])
    m = dict(T=T, D=D, feature=z["m_feature"], bin=z["m_bin"], fthr=z["m_fthr"],
             depth=z["m_depth"], leaves=z["m_leaves"], weight=np.full(T, z["lr"], np.float32),
             base=float(z["m_base"]))
    ens = ctx.ensemble(m, F)
    for r in (517, 13, 7):
        assert np.array_equal(ens.predict(z[f"X{r}"]), z[f"pred{r}"]), r
        ds, _, _ = ctx.quantize_fit(z[f"X{r}"])      # the GPU quantizer: the reference's bins
        assert np.array_equal(ds.bins(), z[f"bins{r}"]), r
        assert np.array_equal(ens.predict_binned(ds), z[f"pbin{r}"]), r


@pytest.mark.parametrize("n", [9, 15, 16, 517, 20000])
def test_predict_binned_matches_oracle(ctx, oracle, n):
    """A17: tc_predict_binned against the oracle's restatement of
    SymmetricEnsemble::predict_batch (FastEnsemble column-major blocks; the
    tree-by-tree fallback under 16 rows; tail rows), NaN rows in bin 255."""
    from turbocat.capi import train_params
    X, y = regression_data(20000, 9, 5)
    X[::17, 4] = np.nan
    ds, edges, ne = ctx.quantize_fit(X)
    m = ds.train(y, train_params(0, n_trees=30, learning_rate=0.1, max_depth=6, lambda_l2=0.0))
    m["weight"] = np.full(30, 0.1, np.float32)
    ens = ctx.ensemble(m, 9)
    Xp = X[:n]
    dp = ctx.dataset(Xp, edges, ne)
    bins = dp.bins()
    assert (bins[4, ::17] == 255).all()
    assert np.array_equal(ens.predict_binned(dp), oracle.predict_binned_f32(m, bins))


def test_predict_binned_matches_float(ctx, oracle):
    """A.3: with thresholds taken from the training edges, bin > t <=> v >= e[t],
    so binned and raw-float predictions agree for non-NaN rows."""
    from turbocat.capi import train_params
    X, y = regression_data(20000, 9, 5)
    ds, edges, ne = ctx.quantize_fit(X)
    m = ds.train(y, train_params(0, n_trees=30, learning_rate=0.1, max_depth=6, lambda_l2=0.0))
    ens = ctx.ensemble(m, 9)
    a = ens.predict(X)
    b = ens.predict_binned(ds)
    assert np.array_equal(a, b)


def test_predict_device_pointers(ctx, oracle):
    import ctypes
    from turbocat.capi import check
    m = random_model(50, 6, 16, seed=3)
    X, _ = regression_data(4096, 16, seed=3)
    ens = ctx.ensemble(m, 16)
    L = ctx.L
    dx, do = ctypes.c_void_p(), ctypes.c_void_p()
    check(L.tc_device_alloc(ctx.h, X.nbytes, ctypes.byref(dx)))
    check(L.tc_device_alloc(ctx.h, 4 * 4096, ctypes.byref(do)))
    check(L.tc_memcpy_h2d(ctx.h, dx, X.ctypes.data_as(ctypes.c_void_p), X.nbytes))
    ens.predict_dev(dx, 4096, do)
    out = np.zeros(4096, np.float32)
    check(L.tc_memcpy_d2h(ctx.h, out.ctypes.data_as(ctypes.c_void_p), do, 4 * 4096))
    check(L.tc_device_free(ctx.h, dx)); check(L.tc_device_free(ctx.h, do))
    assert np.array_equal(out, oracle.predict_f32(m, X))


def test_serving_alone_and_in_the_mix(ctx, oracle):
    """Requests of uneven sizes submitted from several threads are coalesced
    into shared batches (the batch count is below the request count); each
    result is bitwise the same as the request predicted alone."""
    from turbocat.capi import Server
    F = 24
    m = random_model(64, 6, F, seed=9)
    ens = ctx.ensemble(m, F)
    rng = np.random.default_rng(10)
    sizes = [int(s) for s in np.exp(rng.uniform(0, np.log(3000), 120))]
    reqs = [regression_data(s, F, seed=100 + i)[0] for i, s in enumerate(sizes)]
    alone = [ens.predict(x) for x in reqs]
    srv = Server(ens, max_batch_rows=2048, memory_budget=4 << 20, n_streams=3)
    outs = [np.zeros(len(x), np.float32) for x in reqs]
    tickets = [None] * len(reqs)

    def client(k):
        for i in range(k, len(reqs), 4):
            tickets[i] = srv.submit(reqs[i], outs[i])

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    [t.start() for t in th]
    [t.join() for t in th]
    for t in tickets:
        assert srv.wait(t, timeout_us=60_000_000)
    batches, rows = srv.stats()
    srv.stop()
    assert rows == sum(sizes)
    assert batches < len(reqs)
    for a, b in zip(outs, alone):
        assert np.array_equal(a, b)


def test_library_calls_while_a_server_runs(ctx, oracle):
    """hipFree waits for the whole device to go idle, which a server's
    resident kernel never lets it be: the library's frees are deferred while
    a server runs.  Plain predictions whose cached buffers grow (and free the
    old ones), an ensemble and a dataset created and destroyed, and device
    memory allocated and freed, all with a server resident; then the server
    still answers, bit-identically."""
    import ctypes
    from turbocat import capi
    F = 30
    m = random_model(20, 6, F, seed=31)
    ens = ctx.ensemble(m, F)
    srv = capi.Server(ens, max_batch_rows=4096, memory_budget=1 << 24, n_streams=2)
    try:
        for n in (1, 300, 5000, 70000):                   # each size grows the cached buffers
            X, _ = regression_data(n, F, seed=n)
            assert np.array_equal(ens.predict(X), oracle.predict_f32(m, X))
        m2 = random_model(9, 4, F, seed=32)
        e2 = ctx.ensemble(m2, F)
        X, _ = regression_data(999, F, seed=5)
        assert np.array_equal(e2.predict(X), oracle.predict_f32(m2, X))
        e2.close()
        ds, _, _ = ctx.quantize_fit(X)
        ds.close()
        p = ctypes.c_void_p()
        capi.check(ctx.L.tc_device_alloc(ctx.h, 1 << 20, ctypes.byref(p)))
        capi.check(ctx.L.tc_device_free(ctx.h, p))
        X, _ = regression_data(2500, F, seed=6)
        out = np.zeros(len(X), np.float32)
        assert srv.wait(srv.submit(X, out), timeout_us=60_000_000)
        assert np.array_equal(out, ens.predict(X))
    finally:
        srv.stop()
    X, _ = regression_data(100_000, F, seed=7)              # frees work again after the stop
    assert np.array_equal(ens.predict(X), oracle.predict_f32(m, X))


def test_handles_may_be_destroyed_in_any_order(oracle):
    """Every object holds what it uses: a context destroyed before its
    ensemble and dataset, and an ensemble destroyed under a running server,
    stay usable by their holders until those are destroyed in turn."""
    from turbocat import capi
    F = 12
    c = capi.Context(0)
    m = random_model(15, 5, F, seed=41)
    e = c.ensemble(m, F)
    X, _ = regression_data(3000, F, seed=42)
    ds, _, _ = c.quantize_fit(X)
    bins = ds.bins()
    srv = capi.Server(e, max_batch_rows=1024, memory_budget=1 << 22, n_streams=2)
    try:
        c.close()                                       # the ensemble and dataset hold it
        assert np.array_equal(e.predict(X), oracle.predict_f32(m, X))
        assert np.array_equal(ds.bins(), bins)
        ds.close()
        e_ref = e.predict(X)
        e.close()                                       # the server holds the ensemble
        out = np.zeros(len(X), np.float32)
        assert srv.wait(srv.submit(X, out), timeout_us=60_000_000)
        assert np.array_equal(out, e_ref)
    finally:
        srv.stop()                                      # the last holder: all released


def test_serving_c3_shape_alone_and_in_the_mix(ctx):
    """The serving workload's shape: 1000 trees of depth 8 over 200 features,
    requests of 1..65536 rows (log-uniform) submitted from four threads into
    batches of 32768 rows, half of them in device-reachable pinned memory
    (read and written in place), half in ordinary memory (staged).  Every
    result equals the request predicted alone, bit for bit; admitted bytes
    stay within the budget unless one request alone exceeds it."""
    from turbocat.capi import HostBuffer, Server
    F, T, D = 200, 1000, 8
    m = random_model(T, D, F, seed=31)
    ens = ctx.ensemble(m, F)
    rng = np.random.default_rng(32)
    sizes = [int(s) for s in np.floor(np.exp(rng.uniform(0, np.log(65536), 48)))]
    sizes[:3] = [65536, 1, 7]
    bufs, reqs, outs = [], [], []
    for i, n in enumerate(sizes):
        x = rng.standard_normal((n, F), dtype=np.float32)
        x[::97, 3] = np.nan
        if i % 2 == 0:
            hb, ho = HostBuffer((n, F)), HostBuffer((n,))
            hb.array[:] = x
            bufs += [hb, ho]
            reqs.append(hb.array)
            outs.append(ho.array)
        else:
            reqs.append(x)
            outs.append(np.zeros(n, np.float32))
    alone = [ens.predict(x, transform=1) for x in reqs]
    budget = 64 << 20
    srv = Server(ens, max_batch_rows=32768, memory_budget=budget, n_streams=4, transform=1)
    try:
        tickets = [None] * len(reqs)

        def client(k):
            for i in range(k, len(reqs), 4):
                tickets[i] = srv.submit(reqs[i], outs[i])

        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        [t.start() for t in th]
        [t.join() for t in th]
        for t in tickets:
            assert srv.wait(t, timeout_us=120_000_000)
        c = srv.counters()
        lat = srv.latencies_us()
    finally:
        srv.stop()
    assert len(lat) == len(reqs) and (lat > 0).all()
    row = 4 * F + 4
    assert c["rows"] == sum(sizes) and c["batches"] < len(reqs)
    assert c["peak_admitted_bytes"] <= max(budget, max(sizes) * row)
    assert c["max_batches_in_flight"] >= 2 and c["failed"] == 0
    for a, b in zip(outs, alone):
        assert np.array_equal(a, b)
    for b in bufs:
        b.free()


def test_serving_multiclass(ctx, oracle):
    """K-channel ensembles are served too: softmax outputs (n x K) of every
    request equal the request predicted alone."""
    from turbocat.capi import Server
    rng = np.random.default_rng(41)
    T, D, K, F = 60, 6, 4, 17
    depth = np.full(T, D, np.uint16)
    m = dict(T=T, D=D, K=K, feature=rng.integers(0, F, T * D).astype(np.uint16),
             fthr=rng.standard_normal(T * D).astype(np.float32), bin=np.zeros(T * D, np.uint8),
             depth=depth, leaves=(rng.standard_normal((T << D) * K) * 0.1).astype(np.float32),
             weight=np.full(T, 0.05, np.float32), base=0.0)
    ens = ctx.ensemble(m, F)
    reqs = [rng.standard_normal((int(s), F), dtype=np.float32)
            for s in np.exp(rng.uniform(0, np.log(2000), 40))]
    alone = [ens.predict(x, transform=1) for x in reqs]
    srv = Server(ens, max_batch_rows=1024, memory_budget=1 << 24, n_streams=3, transform=1)
    try:
        outs = [np.zeros((len(x), K), np.float32) for x in reqs]
        ts = [srv.submit(x, o) for x, o in zip(reqs, outs)]
        for t in ts:
            assert srv.wait(t, timeout_us=60_000_000)
    finally:
        srv.stop()
    for a, b in zip(outs, alone):
        assert np.array_equal(a, b)
    ref, _ = oracle.predict_mc_f64(m, reqs[0])
    e = np.exp(ref - ref.max(axis=1, keepdims=True))
    assert np.allclose(outs[0], e / e.sum(axis=1, keepdims=True), rtol=1e-5, atol=1e-6)


def test_serving_oversize_request_and_injected_failure(ctx):
    """A request larger than the memory budget is admitted alone and served;
    a batch the host fails on purpose (fail_batch) ends the waits of its
    requests with an error, later submissions are refused, stop returns."""
    from turbocat.capi import Server, TurboCatError
    F = 32
    m = random_model(40, 6, F, seed=51)
    ens = ctx.ensemble(m, F)
    rng = np.random.default_rng(52)
    big = rng.standard_normal((30000, F), dtype=np.float32)
    small = rng.standard_normal((100, F), dtype=np.float32)
    srv = Server(ens, max_batch_rows=8192, memory_budget=1 << 20, n_streams=2)
    try:
        ob, os_ = np.zeros(30000, np.float32), np.zeros(100, np.float32)
        t1, t2 = srv.submit(big, ob), srv.submit(small, os_)
        assert srv.wait(t1, 60_000_000) and srv.wait(t2, 60_000_000)
        c = srv.counters()
    finally:
        srv.stop()
    assert c["peak_admitted_bytes"] == 30000 * (4 * F + 4)
    assert np.array_equal(ob, ens.predict(big)) and np.array_equal(os_, ens.predict(small))
    reqs = [rng.standard_normal((500, F), dtype=np.float32) for _ in range(10)]
    srv = Server(ens, max_batch_rows=500, memory_budget=1 << 24, n_streams=2, fail_batch=3)
    try:
        outs = [np.zeros(500, np.float32) for _ in reqs]
        ts, refused = [], 0
        for x, o in zip(reqs, outs):
            try:
                ts.append(srv.submit(x, o))
            except TurboCatError as e:
                assert e.code == -4
                refused += 1
        failed = 0
        for i, t in enumerate(ts):
            try:
                srv.wait(t, 60_000_000)
                assert np.array_equal(outs[i], ens.predict(reqs[i]))
            except TurboCatError as e:
                assert e.code == -3 and "injected failure" in str(e)
                failed += 1
        assert failed >= 1
    finally:
        srv.stop()
