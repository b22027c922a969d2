"""The serving core's host logic on the CPU (turbo-cat_amd/csrc/tc_serve_core.h),
over a fake device backend (tests/serve_fake_backend.cpp, built here with
g++): continuous batching of uneven requests, each result identical alone and
in the mix, the reference's tree-by-tree rows flagged per request, memory-
budget admission (a request larger than the budget admitted alone, the budget
never exceeded otherwise), cancellation, timeouts, and failures -- an injected
batch failure and a device that dies -- ending every wait with an error and
letting stop return."""
import ctypes
import os
import subprocess
import threading
import time

import numpy as np
import pytest

from conftest import ROOT

OK, EINVAL, EDEVICE, ESTATE, EBUSY, ETIMEOUT = 0, -1, -3, -4, -5, -6


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("fake") / "libfake_serve.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-shared", "-fPIC", "-pthread",
                    "-I", os.path.join(ROOT, "turbo-cat_amd", "csrc"),
                    os.path.join(ROOT, "tests", "serve_fake_backend.cpp"), "-o", str(out)],
                   check=True)
    L = ctypes.CDLL(str(out))
    V, U, U64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    L.fake_start.restype = V
    L.fake_start.argtypes = [U, U, U, U64, U, U64, U64, ctypes.c_int]
    L.fake_hold.argtypes = [V, ctypes.c_int]
    L.fake_published.argtypes = [V]
    L.fake_published.restype = U64
    for fn, args in (("fake_submit", [V, V, U, V, V]), ("fake_wait", [V, U64, ctypes.c_int64]),
                     ("fake_cancel", [V, U64])):
        getattr(L, fn).argtypes = args
        getattr(L, fn).restype = ctypes.c_int32
    L.fake_counters.argtypes = [V, V]
    L.fake_latencies.argtypes = [V, V, U64]
    L.fake_latencies.restype = U64
    L.fake_error.argtypes = [V]
    L.fake_error.restype = ctypes.c_char_p
    L.fake_stop.argtypes = [V]
    return L


def expected(X, K, n):
    """The fake device's output of each row of a request of n rows alone."""
    F = X.shape[1]
    v = (X.astype(np.float32) * np.arange(1, F + 1, dtype=np.float32)).astype(np.float32)
    s = np.zeros(X.shape[0], np.float32)
    for f in range(F):                      # float sum in feature order, as the fake device adds
        s = (s + v[:, f]).astype(np.float32)
    seq = (np.arange(n) >= n // 8 * 8).astype(np.float32) * 0.5
    return (s[:, None] + np.arange(K, dtype=np.float32)[None, :] + seq[:, None]).astype(np.float32)


class Fake:
    def __init__(self, L, F=5, K=1, max_batch_rows=100, budget=1 << 20, n_slots=3, fail_batch=0,
                 device_fail_after=0, hold=False):
        self.L, self.F, self.K = L, F, K
        self.h = L.fake_start(F, K, max_batch_rows, budget, n_slots, fail_batch, device_fail_after,
                              int(hold))
        self.keep = {}

    def submit(self, X, out):
        t = ctypes.c_uint64()
        rc = self.L.fake_submit(self.h, X.ctypes.data, X.shape[0], out.ctypes.data, ctypes.byref(t))
        self.keep[t.value] = (X, out)
        return rc, t.value

    def wait(self, t, timeout_us=-1):
        return self.L.fake_wait(self.h, t, timeout_us)

    def counters(self):
        c = (ctypes.c_uint64 * 6)()
        self.L.fake_counters(self.h, c)
        return dict(zip(("batches", "rows", "peak", "max_inflight", "cancelled", "failed"), c))

    def latencies(self):
        n = self.L.fake_latencies(self.h, None, 0)
        out = np.zeros(n, np.float32)
        self.L.fake_latencies(self.h, out.ctypes.data, n)
        return out

    def stop(self):
        self.L.fake_stop(self.h)
        self.h = None


def requests(sizes, F, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((s, F)).astype(np.float32) for s in sizes]


@pytest.mark.parametrize("K", [1, 3])
def test_alone_and_in_the_mix_from_several_threads(lib, K):
    rng = np.random.default_rng(1)
    sizes = [int(s) for s in np.exp(rng.uniform(0, np.log(700), 150))]
    reqs = requests(sizes, 5, 2)
    srv = Fake(lib, F=5, K=K, max_batch_rows=256, budget=64 << 10, n_slots=3)
    outs = [np.zeros((len(x), K), np.float32) for x in reqs]
    tickets = [None] * len(reqs)

    def client(k):
        for i in range(k, len(reqs), 4):
            rc, tickets[i] = srv.submit(reqs[i], outs[i])
            assert rc == OK

    t0 = time.perf_counter()
    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    [t.start() for t in th]
    [t.join() for t in th]
    for t in tickets:
        assert srv.wait(t, 30_000_000) == OK
    wall_us = (time.perf_counter() - t0) * 1e6
    c = srv.counters()
    lat = srv.latencies()
    srv.stop()
    assert len(lat) == len(reqs)                                    # one per finished request
    assert (lat > 0).all() and lat.max() <= wall_us
    assert c["rows"] == sum(sizes) and c["batches"] < len(reqs)     # coalesced
    assert c["peak"] <= 64 << 10                                    # budget held
    for x, o in zip(reqs, outs):
        assert np.array_equal(o, expected(x, K, len(x)))           # incl. tree-by-tree rows


def test_request_larger_than_the_budget_is_admitted_alone(lib):
    F = 4
    row = 4 * F + 4
    srv = Fake(lib, F=F, max_batch_rows=64, budget=50 * row, n_slots=2)
    big, small = requests([500, 30], F, 3)
    ob, os_ = np.zeros((500, 1), np.float32), np.zeros((30, 1), np.float32)
    rc1, t1 = srv.submit(big, ob)
    rc2, t2 = srv.submit(small, os_)
    assert rc1 == rc2 == OK
    assert srv.wait(t1, 30_000_000) == OK and srv.wait(t2, 30_000_000) == OK
    c = srv.counters()
    srv.stop()
    assert c["peak"] == 500 * row          # admitted alone: never together with the small one
    assert np.array_equal(ob, expected(big, 1, 500)) and np.array_equal(os_, expected(small, 1, 30))


def test_cancel_and_timeout(lib):
    F = 3
    srv = Fake(lib, F=F, max_batch_rows=10, budget=10 * (4 * F + 4), n_slots=1, hold=True)
    a, b, c = requests([10, 10, 10], F, 4)
    oa, ob, oc = (np.zeros((10, 1), np.float32) for _ in range(3))
    _, ta = srv.submit(a, oa)
    _, tb = srv.submit(b, ob)
    _, tc = srv.submit(c, oc)
    for _ in range(200):                    # a is packed into the one slot and published
        if lib.fake_published(srv.h) == 1:
            break
        threading.Event().wait(0.005)
    assert lib.fake_published(srv.h) == 1
    assert lib.fake_cancel(srv.h, ta) == EBUSY          # on the device
    assert lib.fake_cancel(srv.h, tc) == OK             # still waiting for admission
    assert srv.wait(tc) == ESTATE and b"cancelled" in lib.fake_error(srv.h)
    assert srv.wait(tb, 20_000) == ETIMEOUT             # held: b cannot finish
    lib.fake_hold(srv.h, 0)
    assert srv.wait(ta, 30_000_000) == OK and srv.wait(tb, 30_000_000) == OK
    assert lib.fake_cancel(srv.h, ta) == EINVAL         # finished and collected
    cnt = srv.counters()
    n_lat = len(srv.latencies())
    srv.stop()
    assert n_lat == 2                                    # the cancelled request has none
    assert cnt["cancelled"] == 1 and not oc.any()
    assert np.array_equal(ob, expected(b, 1, 10))


def test_injected_batch_failure_fails_waits_and_stop_returns(lib):
    F = 4
    reqs = requests([40] * 12, F, 5)
    srv = Fake(lib, F=F, max_batch_rows=40, budget=1 << 20, n_slots=2, fail_batch=4)
    outs = [np.zeros((40, 1), np.float32) for _ in reqs]
    sub = [srv.submit(x, o) for x, o in zip(reqs, outs)]
    assert all(rc == OK for rc, _ in sub[:4])
    # submissions after the failure are refused: the server stopped itself
    assert all(rc in (OK, ESTATE) for rc, _ in sub)
    rcs = [srv.wait(t, 30_000_000) if rc == OK else None for rc, t in sub]
    assert rcs[3] == EDEVICE                             # batch 3 is request 3
    for rc, x, o in zip(rcs, reqs, outs):
        assert rc in (OK, EDEVICE, None)
        if rc == OK:
            assert np.array_equal(o, expected(x, 1, 40))
    assert srv.submit(reqs[0], outs[0])[0] == ESTATE    # the server stopped itself
    srv.stop()                                          # and stop returns


def test_dead_device_fails_every_unfinished_request(lib):
    F = 2
    reqs = requests([25] * 10, F, 6)
    # the device is held until every request is in: once it fails the server
    # stops itself and refuses later submissions
    srv = Fake(lib, F=F, max_batch_rows=25, budget=1 << 20, n_slots=3, device_fail_after=3,
               hold=True)
    outs = [np.zeros((25, 1), np.float32) for _ in reqs]
    sub = [srv.submit(x, o) for x, o in zip(reqs, outs)]
    assert all(rc == OK for rc, _ in sub)
    lib.fake_hold(srv.h, 0)
    rcs = [srv.wait(t, 30_000_000) for _, t in sub]
    assert rcs[:2] == [OK, OK] and all(rc == EDEVICE for rc in rcs[2:]), rcs
    assert b"fake device failed" in lib.fake_error(srv.h)
    assert srv.counters()["failed"] == 8
    srv.stop()
