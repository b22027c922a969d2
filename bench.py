#!/usr/bin/env python
"""Benchmark of the TurboCat symmetric-tree hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload train|predict|serve|batched]

A step is one boosting iteration (gradients -> per-level histograms ->
oblivious split -> partition -> leaves -> prediction update) over the whole
synthetic training matrix, already resident in HBM.  The default workload is
binary LogLoss on 1,000,000 rows x 100 float32 features (80 N(0,1), 20
U[0,1) with 5% exact duplicates per column; label ~ Bernoulli(sigmoid(w.x)),
10 nonzero weights), 255-bin quantization, depth-6 symmetric trees, lr 0.1,
lambda 1.  `value` is trees/s; the histogram kernel's roofline fraction and
predict rows/s ride along, and the reference's own CPU path (compiled from
its sources into oracle/_ref) is timed beside it on a bounded sample.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "turbo-cat_amd"))
sys.path.insert(0, ROOT)

METRIC = "hist GB/s, trees/s train, rows/s predict + p50/p99 serve latency vs plain"
HBM_PEAK_GBPS = 8000.0
SAMPLING_TEXT = {"none": "", "random": ", row sampling: random 0.8 per tree",
                 "goss": ", row sampling: GOSS top 0.2 / other 0.1 per tree",
                 "mvs": ", row sampling: MVS 0.5 per tree"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_hist_train_1Mx100_d6_traffic.json")
KERNEL_SOURCES = ["tc_common.h", "tc_train_kernels.hip", "tc_train.hip"]


def kernel_sources_hash():
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "turbo-cat_amd", "csrc", name), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic(rows, features, depth, shape):
    """HBM bytes per k_hist launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement (copy-calibrated), used only when it was taken on
    this workload, with the same histogram work-item shape (rows per root
    item, items per deeper level) and these exact kernel sources; otherwise
    None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    same = (t.get("rows"), t.get("features"), t.get("depth"), t.get("shape")) == \
        (rows, features, depth, list(shape))
    if not same or t.get("sources_sha256") != kernel_sources_hash():
        return None
    return t["traffic_bytes_per_launch"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ data
def binary_workload(n, F, seed, n_uniform=None, dup=0.05):
    """80% N(0,1) columns, 20% U[0,1) columns with `dup` of rows overwritten by
    another seeded row's value in the same column; y ~ Bernoulli(sigmoid(w.x))
    with 10 nonzero N(0,1) weights."""
    rng = np.random.default_rng(seed)
    nu = F // 5 if n_uniform is None else n_uniform
    X = np.empty((n, F), np.float32)
    X[:, :F - nu] = rng.standard_normal((n, F - nu), dtype=np.float32)
    X[:, F - nu:] = rng.random((n, nu), dtype=np.float32)
    k = int(n * dup)
    for c in range(F - nu, F):
        dst = rng.choice(n, k, replace=False)
        src = rng.integers(0, n, k)
        X[dst, c] = X[src, c]
    w = np.zeros(F, np.float32)
    nz = rng.choice(F, min(10, F), replace=False)
    w[nz] = rng.standard_normal(len(nz)).astype(np.float32)
    z = X @ w
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-z))).astype(np.float32)
    return X, y


def regression_workload(n, F, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, F), dtype=np.float32)
    y = (X[:, 0] + 2 * np.sin(X[:, 1]) + 0.5 * X[:, 2] * X[:, 3]
         + 0.1 * rng.standard_normal(n)).astype(np.float32)
    return X, y


# ------------------------------------------------------------------ CPU baseline
def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_train(X, y, loss, depth, lr, lam, budget_s, sampling="none"):
    """Time the reference's own boosting loop (compiled from its sources) on
    the same matrix for a bounded number of trees (with the fit's row
    sampling: random 0.8, GOSS 0.2 / 0.1 or MVS 0.5)."""
    from oracle import RefLib
    if not RefLib.available():
        return None

    def body():
        R = RefLib()
        t0 = time.perf_counter()
        ds = R.dataset(X, y)
        q = time.perf_counter() - t0
        def fit(T):  # seconds of the boosting loop
            if sampling in ("none", "random"):
                return ds.fit(loss, T, depth, lr, lam, 1.0, 0.8 if sampling == "random" else 1.0)[2][1]
            t = time.perf_counter()
            if sampling == "goss":
                ds.fit_goss(loss, T, depth, lr, lam, 1.0, 0.2, 0.1)
            else:
                ds.fit_mvs(loss, T, depth, lr, lam, 1.0, 0.5)
            return time.perf_counter() - t
        # one tree to estimate the rate, then as many as fit the budget
        per = max(fit(1), 1e-3)
        T = int(max(2, min(50, (budget_s - q) / per)))
        return dict(quantize_s=q, trees=T, loop_s=fit(T))

    r = RefLib.isolated(body)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=r["trees"] / r["loop_s"], unit="trees/s", cores=cores, kind="reference",
                host_cpu=host_cpu_model(),
                sample=(f"reference hot path (oracle/_ref, -O3 -mavx2 -mfma, OpenMP) on the full "
                        f"matrix: compute_bins {r['quantize_s']:.2f} s, then {r['trees']} trees "
                        f"in {r['loop_s']:.2f} s" + SAMPLING_TEXT[sampling]
                        + (" (the MVS draw on one thread)" if sampling == "mvs" else "")),
                quantize_s=r["quantize_s"])


def cpu_predict_rate(model, X, budget_rows=200_000):
    from oracle import RefLib
    if not RefLib.available():
        return None
    R = RefLib()
    Xs = X[:budget_rows]
    t0 = time.perf_counter()
    R.predict(model, Xs)
    return len(Xs) / (time.perf_counter() - t0)


# ------------------------------------------------------------------ workloads
def run_train(args):
    from turbocat import capi
    n, F, T_timed = args.rows, args.features, args.steps
    mse = args.loss == "mse"
    loss_id, lam = (capi.TC_LOSS_MSE, 0.0) if mse else (capi.TC_LOSS_LOGLOSS, 1.0)
    log(f"[bench] generating {n} x {F} {'regression' if mse else 'binary'} workload")
    X, y = regression_workload(n, F, seed=1000) if mse else binary_workload(n, F, seed=1001)
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    ds, edges, ne = ctx.quantize_fit(X)
    quant_s = time.perf_counter() - t0
    log(f"[bench] quantize_fit {quant_s:.2f} s; bins per feature {int(ne.min()) + 1}..{int(ne.max()) + 1}")
    R = 5  # trees timed kernel by kernel for the roofline
    smp = dict(random=dict(subsample=0.8),
               goss=dict(use_goss=True, goss_top_rate=0.2, goss_other_rate=0.1),
               mvs=dict(use_mvs=True, mvs_subsample=0.5)).get(args.sampling, {})
    prm = capi.train_params(loss_id, n_trees=args.warmup + T_timed + R,
                            learning_rate=0.1, max_depth=6, lambda_l2=lam,
                            hist_chunk=args.hist_chunk, hist_deep_items=args.deep_items, **smp)
    tr = capi.Trainer(ctx, [ds], [y], prm)
    tr.step(args.warmup)
    tr.sync()
    root, deep_items = tr.hist_chunks()
    log(f"[bench] warmup {args.warmup} trees done (histogram work items: {root} rows at the "
        f"root, ~{deep_items} items per deeper level); timing {T_timed}")
    t0 = time.perf_counter()
    tr.step(T_timed)
    tr.sync()
    dt = time.perf_counter() - t0
    trees_per_s = T_timed / dt
    log(f"[bench] {trees_per_s:.1f} trees/s ({1e3 * dt / T_timed:.3f} ms/tree)")
    rows0 = tr.hist_rows()
    hist_ms, launches = tr.step_timed(R)
    rows1 = tr.hist_rows()
    alg_bytes = (rows1 - rows0) * (F + 8)
    achieved = alg_bytes / (hist_ms * 1e-3) / 1e9
    log(f"[bench] histogram: {launches} launches, {hist_ms / launches * 1e3:.1f} us avg, "
        f"{alg_bytes / launches / 1e6:.1f} MB algorithmic per launch -> {achieved:.0f} GB/s")
    model = tr.model()
    tr.close()
    # ------------------------------------------------ predict on the resident rows
    import ctypes
    ens = ctx.ensemble(model, F)
    dx, do = ctypes.c_void_p(), ctypes.c_void_p()
    capi.check(ctx.L.tc_device_alloc(ctx.h, X.nbytes, ctypes.byref(dx)))
    capi.check(ctx.L.tc_device_alloc(ctx.h, 4 * n, ctypes.byref(do)))
    capi.check(ctx.L.tc_memcpy_h2d(ctx.h, dx, X.ctypes.data_as(ctypes.c_void_p), X.nbytes))
    ens.predict_dev(dx, n, do)
    ctx.sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        ens.predict_dev(dx, n, do)
    ctx.sync()
    pred_kernel = reps * n / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    ens.predict(X)
    pred_e2e = n / (time.perf_counter() - t0)
    capi.check(ctx.L.tc_device_free(ctx.h, dx))
    capi.check(ctx.L.tc_device_free(ctx.h, do))
    log(f"[bench] predict {pred_kernel / 1e6:.1f} M rows/s resident, {pred_e2e / 1e6:.1f} M rows/s "
        f"with host copies ({model['T']} trees)")
    cpu = None
    if not args.no_cpu:
        log("[bench] timing the reference CPU path (bounded sample)")
        cpu = cpu_baseline_train(X, y, loss_id, 6, 0.1, lam, args.cpu_budget, args.sampling)
        if cpu:
            cpu["predict_rows_per_s"] = cpu_predict_rate(model, X)
            log(f"[bench] CPU reference {cpu['value']:.3f} trees/s on {cpu['cores']} threads")
    out = {
        "metric": METRIC, "value": round(trees_per_s, 3), "unit": "trees/s", "n_gpus": 1,
        "steps": T_timed, "warmup": args.warmup, "ms_per_step": round(1e3 * dt / T_timed, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ((f"regression MSE {n} rows x {F} float32 N(0,1) features, "
                                f"y = x0 + 2 sin x1 + 0.5 x2 x3 + N(0, 0.1), 255-bin uint8 "
                                f"quantization, symmetric trees depth 6, lr 0.1, lambda 0") if mse
                   else (f"binary logloss {n} rows x {F} float32 features "
                         f"({F - F // 5} N(0,1), {F // 5} U[0,1) with 5% duplicates), "
                         f"255-bin uint8 quantization, symmetric trees depth 6, lr 0.1, "
                         f"lambda 1")) + SAMPLING_TEXT[args.sampling], "rows": n, "features": F,
                   "depth": 6, "trees_timed": T_timed},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": None if mse or args.sampling != "none"
                     else measured_traffic(n, F, 6, (root, deep_items)),
                     "kernel": "k_hist", "avg_launch_us": round(hist_ms / launches * 1e3, 2),
                     "algorithmic_bytes_per_launch": int(alg_bytes / launches)},
        "cpu_baseline": cpu,
        "extra": {"quantize_s": round(quant_s, 3),
                  "hist_work_items": {"root_rows": root, "deep_items": deep_items},
                  "predict_rows_per_s_resident": round(pred_kernel, 1),
                  "predict_rows_per_s_with_copies": round(pred_e2e, 1),
                  "predict_trees": model["T"]},
    }
    print(json.dumps(out), flush=True)


def run_serve(args):
    """Poisson request stream against a resident 1000-tree depth-8 ensemble
    over 200 features; coalesced server vs one launch per request."""
    from turbocat import capi
    F, T, D = 200, args.serve_trees, 8
    rng = np.random.default_rng(3003)
    ctx = capi.Context(0)
    if args.serve_model == "random":
        # a synthetic ensemble of the same shape (every tree full depth, random
        # splits): the same inference cost, no training in the process
        log(f"[bench] synthetic serving ensemble ({T} trees depth {D})")
        mr = np.random.default_rng(1003)
        model = dict(T=T, D=D, feature=mr.integers(0, F, T * D).astype(np.uint16),
                     fthr=mr.standard_normal(T * D).astype(np.float32),
                     bin=np.zeros(T * D, np.uint8), depth=np.full(T, D, np.uint16),
                     leaves=(mr.standard_normal(T << D) * 0.05).astype(np.float32),
                     weight=np.full(T, 0.1, np.float32), base=0.0)
    else:
        log(f"[bench] training the serving ensemble ({args.serve_train_rows} rows, {T} trees depth {D})")
        X, y = binary_workload(args.serve_train_rows, F, seed=1003, n_uniform=40)
        ds, _, _ = ctx.quantize_fit(X)
        tr = capi.Trainer(ctx, [ds], [y], capi.train_params(capi.TC_LOSS_LOGLOSS, n_trees=T,
                                                             learning_rate=0.1, max_depth=D,
                                                             lambda_l2=1.0))
        tr.step(T)
        tr.sync()
        model = tr.model()
        tr.close()
    ens = ctx.ensemble(model, F)
    n_req = args.requests
    sizes = np.floor(np.exp(rng.uniform(0, np.log(65536), n_req))).astype(int)
    pool = rng.standard_normal((int(sizes.max()) + 1024, F), dtype=np.float32)
    offs = rng.integers(0, 1024, n_req)
    reqs = [np.ascontiguousarray(pool[o:o + s]) for o, s in zip(offs, sizes)]
    base_gaps = rng.exponential(1.0, n_req)  # scaled by 1/QPS; QPS 0 = all at once

    def plain(qps):
        lat, call = [], []
        t_start = time.perf_counter()
        t_next = t_start
        for x, g in zip(reqs, base_gaps):
            t_next += g / qps if qps else 0.0
            now = time.perf_counter()
            if now < t_next:
                time.sleep(t_next - now)
            t_arr = t_next if qps else now
            t_call = time.perf_counter()
            ens.predict(x, transform=1)
            t_end = time.perf_counter()
            lat.append(t_end - t_arr)
            call.append(t_end - t_call)
        return np.array(lat), time.perf_counter() - t_start, np.array(call)

    def served(qps, check=False):
        srv = capi.Server(ens, max_batch_rows=args.max_batch_rows, memory_budget=1 << 31,
                          n_streams=args.streams, transform=1)
        outs = [np.empty(len(x), np.float32) for x in reqs]
        lat = np.zeros(n_req)
        done = threading.Event()
        tickets = [None] * n_req
        submit_t = np.zeros(n_req)
        t_start = time.perf_counter()

        def waiter():
            for i in range(n_req):
                while tickets[i] is None:
                    time.sleep(1e-5)
                srv.wait(tickets[i])
                lat[i] = time.perf_counter() - submit_t[i]
            done.set()

        th = threading.Thread(target=waiter)
        th.start()
        t_next = t_start
        for i, (x, g) in enumerate(zip(reqs, base_gaps)):
            t_next += g / qps if qps else 0.0
            now = time.perf_counter()
            if now < t_next:
                time.sleep(t_next - now)
            submit_t[i] = t_next if qps else now
            tickets[i] = srv.submit(x, outs[i])
        done.wait()
        wall = time.perf_counter() - t_start
        b, r = srv.stats()
        internal = srv.latencies_us() * 1e-6  # submit -> result, timed inside the server
        srv.stop()
        if check:  # each request alone and in the mix: identical bits
            for i in range(0, n_req, max(1, n_req // 20)):
                assert np.array_equal(outs[i], ens.predict(reqs[i], transform=1))
        return lat, wall, b, internal

    rows = int(sizes.sum())
    p = lambda a, q: float(np.percentile(a, q) * 1e3)
    ens.predict(reqs[0], transform=1)  # warm the plain path
    lat_p, _, call_p = plain(args.qps)
    lat_s, _, batches, int_s = served(args.qps, check=True)
    log(f"[bench] {args.qps} QPS: plain p50/p99 {p(lat_p, 50):.3f}/{p(lat_p, 99):.3f} ms "
        f"(call alone {p(call_p, 50):.3f}/{p(call_p, 99):.3f}), served "
        f"{p(lat_s, 50):.3f}/{p(lat_s, 99):.3f} ms as the client thread sees it, "
        f"{p(int_s, 50):.3f}/{p(int_s, 99):.3f} ms inside the server ({batches} batches)")
    _, wall_ps, _ = plain(0)
    _, wall_ss, batches_s, _ = served(0)
    log(f"[bench] saturation: plain {rows / wall_ps / 1e6:.2f} M rows/s "
        f"({n_req / wall_ps:.0f} req/s), served {rows / wall_ss / 1e6:.2f} M rows/s "
        f"({n_req / wall_ss:.0f} req/s, {batches_s} batches)")
    res = {
        "metric": METRIC, "value": round(rows / wall_ss, 1), "unit": "rows/s", "n_gpus": 1,
        "steps": n_req, "warmup": 0, "ms_per_step": round(1e3 * wall_ss / n_req, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"serving: {T} symmetric trees depth {D} over {F} features, "
                               f"request rows log-uniform [1, 65536]; latency at Poisson "
                               f"{args.qps} QPS, throughput with every request queued at once",
                   "requests": n_req, "qps": args.qps, "model": args.serve_model},
        "extra": {"served_p50_ms": p(lat_s, 50), "served_p99_ms": p(lat_s, 99),
                  "plain_p50_ms": p(lat_p, 50), "plain_p99_ms": p(lat_p, 99),
                  "served_internal_p50_ms": p(int_s, 50), "served_internal_p99_ms": p(int_s, 99),
                  "plain_call_p50_ms": p(call_p, 50), "plain_call_p99_ms": p(call_p, 99),
                  "saturation_served_rows_per_s": rows / wall_ss,
                  "saturation_plain_rows_per_s": rows / wall_ps,
                  "saturation_served_req_per_s": n_req / wall_ss,
                  "saturation_plain_req_per_s": n_req / wall_ps,
                  "served_vs_plain_throughput": wall_ps / wall_ss,
                  "batches_at_qps": int(batches), "batches_at_saturation": int(batches_s),
                  "requests": n_req, "rows": rows},
    }
    print(json.dumps(res), flush=True)


def run_batched(args):
    """64 independent regression jobs (rows log-uniform in [1e4, 2e6],
    features uniform in [16, 256]) trained in lockstep by one set of launches,
    against the same jobs trained one after another."""
    from turbocat import capi
    rng = np.random.default_rng(2002)
    J = args.jobs
    rows = np.floor(np.exp(rng.uniform(np.log(1e4), np.log(args.max_job_rows), J))).astype(int)
    feats = rng.integers(16, 257, J)
    ctx = capi.Context(0)
    dss, ys = [], []
    for j in range(J):
        X, y = regression_workload(int(rows[j]), int(feats[j]), seed=5000 + j)
        ds, _, _ = ctx.quantize_fit(X)
        dss.append(ds)
        ys.append(y)
        del X
    T = args.steps
    prm = capi.train_params(capi.TC_LOSS_MSE, n_trees=T, learning_rate=0.1, max_depth=6,
                            lambda_l2=0.0)
    log(f"[bench] {J} jobs, {rows.sum()} rows total; batched")
    trb = capi.Trainer(ctx, dss, ys, prm)
    t0 = time.perf_counter()
    trb.step(T)
    trb.sync()
    t_batched = time.perf_counter() - t0
    trb.close()
    log(f"[bench] batched {J * T / t_batched:.1f} trees/s; now one job at a time")
    t_plain = 0.0
    for j in range(J):
        trj = capi.Trainer(ctx, [dss[j]], [ys[j]], prm)
        t0 = time.perf_counter()
        trj.step(T)
        trj.sync()
        t_plain += time.perf_counter() - t0
        trj.close()
    res = {
        "metric": METRIC, "value": round(J * T / t_batched, 2), "unit": "trees/s", "n_gpus": 1,
        "steps": T, "warmup": 0, "ms_per_step": round(1e3 * t_batched / T, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{J} regression jobs, rows log-uniform [1e4, {args.max_job_rows}], "
                               f"features uniform [16, 256], depth 6, {T} trees each",
                   "jobs": J, "total_rows": int(rows.sum())},
        "extra": {"batched_trees_per_s": J * T / t_batched, "plain_trees_per_s": J * T / t_plain,
                  "speedup_vs_plain": t_plain / t_batched},
    }
    print(json.dumps(res), flush=True)


def multiclass_workload(n, F, K, seed, sigma=0.2):
    """X ~ N(0,1); W K x F ~ N(0, sigma); label = argmax(W x + Gumbel(0,1))."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, F), dtype=np.float32)
    W = rng.normal(0.0, sigma, (K, F)).astype(np.float32)
    y = np.empty(n, np.float32)
    for s0 in range(0, n, 1 << 18):  # bounded temporaries
        z = X[s0:s0 + (1 << 18)] @ W.T + rng.gumbel(0.0, 1.0, (min(1 << 18, n - s0), K))
        y[s0:s0 + len(z)] = np.argmax(z, axis=1)
    return X, y


def run_multiclass(args):
    """K-channel symmetric multiclass: training trees/s and histogram GB/s
    (algorithmic bytes rows x (F + 8K)), batched inference over every row
    with a full-size ensemble of random weights, and the FP64 K-channel CPU
    restatement timed on a bounded row sample."""
    from turbocat import capi
    n, F, K, D = args.mc_rows, args.mc_features, args.classes, 8
    log(f"[bench] generating {n} x {F} multiclass workload, {K} classes")
    X, y = multiclass_workload(n, F, K, seed=1004)
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    ds, edges, ne = ctx.quantize_fit(X)
    quant_s = time.perf_counter() - t0
    log(f"[bench] quantize_fit {quant_s:.2f} s")
    R, T_timed = 2, args.steps
    prm = capi.train_params(capi.TC_LOSS_SOFTMAX, n_trees=args.warmup + T_timed + R,
                            learning_rate=0.05, max_depth=D, lambda_l2=1.0, n_classes=K)
    tr = capi.Trainer(ctx, [ds], [y], prm)
    tr.step(args.warmup)
    tr.sync()
    t0 = time.perf_counter()
    tr.step(T_timed)
    tr.sync()
    dt = time.perf_counter() - t0
    rows0 = tr.hist_rows()
    hist_ms, launches = tr.step_timed(R)
    rows1 = tr.hist_rows()
    alg = (rows1 - rows0) * (F + 8 * K)
    achieved = alg / (hist_ms * 1e-3) / 1e9
    log(f"[bench] {T_timed / dt:.2f} trees/s ({1e3 * dt / T_timed:.1f} ms/tree); histogram "
        f"{hist_ms / launches * 1e3:.0f} us avg per launch -> {achieved:.0f} GB/s")
    tr.close()
    # inference: a 1000-tree depth-8 ensemble of this shape, random weights
    rng = np.random.default_rng(7)
    Tm = 1000
    model = dict(T=Tm, D=D, K=K, feature=rng.integers(0, F, Tm * D).astype(np.uint16),
                 fthr=rng.standard_normal(Tm * D).astype(np.float32),
                 depth=np.full(Tm, D, np.uint16),
                 leaves=(rng.standard_normal((Tm << D) * K) * 0.05).astype(np.float32),
                 weight=np.full(Tm, 0.05, np.float32), base=0.0)
    ens = ctx.ensemble(model, F)
    import ctypes
    dx, do = ctypes.c_void_p(), ctypes.c_void_p()
    capi.check(ctx.L.tc_device_alloc(ctx.h, X.nbytes, ctypes.byref(dx)))
    capi.check(ctx.L.tc_device_alloc(ctx.h, 4 * n * K, ctypes.byref(do)))
    capi.check(ctx.L.tc_memcpy_h2d(ctx.h, dx, X.ctypes.data_as(ctypes.c_void_p), X.nbytes))
    ens.predict_dev(dx, n, do, transform=1)
    ctx.sync()
    t0 = time.perf_counter()
    ens.predict_dev(dx, n, do, transform=1)
    ctx.sync()
    pred_s = time.perf_counter() - t0
    capi.check(ctx.L.tc_device_free(ctx.h, dx))
    capi.check(ctx.L.tc_device_free(ctx.h, do))
    log(f"[bench] inference {n / pred_s / 1e6:.1f} M rows/s ({Tm} trees depth {D}, {K} classes)")
    cpu = None
    if not args.no_cpu:
        from oracle import Oracle, tree_params
        o = Oracle()
        ns = min(n, args.mc_cpu_rows)
        bins = ds.bins()[:, :ns].copy() if ns < n else ds.bins()
        g, h = o.gradients_mc_f64(y[:ns], np.zeros((ns, K), np.float32))
        t0 = time.perf_counter()
        o.tree_mc_f64(bins, g, h, edges, ne, tree_params(D, lambda_l2=1.0))
        sec = time.perf_counter() - t0
        cpu = dict(value=1.0 / sec, unit="trees/s", cores=1, kind="port", host_cpu=host_cpu_model(),
                   sample=(f"FP64 K-channel restatement (oracle/, single thread) building one "
                           f"depth-{D} tree on the first {ns} rows: {sec:.2f} s; for all {n} rows "
                           f"expect about {n / ns:.0f}x longer (extrapolated, not measured)"))
        log(f"[bench] CPU K-channel restatement {sec:.2f} s per tree on {ns} rows")
    out = {
        "metric": METRIC, "value": round(T_timed / dt, 3), "unit": "trees/s", "n_gpus": 1,
        "steps": T_timed, "warmup": args.warmup, "ms_per_step": round(1e3 * dt / T_timed, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"multiclass softmax {n} rows x {F} float32 N(0,1) features, "
                               f"{K} classes (argmax(Wx + Gumbel)), K-channel symmetric trees "
                               f"depth {D}, lr 0.05, lambda 1", "rows": n, "features": F,
                   "classes": K, "depth": D},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "k_hist", "avg_launch_us": round(hist_ms / launches * 1e3, 2),
                     "algorithmic_bytes_per_launch": int(alg / launches)},
        "cpu_baseline": cpu,
        "extra": {"quantize_s": round(quant_s, 3), "inference_rows_per_s": round(n / pred_s, 1),
                  "inference_trees": Tm},
    }
    print(json.dumps(out), flush=True)


def run_regular(args):
    """Regular (non-oblivious) depthwise trees, the reference's mode="small"
    (binary LogLoss on the default matrix) or, with --rclasses > 2, its
    multiclass loop (one regular tree per class and iteration, on the
    multiclass matrix): iterations/s, the histogram kernel against the HBM
    roofline (algorithmic bytes rows x (F + 8) per histogrammed node), and
    resident prediction rows/s.  The reference's regular trees cannot be
    compiled here (its tree.cpp needs Eigen), so the CPU column is the
    repository's FP64 restatement (oracle/) building one tree on a bounded
    row sample, single thread."""
    from turbocat import capi
    multi = args.rclasses > 2
    D, lr = args.depth, 0.1
    if multi:
        n, F, K = args.mc_rows, args.mc_features, args.rclasses
        log(f"[bench] generating {n} x {F} multiclass workload, {K} classes")
        X, y = multiclass_workload(n, F, K, seed=1004)
        loss = capi.TC_LOSS_SOFTMAX
    else:
        n, F, K = args.rows, args.features, 1
        log(f"[bench] generating {n} x {F} binary workload")
        X, y = binary_workload(n, F, seed=1001)
        loss = capi.TC_LOSS_LOGLOSS
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    ds, edges, ne = ctx.quantize_fit(X)
    quant_s = time.perf_counter() - t0

    smp = dict(random=dict(subsample=0.8),
               goss=dict(use_goss=True, goss_top_rate=0.2, goss_other_rate=0.1),
               mvs=dict(use_mvs=True, mvs_subsample=0.5)).get(args.sampling, {})

    def fit(iters, time_hist):
        prm = capi.rtrain_params(loss, n_iterations=iters, learning_rate=lr,
                                 n_classes=K if multi else 0, time_hist=time_hist, **smp)
        prm.tree = capi.rtree_params(max_depth=D, max_leaves=args.max_leaves, lambda_l2=1.0,
                                     grow_policy=1 if args.grow_policy == "lossguide" else 0)
        return ds.rtrain(y, prm)

    fit(args.warmup, 0)  # first use: allocations, code objects
    T_timed = args.steps
    m = fit(T_timed, 0)
    loop_ms, _ = ctx.last_train_timing()  # device events around the boosting loop
    its = T_timed / (loop_ms * 1e-3)
    R = 5
    fit(R, 1)
    _, hist_ms = ctx.last_train_timing()
    rows, launches = ctx.last_train_hist_rows()
    alg = rows * (F + 8)
    achieved = alg / (hist_ms * 1e-3) / 1e9
    nodes = np.array([t["n_nodes"] for t in m["trees"]])
    log(f"[bench] {its:.1f} iterations/s ({loop_ms / T_timed:.3f} ms per iteration of {K} "
        f"tree(s), {nodes.mean():.1f} nodes per tree); histogram {launches} launches, "
        f"{hist_ms / launches * 1e3:.1f} us avg -> {achieved:.0f} GB/s")
    import ctypes
    ens = capi.REnsemble(ctx, m["trees"], m["weight"], edges, ne, F, rule=0 if multi else 1, K=K,
                         tree_class=m["tree_class"] if multi else None, base=m["base"])
    dx, do = ctypes.c_void_p(), ctypes.c_void_p()
    capi.check(ctx.L.tc_device_alloc(ctx.h, X.nbytes, ctypes.byref(dx)))
    capi.check(ctx.L.tc_device_alloc(ctx.h, 4 * n * K, ctypes.byref(do)))
    capi.check(ctx.L.tc_memcpy_h2d(ctx.h, dx, X.ctypes.data_as(ctypes.c_void_p), X.nbytes))
    ens.predict_dev(dx, n, do)
    ctx.sync()
    t0 = time.perf_counter()
    ens.predict_dev(dx, n, do)
    ctx.sync()
    pred = n / (time.perf_counter() - t0)
    capi.check(ctx.L.tc_device_free(ctx.h, dx))
    capi.check(ctx.L.tc_device_free(ctx.h, do))
    log(f"[bench] predict {pred / 1e6:.1f} M rows/s resident ({len(m['trees'])} trees)")
    cpu = None
    if not args.no_cpu:
        from oracle import Oracle, rtree_params
        o = Oracle()
        ns = min(n, args.cpu_rows)
        bins = ds.bins()[:, :ns].copy() if ns < n else ds.bins()
        if multi:
            g, h = o.gradients_mc_f32(y[:ns], np.zeros((ns, K), np.float32))
            g, h = np.ascontiguousarray(g[:, 0]), np.ascontiguousarray(h[:, 0])
        else:
            p0 = np.full(ns, o.init_prediction(loss, y[:ns]), np.float32)
            g, h = o.gradients_f32(loss, y[:ns], p0)
        t0 = time.perf_counter()
        o.rtree_f64(bins, g, h, rtree_params(max_depth=D, max_leaves=args.max_leaves,
                                             grow_policy=1 if args.grow_policy == "lossguide" else 0))
        sec = time.perf_counter() - t0
        cpu = dict(value=1.0 / (sec * K), unit="iterations/s", cores=1, kind="port",
                   host_cpu=host_cpu_model(),
                   sample=(f"FP64 regular-tree restatement (oracle/, single thread) building one "
                           f"depth-{D} tree on the first {ns} rows: {sec:.2f} s, x {K} trees per "
                           f"iteration; for all {n} rows expect about {n / ns:.0f}x longer "
                           f"(extrapolated, not measured)"))
    out = {
        "metric": METRIC, "value": round(its, 3), "unit": "iterations/s", "n_gpus": 1,
        "steps": T_timed, "warmup": args.warmup, "ms_per_step": round(loop_ms / T_timed, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ((f"multiclass softmax {n} rows x {F} float32 N(0,1) features, "
                                f"{K} classes, the reference's multiclass loop: one regular "
                                f"{args.grow_policy} tree per class and iteration, depth {D}, max_leaves "
                                f"{args.max_leaves}, lr {lr}, lambda 1") if multi
                   else (f"binary logloss {n} rows x {F} float32 features "
                         f"({F - F // 5} N(0,1), {F // 5} U[0,1) with 5% duplicates), regular "
                         f"{args.grow_policy} trees (mode='small') depth {D}, max_leaves "
                         f"{args.max_leaves}, lr {lr}, lambda 1")) + SAMPLING_TEXT[args.sampling],
                   "rows": n, "features": F, "classes": K, "depth": D},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "k_hist", "avg_launch_us": round(hist_ms / launches * 1e3, 2),
                     "algorithmic_bytes_per_launch": int(alg / launches)},
        "cpu_baseline": cpu,
        "extra": {"quantize_s": round(quant_s, 3), "predict_rows_per_s_resident": round(pred, 1),
                  "mean_nodes_per_tree": round(float(nodes.mean()), 2)},
    }
    print(json.dumps(out), flush=True)


def run_copy(args):
    """Calibration for HBM byte counters: a plain device copy of a known
    size (256 MiB read, 256 MiB written per copy)."""
    import torch
    n = 64 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(1, args.warmup)):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "device copy", "bytes_read_per_copy": 4 * n,
                      "bytes_written_per_copy": 4 * n, "copies": args.steps + max(1, args.warmup),
                      "GB/s": round(8 * n / dt / 1e9, 1)}), flush=True)


def pmc_summary(root, kernel_regex):
    """Average value per dispatch of every counter in the rocprofv3 counter
    collection CSVs under `root`, for the kernels matching kernel_regex."""
    import csv
    import glob
    import re
    pat = re.compile(kernel_regex)
    acc = {}
    for path in glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name", "")
                if not pat.search(name):
                    continue
                key = (name.split("(")[0], r["Counter_Name"])
                d = acc.setdefault(key, {})
                disp = r.get("Dispatch_Id", str(len(d)))
                d[disp] = d.get(disp, 0.0) + float(r["Counter_Value"])
    out = []
    for (name, ctr), d in sorted(acc.items()):
        v = list(d.values())
        out.append({"kernel": name, "counter": ctr, "dispatches": len(v),
                    "mean_per_dispatch": sum(v) / len(v), "min": min(v), "max": max(v)})
    print(json.dumps(out, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="train",
                    choices=["train", "serve", "batched", "multiclass", "regular", "copy"])
    ap.add_argument("--pmc-summary", metavar="DIR", default=None,
                    help="print per-dispatch averages of the rocprofv3 counter CSVs under DIR")
    ap.add_argument("--kernel", default="k_hist", help="kernel name regex for --pmc-summary")
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--features", type=int, default=100)
    ap.add_argument("--hist-chunk", type=int, default=0,
                    help="rows per histogram work item at the root (0: autotuned)")
    ap.add_argument("--deep-items", type=int, default=0,
                    help="with --hist-chunk: work items per deeper level (0: --hist-chunk rows)")
    ap.add_argument("--loss", default="logloss", choices=["logloss", "mse"],
                    help="train workload: binary logloss (default) or the MSE regression set")
    ap.add_argument("--sampling", default="none", choices=["none", "random", "goss", "mvs"],
                    help="train and regular workloads: row sampling per tree (random 0.8, GOSS "
                         "top 0.2 / other 0.1, MVS 0.5; drawn on the host with the reference's "
                         "routines)")
    ap.add_argument("--grow-policy", default="depthwise", choices=["depthwise", "lossguide"],
                    help="regular workload: how the trees grow (TreeConfig::grow_policy)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--qps", type=float, default=200.0)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--serve-trees", type=int, default=1000)
    ap.add_argument("--serve-train-rows", type=int, default=500_000)
    ap.add_argument("--serve-model", default="trained", choices=["trained", "random"],
                    help="serve workload: a trained ensemble, or a synthetic one of the same shape")
    ap.add_argument("--max-batch-rows", type=int, default=32768)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--mc-rows", type=int, default=4_000_000)
    ap.add_argument("--mc-features", type=int, default=256)
    ap.add_argument("--classes", type=int, default=10)
    ap.add_argument("--mc-cpu-rows", type=int, default=200_000)
    ap.add_argument("--max-job-rows", type=int, default=2_000_000)
    ap.add_argument("--depth", type=int, default=6, help="regular workload: max_depth")
    ap.add_argument("--rclasses", type=int, default=0,
                    help="regular workload: > 2 runs the reference's multiclass loop with this many "
                         "classes on the multiclass matrix (--mc-rows x --mc-features)")
    ap.add_argument("--max-leaves", type=int, default=64, help="regular workload: max_leaves")
    ap.add_argument("--cpu-rows", type=int, default=100_000,
                    help="regular workload: rows of the CPU restatement's sample")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.kernel)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 (sharding "
                                   "across GPUs is designed but not implemented)"}))
        sys.exit(2)
    {"train": run_train, "serve": run_serve, "batched": run_batched,
     "multiclass": run_multiclass, "regular": run_regular, "copy": run_copy}[args.workload](args)


if __name__ == "__main__":
    main()
