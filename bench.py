#!/usr/bin/env python3
"""Benchmark of the IRWLS training hot path on one MI355X.

A step is one full-training run (trainFULL through the C entry point) over
the four-Gaussian workload n=50000, d=128 (RBF gamma=1/128, C=1, working
set 2000, eta=0.001), with the dataset resident in HBM before the timed
region.  `value` is seconds per training run (lower is better) of the
mixed-precision trainer (BF16 hi/lo x3 kernel products, FP32 Cholesky with
FP64 diagonal tiles and FP64 refinement); the plain FP64 trainer runs on the
same data in the same process and its time, accuracy and the differences
are reported beside it (--mode plain|mixed runs one of them only).

The roofline object covers the kernel with the largest share of device time
(found in the last warmup step with every timer on); during the timed steps
HIP events are recorded on the library's stream around that kernel's
launches only.  The CPU baseline times the oracle (the FP64 C restatement of
the reference, OpenMP over the host cores) on the same workload: the whole
run by default, or its first --cpu-outer outer iterations extrapolated to the
GPU run's outer-iteration count.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
# in the mixed mode the factorization's tiles (trsm, syrk) are FP32 MFMA
FP32_IN_MIXED = {"syrk", "trsm"}
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
# timer kinds whose bound is the matrix cores vs HBM
MFMA_KINDS = {"kmv", "kmat", "trsm", "syrk", "potf2", "sgma", "syrkd", "bf16"}

WORKLOADS = {
    "four_gaussians_n50000_d128": dict(n=50000, n_test=10000, d=128, gamma=1 / 128, C=1.0, W=2000, eta=0.001,
                                       seed=1),
    "four_gaussians_n10000_d128": dict(n=10000, n_test=2000, d=128, gamma=1 / 128, C=1.0, W=2000, eta=0.001,
                                       seed=1),
    "two_gaussians_n4000_d16": dict(n=4000, n_test=1000, d=16, gamma=1 / 16, C=10.0, W=500, eta=0.001, seed=0),
    # mixed-precision workload: same generator, n=200000, d=256, working set 4000
    "four_gaussians_n200000_d256": dict(n=200000, n_test=20000, d=256, gamma=1 / 256, C=1.0, W=4000, eta=0.001,
                                        seed=1, cpu_outer=1),
    # budgeted workloads (SGMA + IRWLSpar): CSR, d=2000, 1% density, values U[0,1], labels from an RBF
    # expansion over 200 anchors; budget 1000, 59 candidates, gamma 0.5, C 10
    # learned proposal: gains on a fixed subset of 65536 of the 10^6 rows (proposal_rows)
    "sparse_rbf_n1000000_d2000": dict(kind="budgeted", n=1000000, n_test=100000, d=2000, density=0.01, gamma=0.5,
                                      C=10.0, B=1000, cand=59, seed=3, cpu_sgma=3, proposal_rows=65536),
    "sparse_rbf_n100000_d2000": dict(kind="budgeted", n=100000, n_test=20000, d=2000, density=0.01, gamma=0.5,
                                     C=10.0, B=500, cand=59, seed=3, cpu_sgma=3),
    # batched workload: 12 one-vs-rest tasks of a 12-class mixture (n log-uniform in [1000, 30000], d in
    # {16, 32, 64, 128}) x 8 points C in {0.1, 1, 10, 100}, gamma in {0.5/d, 2/d}; W = min(n/4, 1500)
    "ovr12_grid8_batched": dict(kind="batched", n_test=2000, seed=4, n=None, d=None),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(wl):
    from libirwls_amd import synth
    if wl["d"] == 16:
        X, y = synth.two_gaussians(wl["n"], wl["d"], seed=wl["seed"])
        Xt, yt = synth.two_gaussians(wl["n_test"], wl["d"], seed=wl["seed"] + 100)
    else:
        X, y = synth.four_gaussians(wl["n"], wl["d"], seed=wl["seed"], means_seed=wl["seed"])
        Xt, yt = synth.four_gaussians(wl["n_test"], wl["d"], seed=wl["seed"] + 100, means_seed=wl["seed"])
    return X, y, Xt, yt


def cpu_baseline(wl, X, y, gpu_outer, max_outer):
    """Oracle timed on the host cores for `max_outer` outer iterations."""
    from libirwls_amd import abi
    from oracle import pyoracle as po
    props = abi.make_properties(gamma=wl["gamma"], C_=wl["C"], W=wl["W"], eta=wl["eta"])
    ds = po.Dataset.dense(X, y)
    po.set_limits(max(max_outer, 0), 0, 0)
    try:
        t0 = time.time()
        po.train_full(ds, props, seed=0)
        dt = time.time() - t0
        st = po.stats()
    finally:
        po.set_limits(0, 0, 0)
    per_outer = dt / max(st["outer_iters"], 1)
    if max_outer < 0:
        per_outer = dt / max(gpu_outer, 1)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": per_outer * gpu_outer,
        "unit": "s",
        "cores": cores,
        "kind": "port",
        "sample": ("oracle trainFULL, whole run: %d outer / %d inner iterations in %.1f s"
                   % (st["outer_iters"], st["inner_iters"], dt)) if max_outer < 0 else
                  ("oracle trainFULL, first %d outer iterations (%d inner) in %.1f s; %.2f s/outer x %d GPU "
                   "outer iterations (extrapolated)" % (st["outer_iters"], st["inner_iters"], dt, per_outer,
                                                        gpu_outer)),
    }


def run_mode(lib, L, ds, props, precision, warmup, steps):
    """Warmup and timed trainFULL runs in one precision mode.  The last
    warmup step runs with every timer on and names the kernel with the
    largest share of device time; the timed steps record HIP events around
    that kernel's launches only."""
    lib.set_options(precision=precision)
    dom, tw = None, None
    for i in range(warmup):
        prof = (i == warmup - 1)
        if prof:
            lib.timing(True)
        t0 = time.time()
        lib.train_full(ds, props, seed=0)
        log("bench: precision %d warmup %d done in %.2f s" % (precision, i, time.time() - t0))
        if prof:
            tw = lib.timers()
            dom = dominant(tw)
            lib.timing(False)
    lib.timing(True, kinds=None if dom is None else [dom])
    times = []
    beta = None
    for i in range(steps):
        L.irwls_dev_sync()
        t0 = time.perf_counter()
        beta = lib.train_full(ds, props, seed=0)
        L.irwls_dev_sync()
        times.append(time.perf_counter() - t0)
        log("bench: precision %d step %d: %.3f s %s" % (precision, i, times[-1], lib.last_stats()))
    stats = lib.last_stats()
    tm = lib.timers()
    lib.timing(False)
    lib.set_options(precision=0)
    if dom is None:
        dom = dominant(tm)
    return dict(times=times, value=float(np.mean(times)), beta=beta, stats=stats, tm=tm, tw=tw, dom=dom)


# timer kinds that bracket a sequence of launches (a whole CG solve, a
# triangular sweep) rather than one kernel: not candidates for the roofline
COMPOSITE_KINDS = {"pcg", "potrs"}


def dominant(timers):
    """The single-kernel timer kind with the most device time."""
    kinds = [k for k in timers if k not in COMPOSITE_KINDS and timers[k]["launches"]] or list(timers)
    return max(kinds, key=lambda k: timers[k]["ms"])


def measured_traffic(workload, mode, kind):
    """HBM bytes per launch of a kernel kind, from the committed summary of
    the two counter runs (FETCH_SIZE, WRITE_SIZE) for this workload and mode,
    calibrated against the copy probe; None when there is no such summary."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_%s_%s_%s.md" % (kind, mode, workload))
    try:
        with open(path) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def kernel_matrix_probe(reps=10):
    """MFMA utilisation of the kernel-matrix products at the mixed-precision
    workload's shape: K(X, X_W) w for n = 200000 rows, a working set of
    W = 4000 columns, d = 256, RBF; FP64 MFMA tiles and BF16 hi/lo x3 tiles,
    HIP-event timed over `reps` launches each.  Flops are 2 n W d counted
    once; the BF16 path issues three products for each."""
    from libirwls_amd import lib
    L = lib.load()
    L.irwls_dev_init()
    vp, pb = C.c_void_p, C.POINTER(lib.irwls_rows_bf16)
    L.irwls_dev_bf16_split.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, vp]
    L.irwls_dev_kernel_matvec_bf16.argtypes = [pb, pb, C.c_int, C.c_double, vp, C.c_double, C.c_double,
                                               C.c_double, vp, vp]
    n, W, d, g = 200000, 4000, 256, 1.0 / 256
    rng = np.random.default_rng(7)
    X = rng.standard_normal((n, d)) * 0.5

    def dev(a):
        a = np.ascontiguousarray(a)
        ptr = L.irwls_dev_alloc(max(a.nbytes, 8))
        L.irwls_dev_h2d(ptr, a.ctypes.data, a.nbytes)
        return ptr
    Xd, idx, w, out = dev(X), dev(np.arange(W, dtype=np.int32) * (n // W)), dev(rng.standard_normal(W)), \
        dev(np.zeros(n))
    nrm = L.irwls_dev_alloc(n * 8)
    L.irwls_dev_row_norms(Xd, d, n, d, nrm)
    r, c = lib.irwls_rows(), lib.irwls_rows()
    r.X, r.ldx, r.nrm, r.n = Xd, d, nrm, n
    c.X, c.ldx, c.nrm, c.idx, c.n = Xd, d, nrm, idx, W
    ldb = (d + 31) // 32 * 32
    Xh, Xl, n32 = L.irwls_dev_alloc(n * ldb * 2), L.irwls_dev_alloc(n * ldb * 2), L.irwls_dev_alloc(n * 4)
    L.irwls_dev_bf16_split(Xd, d, n, d, Xh, Xl, ldb, n32)
    rb, cb = lib.irwls_rows_bf16(), lib.irwls_rows_bf16()
    rb.Xh, rb.Xl, rb.ld, rb.nrm, rb.idx, rb.n = Xh, Xl, ldb, n32, None, n
    cb.Xh, cb.Xl, cb.ld, cb.nrm, cb.idx, cb.n = Xh, Xl, ldb, n32, idx, W
    flops = 2.0 * n * W * d
    res = {}
    for name, call in [("fp64", lambda: L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 1, g, w, 0.0, 1.0, 0.0,
                                                                   None, out)),
                       ("bf16x3", lambda: L.irwls_dev_kernel_matvec_bf16(C.byref(rb), C.byref(cb), 1, g, w, 0.0,
                                                                         1.0, 0.0, None, out))]:
        call()
        L.irwls_dev_sync()
        lib.timing(True, kinds=["kmv", "bf16"])
        for _ in range(reps):
            call()
        L.irwls_dev_sync()
        t = lib.timers()["kmv" if name == "fp64" else "bf16"]
        lib.timing(False)
        ms = t["ms"] / max(t["launches"], 1)
        tf = flops / (ms * 1e-3) / 1e12
        peak = PEAK_FP64_TFLOPS if name == "fp64" else PEAK_BF16_TFLOPS
        hw = tf * (3.0 if name == "bf16x3" else 1.0)
        res[name] = {"ms_per_launch": ms, "algorithmic_tflops": tf, "peak_tflops": peak,
                     "mfma_utilisation": hw / peak}
        log("kernel-matrix probe %s: %.3f ms per launch, %.1f TFLOP/s algorithmic, MFMA utilisation %.3f"
            % (name, ms, tf, hw / peak))
    print(json.dumps({"kernel_matrix_probe": {"n": n, "W": W, "d": d, "launches": reps, **res}}))


def copy_probe():
    """The calibration run for the traffic counters: five launches of the
    library's plain copy kernel over 512 MiB (each reads and writes that)."""
    from libirwls_amd import lib
    L = lib.load()
    L.irwls_dev_init()
    nbytes = 512 << 20
    src, dst = L.irwls_dev_alloc(nbytes), L.irwls_dev_alloc(nbytes)
    for _ in range(5):
        L.irwls_dev_copy_kernel(dst, src, nbytes)
    L.irwls_dev_sync()
    L.irwls_dev_free(src)
    L.irwls_dev_free(dst)
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 5}))


def roofline(r, steps, workload=None, mode=None):
    t = r["tm"][r["dom"]]
    avg_ms = t["ms"] / max(t["launches"], 1)
    dom = r["dom"]
    if avg_ms <= 0.0:
        return {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None,
                "kernel": dom, "note": "no timed launches of the dominant kernel"}
    if dom in MFMA_KINDS:
        peak = PEAK_BF16_TFLOPS if dom == "bf16" else PEAK_FP64_TFLOPS
        if mode == "mixed" and dom in FP32_IN_MIXED:
            peak = PEAK_FP32_TFLOPS
        ach = (t["flops"] / max(t["launches"], 1)) / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                "traffic": None}
    else:
        ach = (t["bytes"] / max(t["launches"], 1)) / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                "traffic": None}
    if workload is not None:
        roof["traffic"] = measured_traffic(workload, mode, dom)
    roof["kernel"] = dom
    roof["avg_launch_ms"] = avg_ms
    roof["launches_per_step"] = t["launches"] / steps
    roof["share_of_step"] = (t["ms"] / steps) / (r["value"] * 1e3)
    return roof


def run_budgeted(args, wl, lib, abi):
    """SGMA + IRWLSpar (the reference budgeted-train's timed window) with the
    plain and the learned proposal on the same CSR data."""
    from libirwls_amd import synth
    log("bench: generating %s" % args.workload)
    rp, col, val, y = synth.sparse_rbf_labels(wl["n"] + wl["n_test"], wl["d"], wl["density"], gamma=wl["gamma"],
                                              seed=wl["seed"])
    n = wl["n"]
    tr = rp[:n + 1]
    te = rp[n:] - rp[n]
    ds = lib.Dataset.csr(tr, col[:rp[n]], val[:rp[n]], y[:n])
    td = lib.Dataset.csr(te, col[rp[n]:], val[rp[n]:], y[n:])
    yt = y[n:]
    L = lib.load()
    L.irwls_dev_init()
    ds.pin(with_means=True)
    props = abi.make_properties(gamma=wl["gamma"], C_=wl["C"], size=wl["B"])
    pp = abi.make_predict_properties(labels=0)
    modes = {"plain": dict(sgma_candidates=wl["cand"], proposal=0),
             "learned": dict(sgma_candidates=wl["cand"], proposal=1, proposal_keep=8, proposal_pool=4 * wl["cand"],
                             proposal_rows=args.proposal_rows if args.proposal_rows is not None
                             else wl.get("proposal_rows", 0))}
    runs = {}
    sel = ["plain", "learned"] if args.mode == "both" else [{"plain": "plain"}.get(args.mode, "learned")]
    for mode in sel:
        lib.set_options(**modes[mode])
        dom, tw = None, None
        for i in range(args.warmup):
            lib.timing(i == args.warmup - 1)
            t0 = time.time()
            c = lib.sgma(ds, props, seed=1)
            lib.irwls_par(ds, c, props)
            log("bench: %s warmup %d done in %.2f s" % (mode, i, time.time() - t0))
            if i == args.warmup - 1:
                tw = lib.timers()
                dom = dominant(tw)
                lib.timing(False)
        lib.timing(True, kinds=None if dom is None else [dom])
        times, t_sgma = [], []
        for i in range(args.steps):
            L.irwls_dev_sync()
            t0 = time.perf_counter()
            cent = lib.sgma(ds, props, seed=1)
            sst = lib.last_stats()
            t1 = time.perf_counter()
            beta = lib.irwls_par(ds, cent, props)
            L.irwls_dev_sync()
            times.append(time.perf_counter() - t0)
            t_sgma.append(t1 - t0)
            pst = lib.last_stats()
            log("bench: %s step %d: %.3f s (SGMA %.3f s) %s %s" % (mode, i, times[-1], t_sgma[-1], sst, pst))
        tm = lib.timers()
        lib.timing(False)
        lib.set_options(sgma_candidates=64, proposal=0, proposal_keep=8, proposal_pool=0, proposal_rows=0)
        m = lib.budgeted_model(props, ds, cent, beta)
        acc = float(np.mean(lib.hard_test(td, m, pp) == yt))
        if dom is None:
            dom = dominant(tm)
        runs[mode] = dict(value=float(np.mean(times)), sgma_s=float(np.mean(t_sgma)), acc=acc, tm=tm, tw=tw,
                          dom=dom, sgma_stats=sst, par_stats=pst)
    main_mode = "learned" if "learned" in runs else "plain"
    r = runs[main_mode]
    cpu = None
    if wl.get("cpu_sgma", 0) > 0 and args.cpu_outer != 0:
        cpu = cpu_baseline_budgeted(wl, tr, col[:rp[n]], val[:rp[n]], y[:n], props, r)
    out = {
        "metric": "train s to +-0.1% acc of FP64 oracle: mixed/batched/learned vs plain GPU",
        "value": r["value"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["value"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+bf16-mlp" if main_mode == "learned" else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "mode": main_mode, "n": n, "d": wl["d"], "B": wl["B"],
                   "proposal_rows": modes["learned"]["proposal_rows"] if main_mode == "learned" else None,
                   "candidates": wl["cand"], "gamma": wl["gamma"], "C": wl["C"], "n_test": wl["n_test"],
                   "seed": wl["seed"]},
        "test_accuracy": r["acc"],
        "runs": {k: {"s": v["value"], "sgma_s": v["sgma_s"], "test_accuracy": v["acc"],
                     "sgma_steps": v["sgma_stats"]["sgma_steps"], "learned_steps": v["sgma_stats"]["learned_steps"],
                     "mlp_loss0": v["sgma_stats"]["mlp_loss0"], "mlp_loss": v["sgma_stats"]["mlp_loss"],
                     "irwlspar_iters": v["par_stats"]["irwlspar_iters"]} for k, v in runs.items()},
        "kernel_ms_per_step": {k: v["ms"] / args.steps for k, v in r["tm"].items() if v["launches"]},
        "warmup_kernel_ms": None if r["tw"] is None else {k: v["ms"] for k, v in r["tw"].items() if v["launches"]},
        "roofline": roofline(r, args.steps, args.workload, main_mode),
        "cpu_baseline": cpu,
    }
    if "plain" in runs and "learned" in runs:
        out["speedup_vs_plain"] = runs["plain"]["value"] / runs["learned"]["value"]
        out["learned_vs_plain"] = {"accuracy_delta": runs["learned"]["acc"] - runs["plain"]["acc"]}
    print(json.dumps(out), flush=True)


def run_batched(args, wl, lib, abi):
    """96 problems: sequential plain FP64 trainFULL runs against
    trainFULLBatched with BF16 hi/lo kernel products and batched FP64 solves."""
    from libirwls_amd import synth
    log("bench: generating %s" % args.workload)
    tasks = synth.twelve_class_tasks(seed=wl["seed"], n_test=wl["n_test"])
    L = lib.load()
    L.irwls_dev_init()
    probs = []
    for t, (X, y, d, n, Xt, yt) in enumerate(tasks):
        ds = lib.Dataset.dense(X, y)
        ds.pin(with_means=False)
        td = lib.Dataset.dense(Xt, yt)
        for C_ in (0.1, 1.0, 10.0, 100.0):
            for gm in (0.5, 2.0):
                probs.append(dict(ds=ds, td=td, yt=yt, props=abi.make_properties(gamma=gm / d, C_=C_,
                                  W=min(n // 4, 1500), eta=0.001), task=t, n=n, d=d))
    pp = abi.make_predict_properties(labels=0)

    def accuracies(betas):
        out = []
        for pr, b in zip(probs, betas):
            m = lib.full_model(pr["props"], pr["ds"], b)
            out.append(float(np.mean(lib.hard_test(pr["td"], m, pp) == pr["yt"])))
        return np.array(out)

    runs = {}
    sel = ["plain", "batched"] if args.mode == "both" else (["plain"] if args.mode == "plain" else ["batched"])
    for mode in sel:
        lib.set_options(precision=0 if mode == "plain" else 3)
        dom, tw = None, None
        for i in range(args.warmup):
            lib.timing(i == args.warmup - 1)
            t0 = time.time()
            if mode == "plain":
                for pr in probs:
                    lib.train_full(pr["ds"], pr["props"], seed=0)
            else:
                lib.train_full_batched([pr["ds"] for pr in probs], [pr["props"] for pr in probs])
            log("bench: %s warmup %d done in %.2f s" % (mode, i, time.time() - t0))
            if i == args.warmup - 1:
                tw = lib.timers()
                dom = dominant(tw)
                lib.timing(False)
        lib.timing(True, kinds=None if dom is None else [dom])
        times = []
        for i in range(args.steps):
            L.irwls_dev_sync()
            t0 = time.perf_counter()
            if mode == "plain":
                betas = [lib.train_full(pr["ds"], pr["props"], seed=0) for pr in probs]
            else:
                betas = lib.train_full_batched([pr["ds"] for pr in probs], [pr["props"] for pr in probs])
            L.irwls_dev_sync()
            times.append(time.perf_counter() - t0)
            log("bench: %s step %d: %.3f s %s" % (mode, i, times[-1], lib.last_stats() if mode != "plain" else ""))
        st = lib.last_stats() if mode != "plain" else None
        tm = lib.timers()
        lib.timing(False)
        lib.set_options(precision=0)
        if dom is None:
            dom = dominant(tm)
        runs[mode] = dict(value=float(np.mean(times)), betas=betas, acc=accuracies(betas), tm=tm, tw=tw, dom=dom,
                          stats=st)
    main_mode = "batched" if "batched" in runs else "plain"
    r = runs[main_mode]
    nmax = max(min(pr["n"] // 4, 1500) for pr in probs) + 1
    ms = (C.c_double * 3)()
    L.irwls_dev_batched_choice.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    out = {
        "metric": "train s to +-0.1% acc of FP64 oracle: mixed/batched/learned vs plain GPU",
        "value": r["value"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["value"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f64" if main_mode == "batched" else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "mode": main_mode, "problems": len(probs), "tasks": len(tasks),
                   "n": [t[3] for t in tasks], "d": [t[2] for t in tasks], "seed": wl["seed"],
                   "n_test_per_task": wl["n_test"]},
        "test_accuracy": float(r["acc"].mean()),
        "runs": {k: {"s": v["value"], "mean_test_accuracy": float(v["acc"].mean())} for k, v in runs.items()},
        "kernel_ms_per_step": {k: v["ms"] / args.steps for k, v in r["tm"].items() if v["launches"]},
        "warmup_kernel_ms": None if r["tw"] is None else {k: v["ms"] for k, v in r["tw"].items() if v["launches"]},
        "roofline": roofline(r, args.steps, args.workload, main_mode),
        "cpu_baseline": None,
    }
    if r["stats"]:
        out["runs"][main_mode].update({"batch_rounds": r["stats"]["batch_rounds"], "solves": r["stats"]["solves"]})
    if "plain" in runs and "batched" in runs:
        dacc = runs["batched"]["acc"] - runs["plain"]["acc"]
        out["speedup_vs_plain"] = runs["plain"]["value"] / runs["batched"]["value"]
        out["batched_vs_plain"] = {"accuracy_delta_mean": float(dacc.mean()), "accuracy_delta_max_abs":
                                   float(np.abs(dacc).max()), "within_0.1pct": int(np.sum(np.abs(dacc) <= 0.001))}
    print(json.dumps(out), flush=True)


def cpu_baseline_budgeted(wl, rp, col, val, y, props, r):
    """Oracle SGMA for its first cpu_sgma steps and IRWLSpar for one
    iteration on the host cores, extrapolated linearly to the GPU run's step
    and iteration counts (a lower bound: later SGMA steps cost more)."""
    from oracle import pyoracle as po
    ods = po.Dataset.csr(rp, col, val, y)
    k = wl["cpu_sgma"]
    po.set_limits(0, k, 1)
    try:
        t0 = time.time()
        cent = po.sgma(ods, props, ncand=wl["cand"], seed=1)
        t1 = time.time()
        po.irwls_par(ods, np.asarray(cent), props)
        t2 = time.time()
    finally:
        po.set_limits(0, 0, 0)
    steps = max(props.size - 2, 1)
    per_step = (t1 - t0) / k
    per_iter = t2 - t1
    iters = max(r["par_stats"]["irwlspar_iters"], 1)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": per_step * steps + per_iter * iters, "unit": "s", "cores": cores, "kind": "port",
            "sample": "oracle SGMA first %d steps in %.1f s and IRWLSpar first iteration in %.1f s, extrapolated to "
                      "%d steps and %d iterations (lower bound)" % (k, t1 - t0, t2 - t1, steps, iters)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="four_gaussians_n50000_d128", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="both", choices=["plain", "mixed", "learned", "batched", "both"],
                    help="plain FP64; mixed precision (full training) or learned proposal (budgeted); "
                         "both: value = the mixed / learned trainer, plain reported beside it")
    ap.add_argument("--proposal-rows", type=int, default=None,
                    help="budgeted, learned mode: rows on which the shortlist's gains are computed "
                         "(0 = all rows; default: the workload's own setting)")
    ap.add_argument("--kernel-matrix-probe", action="store_true",
                    help="only time the kernel-matrix products (FP64 and BF16 x3 MFMA) at the mixed workload's shape")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the traffic-counter calibration copy (for rocprofv3 --pmc)")
    ap.add_argument("--cpu-outer", type=int, default=None,
                    help="outer iterations in the CPU sample (0 = skip, -1 = the whole run; "
                         "default: the workload's own setting)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.kernel_matrix_probe:
        kernel_matrix_probe()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from libirwls_amd import abi, lib
    wl = WORKLOADS[args.workload]
    if wl.get("kind") == "budgeted":
        run_budgeted(args, wl, lib, abi)
        return
    if wl.get("kind") == "batched":
        run_batched(args, wl, lib, abi)
        return
    cpu_outer = wl.get("cpu_outer", -1) if args.cpu_outer is None else args.cpu_outer
    log("bench: generating %s" % args.workload)
    X, y, Xt, yt = make_data(wl)
    props = abi.make_properties(gamma=wl["gamma"], C_=wl["C"], W=wl["W"], eta=wl["eta"])
    L = lib.load()
    L.irwls_dev_init()
    ds = lib.Dataset.dense(X, y)
    ds.pin(with_means=False)            # inputs resident in HBM before timing
    td = lib.Dataset.dense(Xt, yt)
    pp = abi.make_predict_properties(labels=0)

    def evaluate(r):
        m = lib.full_model(props, ds, r["beta"])
        f = lib.soft_test(td, m, pp)
        r["decision"] = f
        r["acc"] = float(np.mean(np.where(f >= 0, 1.0, -1.0) == yt))
        r["nsv"] = int(np.count_nonzero(r["beta"][:-1]))

    runs = {}
    for mode in (["plain", "mixed"] if args.mode == "both" else [args.mode]):
        runs[mode] = run_mode(lib, L, ds, props, 1 if mode == "mixed" else 0, args.warmup, args.steps)
        evaluate(runs[mode])
    main_mode = "mixed" if "mixed" in runs else "plain"
    r = runs[main_mode]

    cpu = None
    if cpu_outer != 0:
        log("bench: cpu baseline (oracle, %s outer iterations)" % ("all" if cpu_outer < 0 else cpu_outer))
        cpu = cpu_baseline(wl, X, y, runs.get("plain", r)["stats"]["outer_iters"], cpu_outer)

    def summary(x):
        return {"s": x["value"], "test_accuracy": x["acc"], "n_sv": x["nsv"], "outer_iters": x["stats"]["outer_iters"],
                "inner_iters": x["stats"]["inner_iters"], "solve_s": x["stats"]["t_solve"],
                "factorizations": x["stats"]["factorizations"], "pcg_iters": x["stats"]["pcg_iters"],
                "fp64_fallbacks": x["stats"]["fp64_fallbacks"]}

    out = {
        "metric": "train s to +-0.1% acc of FP64 oracle: mixed/batched/learned vs plain GPU",
        "value": r["value"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["value"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+f32+f64" if main_mode == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "mode": main_mode, "n": wl["n"], "d": wl["d"], "W": wl["W"],
                   "gamma": wl["gamma"], "C": wl["C"], "eta": wl["eta"], "n_test": wl["n_test"],
                   "seed": wl["seed"]},
        "test_accuracy": r["acc"],
        "runs": {k: summary(v) for k, v in runs.items()},
        "kernel_ms_per_step": {k: v["ms"] / args.steps for k, v in r["tm"].items() if v["launches"]},
        "warmup_kernel_ms": None if r["tw"] is None else {k: v["ms"] for k, v in r["tw"].items() if v["launches"]},
        "roofline": roofline(r, args.steps, args.workload, main_mode),
        "cpu_baseline": cpu,
    }
    if "plain" in runs and "mixed" in runs:
        p, m = runs["plain"], runs["mixed"]
        out["speedup_vs_plain"] = p["value"] / m["value"]
        out["mixed_vs_plain"] = {
            "accuracy_delta": m["acc"] - p["acc"],
            "weight_rel_err": float(np.linalg.norm(m["beta"] - p["beta"]) / np.linalg.norm(p["beta"])),
            "decision_max_err": float(np.abs(m["decision"] - p["decision"]).max()),
            "decision_rms_err": float(np.sqrt(np.mean((m["decision"] - p["decision"]) ** 2))),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
