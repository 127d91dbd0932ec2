#!/usr/bin/env python3
"""Benchmark of the IRWLS training hot path on one MI355X.

The metric is training seconds of the mixed-precision, learned-proposal
and batched trainers against the plain FP64 GPU trainer, on the
benchmark's three GPU workloads, in one run:

- four_gaussians_n200000_d256 (the headline, `value`): trainFULL through
  the C entry point, n=200000, d=256, RBF gamma=1/256, C=1, working set
  4000, eta=0.001; mixed precision (BF16 hi/lo x3 kernel products, FP32
  Cholesky with FP64 diagonal tiles, FP64 refinement to 1e-10) against
  plain FP64 on the same data resident in HBM.  A step is one training run.
- sparse_rbf_n1000000_d2000: budgeted training (SGMA + IRWLSpar, the
  reference budgeted-train's timed window), 10^6 CSR rows, d=2000, 1%
  density, budget 1000, 59 candidates, gamma 0.5, C 10: the learned
  proposal (gains on 65536 rows, and with exact gains) against plain SGMA.
- ovr12_grid8_batched: 96 problems (12 one-vs-rest tasks x 8 (C, gamma)
  points): trainFULLBatched with BF16 hi/lo kernel products against 96
  sequential plain FP64 trainFULL runs.

Each workload block reports both modes' seconds, the speed-up, the
accuracy difference, a roofline object for its dominant kernel (found in
the last warmup step with every timer on; the timed steps record HIP
events on the library stream around that kernel's launches only) and a
CPU baseline: the oracle (the FP64 C restatement of the reference, OpenMP
over the host cores) on a bounded sample of the same workload,
extrapolated and labelled as such.  The headline uses --steps/--warmup;
the other two blocks use --side-steps/--side-warmup.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload NAME]
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

METRIC = "train s to +-0.1% acc of FP64 oracle: mixed/batched/learned vs plain GPU"
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
                                      C=10.0, B=1000, cand=59, seed=3, cpu_sgma=6, cpu_par_rows=50000,
                                      proposal_rows=65536),
    "sparse_rbf_n100000_d2000": dict(kind="budgeted", n=100000, n_test=20000, d=2000, density=0.01, gamma=0.5,
                                     C=10.0, B=500, cand=59, seed=3, cpu_sgma=6, cpu_par_rows=50000),
    # batched workload: 12 one-vs-rest tasks of a 12-class mixture (n log-uniform in [1000, 30000], d in
    # {16, 32, 64, 128}) x 8 points C in {0.1, 1, 10, 100}, gamma in {0.5/d, 2/d}; W = min(n/4, 1500)
    "ovr12_grid8_batched": dict(kind="batched", n_test=2000, seed=4, n=None, d=None),
}
HEADLINE = "four_gaussians_n200000_d256"
SIDE = ["sparse_rbf_n1000000_d2000", "ovr12_grid8_batched"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def make_data(wl):
    from libirwls_amd import synth
    if wl["d"] == 16:
        X, y = synth.two_gaussians(wl["n"], wl["d"], seed=wl["seed"])
        Xt, yt = synth.two_gaussians(wl["n_test"], wl["d"], seed=wl["seed"] + 100)
    else:
        X, y = synth.four_gaussians(wl["n"], wl["d"], seed=wl["seed"], means_seed=wl["seed"])
        Xt, yt = synth.four_gaussians(wl["n_test"], wl["d"], seed=wl["seed"] + 100, means_seed=wl["seed"])
    return X, y, Xt, yt


def timed_steps(lib, L, fn, warmup, steps, label):
    """W warmup calls (the last with every timer on, to name the dominant
    kernel kind), then K timed calls with events around that kind only.
    Returns (times, last result, timers of the timed calls, warmup timers,
    dominant kind)."""
    dom, tw = None, None
    for i in range(warmup):
        last = (i == warmup - 1)
        if last:
            lib.timing(True)
        t0 = time.time()
        fn()
        log("bench: %s warmup %d done in %.2f s" % (label, i, time.time() - t0))
        if last:
            tw = lib.timers()
            dom = dominant(tw)
            lib.timing(False)
    # events on every stride-th launch of the dominant kind only: each timed
    # launch carries a start and a stop event, which cost a few microseconds
    # (5-7% of a headline training run when every step kernel carried them)
    stride = 1
    if dom is not None and tw is not None and tw[dom]["launches"]:
        stride = max(1, min(8, tw[dom]["launches"] // 64))
    lib.timing(True, kinds=None if dom is None else [dom], stride=stride)
    times, res = [], None
    for i in range(steps):
        L.irwls_dev_sync()
        t0 = time.perf_counter()
        res = fn()
        L.irwls_dev_sync()
        times.append(time.perf_counter() - t0)
        log("bench: %s step %d: %.3f s" % (label, i, times[-1]))
    tm = lib.timers()
    lib.timing(False)
    for v in tm.values():
        v["stride"] = stride
    if dom is None:
        dom = dominant(tm)
    return times, res, tm, tw, dom


def kernel_tables(r, steps):
    return {"kernel_ms_per_step": {k: v["ms"] * v.get("stride", 1) / steps for k, v in r["tm"].items()
                                   if v["launches"]},
            "warmup_kernel_ms": None if r["tw"] is None else {k: v["ms"] for k, v in r["tw"].items()
                                                              if v["launches"]}}


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
    calibrated against the copy probe (its line `traffic_bytes_per_launch
    <kind>: <bytes>`); None when there is no such figure."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_%s_%s.md" % (mode, workload))
    try:
        with open(path) as f:
            m = re.search(r"^traffic_bytes_per_launch %s:\s*([0-9.eE+]+)" % re.escape(kind), f.read(), re.M)
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
    stride = t.get("stride", 1)
    roof["timer_stride"] = stride   # events on every stride-th launch of the kind in the timed runs
    roof["launches_per_step"] = t["launches"] * stride / steps
    roof["share_of_step"] = (t["ms"] * stride / steps) / (r["value"] * 1e3)
    return roof


def predict_figure(lib, L, td, m, pp, n_sv, reps=3):
    """softTest over the test set with the test rows pinned in HBM: time per
    call, and the rates of its kernel-block x vector product (HIP-event
    timed): algorithmic bytes (test rows, support vectors, output) per
    second against the HBM peak and flops per second against the FP64
    MFMA peak."""
    td.pin(with_means=False)
    lib.soft_test(td, m, pp)
    lib.timing(True)
    L.irwls_dev_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.soft_test(td, m, pp)
    L.irwls_dev_sync()
    dt = (time.perf_counter() - t0) / reps
    tm = lib.timers()
    lib.timing(False)
    td.unpin()
    kind = max(("kmv", "csr"), key=lambda k: tm[k]["ms"])
    t = tm[kind]
    if not t["launches"]:
        return None
    sec = t["ms"] * 1e-3
    gbs = t["bytes"] / sec / 1e9
    tfl = t["flops"] / sec / 1e12
    return {"rows": td.ds.l, "n_sv": n_sv, "s_per_call": dt, "rows_per_s": td.ds.l / dt, "kernel": kind,
            "launches_per_call": t["launches"] / reps, "kernel_ms_per_call": t["ms"] / reps,
            "algorithmic_gbs": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
            "tflops": tfl, "fp64_mfma_frac": tfl / PEAK_FP64_TFLOPS}


def cpu_baseline_full(wl, X, y, gpu_outer, max_outer):
    """Oracle trainFULL on the host cores: the whole run (max_outer < 0) or
    its first max_outer outer iterations, extrapolated to the GPU plain
    run's outer-iteration count."""
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
    whole = max_outer < 0
    return {
        "value": dt if whole else per_outer * gpu_outer, "unit": "s", "cores": cores(), "kind": "port",
        "cpu": cpu_model(),
        "sample": ("oracle trainFULL, whole run: %d outer / %d inner iterations in %.1f s"
                   % (st["outer_iters"], st["inner_iters"], dt)) if whole else
                  ("oracle trainFULL, first %d outer iteration(s) (%d inner) in %.1f s; %.2f s/outer x %d GPU "
                   "outer iterations (extrapolated)" % (st["outer_iters"], st["inner_iters"], dt, per_outer,
                                                        gpu_outer)),
    }


def bench_full(args, name, steps, warmup):
    """Plain FP64 and mixed-precision trainFULL on one dense workload, the
    dataset pinned in HBM; then the drop-in call (host dataset converted
    and uploaded inside the call, as a caller of the reference API makes
    it) of the mixed trainer."""
    from libirwls_amd import abi, lib
    wl = WORKLOADS[name]
    log("bench: generating %s" % name)
    X, y, Xt, yt = make_data(wl)
    props = abi.make_properties(gamma=wl["gamma"], C_=wl["C"], W=wl["W"], eta=wl["eta"])
    L = lib.load()
    L.irwls_dev_init()
    ds = lib.Dataset.dense(X, y)
    ds.pin(with_means=False)            # inputs resident in HBM before timing
    td = lib.Dataset.dense(Xt, yt)
    pp = abi.make_predict_properties(labels=0)
    modes = [args.mode] if args.mode in ("plain", "mixed") else ["plain", "mixed"]
    runs = {}
    for mode in modes:
        prec = 1 if mode == "mixed" else 0
        lib.set_options(precision=prec)
        times, beta, tm, tw, dom = timed_steps(lib, L, lambda: lib.train_full(ds, props, seed=0), warmup, steps,
                                               "%s %s" % (name, mode))
        st = lib.last_stats()
        lib.set_options(precision=0)
        m = lib.full_model(props, ds, beta)
        f = lib.soft_test(td, m, pp)
        runs[mode] = dict(times=times, value=float(np.mean(times)), beta=beta, stats=st, tm=tm, tw=tw, dom=dom,
                          decision=f, acc=float(np.mean(np.where(f >= 0, 1.0, -1.0) == yt)),
                          nsv=int(np.count_nonzero(beta[:-1])))
    main_mode = "mixed" if "mixed" in runs else "plain"
    r = runs[main_mode]
    pred = predict_figure(lib, L, td, lib.full_model(props, ds, r["beta"]), pp, r["nsv"])
    log("bench: %s predict: %s" % (name, pred))
    # drop-in: the same call on an unpinned dataset (AoS conversion and upload inside the timed call)
    ds.unpin()
    lib.set_options(precision=1 if main_mode == "mixed" else 0)
    drop = []
    for i in range(2):
        L.irwls_dev_sync()
        t0 = time.perf_counter()
        lib.train_full(ds, props, seed=0)
        L.irwls_dev_sync()
        drop.append(time.perf_counter() - t0)
    lib.set_options(precision=0)
    log("bench: %s drop-in (unpinned) %s: %s" % (name, main_mode, ["%.3f" % t for t in drop]))
    cpu_outer = wl.get("cpu_outer", -1) if args.cpu_outer is None else args.cpu_outer
    cpu = None
    if cpu_outer != 0:
        log("bench: cpu baseline (oracle, %s outer iterations)" % ("all" if cpu_outer < 0 else cpu_outer))
        cpu = cpu_baseline_full(wl, X, y, runs.get("plain", r)["stats"]["outer_iters"], cpu_outer)

    def summary(x):
        return {"s": x["value"], "test_accuracy": x["acc"], "n_sv": x["nsv"], "outer_iters": x["stats"]["outer_iters"],
                "inner_iters": x["stats"]["inner_iters"], "solve_s": x["stats"]["t_solve"],
                "factorizations": x["stats"]["factorizations"], "pcg_iters": x["stats"]["pcg_iters"],
                "fp64_fallbacks": x["stats"]["fp64_fallbacks"]}
    block = {
        "workload": name, "mode": main_mode, "s": r["value"], "steps": steps, "warmup": warmup,
        "dtype": "bf16x3+f32+f64" if main_mode == "mixed" else "f64",
        "config": {"n": wl["n"], "d": wl["d"], "W": wl["W"], "gamma": wl["gamma"], "C": wl["C"], "eta": wl["eta"],
                   "n_test": wl["n_test"], "seed": wl["seed"]},
        "test_accuracy": r["acc"],
        "runs": {k: summary(v) for k, v in runs.items()},
        "drop_in_s": float(np.mean(drop)),
        "predict": pred,
        "roofline": roofline(r, steps, name, main_mode),
        "cpu_baseline": cpu,
    }
    block.update(kernel_tables(r, steps))
    if "plain" in runs and "mixed" in runs:
        p, m = runs["plain"], runs["mixed"]
        block["plain_s"] = p["value"]
        block["speedup_vs_plain"] = p["value"] / m["value"]
        block["accuracy_delta"] = m["acc"] - p["acc"]
        block["mixed_vs_plain"] = {
            "accuracy_delta": m["acc"] - p["acc"],
            "weight_rel_err": float(np.linalg.norm(m["beta"] - p["beta"]) / np.linalg.norm(p["beta"])),
            "decision_max_err": float(np.abs(m["decision"] - p["decision"]).max()),
            "decision_rms_err": float(np.sqrt(np.mean((m["decision"] - p["decision"]) ** 2))),
        }
    return block


def cpu_baseline_budgeted(wl, rp, col, val, y, props, gpu_steps, gpu_iters):
    """Oracle SGMA (the reference's 64-candidate loop at `cand` candidates)
    up to budget size cpu_sgma, and one IRWLSpar iteration on the first
    cpu_par_rows rows, on the host cores; extrapolated linearly to the GPU
    run's SGMA steps and IRWLSpar iterations over all rows.  A lower bound:
    an SGMA step's product with the basis grows with the basis size."""
    from oracle import pyoracle as po
    ods = po.Dataset.csr(rp, col, val, y)
    po.set_limits(0, wl["cpu_sgma"], 1)
    try:
        t0 = time.time()
        po.sgma(ods, props, ncand=wl["cand"], seed=1)
        t1 = time.time()
        ksteps = po.stats()["sgma_steps"]
        m = min(wl.get("cpu_par_rows", len(y)), len(y))
        sub = po.Dataset.csr(rp[:m + 1], col[:rp[m]], val[:rp[m]], y[:m])
        cent = np.arange(props.size, dtype=np.int32) % m
        t2 = time.time()
        po.irwls_par(sub, cent, props)
        t3 = time.time()
    finally:
        po.set_limits(0, 0, 0)
    n = len(y)
    per_step = (t1 - t0) / max(ksteps, 1)
    per_iter = (t3 - t2) * n / m
    return {"value": per_step * gpu_steps + per_iter * gpu_iters, "unit": "s", "cores": cores(), "kind": "port",
            "cpu": cpu_model(),
            "sample": "oracle SGMA %d steps in %.1f s (%.2f s/step x %d steps) and one IRWLSpar iteration on "
                      "%d of %d rows in %.1f s (x %d rows/%d x %d iterations), extrapolated (lower bound)"
                      % (ksteps, t1 - t0, per_step, gpu_steps, m, n, t3 - t2, n, m, gpu_iters)}


def bench_budgeted(args, name, steps, warmup):
    """SGMA + IRWLSpar with the plain and the learned proposal on the same
    CSR data, and the budgeted objectives of each mode's model: the
    held-out Nystrom residual of SGMA's basis and IRWLSpar's primal
    objective on the training rows."""
    from libirwls_amd import abi, lib, objective, synth
    wl = WORKLOADS[name]
    log("bench: generating %s" % name)
    rp, col, val, y = synth.sparse_rbf_labels(wl["n"] + wl["n_test"], wl["d"], wl["density"], gamma=wl["gamma"],
                                              seed=wl["seed"])
    n = wl["n"]
    tr = rp[:n + 1]
    te = rp[n:] - rp[n]
    ds = lib.Dataset.csr(tr, col[:rp[n]], val[:rp[n]], y[:n])
    td = lib.Dataset.csr(te, col[rp[n]:], val[rp[n]:], y[n:])
    yt = y[n:]
    nh = min(2000, wl["n_test"])
    hold = lib.Dataset.csr(te[:nh + 1], col[rp[n]:rp[n] + te[nh]], val[rp[n]:rp[n] + te[nh]], yt[:nh])
    L = lib.load()
    L.irwls_dev_init()
    ds.pin(with_means=True)
    props = abi.make_properties(gamma=wl["gamma"], C_=wl["C"], size=wl["B"])
    pp = abi.make_predict_properties(labels=0)
    prow = args.proposal_rows if args.proposal_rows is not None else wl.get("proposal_rows", 0)
    modes = {"plain": dict(sgma_candidates=wl["cand"], proposal=0),
             "learned": dict(sgma_candidates=wl["cand"], proposal=1, proposal_keep=8,
                             proposal_pool=args.proposal_pool if args.proposal_pool else wl["cand"],
                             proposal_rows=prow)}
    if prow:
        # the network shortlists 16 of plain SGMA's own candidates (the same
        # rand() draws), whose exact gains on all rows decide
        modes["learned_exact_gains"] = dict(modes["learned"], proposal_pool=wl["cand"], proposal_keep=16,
                                            proposal_rows=0)
    if args.mode in ("plain", "learned"):
        modes = {args.mode: modes[args.mode]}
    reset = dict(sgma_candidates=64, proposal=0, proposal_keep=8, proposal_pool=0, proposal_rows=0)
    runs = {}
    for mode, opts in modes.items():
        state = {}

        def step():
            lib.set_options(**opts)
            t0 = time.perf_counter()
            cent = lib.sgma(ds, props, seed=1)
            state["sgma_s"] = time.perf_counter() - t0
            state["sst"] = lib.last_stats()
            beta = lib.irwls_par(ds, cent, props)
            state["pst"] = lib.last_stats()
            return cent, beta
        nst = steps if mode != "learned_exact_gains" else max(1, min(steps, 1))
        times, (cent, beta), tm, tw, dom = timed_steps(lib, L, step, warmup if mode != "learned_exact_gains" else 0,
                                                       nst, "%s %s" % (name, mode))
        lib.set_options(**reset)
        m = lib.budgeted_model(props, ds, cent, beta)
        acc = float(np.mean(lib.hard_test(td, m, pp) == yt))
        ftr = lib.soft_test(ds, m, pp)
        runs[mode] = dict(value=float(np.mean(times)), sgma_s=state["sgma_s"], acc=acc, tm=tm, tw=tw, dom=dom, cent=cent,
                          sgma_stats=state["sst"], par_stats=state["pst"], steps=nst, model=m,
                          resid=objective.nystrom_residual(ds, cent, hold, wl["gamma"]),
                          primal=objective.primal_objective(ds, cent, beta, ftr, wl["C"], wl["gamma"]))
        log("bench: %s %s: %.3f s, accuracy %.5f, held-out Nystrom residual %.6f, primal %.8g"
            % (name, mode, runs[mode]["value"], acc, runs[mode]["resid"], runs[mode]["primal"]))
    ds.unpin()
    main_mode = "learned" if "learned" in runs else "plain"
    r = runs[main_mode]
    pred = predict_figure(lib, L, td, r["model"], pp, wl["B"])
    log("bench: %s predict: %s" % (name, pred))
    cpu = None
    if args.cpu_outer != 0:
        log("bench: cpu baseline (oracle SGMA sample and IRWLSpar row sample)")
        p = runs.get("plain", r)
        cpu = cpu_baseline_budgeted(wl, tr, col[:rp[n]], val[:rp[n]], y[:n], props,
                                    p["sgma_stats"]["sgma_steps"], p["par_stats"]["irwlspar_iters"])
    block = {
        "workload": name, "mode": main_mode, "s": r["value"], "steps": steps, "warmup": warmup,
        "dtype": "f64+f32-mlp+bf16-mlp",
        "config": {"n": n, "d": wl["d"], "B": wl["B"], "candidates": wl["cand"], "proposal_rows": prow,
                   "gamma": wl["gamma"], "C": wl["C"], "n_test": wl["n_test"], "seed": wl["seed"],
                   "heldout_rows_for_residual": nh},
        "test_accuracy": r["acc"],
        "runs": {k: {"s": v["value"], "steps": v["steps"], "sgma_s": v["sgma_s"], "test_accuracy": v["acc"],
                     "heldout_nystrom_residual": v["resid"], "primal_objective": v["primal"],
                     "sgma_steps": v["sgma_stats"]["sgma_steps"], "learned_steps": v["sgma_stats"]["learned_steps"],
                     "mlp_loss0": v["sgma_stats"]["mlp_loss0"], "mlp_loss": v["sgma_stats"]["mlp_loss"],
                     "irwlspar_iters": v["par_stats"]["irwlspar_iters"]} for k, v in runs.items()},
        "predict": pred,
        "roofline": roofline(r, steps, name, main_mode),
        "cpu_baseline": cpu,
    }
    if "plain" in runs:
        block.update(plain_s=runs["plain"]["value"], speedup_vs_plain=runs["plain"]["value"] / r["value"],
                     accuracy_delta=r["acc"] - runs["plain"]["acc"])
    block.update(kernel_tables(r, steps))
    if "learned_exact_gains" in runs and "plain" in runs:
        block["exact_gains_speedup_vs_plain"] = runs["plain"]["value"] / runs["learned_exact_gains"]["value"]
        block["exact_gains_accuracy_delta"] = runs["learned_exact_gains"]["acc"] - runs["plain"]["acc"]
        block["exact_gains_config"] = {"pool": "plain SGMA's %d candidates" % wl["cand"], "keep": 16,
                                       "gains_on_rows": n,
                                       "centroids_shared_with_plain": float(len(
                                           set(np.asarray(runs["learned_exact_gains"]["cent"]).tolist())
                                           & set(np.asarray(runs["plain"]["cent"]).tolist())) / wl["B"])}
    if args.accuracy_seeds > 0:
        block["accuracy_over_seeds"] = budgeted_seed_sweep(wl, args.accuracy_seeds, modes)
    return block


def budgeted_seed_sweep(wl, nseeds, modes):
    """Test accuracy and primal objective of each budgeted mode on nseeds
    further data sets of the workload's generator (seeds seed+1 ...), one
    untimed run per mode and seed."""
    from libirwls_amd import abi, lib, objective, synth
    pp = abi.make_predict_properties(labels=0)
    props = abi.make_properties(gamma=wl["gamma"], C_=wl["C"], size=wl["B"])
    reset = dict(sgma_candidates=64, proposal=0, proposal_keep=8, proposal_pool=0, proposal_rows=0)
    out = {m: [] for m in modes}
    n = wl["n"]
    for k in range(1, nseeds + 1):
        seed = wl["seed"] + k
        log("bench: accuracy sweep, data seed %d" % seed)
        rp, col, val, y = synth.sparse_rbf_labels(n + wl["n_test"], wl["d"], wl["density"], gamma=wl["gamma"],
                                                  seed=seed)
        ds = lib.Dataset.csr(rp[:n + 1], col[:rp[n]], val[:rp[n]], y[:n])
        te = rp[n:] - rp[n]
        td = lib.Dataset.csr(te, col[rp[n]:], val[rp[n]:], y[n:])
        ds.pin(with_means=True)
        for mode, opts in modes.items():
            lib.set_options(**opts)
            cent = lib.sgma(ds, props, seed=1)
            beta = lib.irwls_par(ds, cent, props)
            lib.set_options(**reset)
            m = lib.budgeted_model(props, ds, cent, beta)
            acc = float(np.mean(lib.hard_test(td, m, pp) == y[n:]))
            primal = objective.primal_objective(ds, cent, beta, lib.soft_test(ds, m, pp), wl["C"], wl["gamma"])
            out[mode].append({"seed": seed, "test_accuracy": acc, "primal_objective": primal})
            log("bench: seed %d %s: accuracy %.5f, primal %.8g" % (seed, mode, acc, primal))
        ds.unpin()
    summary = {"seeds": [wl["seed"] + k for k in range(1, nseeds + 1)], "runs": out}
    if "plain" in out:
        for mode in out:
            if mode == "plain":
                continue
            d = [a["test_accuracy"] - b["test_accuracy"] for a, b in zip(out[mode], out["plain"])]
            summary[mode + "_accuracy_delta_mean"] = float(np.mean(d))
            summary[mode + "_accuracy_delta_max_abs"] = float(np.max(np.abs(d)))
    return summary


def cpu_baseline_batched(probs):
    """Oracle trainFULL on the host cores for one problem per task (the
    (C, gamma) = (1, 0.5/d) grid point), whole runs, scaled by the 8 grid
    points per task."""
    from oracle import pyoracle as po
    from libirwls_amd import abi
    t0 = time.time()
    k = 0
    seen = set()
    for pr in probs:
        if pr["task"] in seen or pr["C"] != 1.0 or pr["gm"] != 0.5:
            continue
        seen.add(pr["task"])
        po.train_full(po.Dataset.dense(pr["X"], pr["y"]), pr["props"], seed=0)
        k += 1
    dt = time.time() - t0
    return {"value": dt * len(probs) / max(k, 1), "unit": "s", "cores": cores(), "kind": "port", "cpu": cpu_model(),
            "sample": "oracle trainFULL whole runs of %d problems (one per task, C=1, gamma=0.5/d) in %.1f s, "
                      "x %d/%d (extrapolated)" % (k, dt, len(probs), k)}


def bench_batched(args, name, steps, warmup):
    """96 problems: sequential plain FP64 trainFULL runs against
    trainFULLBatched with BF16 hi/lo kernel products and batched FP64
    solves; each task's dataset pinned once and shared by its 8 problems."""
    from libirwls_amd import abi, lib, synth
    wl = WORKLOADS[name]
    log("bench: generating %s" % name)
    tasks = synth.twelve_class_tasks(seed=wl["seed"], n_test=wl["n_test"])
    L = lib.load()
    L.irwls_dev_init()
    probs = []
    keep = []
    for t, (X, y, d, n, Xt, yt) in enumerate(tasks):
        ds = lib.Dataset.dense(X, y)
        ds.pin(with_means=False)
        td = lib.Dataset.dense(Xt, yt)
        keep.append(ds)
        for C_ in (0.1, 1.0, 10.0, 100.0):
            for gm in (0.5, 2.0):
                probs.append(dict(ds=ds, td=td, yt=yt, X=X, y=y, C=C_, gm=gm, task=t, n=n, d=d,
                                  props=abi.make_properties(gamma=gm / d, C_=C_, W=min(n // 4, 1500), eta=0.001)))
    pp = abi.make_predict_properties(labels=0)

    def accuracies(betas):
        out = []
        for pr, b in zip(probs, betas):
            m = lib.full_model(pr["props"], pr["ds"], b)
            out.append(float(np.mean(lib.hard_test(pr["td"], m, pp) == pr["yt"])))
        return np.array(out)

    runs = {}
    for mode in ("plain", "batched"):
        lib.set_options(precision=0 if mode == "plain" else 3)
        if mode == "plain":
            fn = lambda: [lib.train_full(pr["ds"], pr["props"], seed=0) for pr in probs]
        else:
            fn = lambda: lib.train_full_batched([pr["ds"] for pr in probs], [pr["props"] for pr in probs])
        times, betas, tm, tw, dom = timed_steps(lib, L, fn, warmup, steps, "%s %s" % (name, mode))
        st = lib.last_stats() if mode != "plain" else None
        lib.set_options(precision=0)
        runs[mode] = dict(value=float(np.mean(times)), betas=betas, acc=accuracies(betas), tm=tm, tw=tw, dom=dom,
                          stats=st)
    for ds in keep:
        ds.unpin()
    r = runs["batched"]
    cpu = None
    if args.cpu_outer != 0:
        log("bench: cpu baseline (oracle, one problem per task)")
        cpu = cpu_baseline_batched(probs)
    dacc = r["acc"] - runs["plain"]["acc"]
    block = {
        "workload": name, "mode": "batched", "s": r["value"], "steps": steps, "warmup": warmup,
        "dtype": "bf16x3+f64",
        "config": {"problems": len(probs), "tasks": len(tasks), "n": [t[3] for t in tasks], "d": [t[2] for t in tasks],
                   "seed": wl["seed"], "n_test_per_task": wl["n_test"]},
        "test_accuracy": float(r["acc"].mean()),
        "runs": {k: {"s": v["value"], "mean_test_accuracy": float(v["acc"].mean())} for k, v in runs.items()},
        "roofline": roofline(r, steps, name, "batched"),
        "cpu_baseline": cpu,
        "plain_s": runs["plain"]["value"],
        "speedup_vs_plain": runs["plain"]["value"] / r["value"],
        "accuracy_delta": float(dacc.mean()),
        "batched_vs_plain": {"accuracy_delta_mean": float(dacc.mean()),
                             "accuracy_delta_max_abs": float(np.abs(dacc).max()),
                             "within_0.1pct": int(np.sum(np.abs(dacc) <= 0.001))},
    }
    if r["stats"]:
        block["runs"]["batched"].update({"batch_rounds": r["stats"]["batch_rounds"], "solves": r["stats"]["solves"]})
    block.update(kernel_tables(r, steps))
    return block


def contract_line(block, all_blocks, args):
    """The benchmark line: `value` and `roofline` of the headline block,
    every workload's block beside it."""
    out = {
        "metric": METRIC, "value": block["s"], "unit": "s", "n_gpus": 1, "steps": block["steps"],
        "warmup": block["warmup"], "ms_per_step": block["s"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": block["dtype"], "data": "synthetic",
        "config": dict({"workload": block["workload"], "mode": block["mode"]}, **block["config"]),
        "test_accuracy": block["test_accuracy"],
        "speedup_vs_plain": block.get("speedup_vs_plain"),
        "accuracy_delta": block.get("accuracy_delta"),
        "roofline": block["roofline"],
        "cpu_baseline": block["cpu_baseline"],
        "workloads": {b["workload"]: b for b in all_blocks},
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side-steps", type=int, default=2, help="timed steps of the two other workloads")
    ap.add_argument("--side-warmup", type=int, default=1)
    ap.add_argument("--workload", default="all", choices=["all"] + sorted(WORKLOADS),
                    help="all (default): the headline four_gaussians_n200000_d256 plus the budgeted and batched "
                         "workloads; or one workload alone, as the headline")
    ap.add_argument("--mode", default="both", choices=["plain", "mixed", "learned", "both"],
                    help="dense workloads: plain FP64, mixed precision, or both (value = the mixed trainer); "
                         "budgeted workload: plain or learned proposal alone, or every mode (both)")
    ap.add_argument("--proposal-rows", type=int, default=None,
                    help="budgeted, learned mode: rows on which the shortlist's gains are computed "
                         "(0 = all rows; default: the workload's own setting)")
    ap.add_argument("--proposal-pool", type=int, default=None,
                    help="budgeted, learned mode: candidates the network scores per step "
                         "(default: the workload's candidates, plain SGMA's own draws)")
    ap.add_argument("--kernel-matrix-probe", action="store_true",
                    help="only time the kernel-matrix products (FP64 and BF16 x3 MFMA) at the mixed workload's shape")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the traffic-counter calibration copy (for rocprofv3 --pmc)")
    ap.add_argument("--accuracy-seeds", type=int, default=0,
                    help="budgeted workload: also report accuracy and objective of each mode on this many further "
                         "data seeds (untimed)")
    ap.add_argument("--cpu-outer", type=int, default=None,
                    help="CPU baselines: 0 skips them; dense workloads: outer iterations in the sample "
                         "(-1 = the whole run; default: the workload's own setting)")
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

    def run(name, steps, warmup):
        kind = WORKLOADS[name].get("kind")
        if kind == "budgeted":
            return bench_budgeted(args, name, steps, warmup)
        if kind == "batched":
            return bench_batched(args, name, steps, warmup)
        return bench_full(args, name, steps, warmup)

    if args.workload == "all":
        head = run(HEADLINE, args.steps, args.warmup)
        blocks = [head] + [run(s, args.side_steps, args.side_warmup) for s in SIDE]
    else:
        head = run(args.workload, args.steps, args.warmup)
        blocks = [head]
    print(json.dumps(contract_line(head, blocks, args)), flush=True)


if __name__ == "__main__":
    main()
