"""GPU tests of batched training: the batched factorization and solve
against the single-system solve (same arithmetic, so equal bit for bit),
replays of its captured launch sequence, its autotuned tile shape, and
trainFULLBatched against standalone trainFULL runs (srand(0) each)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import Dev, rbf
from libirwls_amd import abi, lib, synth

pytestmark = pytest.mark.gpu


def _bind():
    L = lib.load()
    pp = C.POINTER(C.c_void_p)
    L.irwls_dev_posv_aug_batched.argtypes = [C.c_int, pp, C.POINTER(C.c_int), C.POINTER(C.c_int), pp, pp,
                                             C.POINTER(C.c_int)]
    L.irwls_dev_batched_choice.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.irwls_dev_posv_aug.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    return L


def _kkt(n, rng):
    """KKT matrix [A y; y^T 0] of size n (A from an RBF block), augmented
    with a zero row n; and a right-hand side."""
    m = n - 1
    H = np.zeros((n + 1, n))
    if m > 0:
        X = rng.standard_normal((m, 5))
        y = rng.choice([-1.0, 1.0], m)
        H[:m, :m] = rbf(X, X, 0.3) * np.outer(y, y) + np.diag(rng.uniform(0.01, 1.0, m))
        H[:m, m] = y
        H[m, :m] = y
    return np.asfortranarray(H), rng.standard_normal(n)


def _single(L, H, r):
    n = H.shape[1]
    Ad, rd, xd = Dev(H.ravel(order="F")), Dev(r), Dev(np.zeros(n))
    L.irwls_dev_posv_aug(Ad.p, n, n + 1, rd.p, xd.p)
    return xd.get()


def test_posv_aug_batched_equals_single_and_replays():
    sizes = [2, 50, 64, 65, 129, 300, 700, 1001]
    L = _bind()
    rng = np.random.default_rng(800)
    for rep in range(3):   # three replays of the same batch shape with new data, eager solves in between
        probs = [_kkt(n, rng) for n in sizes]
        want = [_single(L, H, r) for H, r in probs]
        devs = [(Dev(H.ravel(order="F")), Dev(r), Dev(np.zeros(H.shape[1]))) for H, r in probs]
        k = len(sizes)
        A = (C.c_void_p * k)(*[d[0].p for d in devs])
        R = (C.c_void_p * k)(*[d[1].p for d in devs])
        X = (C.c_void_p * k)(*[d[2].p for d in devs])
        nn = (C.c_int * k)(*sizes)
        lda = (C.c_int * k)(*[s + 1 for s in sizes])
        info = (C.c_int * k)()
        L.irwls_dev_posv_aug_batched(k, A, nn, lda, R, X, info)
        for i, n in enumerate(sizes):
            got = devs[i][2].get()
            assert info[i] == n, (n, info[i])
            assert np.array_equal(got, want[i]), (rep, n, np.abs(got - want[i]).max())
    ms = (C.c_double * 3)()
    choice = L.irwls_dev_batched_choice(len(sizes), max(sizes), ms)
    print("autotuned rows per work-group (64-row tiles):", choice, "timings ms:", list(ms))
    assert choice in (1, 2, 4)


@pytest.mark.parametrize("precision", [0, 3])
@pytest.mark.parametrize("wait_us,alone", [(None, None), ("50", None), (None, "0")])
def test_train_full_batched_equals_standalone(wait_us, alone, precision, monkeypatch, tmp_path):
    """Six problems of uneven size (two datasets, three C/gamma points):
    every batched result equals its standalone trainFULL run, in FP64 and
    with BF16 hi/lo kernel products (precision 3, whose BF16 copies of a
    shared dataset the worker threads build once between them).  With a
    50 us trigger, batches start as soon as a problem has waited that long,
    so several run at once on different round slots.  By default an
    iteration that no other problem would join runs the single-problem
    kernels; with IRWLS_BATCH_ALONE=0 every iteration goes through a round.
    The round trace accounts for every batched iteration."""
    if wait_us is not None:
        monkeypatch.setenv("IRWLS_BATCH_WAIT_US", wait_us)
    if alone is not None:
        monkeypatch.setenv("IRWLS_BATCH_ALONE", alone)
    trace = tmp_path / "rounds.txt"
    monkeypatch.setenv("IRWLS_BATCH_TRACE", str(trace))
    tasks = []
    for seed, (n, d) in enumerate([(900, 8), (1600, 16)]):
        X, y = synth.two_gaussians(n, d, seed=900 + seed)
        ds = lib.Dataset.dense(X, y)
        for C_, g in [(1.0, 0.5 / d), (10.0, 2.0 / d), (100.0, 0.5 / d)]:
            tasks.append((ds, abi.make_properties(gamma=g, C_=C_, W=min(n // 4, 1500), eta=0.001)))
    try:
        lib.set_options(precision=precision)
        want = [lib.train_full(ds, p, seed=0) for ds, p in tasks]
        if precision:
            for ds, _ in tasks:
                ds.pin(with_means=False)   # one shared device copy (and its BF16 split) per dataset
        got = lib.train_full_batched([t[0] for t in tasks], [t[1] for t in tasks])
        st = lib.last_stats()
    finally:
        lib.set_options(precision=0)
        for ds, _ in tasks:
            ds.unpin()
    print("batched stats", st)
    assert st["batch_rounds"] > 0
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), (i, np.abs(g - w).max())
    lines = trace.read_text().splitlines()
    head = [ln for ln in lines if ln.startswith("# call:")]
    rounds = [ln.split() for ln in lines if ln and not ln.startswith("#")]
    probs = [ln for ln in lines if ln.startswith("# problem")]
    assert len(head) == 1 and len(rounds) == st["batch_rounds"] and len(probs) == len(tasks)
    n_alone = int(head[0].split(", ")[2].split()[0])
    if alone == "0":
        assert n_alone == 0
    # every iteration ran either in a round or alone
    assert sum(int(r[2]) for r in rounds) + n_alone == st["inner_iters"]
    for r in rounds:
        assert float(r[0]) <= float(r[1]) and int(r[2]) >= 1 and int(r[3]) in (0, 1)
