"""GPU tests of entry points that the end-to-end tests do not reach:
ParallelVectorMatrixT / ParallelVectorMatrix against numpy, softTestMany
model by model against the oracle's predictor, the pinned-dataset cache
(pinned and unpinned runs equal bit for bit, labels changed in place, a
freed and rebuilt dataset of the same size), and a compiled C program that
uses the public headers and links libirwls.so."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from libirwls_amd import abi, lib, synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("size", [1, 7, 300, 1025])
def test_parallel_vector_matrix(size):
    """ParallelVectorMatrixT: result = m2 m1; ParallelVectorMatrix:
    result = m1^T m2 (m2 column-major size x size), as the reference's
    dgemm_ calls (src/ParallelAlgorithms.c:1318-1371)."""
    rng = np.random.default_rng(size)
    M = rng.standard_normal((size, size))
    v = rng.standard_normal(size)
    m2 = np.asfortranarray(M).ravel(order="F").copy()
    L = lib.load()
    r1, r2 = np.zeros(size), np.zeros(size)
    L.ParallelVectorMatrixT(_dp(v), size, _dp(m2), _dp(r1), 4)
    L.ParallelVectorMatrix(_dp(v), size, _dp(m2), _dp(r2), 4)
    np.testing.assert_allclose(r1, M @ v, rtol=1e-12, atol=1e-12 * np.abs(M).sum(1).max())
    np.testing.assert_allclose(r2, M.T @ v, rtol=1e-12, atol=1e-12 * np.abs(M).sum(0).max())


def test_soft_test_many_against_oracle():
    """Three models (two full, one budgeted) on one test set: every row of
    softTestMany equals the oracle's softTest of that model to 1e-9 and the
    library's softTest bit for bit."""
    X, y = synth.two_gaussians(1200, 12, seed=61)
    Xt, yt = synth.two_gaussians(700, 12, seed=161)
    gds, ods = lib.Dataset.dense(X, y), po.Dataset.dense(X, y)
    gtd, otd = lib.Dataset.dense(Xt, yt), po.Dataset.dense(Xt, yt)
    pp = abi.make_predict_properties(labels=0)
    gms, oms = [], []
    for g, c in [(1 / 12, 1.0), (0.3, 10.0)]:
        props = abi.make_properties(gamma=g, C_=c, W=200)
        beta = po.train_full(ods, props, seed=0)
        gms.append(lib.full_model(props, gds, beta))
        oms.append(po.full_model(props, ods, beta))
    bp = abi.make_properties(gamma=0.2, C_=10.0, size=40)
    cent = po.sgma(ods, bp, ncand=64, seed=1)
    bb = po.irwls_par(ods, cent, bp)
    gms.append(lib.budgeted_model(bp, gds, cent, bb))
    oms.append(po.budgeted_model(bp, ods, cent, bb))
    got = lib.soft_test_many(gtd, gms, pp)
    for g in range(len(gms)):
        np.testing.assert_allclose(got[g], po.soft_test(otd, oms[g]), rtol=1e-9, atol=1e-12)
        assert np.array_equal(got[g], lib.soft_test(gtd, gms[g], pp))


def test_pinned_and_unpinned_equal():
    """trainFULL, SGMA and IRWLSpar on a pinned dataset equal the unpinned
    runs bit for bit; labels flipped in place on a pinned dataset are seen
    by the next call."""
    X, y = synth.two_gaussians(1500, 10, seed=71)
    props = abi.make_properties(gamma=0.1, C_=10.0, W=200)
    bprops = abi.make_properties(gamma=0.1, C_=10.0, size=30)
    ref = lib.Dataset.dense(X, y)
    want = lib.train_full(ref, props, seed=0)
    cw = lib.sgma(ref, bprops, seed=2)
    bw = lib.irwls_par(ref, cw, bprops)
    ds = lib.Dataset.dense(X, y)
    ds.pin(with_means=True)
    assert np.array_equal(lib.train_full(ds, props, seed=0), want)
    cg = lib.sgma(ds, bprops, seed=2)
    assert list(cg) == list(cw)
    assert np.array_equal(lib.irwls_par(ds, cg, bprops), bw)
    # flip labels in place (one-vs-rest on a shared dataset): the pinned copy follows
    y2 = y.copy()
    y2[::7] = -y2[::7]
    yv = np.ctypeslib.as_array(ds.ds.y, shape=(ds.ds.l,))
    yv[:] = y2
    ref2 = lib.Dataset.dense(X, y2)
    assert np.array_equal(lib.train_full(ds, props, seed=0), lib.train_full(ref2, props, seed=0))
    ds.free()


def test_freed_pinned_dataset_is_not_reused():
    """A pinned dataset released with freeDataset alone (as a C caller
    does), then a new dataset of the same size: its results are its own."""
    X, y = synth.two_gaussians(900, 8, seed=81)
    X2, y2 = synth.two_gaussians(900, 8, seed=82)
    props = abi.make_properties(gamma=0.2, C_=10.0, W=150)
    want2 = lib.train_full(lib.Dataset.dense(X2, y2), props, seed=0)
    a = lib.Dataset.dense(X, y)
    a.pin(with_means=False)
    lib.train_full(a, props, seed=0)
    lib.load().freeDataset(a.ds)   # no unpin call
    a.ds = None
    b = lib.Dataset.dense(X2, y2)
    assert np.array_equal(lib.train_full(b, props, seed=0), want2)
    b.pin(with_means=False)
    assert np.array_equal(lib.train_full(b, props, seed=0), want2)
    b.free()


def _write_libsvm(path, X, y):
    with open(path, "w") as f:
        for xi, yi in zip(X, y):
            f.write("%d %s\n" % (int(yi), " ".join("%d:%.17g" % (k + 1, v) for k, v in enumerate(xi))))


def _build_caller(tmp_path):
    exe = str(tmp_path / "caller")
    subprocess.check_call(["gcc", "-O2", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c_caller", "caller.c"),
                           "-L", os.path.join(ROOT, "libirwls_amd"), "-lirwls",
                           "-Wl,-rpath," + os.path.join(ROOT, "libirwls_amd"), "-lm", "-o", exe])
    return exe


@pytest.mark.parametrize("kind", ["full", "budgeted"])
def test_compiled_c_caller(tmp_path, kind):
    """The C program trains from a libsvm file through the exported parsers
    and trainers, stores and reloads the model and predicts; its model file
    equals the one the same calls make through ctypes, byte for byte, and
    its labels equal the library's."""
    exe = _build_caller(tmp_path)
    X, y = synth.two_gaussians(800, 6, seed=91)
    Xt, yt = synth.two_gaussians(300, 6, seed=191)
    tr, te = str(tmp_path / "train.txt"), str(tmp_path / "test.txt")
    _write_libsvm(tr, X, y)
    _write_libsvm(te, Xt, yt)
    flags = ["-g", "0.25", "-c", "10", "-v", "0"] + (["-s", "40"] if kind == "budgeted" else ["-w", "150"])
    mf, lf = str(tmp_path / "model.bin"), str(tmp_path / "labels.txt")
    out = subprocess.run([exe, kind, tr, te, mf, lf] + flags, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0
    ds = lib.Dataset.read(tr)
    if kind == "full":
        props = abi.make_properties(gamma=0.25, C_=10.0, W=150)
        m = lib.full_model(props, ds, lib.train_full(ds, props, seed=0))
    else:
        props = abi.make_properties(gamma=0.25, C_=10.0, size=40)
        c = lib.sgma(ds, props, seed=0)
        m = lib.budgeted_model(props, ds, c, lib.irwls_par(ds, c, props))
    with open(mf, "rb") as f:
        assert f.read() == m.bytes(str(tmp_path))
    labels = np.loadtxt(lf)
    pp = abi.make_predict_properties(labels=0)
    assert np.array_equal(labels, lib.hard_test(lib.Dataset.read(te), m, pp))


@pytest.mark.parametrize("kind", ["csr", "dense"])
def test_upload_across_chunks(kind):
    """The device upload writes the rows through 32 MiB pinned chunks on
    several host threads: a test set of several chunks (CSR with empty rows;
    dense) predicts as the same rows uploaded a piece of one chunk at a time
    (CSR bit for bit: its row products do not depend on the row count;
    dense to 1e-13, whose column splits follow the row count)."""
    rng = np.random.default_rng(5)
    d = 300 if kind == "csr" else 40
    if kind == "csr":
        rp, col, val = synth.sparse_uniform(3000, d, 0.03, rng)
        ds = lib.Dataset.csr(rp, col, val, np.where(rng.random(3000) < 0.5, 1.0, -1.0))
        n = 1_500_000
        lens0 = rng.integers(0, 13, 20000)
        lens0[::97] = 0                                 # empty rows
        cols0 = np.concatenate([np.sort(rng.choice(d, k, replace=False)) for k in lens0]).astype(np.int32)
        lens = np.tile(lens0, n // 20000)
        trp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        tcol = np.tile(cols0, n // 20000)
        tval = rng.random(len(tcol))
        assert trp[-1] == len(tcol) and trp[-1] * 8 > 2 * (32 << 20)

        def piece(a, b):
            return lib.Dataset.csr(trp[a:b + 1] - trp[a], tcol[trp[a]:trp[b]], tval[trp[a]:trp[b]], np.zeros(b - a))
        big = piece(0, n)
    else:
        X, y = synth.two_gaussians(3000, d, seed=12)
        ds = lib.Dataset.dense(X, y)
        n = 220_000
        Xt = rng.standard_normal((n, d)).astype(np.float64)
        assert n * d * 8 > 32 << 20

        def piece(a, b):
            return lib.Dataset.dense(Xt[a:b], np.zeros(b - a))
        big = piece(0, n)
    props = abi.make_properties(gamma=0.2, C_=10.0, W=300)
    m = lib.full_model(props, ds, lib.train_full(ds, props, seed=0))
    pp = abi.make_predict_properties(labels=0)
    got = lib.soft_test(big, m, pp)
    step = 100_000 if kind == "csr" else 20_000
    want = np.concatenate([lib.soft_test(piece(a, min(a + step, n)), m, pp) for a in range(0, n, step)])
    if kind == "csr":
        assert np.array_equal(got, want)
    else:
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
