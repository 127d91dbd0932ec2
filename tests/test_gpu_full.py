"""GPU parity of the full-training path: kernel tiles, the device Cholesky and
the end-to-end trainFULL / predict entry points against the oracle and
against closed forms computed with numpy."""
import ctypes as C

import numpy as np
import pytest

from libirwls_amd import abi, lib, synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


from gpu_util import Dev, dense_rows, rbf


# ---- kernel tiles -------------------------------------------------------------
@pytest.mark.parametrize("kernel", [1, 0])
def test_kernel_matrix_dense(kernel):
    rng = np.random.default_rng(11)
    n, d, m = 150, 37, 70
    X = rng.standard_normal((n, d))
    idx = np.sort(rng.choice(n, m, replace=False)).astype(np.int32)
    g = 0.05
    L = lib.load()
    Xd, nd, idd = Dev(X), Dev((X * X).sum(1)), Dev(idx)
    K = Dev(np.zeros((n, m)))
    r = dense_rows(Xd, nd, n, d)
    c = dense_rows(Xd, nd, m, d, idd)
    L.irwls_dev_kernel_matrix(C.byref(r), C.byref(c), d, kernel, g, K.p, m)
    got = K.get()
    want = rbf(X, X[idx], g) if kernel == 1 else X @ X[idx].T
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_kernel_matvec_dense_epilogue():
    rng = np.random.default_rng(12)
    n, d, m = 777, 130, 300
    X = rng.standard_normal((n, d))
    Z = rng.standard_normal((m, d))
    w = rng.standard_normal(m)
    out0 = rng.standard_normal(n)
    rs = rng.choice([-1.0, 1.0], n)
    g = 1.0 / d
    L = lib.load()
    Xd, Zd = Dev(X), Dev(Z)
    xn, zn = Dev((X * X).sum(1)), Dev((Z * Z).sum(1))
    wd, od, rsd = Dev(w), Dev(out0), Dev(rs)
    r = dense_rows(Xd, xn, n, d)
    c = dense_rows(Zd, zn, m, d)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 1, g, wd.p, 1.0, -1.0, 0.25, rsd.p, od.p)
    want = out0 + rs * (-(rbf(X, Z, g) @ w)) + 0.25
    np.testing.assert_allclose(od.get(), want, rtol=1e-11, atol=1e-12)


def test_kernel_matvec_many_columns():
    """Column split path (partials + deterministic reduction)."""
    rng = np.random.default_rng(13)
    n, d, m = 96, 16, 5000
    X = rng.standard_normal((n, d))
    Z = rng.standard_normal((m, d))
    w = rng.standard_normal(m)
    g = 1.0 / 16
    L = lib.load()
    Xd, Zd = Dev(X), Dev(Z)
    xn, zn = Dev((X * X).sum(1)), Dev((Z * Z).sum(1))
    wd, od = Dev(w), Dev(np.zeros(n))
    r = dense_rows(Xd, xn, n, d)
    c = dense_rows(Zd, zn, m, d)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 1, g, wd.p, 0.0, 1.0, 0.0, None, od.p)
    first = od.get()
    np.testing.assert_allclose(first, rbf(X, Z, g) @ w, rtol=1e-11, atol=1e-12)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 1, g, wd.p, 0.0, 1.0, 0.0, None, od.p)
    assert np.array_equal(first, od.get())   # bitwise reproducible


def test_kernel_matvec_linear_kernel():
    """Linear kernel: K w formed as X (Z^T w) (no kernel block), for row and
    column selections of one matrix, 5000 columns (all column ranges of the
    first pass), the full output epilogue, and bitwise reproducible."""
    rng = np.random.default_rng(15)
    n, d, m = 6000, 70, 5000
    X = rng.standard_normal((n, d))
    ridx = rng.permutation(n)[:900].astype(np.int32)
    cidx = rng.permutation(n)[:m].astype(np.int32)
    w = rng.standard_normal(m)
    out0 = rng.standard_normal(900)
    rs = rng.choice([-1.0, 1.0], 900)
    L = lib.load()
    Xd, nd = Dev(X), Dev((X * X).sum(1))
    ri, ci = Dev(ridx), Dev(cidx)
    r = dense_rows(Xd, nd, 900, d, ri)
    c = dense_rows(Xd, nd, m, d, ci)
    wd, od, rsd = Dev(w), Dev(out0), Dev(rs)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 0, 0.0, wd.p, 0.5, -2.0, 0.25, rsd.p, od.p)
    got = od.get()
    want = 0.5 * out0 + rs * (-2.0 * (X[ridx] @ (X[cidx].T @ w))) + 0.25
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(w).sum())
    od2 = Dev(out0)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 0, 0.0, wd.p, 0.5, -2.0, 0.25, rsd.p, od2.p)
    assert np.array_equal(od2.get(), got)


def test_kernel_matvec_csr_rows():
    rng = np.random.default_rng(14)
    n, d, m = 500, 300, 90
    rp, col, val = synth.sparse_uniform(n, d, 0.05, rng)
    Xs = np.zeros((n, d))
    for i in range(n):
        Xs[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    Z = rng.random((m, d)) * (rng.random((m, d)) < 0.05)
    w = rng.standard_normal(m)
    g = 0.5
    L = lib.load()
    rpd, cd, vd = Dev(rp), Dev(col), Dev(val)
    xn = Dev((Xs * Xs).sum(1))
    Zd, zn = Dev(Z), Dev((Z * Z).sum(1))
    wd, od = Dev(w), Dev(np.zeros(n))
    r = lib.irwls_rows()
    r.rowptr, r.col, r.val, r.nrm, r.n = rpd.p, cd.p, vd.p, xn.p, n
    c = dense_rows(Zd, zn, m, d)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, 1, g, wd.p, 0.0, 1.0, 0.0, None, od.p)
    np.testing.assert_allclose(od.get(), rbf(Xs, Z, g) @ w, rtol=1e-11, atol=1e-12)
    K = Dev(np.zeros((n, m)))
    L.irwls_dev_kernel_matrix(C.byref(r), C.byref(c), d, 1, g, K.p, m)
    np.testing.assert_allclose(K.get(), rbf(Xs, Z, g), rtol=1e-12, atol=1e-13)


# ---- device Cholesky ------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 7, 64, 65, 200, 333])
def test_potrf_potrs_spd(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    b = rng.standard_normal(n)
    L = lib.load()
    Ad = Dev(np.asfortranarray(A).ravel(order="F"))
    info = L.irwls_dev_potrf(Ad.p, n, n)
    assert info == 0
    F = Ad.get().reshape((n, n), order="F")
    np.testing.assert_allclose(np.tril(F), np.linalg.cholesky(A), rtol=1e-10, atol=1e-12)
    bd = Dev(b)
    L.irwls_dev_potrs(Ad.p, n, n, bd.p)
    np.testing.assert_allclose(bd.get(), np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("n,f", [(100, 37), (64, 16), (130, 70), (200, 127), (40, 0)])
def test_potrf_pivot_failing_inside_a_tile(n, f):
    """A matrix whose pivot f has Schur value -1 (LAPACK potrf semantics):
    info = f + 1, the leading f x f factor is the Cholesky factor of A's
    leading block, row f holds l_f = L^-1 a_f and the failing diagonal keeps
    the Schur value.  f falls inside a 64-row tile and inside one of its
    16-column stages (or on a stage or tile boundary)."""
    rng = np.random.default_rng(300 + n + f)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    lead = A[:f, :f]
    A[f, f] = (A[f, :f] @ np.linalg.solve(lead, A[:f, f]) if f else 0.0) - 1.0
    L = lib.load()
    Ad = Dev(np.asfortranarray(A).ravel(order="F"))
    info = L.irwls_dev_potrf(Ad.p, n, n)
    assert info == f + 1
    F = Ad.get().reshape((n, n), order="F")
    if f:
        Ll = np.linalg.cholesky(lead)
        np.testing.assert_allclose(np.tril(F[:f, :f]), Ll, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(F[f, :f], np.linalg.solve(Ll, A[:f, f]), rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(F[f, f], -1.0, rtol=1e-8)


@pytest.mark.parametrize("n", [2, 50, 129, 257])
def test_kkt_failed_last_pivot(n):
    """Indefinite KKT [A y; y^T 0]: the factorization stops at the last pivot
    with the Schur value -s there; the solve then matches the closed form
    b = (r_n - y^T A^-1 r_1)/s^2, beta = A^-1 (r_1 - y b) (s = y^T A^-1 y)."""
    rng = np.random.default_rng(100 + n)
    m = n - 1
    X = rng.standard_normal((m, 4))
    y = rng.choice([-1.0, 1.0], m)
    A = rbf(X, X, 0.3) * np.outer(y, y) + np.diag(rng.uniform(0.1, 1.0, m))
    H = np.zeros((n, n))
    H[:m, :m] = A
    H[:m, m] = y
    H[m, :m] = y
    r = rng.standard_normal(n)
    L = lib.load()
    Hd = Dev(np.asfortranarray(H).ravel(order="F"))
    info = L.irwls_dev_potrf(Hd.p, n, n)
    assert info == n
    F = Hd.get().reshape((n, n), order="F")
    s = y @ np.linalg.solve(A, y)
    np.testing.assert_allclose(F[m, m], -s, rtol=1e-10)
    rd = Dev(r)
    L.irwls_dev_potrs(Hd.p, n, n, rd.p)
    x = rd.get()
    u = np.linalg.solve(A, r[:m])
    b = (r[m] - y @ u) / s ** 2
    beta = u - np.linalg.solve(A, y) * b
    np.testing.assert_allclose(x[m], b, rtol=1e-8)
    np.testing.assert_allclose(x[:m], beta, rtol=1e-8, atol=1e-10)
    # and the oracle's restatement agrees
    np.testing.assert_allclose(x, po.linear_system(H, r), rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 64, 65, 130, 300])
def test_posv_aug_kkt(n):
    """Factorization with the right-hand side riding in row n (the forward
    sweep done by the factorization steps) plus one backward sweep, on the
    KKT matrix with its failed last pivot."""
    rng = np.random.default_rng(200 + n)
    m = n - 1
    H = np.zeros((n, n))
    if m > 0:
        X = rng.standard_normal((m, 5))
        y = rng.choice([-1.0, 1.0], m)
        H[:m, :m] = rbf(X, X, 0.2) * np.outer(y, y) + np.diag(rng.uniform(0.01, 1.0, m))
        H[:m, m] = y
        H[m, :m] = y
    r = rng.standard_normal(n)
    L = lib.load()
    buf = np.zeros((n + 1, n), order="F")
    buf[:n, :n] = H
    Ad = Dev(buf.ravel(order="F"))
    rd, xd = Dev(r), Dev(np.zeros(n))
    info = L.irwls_dev_posv_aug(Ad.p, n, n + 1, rd.p, xd.p)
    assert info == n
    want = po.linear_system(H, r)
    if m == 0:
        assert not np.isfinite(xd.get()[0]) and not np.isfinite(want[0])
    else:
        np.testing.assert_allclose(xd.get(), want, rtol=1e-8, atol=1e-10)


def test_parallel_linear_system_entry_point():
    rng = np.random.default_rng(5)
    n = 150
    B = rng.standard_normal((n, n))
    A = B @ B.T + np.eye(n)
    b = rng.standard_normal(n)
    x, F = lib.linear_system(A, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    assert np.all(np.triu(F, 1) == 0.0)


# ---- end-to-end full training ---------------------------------------------------
def _parity(n, d, W, gamma, C_, seed, n_test=500, eta=0.001, kernel=1):
    X, y = synth.two_gaussians(n, d, seed=seed)
    Xt, yt = synth.two_gaussians(n_test, d, seed=seed + 100)
    props = abi.make_properties(gamma=gamma, C_=C_, W=W, eta=eta, kernel=kernel)
    ods = po.Dataset.dense(X, y)
    want = po.train_full(ods, props, seed=0)
    ostats = po.stats()
    gds = lib.Dataset.dense(X, y)
    got = lib.train_full(gds, props, seed=0)
    gstats = lib.last_stats()
    print("oracle", ostats, "gpu", gstats)
    assert gstats["outer_iters"] == ostats["outer_iters"]
    assert gstats["inner_iters"] == ostats["inner_iters"]
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("relative weight error", rel, "max abs", np.abs(got - want).max())
    assert rel < 1e-6
    assert np.array_equal(got[:-1] != 0, want[:-1] != 0)
    # predictions on a test set: oracle model/oracle predictor vs GPU model/GPU predictor
    om = po.full_model(props, ods, want)
    gm = lib.full_model(props, gds, got)
    assert om.bytes()[:16] != b"" and gm.m.nSVs == om.m.nSVs
    otd = po.Dataset.dense(Xt, yt)
    gtd = lib.Dataset.dense(Xt, yt)
    pp = abi.make_predict_properties(labels=0)
    osoft = po.soft_test(otd, om)
    gsoft = lib.soft_test(gtd, gm, pp)
    np.testing.assert_allclose(gsoft, osoft, rtol=1e-6, atol=1e-8)
    assert np.array_equal(lib.hard_test(gtd, gm, pp), po.hard_test(otd, om))
    return got, want


def test_train_full_small():
    _parity(n=1000, d=16, W=150, gamma=1 / 16, C_=10.0, seed=3)


def test_linear_kernel_sub_irwls_and_predict():
    """Linear kernel through the training and prediction entry points.
    The reference's full trainer is unstable with this kernel at such
    shapes: input perturbations of 1e-14 (relative) change the oracle's own
    trainFULL result on this problem from test accuracy 0.93 to between
    0.08 and 0.93 (4 of 8 runs blow up to |beta| ~ 1e6), so the end-to-end
    trajectory cannot be compared.  What is deterministic is: one working-set
    sub-problem (subIRWLS, whose error updates and G13 run on the linear
    block) and the predictor (X (Z^T w) with the model's vectors) on the
    oracle's trained weights."""
    n, d = 150, 16
    X, y = synth.two_gaussians(n, d, seed=21)
    rng = np.random.default_rng(21)
    props = abi.make_properties(gamma=1 / 16, C_=1.0, W=n, kernel=0)
    GIN = rng.standard_normal(n + 1) * 0.1
    ods, gds = po.Dataset.dense(X, y), lib.Dataset.dense(X, y)
    bw, ew, _ = po.sub_irwls(ods, props, GIN, y.copy(), np.zeros(n + 1))
    bg, eg, _ = lib.sub_irwls(gds, props, GIN, y.copy(), np.zeros(n + 1))
    np.testing.assert_allclose(bg, bw, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(eg, ew, rtol=1e-9, atol=1e-10)
    Xf, yf = synth.two_gaussians(1000, 16, seed=5)
    Xt, yt = synth.two_gaussians(1000, 16, seed=105)
    pf = abi.make_properties(gamma=1 / 16, C_=1.0, W=150, eta=0.001, kernel=0)
    of = po.Dataset.dense(Xf, yf)
    beta = po.train_full(of, pf, seed=0)
    fo = po.soft_test(po.Dataset.dense(Xt, yt), po.full_model(pf, of, beta))
    pp = abi.make_predict_properties(labels=0)
    fg = lib.soft_test(lib.Dataset.dense(Xt, yt), lib.full_model(pf, lib.Dataset.dense(Xf, yf), beta), pp)
    np.testing.assert_allclose(fg, fo, rtol=1e-10, atol=1e-11)
    assert np.mean(np.sign(fo) == yt) > 0.9


def test_train_full_two_gaussians_n4000():
    """The CPU-reference workload: n=4000, d=16, gamma=1/16, C=10, W=500."""
    _parity(n=4000, d=16, W=500, gamma=1 / 16, C_=10.0, seed=0, n_test=1000)


def test_sub_irwls_entry_point():
    rng = np.random.default_rng(21)
    n, d = 300, 8
    X, y = synth.two_gaussians(n, d, seed=21)
    props = abi.make_properties(gamma=0.2, C_=5.0, W=n)
    GIN = rng.standard_normal(n + 1) * 0.1
    e = y.copy()
    beta = np.zeros(n + 1)
    ods = po.Dataset.dense(X, y)
    gds = lib.Dataset.dense(X, y)
    bw, ew, _ = po.sub_irwls(ods, props, GIN, e, beta)
    bg, eg, _ = lib.sub_irwls(gds, props, GIN, e, beta)
    np.testing.assert_allclose(bg, bw, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(eg, ew, rtol=1e-7, atol=1e-9)
