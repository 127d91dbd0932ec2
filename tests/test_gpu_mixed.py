"""GPU tests of the mixed-precision solver: the FP32 Cholesky with FP64
diagonal tiles and its L^-T, preconditioned CG in FP64 to a 1e-10 relative
residual (fresh and stale preconditioner), and subIRWLS / trainFULL in mixed
mode against the FP64 path.

The mixed mode has no reference counterpart.  Its pins are the residual of
the FP64 matrix (here against numpy in FP64) and the end-to-end agreement
with the FP64 path, which is itself pinned to the oracle."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import Dev, rbf
from libirwls_amd import abi, lib, synth

pytestmark = pytest.mark.gpu


def _bind():
    L = lib.load()
    L.irwls_dev_potrf_inv_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.irwls_dev_pcg_spd2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_int, C.c_double, C.c_int, C.c_void_p, C.POINTER(C.c_double)]
    return L


def _kernel_spd(n, seed, dmin=1e-3):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 6))
    y = rng.choice([-1.0, 1.0], n)
    return rbf(X, X, 0.3) * np.outer(y, y) + np.diag(rng.uniform(dmin, 1.0, n)), rng


@pytest.mark.parametrize("n", [64, 200, 513, 3000])
def test_potrf_inv_f32(n):
    """FP32 factor and L^-T from the identity rows; n = 3000 also takes the
    two-launch steps (panel solve, then trailing update) before the fused ones,
    with each step's rows bounded where the identity rows are still zero."""
    A, _ = _kernel_spd(n, 300 + n, dmin=0.05)
    L = _bind()
    Ad = Dev(np.asfortranarray(A).ravel(order="F"))
    Fd = Dev(np.zeros(2 * n * n, np.float32))
    Zd = Dev(np.zeros(n * n, np.float32))
    info = L.irwls_dev_potrf_inv_f32(Ad.p, n, n, Fd.p, 2 * n, Zd.p, n)
    assert info == 0
    F = Fd.get().reshape((2 * n, n), order="F").astype(np.float64)
    Lf = np.tril(F[:n])
    Z = F[n:]
    ZT = Zd.get().reshape((n, n), order="F").astype(np.float64)
    Lx = np.linalg.cholesky(A)
    rel = np.linalg.norm(Lf - Lx) / np.linalg.norm(Lx)
    print("n", n, "factor rel err", rel)
    assert rel < 1e-5
    # Z = L^-T (upper triangular), ZT its transpose
    np.testing.assert_allclose(np.triu(Z), Z, atol=0)
    zi = np.linalg.norm(Z.T @ Lf - np.eye(n)) / np.sqrt(n)
    print("n", n, "|Z^T L - I|", zi)
    assert zi < 1e-4
    np.testing.assert_array_equal(ZT, Z.T)


@pytest.mark.parametrize("n,dmin", [(300, 1e-2), (1000, 1e-4)])
def test_pcg_two_rhs_to_1e10(n, dmin):
    A, rng = _kernel_spd(n, 400 + n, dmin=dmin)
    print("cond(A) = %.3g" % np.linalg.cond(A))
    B = rng.standard_normal((2, n))
    L = _bind()
    Ad = Dev(np.asfortranarray(A).ravel(order="F"))
    Fd = Dev(np.zeros(2 * n * n, np.float32))
    Zd = Dev(np.zeros(n * n, np.float32))
    assert L.irwls_dev_potrf_inv_f32(Ad.p, n, n, Fd.p, 2 * n, Zd.p, n) == 0
    Bd, Xd = Dev(B), Dev(np.zeros((2, n)))
    rel = (C.c_double * 2)()
    it = L.irwls_dev_pcg_spd2(Ad.p, n, n, Bd.p, Fd.p, 2 * n, Zd.p, n, 1e-10, 200, Xd.p, rel)
    X = Xd.get()
    print("iterations", it, "relres", list(rel))
    assert it > 0
    for c in range(2):
        true_rel = np.linalg.norm(B[c] - A @ X[c]) / np.linalg.norm(B[c])
        print("column", c, "true relative residual", true_rel)
        assert true_rel <= 2e-10
        np.testing.assert_allclose(X[c], np.linalg.solve(A, B[c]), rtol=1e-6, atol=1e-8 * np.abs(X[c]).max())


def test_pcg_stale_preconditioner():
    """A factor of an earlier matrix (other diagonal) still preconditions CG
    to 1e-10 on the new one."""
    n = 400
    A1, rng = _kernel_spd(n, 501, dmin=1e-2)
    A2 = A1 + np.diag(rng.uniform(0.0, 0.2, n) * np.diag(A1))   # stays SPD
    L = _bind()
    A1d = Dev(np.asfortranarray(A1).ravel(order="F"))
    A2d = Dev(np.asfortranarray(A2).ravel(order="F"))
    Fd = Dev(np.zeros(2 * n * n, np.float32))
    Zd = Dev(np.zeros(n * n, np.float32))
    assert L.irwls_dev_potrf_inv_f32(A1d.p, n, n, Fd.p, 2 * n, Zd.p, n) == 0
    B = rng.standard_normal((2, n))
    Bd, Xd = Dev(B), Dev(np.zeros((2, n)))
    rel = (C.c_double * 2)()
    it = L.irwls_dev_pcg_spd2(A2d.p, n, n, Bd.p, Fd.p, 2 * n, Zd.p, n, 1e-10, 200, Xd.p, rel)
    X = Xd.get()
    print("iterations with the stale factor", it)
    assert it > 0
    for c in range(2):
        assert np.linalg.norm(B[c] - A2 @ X[c]) / np.linalg.norm(B[c]) <= 2e-10


def test_tile_potrf_inv():
    """One SPD tile (n <= 64) factored in place with L^-1 and its transpose
    beside it, against numpy."""
    L = _bind()
    L.irwls_dev_tile_potrf_inv.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]
    for n in (1, 7, 37, 64):
        A, _ = _kernel_spd(n, 700 + n, dmin=0.05)
        lda = 70
        Ah = np.zeros((lda, n))
        Ah[:n, :n] = A
        Ad = Dev(Ah.ravel(order="F"))
        Dd = Dev(np.zeros(2 * 64 * 64))
        assert L.irwls_dev_tile_potrf_inv(Ad.p, lda, n, Dd.p, 1) == 0
        Lx = np.linalg.cholesky(A)
        Lg = np.tril(Ad.get().reshape((lda, n), order="F")[:n])
        D = Dd.get()
        X = D[:4096].reshape((64, 64), order="F")[:n, :n]
        XT = D[4096:].reshape((64, 64), order="F")[:n, :n]
        np.testing.assert_allclose(Lg, Lx, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(X, np.linalg.inv(Lx), rtol=1e-10, atol=1e-11)
        np.testing.assert_array_equal(XT, X.T)


@pytest.mark.parametrize("n,nrm", [(700, 40), (900, 150)])
def test_pcg_schur_removed_rows(n, nrm):
    """Rows removed from a factored matrix (and the diagonal changed): the
    mapped preconditioner with the exact correction for the removed rows
    (the inverse of the factored matrix's kept block; S^-1 from one tile for
    40 rows, from the identity-row factor for 150) takes CG to 1e-10 in a few
    iterations, far fewer than the mapped form without it; U equals
    (Z Z^T)(:, R) and S^-1 the inverse of (Z Z^T)(R, R)."""
    A1, rng = _kernel_spd(n, 601 + n, dmin=1e-3)
    L = _bind()
    L.irwls_dev_pcg_schur_setup.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_int]
    L.irwls_dev_pcg_spd2_schur.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.POINTER(C.c_double)]
    L.irwls_dev_pcg_spd2_mapped.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                            C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                            C.c_void_p, C.POINTER(C.c_double)]
    A1d = Dev(np.asfortranarray(A1).ravel(order="F"))
    Fd = Dev(np.zeros(2 * n * n, np.float32))
    Zd = Dev(np.zeros(n * n, np.float32))
    assert L.irwls_dev_potrf_inv_f32(A1d.p, n, n, Fd.p, 2 * n, Zd.p, n) == 0
    removed = np.sort(rng.choice(n, nrm, replace=False)).astype(np.int32)
    keep = np.setdiff1d(np.arange(n), removed).astype(np.int32)
    m = keep.size
    A2 = A1[np.ix_(keep, keep)] + np.diag(rng.uniform(0.0, 0.05, m) * np.diag(A1)[keep])
    ld2 = m + 2
    A2h = np.zeros((ld2, m))
    A2h[:m] = A2
    A2d = Dev(A2h.ravel(order="F"))
    mp = Dev(np.arange(n, dtype=np.int32)[keep])              # row i of the new system -> factor position
    inv = np.full(n, -1, np.int32)
    inv[keep] = np.arange(m, dtype=np.int32)
    ivd = Dev(inv)
    Rd = Dev(removed)
    Ud = Dev(np.zeros(n * nrm))
    Sd = Dev(np.zeros(nrm * nrm))
    assert L.irwls_dev_pcg_schur_setup(Zd.p, n, n, Rd.p, removed.size, Ud.p, Sd.p, 1) == 0
    ZT = Zd.get().reshape((n, n), order="F").astype(np.float64)
    Z = ZT.T
    Bfull = Z @ Z.T
    U = Ud.get().reshape((n, nrm), order="F")
    np.testing.assert_allclose(U, Bfull[:, removed], rtol=1e-12, atol=1e-14 * np.abs(Bfull).max())
    Sinv = Sd.get().reshape((nrm, nrm), order="F")
    Sx = np.linalg.inv(Bfull[np.ix_(removed, removed)])
    print("S^-1 relative error", np.abs(Sinv - Sx).max() / np.abs(Sx).max())
    assert np.abs(Sinv - Sx).max() <= 1e-6 * np.abs(Sx).max()
    B = rng.standard_normal((2, m))
    rel = (C.c_double * 2)()
    Bd, Xd = Dev(B), Dev(np.zeros((2, m)))
    it_s = L.irwls_dev_pcg_spd2_schur(A2d.p, m, ld2, Bd.p, Fd.p, 2 * n, Zd.p, n, n, mp.p, ivd.p, Rd.p,
                                      removed.size, Ud.p, Sd.p, 1e-10, 200, Xd.p, rel)
    X = Xd.get()
    Xm = Dev(np.zeros((2, m)))
    it_m = L.irwls_dev_pcg_spd2_mapped(A2d.p, m, ld2, Bd.p, Fd.p, 2 * n, Zd.p, n, n, mp.p, ivd.p, 1e-10, 200, Xm.p,
                                       rel)
    print("iterations: corrected", it_s, "mapped without the correction", it_m)
    assert 0 < it_s <= 12
    assert it_m < 0 or it_m >= it_s + 10
    for c in range(2):
        assert np.linalg.norm(B[c] - A2 @ X[c]) / np.linalg.norm(B[c]) <= 2e-10
        np.testing.assert_allclose(X[c], np.linalg.solve(A2, B[c]), rtol=1e-6, atol=1e-8 * np.abs(X[c]).max())


def _sub_problem(n, seed):
    rng = np.random.default_rng(seed)
    X, y = synth.two_gaussians(n, 8, seed=seed)
    GIN = rng.standard_normal(n + 1) * 0.1
    return X, y, GIN


def test_sub_irwls_mixed_solver_matches_fp64():
    """subIRWLS with only the mixed solver (closed form of the failed-pivot
    solve from two PCG solves, FP64 kernel block) against FP64 subIRWLS on
    the same problem: weights within 1e-7 relative."""
    n = 600
    X, y, GIN = _sub_problem(n, 41)
    props = abi.make_properties(gamma=0.2, C_=5.0, W=n)
    gds = lib.Dataset.dense(X, y)
    beta0 = np.zeros(n + 1)
    lib.set_options(precision=0)
    b64, e64, _ = lib.sub_irwls(gds, props, GIN, y.copy(), beta0)
    s64 = lib.last_stats()
    try:
        lib.set_options(precision=2)
        bmx, emx, _ = lib.sub_irwls(gds, props, GIN, y.copy(), beta0)
        smx = lib.last_stats()
    finally:
        lib.set_options(precision=0)
    print("fp64", s64)
    print("mixed", smx)
    assert smx["fp64_fallbacks"] == 0 and smx["pcg_iters"] > 0
    rel = np.linalg.norm(bmx - b64) / np.linalg.norm(b64)
    print("relative weight difference", rel)
    assert rel < 1e-7
    np.testing.assert_allclose(emx, e64, rtol=1e-6, atol=1e-8)


def test_train_full_mixed_accuracy():
    """Mixed trainFULL against FP64 trainFULL: test accuracy within 0.1%
    (absolute), weights and decision values close."""
    n, d = 4000, 32
    X, y = synth.four_gaussians(n, d, seed=7, means_seed=7)
    Xt, yt = synth.four_gaussians(2000, d, seed=107, means_seed=7)
    props = abi.make_properties(gamma=1.0 / d, C_=1.0, W=1000, eta=0.001)
    ds = lib.Dataset.dense(X, y)
    td = lib.Dataset.dense(Xt, yt)
    pp = abi.make_predict_properties(labels=0)
    lib.set_options(precision=0)
    b64 = lib.train_full(ds, props, seed=0)
    m64 = lib.full_model(props, ds, b64)
    f64 = lib.soft_test(td, m64, pp)
    try:
        lib.set_options(precision=1)
        bmx = lib.train_full(ds, props, seed=0)
        st = lib.last_stats()
        mmx = lib.full_model(props, ds, bmx)
        fmx = lib.soft_test(td, mmx, pp)
    finally:
        lib.set_options(precision=0)
    acc64 = np.mean(np.where(f64 >= 0, 1.0, -1.0) == yt)
    accmx = np.mean(np.where(fmx >= 0, 1.0, -1.0) == yt)
    wrel = np.linalg.norm(bmx - b64) / np.linalg.norm(b64)
    print("mixed stats", st)
    print("accuracy fp64 %.4f mixed %.4f; weight rel err %.3g; decision max err %.3g rms %.3g"
          % (acc64, accmx, wrel, np.abs(fmx - f64).max(), np.sqrt(np.mean((fmx - f64) ** 2))))
    assert abs(accmx - acc64) <= 0.001
    assert wrel < 5e-2


def _bf16_rows(Xd, n, d):
    """Device BF16 hi/lo copy of n dense FP64 rows (width d), padded to 32."""
    L = lib.load()
    ldb = (d + 31) // 32 * 32
    Xh = Dev(np.zeros(n * ldb, np.uint16))
    Xl = Dev(np.zeros(n * ldb, np.uint16))
    nrm = Dev(np.zeros(n, np.float32))
    L.irwls_dev_bf16_split.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_void_p]
    L.irwls_dev_bf16_split(Xd.p, d, n, d, Xh.p, Xl.p, ldb, nrm.p)
    r = lib.irwls_rows_bf16()
    r.Xh, r.Xl, r.ld, r.nrm, r.idx, r.n = Xh.p, Xl.p, ldb, nrm.p, None, n
    return r, (Xh, Xl, nrm)


@pytest.mark.parametrize("n,m,d", [(300, 200, 256), (257, 129, 100), (64, 1000, 16)])
def test_bf16_kernel_matrix_error_bound(n, m, d):
    """RBF block from BF16 hi/lo x3 products (FP32 accumulation) against the
    FP64 kernel: |x|^2 + |z|^2 - 2 x.z is off by at most about
    2^-15 (|x|^2 + |z|^2) (split error 2^-16 per factor, FP32 sums), so
    |k - k64| <= gamma * 2^-15 * (|x|^2 + |z|^2) (k <= 1)."""
    rng = np.random.default_rng(600 + d)
    X = rng.standard_normal((n, d)) + 0.5
    Z = rng.standard_normal((m, d)) - 0.25
    g = 1.0 / d
    L = lib.load()
    L.irwls_dev_kernel_matrix_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int]
    Xd, Zd = Dev(X), Dev(Z)
    r, keep_r = _bf16_rows(Xd, n, d)
    c, keep_c = _bf16_rows(Zd, m, d)
    Kd = Dev(np.zeros(n * m))
    L.irwls_dev_kernel_matrix_bf16(C.byref(r), C.byref(c), 1, g, Kd.p, m)
    K = Kd.get().reshape(n, m)
    K64 = rbf(X, Z, g)
    nx = (X ** 2).sum(1)[:, None] + (Z ** 2).sum(1)[None, :]
    bound = g * 2.0 ** -15 * nx
    err = np.abs(K - K64)
    print("n m d", n, m, d, "max err %.3g, max err/bound %.3f" % (err.max(), (err / bound).max()))
    assert np.all(err <= bound)
    # the same rows against themselves: k(x, x) = 1 up to the accumulation error
    Kd2 = Dev(np.zeros(n * n))
    L.irwls_dev_kernel_matrix_bf16(C.byref(r), C.byref(r), 1, g, Kd2.p, n)
    diag = np.diag(Kd2.get().reshape(n, n))
    assert np.all(np.abs(diag - 1.0) <= 1e-5)


@pytest.mark.parametrize("n,m,d,kernel", [(5000, 700, 128, 1), (300, 3000, 64, 1), (2000, 600, 100, 1),
                                          (2000, 600, 100, 0)])
def test_bf16_kernel_matvec(n, m, d, kernel):
    """out = a out + rs (b K w) + cst with the BF16 block, against FP64
    numpy; also covers the split-column path (many column tiles), ragged
    row and column tiles and zero-padded features (d = 100 in rows of 128).
    RBF: |error| <= 1e-4 sum |w| (k <= 1).  Linear: the dot products are off
    by at most about 2^-15 |x||z|, so |error_i| <= 2^-14 |x_i| sum_j |w_j||z_j|."""
    rng = np.random.default_rng(700 + n + d + kernel)
    X = rng.standard_normal((n, d))
    Z = rng.standard_normal((m, d))
    w = rng.standard_normal(m)
    rs = rng.choice([-1.0, 1.0], n)
    out0 = rng.standard_normal(n)
    g = 0.5 / d
    L = lib.load()
    L.irwls_dev_kernel_matvec_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_double,
                                               C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    Xd, Zd = Dev(X), Dev(Z)
    r, keep_r = _bf16_rows(Xd, n, d)
    c, keep_c = _bf16_rows(Zd, m, d)
    wd, rsd, outd = Dev(w), Dev(rs), Dev(out0)
    L.irwls_dev_kernel_matvec_bf16(C.byref(r), C.byref(c), kernel, g, wd.p, 0.5, -1.0, 0.25, rsd.p, outd.p)
    K = rbf(X, Z, g) if kernel == 1 else X @ Z.T
    want = 0.5 * out0 + rs * (-(K @ w)) + 0.25
    err = np.abs(outd.get() - want)
    if kernel == 1:
        bound = 1e-4 * np.abs(w).sum()
    else:
        bound = 2.0 ** -14 * np.linalg.norm(X, axis=1) * (np.abs(w) * np.linalg.norm(Z, axis=1)).sum()
    print("matvec max abs err", err.max(), "max err / bound", (err / bound).max())
    assert np.all(err <= bound)


@pytest.mark.parametrize("kernel", [1, 0])
def test_bf16_matvec_uses_the_blocks_kernel_values(kernel):
    """The BF16 matvec (256 x 256 tiles, columns on the MFMA A side) and the
    materialised BF16 block (128 x 128 tiles, rows on the A side) produce the
    same FP32 kernel values, so K w from the matvec equals the FP64 sum of
    w_j K_ij over the block up to FP64 summation order.  Training relies on
    it: the working-set block and the error updates must see one kernel.
    Rows and columns are index selections of one data set (as in training),
    with ragged tiles and zero-padded features."""
    rng = np.random.default_rng(900 + kernel)
    N, d = 3000, 100
    X = rng.standard_normal((N, d)) * 0.7
    ridx = rng.permutation(N)[:1100].astype(np.int32)
    cidx = rng.permutation(N)[:700].astype(np.int32)
    w = rng.standard_normal(700) * np.where(rng.random(700) < 0.1, 100.0, 1.0)
    g = 0.5 / d
    L = lib.load()
    L.irwls_dev_kernel_matrix_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int]
    L.irwls_dev_kernel_matvec_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_double,
                                               C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    Xd = Dev(X)
    base, keep = _bf16_rows(Xd, N, d)
    rd, cd = Dev(ridx), Dev(cidx)
    r, c = lib.irwls_rows_bf16(), lib.irwls_rows_bf16()
    for s, ix, n in ((r, rd, 1100), (c, cd, 700)):
        s.Xh, s.Xl, s.ld, s.nrm, s.idx, s.n = base.Xh, base.Xl, base.ld, base.nrm, ix.p, n
    Kd = Dev(np.zeros(1100 * 700))
    L.irwls_dev_kernel_matrix_bf16(C.byref(r), C.byref(c), kernel, g, Kd.p, 700)
    K = Kd.get().reshape(1100, 700)
    wd, od = Dev(w), Dev(np.zeros(1100))
    L.irwls_dev_kernel_matvec_bf16(C.byref(r), C.byref(c), kernel, g, wd.p, 0.0, 1.0, 0.0, None, od.p)
    got, want = od.get(), K @ w
    scale = np.abs(K) @ np.abs(w)
    rel = np.abs(got - want) / scale
    print("kernel", kernel, "max |matvec - block w| / sum |w k|: %.3g" % rel.max())
    assert rel.max() < 1e-13
