"""GPU tests of the budgeted path's device kernels against numpy FP64:
kernel blocks and matrix-vector products of CSR rows with CSR or dense
columns (the one-thread-per-row kernel for at most 16 columns, the
one-wave-per-row kernel with feature masks over 64-column chunks), the SGMA
gain reduction (the narrow kernel for at most 16 candidates, the 128 x 64
tile kernel above that) with its half-range sums, and the weighted Gram
matrix of IRWLSpar.  These restate the reference's arithmetic
(src/kernels.c:101-116, src/budgeted-train.c:195-198, :401-416) in numpy;
the differences allowed are those of a different summation order."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import Dev, dense_rows
from libirwls_amd import lib, synth

pytestmark = pytest.mark.gpu


def _csr_rows(rp, col, val, nrm, idx=None):
    """irwls_rows over device CSR arrays (kept alive by the returned list)."""
    keep = [Dev(rp), Dev(col), Dev(val), Dev(nrm)]
    r = lib.irwls_rows()
    r.rowptr, r.col, r.val, r.nrm = keep[0].p, keep[1].p, keep[2].p, keep[3].p
    r.n = len(rp) - 1
    if idx is not None:
        keep.append(Dev(np.asarray(idx, np.int32)))
        r.idx = keep[-1].p
        r.n = len(idx)
    r.nnz_per_row = len(val) / max(len(rp) - 1, 1)
    return r, keep


def _kfun(kind, g, dot, xn, zn):
    return np.exp(-g * (xn[:, None] + zn[None, :] - 2.0 * dot)) if kind == 1 else dot


@pytest.mark.parametrize("ncols", [1, 8, 16, 17, 64, 65, 150])
@pytest.mark.parametrize("kind", [0, 1])
def test_csr_kernel_block_and_matvec(ncols, kind):
    rng = np.random.default_rng(1000 + ncols + 7 * kind)
    n, d, g = 700, 400, 0.3
    rp, col, val = synth.sparse_uniform(n, d, 0.04, rng)
    # rows 5 and 6 empty, row 7 with one nonzero
    X = np.zeros((n, d))
    for i in range(n):
        X[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    X[5:7] = 0.0
    X[7] = 0.0
    X[7, 3] = 0.5
    rp = np.concatenate([[0], np.cumsum((X != 0).sum(1))]).astype(np.int64)
    col = np.concatenate([np.nonzero(X[i])[0] for i in range(n)]).astype(np.int32)
    val = np.concatenate([X[i, np.nonzero(X[i])[0]] for i in range(n)])
    xn = (X * X).sum(1)
    L = lib.load()
    rows, k1 = _csr_rows(rp, col, val, xn)
    cidx = rng.choice(n, ncols, replace=False).astype(np.int32)
    cidx[0] = 5   # an empty column
    cols, k2 = _csr_rows(rp, col, val, xn, cidx)
    Z = X[cidx]
    want = _kfun(kind, g, X @ Z.T, xn, xn[cidx])
    K = Dev(np.zeros((n, ncols)))
    L.irwls_dev_kernel_matrix(C.byref(rows), C.byref(cols), d, kind, g, K.p, ncols)
    np.testing.assert_allclose(K.get(), want, rtol=1e-13, atol=1e-14)
    # a row selection on the left
    ridx = rng.choice(n, 123, replace=False).astype(np.int32)
    rsel, k3 = _csr_rows(rp, col, val, xn, ridx)
    Ks = Dev(np.zeros((123, ncols)))
    L.irwls_dev_kernel_matrix(C.byref(rsel), C.byref(cols), d, kind, g, Ks.p, ncols)
    np.testing.assert_allclose(Ks.get(), want[ridx], rtol=1e-13, atol=1e-14)
    # dense columns against CSR rows give the same block
    Zd, zn = Dev(Z), Dev(xn[cidx])
    dc = dense_rows(Zd, zn, ncols, d)
    Kd = Dev(np.zeros((n, ncols)))
    L.irwls_dev_kernel_matrix(C.byref(rows), C.byref(dc), d, kind, g, Kd.p, ncols)
    np.testing.assert_allclose(Kd.get(), K.get(), rtol=1e-13, atol=1e-14)
    # matrix-vector product with the same columns
    w = rng.standard_normal(ncols)
    wd, od = Dev(w), Dev(np.ones(n))
    L.irwls_dev_kernel_matvec(C.byref(rows), C.byref(cols), d, kind, g, wd.p, 0.5, 2.0, 0.25, None, od.p)
    np.testing.assert_allclose(od.get(), 0.5 + 2.0 * (want @ w) + 0.25, rtol=1e-12, atol=1e-12)


def _bind_sgma(L):
    vp = C.c_void_p
    L.irwls_dev_sgma_gain.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp]
    L.irwls_dev_sgma_gain_half.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, vp]
    L.irwls_dev_wgram.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, C.c_int]


@pytest.mark.parametrize("ncand", [1, 5, 16, 17, 59, 64, 65, 130])
@pytest.mark.parametrize("n,s", [(1037, 0), (1037, 1), (1000, 13), (1037, 100), (1037, 300), (4096, 500)])
def test_sgma_gain_against_numpy(ncand, n, s):
    """q[c] = sum_i (KSM[i, c] - (KSC Z)[i, c])^2 (KSC n x s column-major,
    Z s x ncand column-major, KSM row-major), and the same sums over the
    first half of the row tiles (64 rows per tile up to 16 candidates, 128
    above).  With few row tiles and s > 128 the k range is split over
    work-groups (3 ranges at s = 300, 4 at s = 500) and the ranges' products
    are added before the squares."""
    rng = np.random.default_rng(n + 31 * s + ncand)
    L = lib.load()
    _bind_sgma(L)
    KSC = rng.standard_normal((n, s))
    Z = rng.standard_normal((s, ncand))
    KSM = rng.standard_normal((n, ncand)) + KSC @ Z   # residuals of moderate size
    ldz, ldm = s + 3, ncand + 1
    Zp = np.zeros((ncand, ldz))
    Zp[:, :s] = Z.T
    Mp = np.zeros((n, ldm))
    Mp[:, :ncand] = KSM
    KSCd = Dev(np.asfortranarray(KSC).ravel(order="F") if s else np.zeros(1))
    Zd, Md = Dev(Zp.ravel()), Dev(Mp.ravel())
    qd, hd = Dev(np.zeros(ncand)), Dev(np.zeros(ncand))
    L.irwls_dev_sgma_gain_half(n, s, ncand, KSCd.p, Zd.p, ldz, Md.p, ldm, qd.p, hd.p)
    R = KSM - KSC @ Z
    np.testing.assert_allclose(qd.get(), (R * R).sum(0), rtol=1e-12)
    tile = 64 if ncand <= 16 else 128
    rt = (n + tile - 1) // tile
    h = max(rt // 2, 1)
    np.testing.assert_allclose(hd.get(), (R[:min(n, h * tile)] ** 2).sum(0), rtol=1e-12)
    q2 = Dev(np.zeros(ncand))
    L.irwls_dev_sgma_gain(n, s, ncand, KSCd.p, Zd.p, ldz, Md.p, ldm, q2.p)
    assert np.array_equal(q2.get(), qd.get())


@pytest.mark.parametrize("B,nact,weighted", [(70, 2999, True), (150, 1000, False), (300, 5, True), (129, 0, True),
                                              (257, 3000, True)])
def test_wgram_against_numpy(B, nact, weighted):
    """K1 = KC + sum_q (sw_q Phi[a_q])^T (sw_q Phi[a_q]) over the active rows
    (row-major Phi), written with leading dimension B + 1."""
    rng = np.random.default_rng(B + nact)
    n = 3000
    L = lib.load()
    _bind_sgma(L)
    Phi = rng.standard_normal((n, B))
    act = np.sort(rng.choice(n, nact, replace=False)).astype(np.int32) if nact else np.zeros(1, np.int32)
    sw = rng.uniform(0.1, 2.0, max(nact, 1))
    M = rng.standard_normal((B, B))
    KC = M + M.T
    ldk = B + 1
    Phid, actd, swd, KCd = Dev(Phi), Dev(act), Dev(sw), Dev(np.asfortranarray(KC).ravel(order="F"))
    K1d = Dev(np.zeros(ldk * B))
    L.irwls_dev_wgram(B, Phid.p, actd.p, swd.p if weighted else None, nact, KCd.p, K1d.p, ldk)
    P = Phi[act[:nact]] * (sw[:nact, None] if weighted else 1.0)
    want = KC + P.T @ P
    got = K1d.get().reshape((B, ldk))[:, :B].T   # column c at offset c * ldk
    scale = np.abs(P).max() ** 2 * max(nact, 1) if nact else 1.0
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * scale)
    assert np.array_equal(got, got.T)


@pytest.mark.parametrize("with_idx", [False, True])
def test_csr_kernel_high_dimensional_column_chunks(with_idx):
    """d = 3 x 10^6 features, 20 nonzeros per row: a transposed column block
    of 150 x d doubles would take 3.6 GB, so the CSR row kernels take the
    columns 64 at a time (1 GiB budget per block) and add the chunks' sums.
    Block and product against scipy in FP64, with and without a column
    selection."""
    from scipy.sparse import csr_matrix
    rng = np.random.default_rng(77)
    n, d, g = 400, 3_000_000, 0.4
    nnz_row = 20
    col = np.concatenate([np.sort(rng.choice(d, nnz_row, replace=False)) for _ in range(n)]).astype(np.int32)
    rp = np.arange(0, n * nnz_row + 1, nnz_row, dtype=np.int64)
    # a few shared features so that dot products are not all zero
    col[::7] = rng.integers(0, 50, size=len(col[::7]))
    for i in range(n):
        col[rp[i]:rp[i + 1]] = np.sort(np.unique(np.concatenate(
            [col[rp[i]:rp[i + 1]], rng.choice(d, nnz_row, replace=False)]))[:nnz_row])
    val = rng.random(len(col))
    A = csr_matrix((val, col, rp), shape=(n, d))
    xn = np.asarray(A.multiply(A).sum(1)).ravel()
    cidx = rng.choice(n, 150, replace=False).astype(np.int32) if with_idx else np.arange(150, dtype=np.int32)
    L = lib.load()
    rows, k1 = _csr_rows(rp, col, val, xn)
    if with_idx:
        cols, k2 = _csr_rows(rp, col, val, xn, cidx)
    else:
        rp2 = rp[:151].copy()
        cols, k2 = _csr_rows(rp2, col[:rp2[-1]], val[:rp2[-1]], xn[:150])
    dots = (A @ A[cidx].T).toarray()
    Kw = np.exp(-g * (xn[:, None] + xn[cidx][None, :] - 2.0 * dots))
    Kd = Dev(np.zeros((n, 150)))
    L.irwls_dev_kernel_matrix(C.byref(rows), C.byref(cols), d, 1, g, Kd.p, 150)
    np.testing.assert_allclose(Kd.get(), Kw, rtol=1e-12, atol=1e-14)
    w = rng.standard_normal(150)
    out0 = rng.standard_normal(n)
    rs = rng.choice([-1.0, 1.0], n)
    wd, od, rsd = Dev(w), Dev(out0), Dev(rs)
    L.irwls_dev_kernel_matvec(C.byref(rows), C.byref(cols), d, 1, g, wd.p, 0.5, -2.0, 0.125, rsd.p, od.p)
    np.testing.assert_allclose(od.get(), 0.5 * out0 + rs * (-2.0 * (Kw @ w)) + 0.125, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("kind", [0, 1])
def test_growing_column_table_equals_kernel_matrix(kind):
    """The kept centroid table, grown a few columns at a time across the
    narrow (at most 16) and wide (several 64-column chunks) kernels, gives
    the same kernel blocks bit for bit as the block built for each call, on
    all rows and on a row selection."""
    rng = np.random.default_rng(77 + kind)
    n, d, g = 900, 300, 0.25
    rp, col, val = synth.sparse_uniform(n, d, 0.05, rng)
    xn = np.array([np.sum(val[rp[i]:rp[i + 1]] ** 2) for i in range(n)])
    L = lib.load()
    vp = C.c_void_p
    L.irwls_dev_coltab_create.restype = vp
    L.irwls_dev_coltab_create.argtypes = [C.c_int, C.c_int]
    L.irwls_dev_coltab_destroy.argtypes = [vp]
    L.irwls_dev_coltab_size.argtypes = [vp]
    L.irwls_dev_coltab_extend.argtypes = [vp, C.POINTER(lib.irwls_rows)]
    L.irwls_dev_kernel_matrix_coltab.argtypes = [C.POINTER(lib.irwls_rows), vp, C.c_int, C.c_double, vp, C.c_int]
    cap = 160
    cidx = rng.choice(n, 150, replace=False).astype(np.int32)
    rows, k1 = _csr_rows(rp, col, val, xn)
    ridx = rng.choice(n, 211, replace=False).astype(np.int32)
    rsel, k2 = _csr_rows(rp, col, val, xn, ridx)
    cols, k3 = _csr_rows(rp, col, val, xn, cidx)
    t = L.irwls_dev_coltab_create(d, cap)
    assert t
    try:
        for m in [1, 5, 16, 17, 64, 65, 129, 150]:
            cols.n = m
            L.irwls_dev_coltab_extend(t, C.byref(cols))
            assert L.irwls_dev_coltab_size(t) == m
            for r, nr in ((rows, n), (rsel, len(ridx))):
                want, got = Dev(np.zeros((nr, cap))), Dev(np.zeros((nr, cap)))
                L.irwls_dev_kernel_matrix(C.byref(r), C.byref(cols), d, kind, g, want.p, cap)
                L.irwls_dev_kernel_matrix_coltab(C.byref(r), t, kind, g, got.p, cap)
                assert np.array_equal(got.get()[:, :m], want.get()[:, :m]), (m, nr)
    finally:
        L.irwls_dev_coltab_destroy(t)
