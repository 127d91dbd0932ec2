"""The allocation-free kernel-block x vector product
(irwls_dev_kernel_matvec_noalloc) against irwls_dev_kernel_matvec through the
C-ABI: bit-equal results on two calls in a row, the second with a smaller
input (the fixed partial-sum array keeps the first call's contents), for the
column-split RBF path, the unsplit path and the linear kernel; inputs that
need more than the array, or CSR columns, are rejected."""
import ctypes as C

import numpy as np
import pytest

from libirwls_amd import lib

from gpu_util import Dev, dense_rows

pytestmark = pytest.mark.gpu

ROWS = lib.irwls_rows


def _bind(L):
    f = L.irwls_dev_kernel_matvec_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(ROWS), C.POINTER(ROWS), C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_double,
                  C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    return f


def _case(rng, n, m, d):
    X = rng.standard_normal((n, d)) * 0.3
    Z = rng.standard_normal((m, d)) * 0.3
    return X, Z, rng.standard_normal(m), rng.standard_normal(n), rng.choice([-1.0, 1.0], n)


def _both(L, f, X, Z, w, out0, rs, kernel, g, a=1.0, b=-1.0, cst=0.25):
    n, d = X.shape
    m = Z.shape[0]
    Xd, Zd = Dev(X), Dev(Z)
    xn, zn = Dev((X * X).sum(1)), Dev((Z * Z).sum(1))
    wd, rsd = Dev(w), Dev(rs)
    o1, o2 = Dev(out0), Dev(out0)
    r, c = dense_rows(Xd, xn, n, d), dense_rows(Zd, zn, m, d)
    L.irwls_dev_kernel_matvec(C.byref(r), C.byref(c), d, kernel, g, wd.p, a, b, cst, rsd.p, o1.p)
    rc = f(C.byref(r), C.byref(c), d, kernel, g, wd.p, a, b, cst, rsd.p, o2.p)
    L.irwls_dev_sync()
    return rc, o1.get(), o2.get()


@pytest.mark.parametrize("kernel", [1, 0])
def test_noalloc_equals_matvec_two_calls(kernel):
    L = lib.load()
    L.irwls_dev_init()
    f = _bind(L)
    rng = np.random.default_rng(41)
    # first call: 40000 rows x 3000 columns splits the columns (partials in
    # the fixed array); second call: a smaller input over the same array
    for (n, m, d) in [(40000, 3000, 64), (5000, 700, 33)]:
        X, Z, w, out0, rs = _case(rng, n, m, d)
        rc, want, got = _both(L, f, X, Z, w, out0, rs, kernel, 1.0 / d)
        assert rc == 0
        assert np.array_equal(got, want), (n, m, d, np.abs(got - want).max())
        assert np.all(np.isfinite(got))


def test_noalloc_unsplit_and_empty_columns():
    L = lib.load()
    L.irwls_dev_init()
    f = _bind(L)
    rng = np.random.default_rng(42)
    X, Z, w, out0, rs = _case(rng, 300, 100, 20)      # one column tile: no partial sums
    rc, want, got = _both(L, f, X, Z, w, out0, rs, 1, 0.05)
    assert rc == 0 and np.array_equal(got, want)
    # no columns: out = a out + cst
    Xd, xn, od = Dev(X), Dev((X * X).sum(1)), Dev(out0)
    r = dense_rows(Xd, xn, 300, 20)
    c = dense_rows(Xd, xn, 0, 20)
    assert f(C.byref(r), C.byref(c), 20, 1, 0.05, None, 2.0, 1.0, 0.5, None, od.p) == 0
    np.testing.assert_array_equal(od.get(), 2.0 * out0 + 0.5)


def test_noalloc_rejects_csr_and_oversize():
    L = lib.load()
    L.irwls_dev_init()
    f = _bind(L)
    rng = np.random.default_rng(43)
    # linear kernel with d = 40000 features over 64 columns needs 65 x 40000 doubles > 16 MiB
    d = 40000
    X = rng.standard_normal((4, d))
    Xd, xn, od, wd = Dev(X), Dev((X * X).sum(1)), Dev(np.zeros(4)), Dev(np.ones(4))
    r = dense_rows(Xd, xn, 4, d)
    c = dense_rows(Xd, xn, 4, d)
    assert f(C.byref(r), C.byref(c), d, 0, 0.0, wd.p, 0.0, 1.0, 0.0, None, od.p) == 0   # 5 x d: fits
    Z = rng.standard_normal((64, d))
    Zd, zn, w64 = Dev(Z), Dev((Z * Z).sum(1)), Dev(np.ones(64))
    c64 = dense_rows(Zd, zn, 64, d)
    assert f(C.byref(r), C.byref(c64), d, 0, 0.0, w64.p, 0.0, 1.0, 0.0, None, od.p) == -1
    # CSR columns
    cc = ROWS()
    cc.X, cc.n = None, 4
    assert f(C.byref(r), C.byref(cc), d, 1, 0.1, wd.p, 0.0, 1.0, 0.0, None, od.p) == -1
