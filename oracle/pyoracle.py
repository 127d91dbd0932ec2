"""ORACLE -- test infrastructure only.

ctypes wrapper over liboracle.so (the FP64 CPU restatement of the reference
path in irwls_oracle.c).  Imported by the tests, the smoke check and the CPU
baseline leg of bench.py; never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from libirwls_amd.abi import (svm_dataset, properties, model, dataset_arrays)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class orc_stats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters", C.c_int), ("solves", C.c_int),
                ("sgma_steps", C.c_int), ("irwlspar_iters", C.c_int), ("t_total", C.c_double),
                ("m_decays", C.c_int)]


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", HERE, "liboracle.so"])
    # the reference's reader / model-file source, where the reference tree is present
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-B", "-C", HERE, "_ref/libref_io.so", "_ref/libref_predict.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp, ip = P(C.c_double), P(C.c_int)
        L.orc_dataset_dense.restype = svm_dataset
        L.orc_dataset_dense.argtypes = [dp, dp, C.c_int, C.c_int]
        L.orc_dataset_csr.restype = svm_dataset
        L.orc_dataset_csr.argtypes = [P(C.c_int64), P(C.c_int32), dp, dp, C.c_int]
        L.orc_dataset_dense_unlabeled.restype = svm_dataset
        L.orc_dataset_dense_unlabeled.argtypes = [dp, C.c_int, C.c_int]
        L.orc_free_dataset.argtypes = [svm_dataset]
        L.orc_kernel.restype = C.c_double
        L.orc_kernel.argtypes = [svm_dataset, C.c_int, C.c_int, properties]
        L.orc_kernel_test.restype = C.c_double
        L.orc_kernel_test.argtypes = [svm_dataset, C.c_int, model, C.c_int]
        L.orc_potrf.restype = C.c_int
        L.orc_potrf.argtypes = [dp, C.c_int, C.c_int]
        L.orc_potrs.argtypes = [dp, C.c_int, C.c_int, dp]
        L.orc_linear_system.argtypes = [dp, C.c_int, dp, dp]
        L.orc_rpermute.restype = ip
        L.orc_rpermute.argtypes = [C.c_int]
        L.orc_subIRWLS.restype = dp
        L.orc_subIRWLS.argtypes = [svm_dataset, properties, dp, dp, dp]
        L.orc_trainFULL.restype = dp
        L.orc_trainFULL.argtypes = [svm_dataset, properties]
        L.orc_fullModel.restype = model
        L.orc_fullModel.argtypes = [properties, svm_dataset, dp]
        L.orc_randomCentroids.restype = ip
        L.orc_randomCentroids.argtypes = [svm_dataset, properties]
        L.orc_SGMA.restype = ip
        L.orc_SGMA.argtypes = [svm_dataset, properties, C.c_int, dp]
        L.orc_IRWLSpar.restype = dp
        L.orc_IRWLSpar.argtypes = [svm_dataset, ip, properties]
        L.orc_budgetedModel.restype = model
        L.orc_budgetedModel.argtypes = [properties, svm_dataset, ip, dp]
        L.orc_softTest.restype = dp
        L.orc_softTest.argtypes = [svm_dataset, model]
        L.orc_test.restype = dp
        L.orc_test.argtypes = [svm_dataset, model]
        L.orc_model_serialize.restype = C.c_size_t
        L.orc_model_serialize.argtypes = [P(model), C.c_void_p]
        L.orc_free_model.argtypes = [model]
        L.orc_srand.argtypes = [C.c_uint]
        L.orc_last_stats.restype = P(orc_stats)
        L.orc_set_limits.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_inner_continue.restype = C.c_int
        L.orc_inner_continue.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                         C.c_double]
        L.orc_regroup.argtypes = [C.c_int, dp, dp, dp, C.c_double, dp, ip]
        L.orc_initial_groups.argtypes = [C.c_int, dp, dp, dp, C.c_double, dp, ip]
        L.orc_outer_track.restype = C.c_int
        L.orc_outer_track.argtypes = [C.c_double, C.c_double, C.c_double, dp, ip]
        L.orc_ws_rebuild.restype = C.c_int
        L.orc_ws_rebuild.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_int, ip, ip, ip, ip, ip, ip]
        L.orc_par_continue.restype = C.c_int
        L.orc_par_continue.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_par_track.restype = C.c_int
        L.orc_par_track.argtypes = [C.c_int, C.c_double, dp, dp, dp, ip]
        _lib = L
    return _lib


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def take_doubles(ptr, n):
    """Copy n doubles from a malloc-ed pointer and free it."""
    out = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)
    _libc.free(C.cast(ptr, C.c_void_p))
    return out


def take_ints(ptr, n):
    out = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, np.int32)
    _libc.free(C.cast(ptr, C.c_void_p))
    return out


class Dataset:
    """Owns an svm_dataset built by the oracle's reader restatement."""

    def __init__(self, ds):
        self.ds = ds

    @classmethod
    def dense(cls, X, y):
        X = np.ascontiguousarray(X, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        n, d = X.shape
        return cls(lib().orc_dataset_dense(_dptr(X), _dptr(y), n, d))

    @classmethod
    def dense_unlabeled(cls, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        n, d = X.shape
        return cls(lib().orc_dataset_dense_unlabeled(_dptr(X), n, d))

    @classmethod
    def csr(cls, rowptr, col, val, y):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        return cls(lib().orc_dataset_csr(rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                         col.ctypes.data_as(C.POINTER(C.c_int32)),
                                         _dptr(val), _dptr(y), len(rowptr) - 1))

    def arrays(self, with_means=True):
        return dataset_arrays(self.ds, with_means)

    def free(self):
        if self.ds is not None:
            lib().orc_free_dataset(self.ds)
            self.ds = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def srand(seed):
    lib().orc_srand(seed)


def kernel(ds, i, j, props):
    return lib().orc_kernel(ds.ds, i, j, props)


def train_full(ds, props, seed=0):
    srand(seed)
    p = lib().orc_trainFULL(ds.ds, props)
    return take_doubles(p, ds.ds.l + 1)


def sub_irwls(ds, props, GIN, e, beta):
    GIN = np.ascontiguousarray(GIN, dtype=np.float64).copy()
    e = np.ascontiguousarray(e, dtype=np.float64).copy()
    beta = np.ascontiguousarray(beta, dtype=np.float64).copy()
    p = lib().orc_subIRWLS(ds.ds, props, _dptr(GIN), _dptr(e), _dptr(beta))
    return take_doubles(p, ds.ds.l + 1), e, beta


def sgma(ds, props, ncand=64, seed=1, log_gains=False):
    srand(seed)
    g = None
    if log_gains:
        steps = max(props.size - 2, 0)
        g = np.zeros(max(steps * ncand * 2, 1))
    p = lib().orc_SGMA(ds.ds, props, ncand, _dptr(g) if g is not None else None)
    c = take_ints(p, props.size)
    if log_gains:
        return c, g[:max(props.size - 2, 0) * ncand * 2].reshape(-1, ncand, 2)
    return c


def random_centroids(ds, props, seed=1):
    srand(seed)
    return take_ints(lib().orc_randomCentroids(ds.ds, props), props.size)


def irwls_par(ds, centroids, props):
    c = np.ascontiguousarray(centroids, dtype=np.int32)
    p = lib().orc_IRWLSpar(ds.ds, c.ctypes.data_as(C.POINTER(C.c_int)), props)
    return take_doubles(p, props.size)


class Model:
    def __init__(self, m):
        self.m = m

    def bytes(self):
        L = lib()
        n = L.orc_model_serialize(C.byref(self.m), None)
        buf = (C.c_ubyte * n)()
        L.orc_model_serialize(C.byref(self.m), buf)
        return bytes(buf)

    def free(self):
        if self.m is not None:
            lib().orc_free_model(self.m)
            self.m = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def full_model(props, ds, beta):
    b = np.ascontiguousarray(beta, dtype=np.float64)
    return Model(lib().orc_fullModel(props, ds.ds, _dptr(b)))


def budgeted_model(props, ds, centroids, beta):
    c = np.ascontiguousarray(centroids, dtype=np.int32)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    return Model(lib().orc_budgetedModel(props, ds.ds, c.ctypes.data_as(C.POINTER(C.c_int)), _dptr(b)))


def soft_test(ds, m):
    return take_doubles(lib().orc_softTest(ds.ds, m.m), ds.ds.l)


def hard_test(ds, m):
    return take_doubles(lib().orc_test(ds.ds, m.m), ds.ds.l)


def potrf(A):
    """Column-major factor of a square numpy array (returns (L_colmajor, info))."""
    n = A.shape[0]
    F = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    buf = F.ravel(order="F").copy()
    info = lib().orc_potrf(_dptr(buf), n, n)
    return buf.reshape((n, n), order="F"), info


def linear_system(A, b):
    n = A.shape[0]
    buf = np.asfortranarray(A, dtype=np.float64).ravel(order="F").copy()
    rhs = np.ascontiguousarray(b, dtype=np.float64).copy()
    x = np.zeros(n)
    lib().orc_linear_system(_dptr(buf), n, _dptr(rhs), _dptr(x))
    return x


def stats():
    s = lib().orc_last_stats().contents
    return {k: getattr(s, k) for k, _ in orc_stats._fields_}


def set_limits(max_outer=0, max_sgma=0, max_par=0):
    lib().orc_set_limits(max_outer, max_sgma, max_par)
