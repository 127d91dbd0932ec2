"""ctypes binding of libirwls.so, the GPU library behind the reference C API.

The Python side only marshals arguments: every call lands in the C entry
points declared in include/*.h (trainFULL, SGMA, IRWLSpar, test, ...).  If
the shared library is missing this module raises at import time; there is no
Python or CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .abi import svm_dataset, properties, predictProperties, model, svm_sample, dataset_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libirwls.so")

# timer kinds (enum irwls_timer_kind in include/irwls_device.h)
TIMER_KINDS = ["kmv", "kmat", "potf2", "trsm", "syrk", "potrs", "kkt", "gemv", "sgma", "csr",
               "syrkd", "matvec", "pcg", "bf16", "mlp", "gain16"]


class irwls_gpu_options(C.Structure):
    _fields_ = [("precision", C.c_int), ("sgma_candidates", C.c_int), ("proposal", C.c_int),
                ("proposal_keep", C.c_int), ("refine_tol", C.c_double), ("proposal_pool", C.c_int),
                ("proposal_warmup", C.c_int), ("proposal_rows", C.c_int), ("par_max_iters", C.c_int)]


class irwls_gpu_stats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters", C.c_int), ("solves", C.c_int),
                ("sgma_steps", C.c_int), ("irwlspar_iters", C.c_int), ("t_total", C.c_double),
                ("t_solve", C.c_double), ("factorizations", C.c_int), ("pcg_iters", C.c_int),
                ("fp64_fallbacks", C.c_int), ("learned_steps", C.c_int), ("mlp_loss0", C.c_double),
                ("mlp_loss", C.c_double), ("batch_rounds", C.c_int), ("t_wait", C.c_double),
                ("par_m_decays", C.c_int)]


class irwls_rows(C.Structure):
    _fields_ = [("X", C.c_void_p), ("ldx", C.c_int), ("rowptr", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p), ("nrm", C.c_void_p), ("idx", C.c_void_p), ("n", C.c_int),
                ("nnz_per_row", C.c_double)]


class irwls_rows_bf16(C.Structure):
    _fields_ = [("Xh", C.c_void_p), ("Xl", C.c_void_p), ("ld", C.c_int), ("nrm", C.c_void_p),
                ("idx", C.c_void_p), ("n", C.c_int)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libirwls.so is not built (run python -m libirwls_amd.build): %s" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp, ip = P(C.c_double), P(C.c_int)
    vp = C.c_void_p
    sig = {
        "trainFULL": (dp, [svm_dataset, properties]),
        "subIRWLS": (dp, [svm_dataset, properties, dp, dp, dp]),
        "rpermute": (ip, [C.c_int]),
        "calculateFULLModel": (model, [properties, svm_dataset, dp]),
        "SGMA": (ip, [svm_dataset, properties]),
        "randomCentroids": (ip, [svm_dataset, properties]),
        "IRWLSpar": (dp, [svm_dataset, ip, properties]),
        "calculateBudgetedModel": (model, [properties, svm_dataset, ip, dp]),
        "test": (dp, [svm_dataset, model, predictProperties]),
        "softTest": (dp, [svm_dataset, model, predictProperties]),
        "softTestMany": (C.c_int, [svm_dataset, P(model), C.c_int, predictProperties, dp]),
        "kernelFunction": (C.c_double, [svm_dataset, C.c_int, C.c_int, properties]),
        "kernelTest": (C.c_double, [svm_dataset, C.c_int, model, C.c_int]),
        "initMemory": (None, [C.c_int, C.c_int]),
        "freeMemory": (None, [C.c_int]),
        "ParallelLinearSystem": (None, [dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int]),
        "ParallelVectorMatrixT": (None, [dp, C.c_int, dp, dp, C.c_int]),
        "ParallelVectorMatrix": (None, [dp, C.c_int, dp, dp, C.c_int]),
        "freeDataset": (None, [svm_dataset]),
        "freeModel": (None, [model]),
        "readTrainFile": (svm_dataset, [C.c_char_p]),
        "readTrainFileCSV": (svm_dataset, [C.c_char_p, C.c_char_p]),
        "readUnlabeledFile": (svm_dataset, [C.c_char_p]),
        "readUnlabeledFileCSV": (svm_dataset, [C.c_char_p, C.c_char_p]),
        "storeModel": (None, [P(model), vp]),
        "readModel": (None, [P(model), vp]),
        "writeOutput": (None, [C.c_char_p, dp, C.c_int]),
        "irwls_dataset_from_dense": (svm_dataset, [dp, dp, C.c_int, C.c_int]),
        "irwls_dataset_from_csr": (svm_dataset, [P(C.c_longlong), ip, dp, dp, C.c_int]),
        "irwls_dataset_from_numpy": (svm_dataset, [dp, dp, C.c_int, C.c_int, C.c_int]),
        "irwls_dataset_save": (C.c_int, [P(svm_dataset), C.c_int, C.c_char_p]),
        "irwls_dataset_load": (svm_dataset, [C.c_char_p]),
        "irwls_gpu_get_options": (None, [P(irwls_gpu_options)]),
        "irwls_gpu_set_options": (None, [P(irwls_gpu_options)]),
        "irwls_gpu_pin_dataset": (None, [svm_dataset, C.c_int]),
        "irwls_gpu_unpin_dataset": (None, [svm_dataset]),
        "irwls_gpu_timing": (None, [C.c_int]),
        "irwls_gpu_timing_mask": (None, [C.c_uint]),
        "irwls_gpu_timer": (None, [C.c_int, P(C.c_long), dp, dp, dp]),
        "irwls_gpu_last_stats": (None, [P(irwls_gpu_stats)]),
        "irwls_gpu_shutdown": (None, []),
        "irwls_dev_init": (C.c_int, []),
        "irwls_dev_alloc": (vp, [C.c_size_t]),
        "irwls_dev_free": (None, [vp]),
        "irwls_dev_h2d": (None, [vp, vp, C.c_size_t]),
        "irwls_dev_d2h": (None, [vp, vp, C.c_size_t]),
        "irwls_dev_sync": (None, []),
        "irwls_dev_trim": (None, []),
        "irwls_dev_copy_kernel": (None, [vp, vp, C.c_size_t]),
        "irwls_dev_potrf": (C.c_int, [vp, C.c_int, C.c_int]),
        "irwls_dev_potrs": (None, [vp, C.c_int, C.c_int, vp]),
        "irwls_dev_posv_aug": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
        "irwls_dev_kernel_matrix": (None, [P(irwls_rows), P(irwls_rows), C.c_int, C.c_int, C.c_double,
                                           vp, C.c_int]),
        "irwls_dev_kernel_matvec": (None, [P(irwls_rows), P(irwls_rows), C.c_int, C.c_int, C.c_double,
                                           vp, C.c_double, C.c_double, C.c_double, vp, vp]),
        "irwls_dev_row_norms": (None, [vp, C.c_int, C.c_int, C.c_int, vp]),
    }
    for name, (res, args) in sig.items():
        try:
            f = getattr(L, name)
        except AttributeError:
            continue   # reported by tests/test_abi.py
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.srand.argtypes = [C.c_uint]
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]


def srand(seed):
    _libc.srand(seed)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def take_doubles(ptr, n):
    out = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)
    _libc.free(C.cast(ptr, C.c_void_p))
    return out


def take_ints(ptr, n):
    out = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, np.int32)
    _libc.free(C.cast(ptr, C.c_void_p))
    return out


class Dataset:
    """An svm_dataset built by the library's reader-semantics builders."""

    def __init__(self, ds):
        self.ds = ds

    @classmethod
    def dense(cls, X, y):
        X = np.ascontiguousarray(X, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        n, d = X.shape
        return cls(load().irwls_dataset_from_dense(_dptr(X), _dptr(y), n, d))

    @classmethod
    def numpy(cls, X, y, with_average=False):
        """The conversion of the reference's Python extension (zeros dropped,
        sparse = 1, optional mean rows exactly as it builds them)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        n, d = X.shape
        return cls(load().irwls_dataset_from_numpy(_dptr(X), _dptr(y), n, d, 1 if with_average else 0))

    @classmethod
    def csr(cls, rowptr, col, val, y):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        return cls(load().irwls_dataset_from_csr(rowptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                                 col.ctypes.data_as(C.POINTER(C.c_int)),
                                                 _dptr(val), _dptr(y), len(rowptr) - 1))

    @classmethod
    def read(cls, path, csv=False, labelled=True, sep=b","):
        L = load()
        p = path.encode() if isinstance(path, str) else path
        if labelled:
            ds = L.readTrainFileCSV(p, sep) if csv else L.readTrainFile(p)
        else:
            ds = L.readUnlabeledFileCSV(p, sep) if csv else L.readUnlabeledFile(p)
        return cls(ds)

    @classmethod
    def load(cls, path):
        """A dataset saved with save() (the library's binary cache)."""
        return cls(load().irwls_dataset_load(os.fspath(path).encode()))

    def save(self, path, with_means=True):
        if load().irwls_dataset_save(C.byref(self.ds), 1 if with_means else 0, os.fspath(path).encode()) != 0:
            raise OSError("cannot write %s" % path)

    def arrays(self, with_means=True):
        return dataset_arrays(self.ds, with_means)

    def pin(self, with_means=True):
        load().irwls_gpu_pin_dataset(self.ds, 1 if with_means else 0)

    def unpin(self):
        load().irwls_gpu_unpin_dataset(self.ds)

    def free(self):
        if self.ds is not None:
            load().irwls_gpu_unpin_dataset(self.ds)
            load().freeDataset(self.ds)
            self.ds = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Model:
    def __init__(self, m):
        self.m = m

    def bytes(self, tmpdir=None):
        import tempfile
        d = tmpdir or tempfile.mkdtemp()
        path = os.path.join(d, "model.bin")
        self.store(path)
        with open(path, "rb") as f:
            return f.read()

    def store(self, path):
        f = _libc.fopen(path.encode(), b"wb")
        load().storeModel(C.byref(self.m), f)
        _libc.fclose(f)

    @classmethod
    def read(cls, path):
        m = model()
        f = _libc.fopen(path.encode(), b"rb")
        if not f:
            raise FileNotFoundError(path)
        load().readModel(C.byref(m), f)
        _libc.fclose(f)
        return cls(m)

    def free(self):
        if self.m is not None:
            load().freeModel(self.m)
            self.m = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def train_full(ds, props, seed=0):
    """trainFULL with srand(seed) first, as the full-train CLI does
    (src/Exec-full-train.c:59).  Returns l+1 doubles (last = bias)."""
    srand(seed)
    return take_doubles(load().trainFULL(ds.ds, props), ds.ds.l + 1)


def sub_irwls(ds, props, GIN, e, beta):
    GIN = np.ascontiguousarray(GIN, dtype=np.float64).copy()
    e = np.ascontiguousarray(e, dtype=np.float64).copy()
    beta = np.ascontiguousarray(beta, dtype=np.float64).copy()
    p = load().subIRWLS(ds.ds, props, _dptr(GIN), _dptr(e), _dptr(beta))
    return take_doubles(p, ds.ds.l + 1), e, beta


def full_model(props, ds, beta):
    b = np.ascontiguousarray(beta, dtype=np.float64)
    return Model(load().calculateFULLModel(props, ds.ds, _dptr(b)))


def sgma(ds, props, seed=1):
    srand(seed)
    return take_ints(load().SGMA(ds.ds, props), props.size)


def random_centroids(ds, props, seed=1):
    srand(seed)
    return take_ints(load().randomCentroids(ds.ds, props), props.size)


def irwls_par(ds, centroids, props):
    c = np.ascontiguousarray(centroids, dtype=np.int32)
    return take_doubles(load().IRWLSpar(ds.ds, c.ctypes.data_as(C.POINTER(C.c_int)), props), props.size)


def budgeted_model(props, ds, centroids, beta):
    c = np.ascontiguousarray(centroids, dtype=np.int32)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    return Model(load().calculateBudgetedModel(props, ds.ds, c.ctypes.data_as(C.POINTER(C.c_int)), _dptr(b)))


def soft_test(ds, m, pprops):
    return take_doubles(load().softTest(ds.ds, m.m, pprops), ds.ds.l)


def soft_test_many(ds, models, pprops):
    """Decision values of several models on one dataset (softTestMany):
    array (len(models), rows), row g equal to soft_test(ds, models[g])."""
    arr = (model * len(models))(*[m.m for m in models])
    out = np.zeros((len(models), ds.ds.l))
    load().softTestMany(ds.ds, arr, len(models), pprops, _dptr(out))
    return out


def one_vs_rest(ds, models, pprops, classes=None):
    """Class of each row from a one-vs-rest bundle: the model with the largest
    decision value (classes[g] for model g, or g)."""
    g = np.argmax(soft_test_many(ds, models, pprops), axis=0)
    return g if classes is None else np.asarray(classes)[g]


def hard_test(ds, m, pprops):
    return take_doubles(load().test(ds.ds, m.m, pprops), ds.ds.l)


def get_options():
    o = irwls_gpu_options()
    load().irwls_gpu_get_options(C.byref(o))
    return o


def set_options(**kw):
    o = get_options()
    for k, v in kw.items():
        setattr(o, k, v)
    load().irwls_gpu_set_options(C.byref(o))
    return o


def timing(enable=True, kinds=None, stride=1):
    """Per-kernel HIP-event timers; `kinds` restricts them to those timer
    names, `stride` times only every stride-th launch of each kind."""
    L = load()
    mask = 0xFFFFFFFF if kinds is None else sum(1 << TIMER_KINDS.index(k) for k in kinds)
    L.irwls_gpu_timing_mask(mask)
    L.irwls_gpu_timing_stride(int(stride))
    L.irwls_gpu_timing(1 if enable else 0)


def timers():
    out = {}
    L = load()
    for k, name in enumerate(TIMER_KINDS):
        n = C.c_long()
        ms, fl, by = C.c_double(), C.c_double(), C.c_double()
        L.irwls_gpu_timer(k, C.byref(n), C.byref(ms), C.byref(fl), C.byref(by))
        out[name] = dict(launches=n.value, ms=ms.value, flops=fl.value, bytes=by.value)
    return out


def last_stats():
    s = irwls_gpu_stats()
    load().irwls_gpu_last_stats(C.byref(s))
    return {k: getattr(s, k) for k, _ in irwls_gpu_stats._fields_}


def linear_system(A, b):
    """ParallelLinearSystem on a column-major copy of A; returns (x, factor)."""
    n = A.shape[0]
    buf = np.asfortranarray(A, dtype=np.float64).ravel(order="F").copy()
    rhs = np.ascontiguousarray(b, dtype=np.float64).copy()
    x = np.zeros(n)
    load().ParallelLinearSystem(_dptr(buf), n, n, 0, 0, _dptr(rhs), n, 1, 0, 0, n, 1, _dptr(x), n, 1, 0, 0, 1)
    return x, buf.reshape((n, n), order="F")


def train_full_batched(datasets, props_list):
    """trainFULLBatched over [(Dataset, properties)]: one beta array (l+1) per
    problem, each equal to a standalone trainFULL run after srand(0)."""
    n = len(datasets)
    L = load()
    DS = (svm_dataset * n)(*[d.ds for d in datasets])
    P = (properties * n)(*props_list)
    out = (C.POINTER(C.c_double) * n)()
    L.trainFULLBatched.argtypes = [C.POINTER(svm_dataset), C.POINTER(properties), C.c_int,
                                   C.POINTER(C.POINTER(C.c_double))]
    L.trainFULLBatched.restype = C.c_int
    L.trainFULLBatched(DS, P, n, out)
    return [take_doubles(out[i], datasets[i].ds.l + 1) for i in range(n)]
