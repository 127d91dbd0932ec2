"""Device-buffer helpers for the GPU tests (over the library's C-ABI)."""
import numpy as np

from libirwls_amd import lib


class Dev:
    def __init__(self, arr):
        self.L = lib.load()
        arr = np.ascontiguousarray(arr)
        self.nbytes = arr.nbytes
        self.dtype, self.shape = arr.dtype, arr.shape
        self.p = self.L.irwls_dev_alloc(max(arr.nbytes, 8))
        if arr.nbytes:
            self.L.irwls_dev_h2d(self.p, arr.ctypes.data, arr.nbytes)

    def get(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            self.L.irwls_dev_d2h(out.ctypes.data, self.p, self.nbytes)
        return out

    def __del__(self):
        try:
            self.L.irwls_dev_free(self.p)
        except Exception:
            pass


def dense_rows(Xd, nrm, n, ldx, idx=None):
    r = lib.irwls_rows()
    r.X, r.ldx, r.nrm, r.n = Xd.p, ldx, nrm.p, n
    r.idx = idx.p if idx is not None else None
    return r


def rbf(X, Z, g):
    d2 = ((X[:, None, :] - Z[None, :, :]) ** 2).sum(-1)
    return np.exp(-g * d2)


