"""Objectives of a trained budgeted model, evaluated on the host (numpy).

SGMA chooses centroids C to shrink the Nystrom residual of the kernel,
r(x) = k(x, x) - k_C(x)^T K_C^-1 k_C(x), with k(x, x) = 1 for the RBF kernel
(its gain of a candidate is the drop of sum_i r(x_i),
src/budgeted-train.c:176-203).  IRWLSpar then minimises the SVM primal in
the span of those centroids, L_P = 1/2 beta^T K_C beta + C sum_i
max(0, 1 - y_i f(x_i)) with f(x) = sum_j beta_j k(x, c_j) and no bias (its
reweighting D_i = C / (y_i e_i), src/budgeted-train.c:448-457, is the
fixed point of that objective's subgradient condition).

Kernel values use the dataset's stored squared norms (quadratic_value,
including the reader's norm of the negative-class mean row), which is the
kernel the trainers and the predictor evaluate on sparse data
(src/kernels.c:99-116, 162-181).  These helpers only evaluate results; the
trainers never call them.
"""
import ctypes as C

import numpy as np

from .abi import SAMPLE_DTYPE


def rows_csr(ds, rows, d=None):
    """CSR copy (rowptr, col 0-based, val, stored norms, width) of the given
    rows of a dataset built by the library (mean rows l and l+1 included)."""
    rows = np.asarray(rows, dtype=np.int64)
    q = np.ctypeslib.as_array(ds.ds.quadratic_value, shape=(ds.ds.l + 2,))
    cols, vals, rp = [], [], [0]
    for r in rows:
        x = ds.ds.x[int(r)]
        k = 0
        while x[k].index != -1:
            k += 1
        ent = np.frombuffer(C.string_at(C.addressof(x.contents), 16 * k), dtype=SAMPLE_DTYPE) if k else None
        cols.append(ent["index"].astype(np.int64) - 1 if k else np.zeros(0, np.int64))
        vals.append(ent["value"].copy() if k else np.zeros(0))
        rp.append(rp[-1] + k)
    col = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    v = np.concatenate(vals) if vals else np.zeros(0)
    dmax = int(col.max()) + 1 if len(col) else 1
    return np.asarray(rp, np.int64), col, v, q[rows].copy(), max(dmax, d or 0)


def _dense(rp, col, val, d):
    out = np.zeros((len(rp) - 1, d))
    for i in range(len(rp) - 1):
        out[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return out


def kernel_block(A, qa, B, qb, gamma):
    """exp(-gamma (qa_i + qb_j - 2 a_i . b_j)) for dense row blocks."""
    return np.exp(-gamma * ((qa[:, None] + qb[None, :]) - 2.0 * (A @ B.T)))


def nystrom_residual(ds, centroids, sample_ds, gamma, jitter=1e-6):
    """Mean Nystrom residual 1 - k_C(x)^T (K_C + jitter I)^-1 k_C(x) over the
    rows of sample_ds (a held-out set), C = the centroid rows of ds."""
    rp, col, val, qc, d = rows_csr(ds, centroids)
    srp, scol, sval, qs, ds_ = rows_csr(sample_ds, np.arange(sample_ds.ds.l))
    d = max(d, ds_)
    V = _dense(rp, col, val, d)
    S = _dense(srp, scol, sval, d)
    KC = kernel_block(V, qc, V, qc, gamma) + jitter * np.eye(len(qc))
    Ks = kernel_block(S, qs, V, qc, gamma)
    Lc = np.linalg.cholesky(KC)
    Z = np.linalg.solve(Lc, Ks.T)
    return float(np.mean(1.0 - (Z * Z).sum(0)))


def primal_objective(ds, centroids, beta, f_train, C, gamma):
    """1/2 beta^T K_C beta + C sum_i max(0, 1 - y_i f_i), f_i = the model's
    decision values on the training rows (softTest)."""
    rp, col, val, qc, d = rows_csr(ds, centroids)
    V = _dense(rp, col, val, d)
    KC = kernel_block(V, qc, V, qc, gamma)
    beta = np.asarray(beta, dtype=np.float64)
    y = np.ctypeslib.as_array(ds.ds.y, shape=(ds.ds.l,))
    hinge = np.maximum(0.0, 1.0 - y * np.asarray(f_train))
    return float(0.5 * beta @ KC @ beta + C * hinge.sum())
