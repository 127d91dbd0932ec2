"""Seeded synthetic workloads for the IRWLS trainers.

Each generator returns plain numpy arrays.  Choices the workload text leaves
open are fixed here and recorded in the docstrings: class priors are
balanced coin flips, seeds are fixed per workload, anchor weights are N(0,1),
the test sets are drawn from the same generator with the next seed.
"""
import numpy as np


def two_gaussians(n, d=16, seed=0, flip=0.05, shift=0.5):
    """Two-class Gaussian mixture: y = +-1 with P=1/2, x ~ N(shift*y*1, I),
    then a fraction `flip` of the labels flipped uniformly at random."""
    rng = np.random.default_rng(seed)
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    X = rng.standard_normal((n, d)) + shift * y[:, None]
    nf = int(round(flip * n))
    if nf:
        idx = rng.choice(n, size=nf, replace=False)
        y[idx] = -y[idx]
    return np.ascontiguousarray(X), y


def four_gaussians(n, d=128, seed=1, mean_var=0.25, means_seed=None):
    """Four-component mixture: component means mu_k ~ N(0, mean_var*I) drawn
    once from `means_seed` (default: `seed`), components 0,1 -> y=+1 and
    2,3 -> y=-1 with equal weights, x ~ N(mu_k, I)."""
    mrng = np.random.default_rng(seed if means_seed is None else means_seed)
    mu = mrng.standard_normal((4, d)) * np.sqrt(mean_var)
    rng = np.random.default_rng(seed + 1000)
    comp = rng.integers(0, 4, size=n)
    X = rng.standard_normal((n, d)) + mu[comp]
    y = np.where(comp < 2, 1.0, -1.0)
    return np.ascontiguousarray(X), y


def sparse_uniform(n, d, density, rng):
    """CSR rows: each entry nonzero with probability `density` (independently),
    values U[0,1].  Returns (rowptr int64, col int32 0-based sorted per row,
    val float64).  The nonzero positions of the flattened n x d index space
    are drawn as cumulative sums of geometric gaps, which is the Bernoulli
    process itself, in one vectorised pass."""
    total = n * d
    pos = []
    start = -1
    while True:
        m = max(1024, int((total - start) * density * 1.05) + 64)
        gaps = rng.geometric(density, size=m)
        p = start + np.cumsum(gaps)
        p = p[p < total]
        pos.append(p)
        if len(p) < m:
            break
        start = int(p[-1])
    pos = np.concatenate(pos) if pos else np.zeros(0, dtype=np.int64)
    rows = (pos // d).astype(np.int64)
    col = (pos % d).astype(np.int32)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rowptr[1:])
    val = rng.random(len(col))
    return rowptr, col, val


def _csr_sqnorm(rowptr, val):
    sq = val * val
    out = np.add.reduceat(sq, rowptr[:-1]) if len(sq) else np.zeros(len(rowptr) - 1)
    out[np.diff(rowptr) == 0] = 0.0
    return out


def _csr_dense_dot(rowptr, col, val, Z):
    """rows(CSR) @ Z^T for a dense Z (m x d)."""
    from scipy.sparse import csr_matrix
    n = len(rowptr) - 1
    A = csr_matrix((val, col, rowptr), shape=(n, Z.shape[1]))
    return np.asarray(A @ Z.T)


def sparse_rbf_labels(n, d=2000, density=0.01, anchors=200, gamma=0.5, seed=3, chunk=200000):
    """CSR points with labels sign(f(x) - median f), f(x) = sum_a w_a
    exp(-gamma |x - z_a|^2); anchors z_a from the same generator, w_a ~ N(0,1).
    At 1% density two points rarely share a column, so every k(x, z_a) is
    close to exp(-gamma (|x|^2 + |z_a|^2)) and the plain sign of f would be
    the sign of sum_a w_a for almost every point; the median offset (a
    constant term in the expansion) makes the two classes equal in size.
    Ties map to +1."""
    arng = np.random.default_rng(seed + 7)
    zr, zc, zv = sparse_uniform(anchors, d, density, arng)
    Z = np.zeros((anchors, d))
    for a in range(anchors):
        Z[a, zc[zr[a]:zr[a + 1]]] = zv[zr[a]:zr[a + 1]]
    wa = arng.standard_normal(anchors)
    zn = (Z * Z).sum(1)
    rng = np.random.default_rng(seed)
    rowptr, col, val = sparse_uniform(n, d, density, rng)
    f = np.empty(n)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        rp = rowptr[s:e + 1] - rowptr[s]
        cc = col[rowptr[s]:rowptr[e]]
        vv = val[rowptr[s]:rowptr[e]]
        xn = _csr_sqnorm(rp, vv)
        dots = _csr_dense_dot(rp, cc, vv, Z)
        dist = xn[:, None] + zn[None, :] - 2.0 * dots
        f[s:e] = np.exp(-gamma * dist) @ wa
    y = np.where(f >= np.median(f), 1.0, -1.0)
    return rowptr, col, val, y


def twelve_class_tasks(seed=4, n_tasks=12, n_test=0):
    """Per one-vs-rest task: n log-uniform in [1000, 30000], d from
    {16, 32, 64, 128}; a 12-class mixture with means N(0, I_d), covariance
    0.5*I; task t relabels class t as +1 and the rest -1.  Returns a list of
    (X, y, d, n, Xt, yt) tuples; Xt, yt are n_test held-out points of the
    same task (empty when n_test is 0)."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n_tasks):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(30000))))
        d = int(rng.choice([16, 32, 64, 128]))
        mu = rng.standard_normal((12, d))
        cls = rng.integers(0, 12, size=n + n_test)
        X = mu[cls] + np.sqrt(0.5) * rng.standard_normal((n + n_test, d))
        y = np.where(cls == t, 1.0, -1.0)
        out.append((np.ascontiguousarray(X[:n]), y[:n], d, n, np.ascontiguousarray(X[n:]), y[n:]))
    return out


WORKLOADS = {
    # name: (generator kwargs, train params)
    "two_gaussians_n4000_d16": dict(n=4000, n_test=1000, d=16, gamma=1 / 16, C=10.0, W=500),
    "four_gaussians_n50000_d128": dict(n=50000, n_test=10000, d=128, gamma=1 / 128, C=1.0, W=2000),
    "four_gaussians_n200000_d256": dict(n=200000, n_test=20000, d=256, gamma=1 / 256, C=1.0, W=4000),
}
