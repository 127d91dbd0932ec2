"""Pins of the proposal network's CPU oracle (oracle/mlp_ref.py): its
analytic gradient against central finite differences of its own loss, the
layer sizes, and the Adam step against the textbook update written out for
one coordinate."""
import numpy as np

from oracle import mlp_ref as M


def _problem(seed, nb=40):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(M.NPARAMS) * 0.2
    X = rng.standard_normal((nb, M.NIN))
    T = rng.standard_normal(nb)
    return p, X, T


def test_parameter_count():
    assert M.NPARAMS == 32 * 128 + 128 + 128 * 128 + 128 + 128 + 1 == 20865


def test_gradient_finite_differences():
    p, X, T = _problem(1)
    g = M.grad(p, X, T)
    rng = np.random.default_rng(2)
    # every block of the layout is probed
    idx = np.concatenate([rng.choice(np.arange(lo, hi), 6, replace=False) for lo, hi in
                          [(M.OW1, M.OB1), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3), (M.OW3, M.OB3)]]
                         + [np.array([M.OB3])])
    h = 1e-6
    for k in idx:
        e = np.zeros(M.NPARAMS)
        e[k] = h
        fd = (M.loss(p + e, X, T) - M.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_adam_single_coordinate():
    p = np.array([0.5])
    g = np.array([0.2])
    m = np.zeros(1)
    v = np.zeros(1)
    p1, m1, v1 = M.adam(p, g, m, v, 1, 0.01)
    # first step: m_hat = g, v_hat = g^2, update = lr * sign(g) (up to eps)
    assert np.isclose(p1[0], 0.5 - 0.01 * 0.2 / (0.2 + 1e-8))
    assert np.isclose(m1[0], 0.1 * 0.2) and np.isclose(v1[0], 0.001 * 0.04)
