"""CPU oracle of the SGMA proposal network (test infrastructure only).

The learned proposal has no counterpart in the reference: this restates the
network the library trains and evaluates, a 32 -> 128 -> 128 -> 1 MLP with
ReLU hidden layers and mean-squared-error loss, in float64 numpy, with the
library's flat parameter layout W1[32][128], b1[128], W2[128][128], b2[128],
w3[128], b3.  Its gradient is pinned by central finite differences in the
CPU tests; the GPU kernels are checked against it.
"""
import numpy as np

NIN, NH = 32, 128
OW1 = 0
OB1 = OW1 + NIN * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NPARAMS = OB3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(NIN, NH), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3], p[OW3:OB3],
            p[OB3])


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ w3 + b3, (h1, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - T) ** 2))


def grad(p, X, T):
    """Gradient of the mean squared error, flat (NPARAMS,)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    nb = X.shape[0]
    y, (h1, h2) = forward(p, X)
    dy = 2.0 * (y - T) / nb
    g = np.zeros(NPARAMS)
    g[OW3:OB3] = h2.T @ dy
    g[OB3] = dy.sum()
    dh2 = np.outer(dy, w3) * (h2 > 0)
    g[OW2:OB2] = (h1.T @ dh2).ravel()
    g[OB2:OW3] = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    g[OW1:OB1] = (X.T @ dh1).ravel()
    g[OB1:OW2] = dh1.sum(0)
    return g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
