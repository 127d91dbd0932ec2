"""GPU tests of the device memory cache behind irwls_dev_alloc / irwls_dev_free:
a freed block is handed out again for a request of the same size class and
comes back zero-filled, a block in use is never handed out twice, blocks above
the cache limit go back to the driver, and trim empties the cache (later
allocations still work)."""
import numpy as np
import pytest

from libirwls_amd import lib

pytestmark = pytest.mark.gpu


def _fill(L, p, nbytes, v):
    a = np.full(nbytes // 8, v)
    L.irwls_dev_h2d(p, a.ctypes.data, nbytes)


def _get(L, p, nbytes):
    out = np.empty(nbytes // 8)
    L.irwls_dev_d2h(out.ctypes.data, p, nbytes)
    return out


def test_freed_block_is_reused_and_zero_filled():
    L = lib.load()
    n = 3 << 20                      # size class 3 MiB (2 MiB x 1.5)
    p = L.irwls_dev_alloc(n)
    _fill(L, p, n, 7.0)
    L.irwls_dev_free(p)
    q = L.irwls_dev_alloc(n - 4096)  # same class: the block just freed
    assert q == p
    assert not _get(L, q, n - 4096).any()
    r = L.irwls_dev_alloc(n)         # q is in use: another block
    assert r != q
    _fill(L, q, n - 4096, 1.0)
    _fill(L, r, n, 2.0)
    assert (_get(L, q, n - 4096) == 1.0).all() and (_get(L, r, n) == 2.0).all()
    L.irwls_dev_free(q)
    L.irwls_dev_free(r)


def test_large_blocks_and_trim():
    L = lib.load()
    big = (1 << 30) + (1 << 28)      # above the 1 GiB cache limit: freed at once
    p = L.irwls_dev_alloc(big)
    _fill(L, p, 1 << 20, 3.0)
    L.irwls_dev_free(p)
    q = L.irwls_dev_alloc(big)
    assert not _get(L, q, 1 << 20).any()
    L.irwls_dev_free(q)
    small = [L.irwls_dev_alloc(5000 * (k + 1)) for k in range(8)]
    for s in small:
        L.irwls_dev_free(s)
    L.irwls_dev_trim()
    t = L.irwls_dev_alloc(40000)
    assert not _get(L, t, 40000).any()
    L.irwls_dev_free(t)
