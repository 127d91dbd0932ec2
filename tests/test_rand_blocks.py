"""CPU tests of the library's lock-free blocks of libc rand() draws
(rpermute, randomCentroids): the permutations equal the reference's
Fisher-Yates over libc rand() (src/full-train.c:71, src/budgeted-train.c:84),
and libc's generator continues afterwards exactly as if rand() had been
called once per draw, also after initstate() with the caller's own buffer
of each generator size."""
import ctypes as C

import numpy as np
import pytest

from libirwls_amd import abi, lib

libc = C.CDLL(None)
libc.rand.restype = C.c_int
libc.srand.argtypes = [C.c_uint]
libc.initstate.restype = C.c_void_p
libc.initstate.argtypes = [C.c_uint, C.c_void_p, C.c_size_t]
libc.setstate.restype = C.c_void_p
libc.setstate.argtypes = [C.c_void_p]


def _fisher_yates_libc(n):
    p = list(range(n))
    for k in range(n - 1, 0, -1):
        j = libc.rand() % (k + 1)
        p[j], p[k] = p[k], p[j]
    return p


@pytest.mark.parametrize("n", [1, 2, 17, 5000])
def test_rpermute_equals_libc_and_stream_continues(n):
    L = lib.load()
    L.rpermute.restype = C.POINTER(C.c_int)
    L.rpermute.argtypes = [C.c_int]
    libc.srand(1234)
    want = _fisher_yates_libc(n)
    after = [libc.rand() for _ in range(5)]
    libc.srand(1234)
    got = L.rpermute(n)
    assert [got[i] for i in range(n)] == want
    assert [libc.rand() for _ in range(5)] == after


@pytest.mark.parametrize("size", [8, 32, 64, 128, 256])
def test_draw_block_on_a_caller_state_buffer(size):
    """initstate() with the caller's buffer (TYPE_0 at 8 bytes up to TYPE_4
    at 256): the library's draws and the stream after them match plain
    rand() calls."""
    L = lib.load()
    L.rpermute.restype = C.POINTER(C.c_int)
    L.rpermute.argtypes = [C.c_int]
    buf = C.create_string_buffer(size)
    prev = libc.initstate(77, C.cast(buf, C.c_void_p), size)
    try:
        snap = C.create_string_buffer(buf.raw, size)
        libc.setstate(C.cast(snap, C.c_void_p))      # a copy of the fresh state, to replay from
        want = _fisher_yates_libc(300)
        after = [libc.rand() for _ in range(7)]
        libc.setstate(C.cast(buf, C.c_void_p))
        got = L.rpermute(300)
        assert [got[i] for i in range(300)] == want
        assert [libc.rand() for _ in range(7)] == after
    finally:
        libc.setstate(prev)


def test_random_centroids_equal_libc():
    X = np.random.default_rng(0).standard_normal((400, 3))
    y = np.where(X[:, 0] > 0, 1.0, -1.0)
    ds = lib.Dataset.dense(X, y)
    props = abi.make_properties(size=25)
    libc.srand(9)
    perm = list(range(400))
    for i in range(399, -1, -1):
        j = libc.rand() % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    after = libc.rand()
    got = lib.random_centroids(ds, props, seed=9)
    assert list(got) == perm[:25]
    assert libc.rand() == after
