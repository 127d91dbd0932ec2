"""CPU tests of the numpy conversion behind the Python 3 module
(irwls_dataset_from_numpy): the reference extension's numpy2dataset and
numpy2datasetWithAverage (python-package/pythonmodule.c:210-330).

A hand-worked 3 x 2 example fixes every field, including the extension's
misplaced mean-row terminator (the first mean row gains an index-0 entry,
the second mean row is empty) and its mean-row "norms" (sums of values);
a larger random case checks the same rules row by row."""
import ctypes as C

import numpy as np
import pytest

from libirwls_amd import lib


def _rows(ds, count):
    out = []
    for i in range(count):
        row, k = [], 0
        while ds.x[i][k].index != -1:
            row.append((ds.x[i][k].index, ds.x[i][k].value))
            k += 1
        out.append(row)
    return out


def _vec(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def test_hand_worked_example_without_average():
    X = np.array([[1.0, 0.0], [0.0, 2.0], [3.0, 4.0]])
    y = np.array([1.0, -1.0, 1.0])
    ds = lib.Dataset.numpy(X, y)
    d = ds.ds
    assert (d.l, d.sparse, d.maxdim) == (3, 1, 2)
    assert _rows(d, 3) == [[(1, 1.0)], [(2, 2.0)], [(1, 3.0), (2, 4.0)]]
    assert list(_vec(d.y, 3)) == [1.0, -1.0, 1.0]
    assert list(_vec(d.quadratic_value, 3)) == [1.0, 4.0, 25.0]
    ds.free()


def test_hand_worked_example_with_average():
    X = np.array([[1.0, 0.0], [0.0, 2.0], [3.0, 4.0]])
    y = np.array([1.0, -1.0, 1.0])
    ds = lib.Dataset.numpy(X, y, with_average=True)
    d = ds.ds
    assert (d.l, d.sparse, d.maxdim) == (3, 1, 2)
    rows = _rows(d, 5)
    assert rows[:3] == [[(1, 1.0)], [(2, 2.0)], [(1, 3.0), (2, 4.0)]]
    # two positive rows: means (1 + 3)/2 = 2 and 4/2 = 2, then the extension's stray (0, 0.0) entry
    assert rows[3] == [(1, 2.0), (2, 2.0), (0, 0.0)]
    assert rows[4] == []   # its first entry was overwritten by the misplaced terminator
    assert list(_vec(d.y, 5)) == [1.0, -1.0, 1.0, 1.0, -1.0]
    # norms of the mean rows: plain sums of the accumulated means (negative: 0 + 2/1)
    assert list(_vec(d.quadratic_value, 5)) == [1.0, 4.0, 25.0, 4.0, 2.0]
    ds.free()


def test_all_zero_matrix_is_not_sparse():
    ds = lib.Dataset.numpy(np.zeros((4, 3)), np.array([1.0, -1.0, 1.0, -1.0]))
    assert ds.ds.sparse == 0 and ds.ds.maxdim == 3
    assert _rows(ds.ds, 4) == [[], [], [], []]
    assert list(_vec(ds.ds.quadratic_value, 4)) == [0.0] * 4
    ds.free()


@pytest.mark.parametrize("with_average", [False, True])
def test_random_matrix_row_by_row(with_average):
    rng = np.random.default_rng(5)
    n, d = 40, 7
    X = rng.standard_normal((n, d))
    X[rng.random((n, d)) < 0.4] = 0.0
    y = rng.choice([1.0, -1.0, 2.0], n)   # 2.0 counts with the negatives (not exactly 1.0)
    ds = lib.Dataset.numpy(X, y, with_average=with_average)
    rows = _rows(ds.ds, n)
    q = _vec(ds.ds.quadratic_value, n + (2 if with_average else 0))
    for i in range(n):
        nz = np.nonzero(X[i])[0]
        assert rows[i] == [(int(k) + 1, X[i, k]) for k in nz]
        s = 0.0
        for k in nz:
            s += X[i, k] ** 2
        assert q[i] == s
    if with_average:
        pos = int(np.sum(y == 1.0))
        neg = n - pos
        mp, mn = np.zeros(d), np.zeros(d)
        for i in range(n):
            for k in np.nonzero(X[i])[0]:
                if y[i] == 1.0:
                    mp[k] += X[i, k] / pos
                else:
                    mn[k] += X[i, k] / neg
        extra = _rows(ds.ds, n + 2)[n:]
        assert extra[0] == [(k + 1, mp[k]) for k in range(d)] + [(0, 0.0)]
        assert extra[1] == []
        assert q[n] == sum(mp[k] for k in range(d)) and q[n + 1] == sum(mn[k] for k in range(d))
        # the overwritten terminator still carries the first negative mean (reference behaviour)
        assert ds.ds.x[n + 1][0].value == mn[0]
    ds.free()


def test_python_module_exports_the_reference_names():
    from libirwls_amd import LIBIRWLS
    for name in ("full_train", "budgeted_train", "predict", "save", "load"):
        assert callable(getattr(LIBIRWLS, name))
    import inspect
    assert list(inspect.signature(LIBIRWLS.full_train).parameters) == [
        "data", "labels", "gamma", "C", "threads", "workingSet", "eta", "kernel", "verbose"]
    assert list(inspect.signature(LIBIRWLS.budgeted_train).parameters) == [
        "data", "labels", "gamma", "C", "threads", "size", "algorithm", "kernel", "verbose"]
    assert list(inspect.signature(LIBIRWLS.predict).parameters) == [
        "classifier", "data", "labels", "threads", "Soft", "verbose"]
