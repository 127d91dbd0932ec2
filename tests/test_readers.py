"""CPU tests of the dataset readers and the binary model file against the
reference's own code (oracle/_ref/libref_io.so, compiled from
src/IOStructures.c by oracle/Makefile).  Every field of the datasets is
compared exactly: row count, `sparse`, `maxdim`, labels, |x|^2 and every
(index, value) entry, including the two class-mean rows the labelled
readers append (src/IOStructures.c:138-265, :349-483).  Model files written
by either side are read by the other and must be byte-identical.
Skipped where the reference tree was not present to build the library."""
import ctypes as C
import os

import numpy as np
import pytest

from libirwls_amd import abi, lib
from oracle import refio

pytestmark = pytest.mark.skipif(not refio.available(), reason="oracle/_ref not built (needs the reference tree)")


def _dump(ds, rows):
    return dict(
        l=ds.l, sparse=ds.sparse, maxdim=ds.maxdim,
        y=[ds.y[i] for i in range(rows)] if ds.y else None,
        qv=[ds.quadratic_value[i] for i in range(rows)],
        rows=[_row(ds.x[i]) for i in range(rows)])


def _row(p):
    out, k = [], 0
    while p[k].index != -1:
        out.append((p[k].index, p[k].value))
        k += 1
    return out


def _fmt(rng, v):
    return rng.choice(["%.17g", "%.6g", "%g", "%.3e"]) % v


def _libsvm(path, n, d, rng, dense=False, labelled=True):
    with open(path, "w") as f:
        for i in range(n):
            if dense:
                idx = np.arange(1, d + 1)
            else:
                idx = np.sort(rng.choice(np.arange(1, d + 1), rng.integers(1, d + 1), replace=False))
            vals = rng.standard_normal(len(idx)) * rng.choice([1.0, 1e-3, 100.0])
            toks = ["%d:%s" % (k, _fmt(rng, v)) for k, v in zip(idx, vals)]
            if labelled:
                toks.insert(0, rng.choice(["1", "-1", "+1", "1.0"]) if i % 7 else ("1" if i % 2 else "-1"))
            f.write(" ".join(toks) + "\n")


def _csv(path, n, d, rng, sep, labelled=True):
    with open(path, "w") as f:
        for i in range(n):
            v = rng.standard_normal(d)
            v[rng.random(d) < 0.3] = 0.0
            toks = [("0" if x == 0.0 else _fmt(rng, x)) for x in v]
            if labelled:
                toks.insert(0, "1" if rng.random() < 0.5 else "-1")
            f.write(sep.join(toks) + "\n")


def _compare(mine, ref, rows):
    a, b = _dump(mine, rows), _dump(ref, rows)
    for key in ("l", "sparse", "maxdim", "y", "qv"):
        assert a[key] == b[key], key
    for i, (ra, rb) in enumerate(zip(a["rows"], b["rows"])):
        assert ra == rb, "row %d" % i


@pytest.mark.parametrize("dense", [False, True])
def test_libsvm_labelled(tmp_path, dense):
    rng = np.random.default_rng(10 + dense)
    path = str(tmp_path / "train.libsvm")
    _libsvm(path, 60, 25, rng, dense=dense)
    mine = lib.Dataset.read(path)
    ref = refio.lib().readTrainFile(path.encode())
    assert ref.sparse == (0 if dense else 1)
    _compare(mine.ds, ref, ref.l + 2)
    refio.lib().freeDataset(ref)


def test_libsvm_unlabelled(tmp_path):
    rng = np.random.default_rng(12)
    path = str(tmp_path / "test.libsvm")
    _libsvm(path, 40, 18, rng, labelled=False)
    mine = lib.Dataset.read(path, labelled=False)
    ref = refio.lib().readUnlabeledFile(path.encode())
    a, b = _dump(mine.ds, ref.l), _dump(ref, ref.l)
    for key in ("l", "sparse", "maxdim", "qv", "rows"):
        assert a[key] == b[key], key
    refio.lib().freeDataset(ref)


@pytest.mark.parametrize("sep", [",", ";"])
def test_csv_labelled(tmp_path, sep):
    rng = np.random.default_rng(13 + len(sep) + ord(sep))
    path = str(tmp_path / "train.csv")
    _csv(path, 50, 12, rng, sep)
    mine = lib.Dataset.read(path, csv=True, sep=sep.encode())
    ref = refio.lib().readTrainFileCSV(path.encode(), sep.encode())
    _compare(mine.ds, ref, ref.l + 2)
    refio.lib().freeDataset(ref)


def test_csv_unlabelled(tmp_path):
    """The reference's readUnlabeledFileCSV cannot read such a file: its
    second pass continues strtok from the first pass's last position instead
    of tokenizing the new line, and rejects a last value followed by a
    newline, exiting the process (src/IOStructures.c:694-721).  So the
    unlabelled reader here is checked against the labelled one (itself
    pinned to the reference above) on the same values with a label column."""
    rng = np.random.default_rng(15)
    pu, pl = str(tmp_path / "test.csv"), str(tmp_path / "train.csv")
    _csv(pu, 30, 9, rng, ",", labelled=False)
    with open(pu) as f, open(pl, "w") as g:
        for i, line in enumerate(f):
            g.write(("1," if i % 2 else "-1,") + line)
    mine = lib.Dataset.read(pu, csv=True, labelled=False)
    ref = refio.lib().readTrainFileCSV(pl.encode(), b",")
    _compare_features(mine.ds, ref, ref.l)
    refio.lib().freeDataset(ref)


def _compare_features(mine, ref, rows):
    a, b = _dump(mine, rows), _dump(ref, rows)
    for key in ("l", "sparse", "maxdim", "qv", "rows"):
        assert a[key] == b[key], key


@pytest.mark.parametrize("fmt", ["libsvm", "csv"])
def test_large_file_read_in_parallel_chunks(tmp_path, fmt):
    """Files of a few MB are parsed in chunks on several threads; the
    joined result equals the reference's single pass."""
    rng = np.random.default_rng(17)
    path = str(tmp_path / ("big." + fmt))
    if fmt == "libsvm":
        _libsvm(path, 5000, 60, rng)
    else:
        _csv(path, 5000, 80, rng, ",")
    assert os.path.getsize(path) > 1 << 20   # at least four chunks of 256 KiB
    mine = lib.Dataset.read(path, csv=(fmt == "csv"))
    ref = (refio.lib().readTrainFileCSV(path.encode(), b",") if fmt == "csv"
           else refio.lib().readTrainFile(path.encode()))
    _compare(mine.ds, ref, ref.l + 2)
    refio.lib().freeDataset(ref)


@pytest.mark.parametrize("labelled", [True, False])
def test_binary_cache_round_trip(tmp_path, labelled):
    """A parsed dataset saved to the binary cache loads back equal to the
    reference reader's result, class-mean rows included; a file that is not
    a cache is rejected the way the text readers reject bad input."""
    rng = np.random.default_rng(18)
    path, cache = str(tmp_path / "d.libsvm"), tmp_path / "d.irwlsds"
    _libsvm(path, 300, 40, rng, labelled=labelled)
    lib.Dataset.read(path, labelled=labelled).save(cache, with_means=labelled)
    back = lib.Dataset.load(cache)
    if labelled:
        ref = refio.lib().readTrainFile(path.encode())
        _compare(back.ds, ref, ref.l + 2)
    else:
        ref = refio.lib().readUnlabeledFile(path.encode())
        _compare_features(back.ds, ref, ref.l)
    refio.lib().freeDataset(ref)
    import subprocess
    import sys
    code = "from libirwls_amd import lib; lib.Dataset.load(%r)" % path
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 2 and "Wrong file format" in r.stderr


def test_binary_cache_of_an_empty_dataset(tmp_path):
    """An empty unlabelled dataset round-trips through the cache."""
    path, cache = tmp_path / "empty.libsvm", tmp_path / "empty.irwlsds"
    path.write_text("")
    lib.Dataset.read(str(path), labelled=False).save(cache, with_means=False)
    back = lib.Dataset.load(cache)
    assert back.ds.l == 0


def test_model_file_both_ways(tmp_path):
    """A model built here, written by this library and by the reference,
    gives the same bytes; each side reads the other's file back."""
    rng = np.random.default_rng(16)
    path = str(tmp_path / "train.libsvm")
    _libsvm(path, 50, 20, rng)
    ds = lib.Dataset.read(path)
    beta = rng.standard_normal(ds.ds.l + 1)
    beta[:-1][rng.random(ds.ds.l) < 0.5] = 0.0
    props = abi.make_properties(gamma=0.25, C_=3.0)
    m = lib.full_model(props, ds, beta)
    mine_path, ref_path = str(tmp_path / "mine.model"), str(tmp_path / "ref.model")
    m.store(mine_path)
    libc = lib._libc
    f = libc.fopen(ref_path.encode(), b"wb")
    refio.lib().storeModel(C.byref(m.m), f)
    libc.fclose(f)
    assert open(mine_path, "rb").read() == open(ref_path, "rb").read()
    back = abi.model()
    f = libc.fopen(mine_path.encode(), b"rb")
    refio.lib().readModel(C.byref(back), f)
    libc.fclose(f)
    assert (back.nSVs, back.nElem, back.bias, back.Kgamma) == (m.m.nSVs, m.m.nElem, m.m.bias, m.m.Kgamma)
    for i in range(m.m.nSVs):
        assert _row(back.x[i]) == _row(m.m.x[i])
        assert back.weights[i] == m.m.weights[i]
    mine_back = lib.Model.read(ref_path)
    assert mine_back.bytes(str(tmp_path)) == open(mine_path, "rb").read()
