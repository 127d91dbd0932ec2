"""GPU tests of the Python 3 module with the reference extension's interface
(libirwls_amd.LIBIRWLS).  Each trainer and the predictor run through the
module and are compared with the oracle run on the same converted dataset
(the extension's numpy conversion: zeros dropped, sparse rows, and for the
budgeted trainer its two mean rows as it builds them), with rand() seeded
alike on both sides.  Inputs have 30% exact zeros so the conversion drops
entries."""
import types

import numpy as np
import pytest

from libirwls_amd import LIBIRWLS, abi, lib, synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _data(n, d, seed, zero_frac=0.3):
    X, y = synth.two_gaussians(n, d, seed=seed)
    rng = np.random.default_rng(seed + 1)
    X[rng.random(X.shape) < zero_frac] = 0.0
    return X, y


def _view(ds):
    return types.SimpleNamespace(ds=ds.ds)


def test_full_train_predict_save_load_against_oracle(tmp_path):
    X, y = _data(800, 8, 21)
    Xt, yt = _data(300, 8, 22)
    g, C_, W = 0.1, 10.0, 120
    lib.srand(0)
    clf = LIBIRWLS.full_train(X, y, gamma=g, C=C_, workingSet=W, verbose=0)
    props = abi.make_properties(gamma=g, C_=C_, W=W, eta=0.001)
    ods = lib.Dataset.numpy(X, y)
    want = po.train_full(_view(ods), props, seed=0)
    om = po.full_model(props, _view(ods), want)
    print("nSVs gpu", clf.nSVs, "oracle", om.m.nSVs, "bias", clf.bias, om.m.bias)
    assert clf.nSVs == om.m.nSVs
    otd = lib.Dataset.numpy(Xt, yt)
    soft = LIBIRWLS.predict(clf, Xt, Soft=1, verbose=0)
    np.testing.assert_allclose(soft, po.soft_test(_view(otd), om), rtol=1e-6, atol=1e-8)
    hard = LIBIRWLS.predict(clf, Xt, yt, verbose=0)
    assert np.array_equal(hard, po.hard_test(_view(otd), om))
    assert np.mean(hard == yt) > 0.8
    path = tmp_path / "full.model"
    LIBIRWLS.save(clf, path)
    clf2 = LIBIRWLS.load(path)
    assert clf2.nSVs == clf.nSVs
    assert np.array_equal(LIBIRWLS.predict(clf2, Xt, Soft=1, verbose=0), soft)


@pytest.mark.parametrize("algorithm", [0, 1])
def test_budgeted_train_against_oracle(algorithm):
    """algorithm 0: randomCentroids; 1: SGMA, whose first two centroids are
    the extension's mean rows (the second one empty)."""
    X, y = _data(1000, 6, 31)
    Xt, yt = _data(300, 6, 32)
    g, C_, B = 0.2, 10.0, 40
    lib.srand(7)
    clf = LIBIRWLS.budgeted_train(X, y, gamma=g, C=C_, size=B, algorithm=algorithm, verbose=0)
    props = abi.make_properties(gamma=g, C_=C_, size=B, algorithm=algorithm)
    ods = lib.Dataset.numpy(X, y, with_average=True)
    if algorithm == 0:
        cent = po.random_centroids(_view(ods), props, seed=7)
    else:
        cent = po.sgma(_view(ods), props, ncand=64, seed=7)
        assert list(cent[:2]) == [1000, 1001]
    beta = po.irwls_par(_view(ods), cent, props)
    om = po.budgeted_model(props, _view(ods), cent, beta)
    assert clf.nSVs == om.m.nSVs
    otd = lib.Dataset.numpy(Xt, yt)
    soft = LIBIRWLS.predict(clf, Xt, Soft=1, verbose=0)
    np.testing.assert_allclose(soft, po.soft_test(_view(otd), om), rtol=1e-6, atol=1e-8)
    hard = LIBIRWLS.predict(clf, Xt, verbose=0)
    assert np.array_equal(hard, po.hard_test(_view(otd), om))
    print("algorithm", algorithm, "accuracy", np.mean(hard == yt))
    assert np.mean(hard == yt) > 0.75
