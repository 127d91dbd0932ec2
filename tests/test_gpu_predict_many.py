"""GPU tests of multi-model prediction (softTestMany): a bundle of models on
one test set gives, row for row, exactly what softTest gives for each model
(the same launches, one upload of the test rows), and a one-vs-rest bundle
classifies a four-class mixture by the largest decision value."""
import numpy as np
import pytest

from libirwls_amd import abi, lib

pytestmark = pytest.mark.gpu


def _mixture(n, d, seed):
    rng = np.random.default_rng(seed)
    means = np.random.default_rng(99).standard_normal((4, d)) * 1.5
    c = rng.integers(0, 4, n)
    return means[c] + rng.standard_normal((n, d)), c


def test_bundle_equals_single_model_predictions_and_ovr():
    d = 10
    X, c = _mixture(2000, d, 1)
    Xt, ct = _mixture(600, d, 2)
    models = []
    for k in range(4):
        y = np.where(c == k, 1.0, -1.0)
        ds = lib.Dataset.dense(X, y)
        props = abi.make_properties(gamma=0.1, C_=10.0, W=200)
        models.append(lib.full_model(props, ds, lib.train_full(ds, props, seed=0)))
    # a budgeted model with another width, and a model without support vectors
    y0 = np.where(c == 0, 1.0, -1.0)
    ds0 = lib.Dataset.dense(X, y0)
    pb = abi.make_properties(gamma=0.3, C_=1.0, size=60, algorithm=0)
    cent = lib.random_centroids(ds0, pb, seed=3)
    models.append(lib.budgeted_model(pb, ds0, cent, lib.irwls_par(ds0, cent, pb)))
    pz = abi.make_properties(gamma=0.1, C_=1.0)
    models.append(lib.full_model(pz, ds0, np.r_[np.zeros(2000), 0.25]))
    assert models[-1].m.nSVs == 0
    td = lib.Dataset.dense(Xt, np.zeros(len(Xt)))
    pp = abi.make_predict_properties(labels=0)
    many = lib.soft_test_many(td, models, pp)
    assert many.shape == (len(models), len(Xt))
    for g, m in enumerate(models):
        assert np.array_equal(many[g], lib.soft_test(td, m, pp)), g
    assert np.all(many[-1] == 0.25)
    cls = lib.one_vs_rest(td, models[:4], pp)
    acc = float(np.mean(cls == ct))
    print("one-vs-rest accuracy %.4f" % acc)
    assert acc > 0.85
    assert np.array_equal(lib.one_vs_rest(td, models[:4], pp, classes=[10, 11, 12, 13]), cls + 10)
