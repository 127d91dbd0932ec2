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


@pytest.mark.parametrize("kind", ["dense", "csr"])
def test_pinned_test_rows_predict_as_unpinned(kind):
    """softTest and softTestMany with the test rows pinned in HBM (their
    feature range then comes from the device copy, not a scan of the rows)
    equal the unpinned calls bit for bit, also for a model whose support
    vectors reach a higher feature index than the test rows."""
    from libirwls_amd import synth
    rng = np.random.default_rng(11)
    if kind == "dense":
        X, c = _mixture(1500, 12, 5)
        y = np.where(c == 0, 1.0, -1.0)
        ds = lib.Dataset.dense(X, y)
        Xt, _ = _mixture(400, 9, 6)   # fewer features than the model's rows
        td = lib.Dataset.dense(Xt, np.zeros(len(Xt)))
    else:
        rp, col, val = synth.sparse_uniform(1500, 60, 0.1, rng)
        y = np.where(rng.random(1500) < 0.5, 1.0, -1.0)
        ds = lib.Dataset.csr(rp, col, val, y)
        rpt, colt, valt = synth.sparse_uniform(400, 40, 0.1, rng)
        td = lib.Dataset.csr(rpt, colt, valt, np.zeros(400))
    props = abi.make_properties(gamma=0.2, C_=10.0, W=200)
    m = lib.full_model(props, ds, lib.train_full(ds, props, seed=0))
    pb = abi.make_properties(gamma=0.2, C_=1.0, size=40, algorithm=0)
    cent = lib.random_centroids(ds, pb, seed=2)
    mb = lib.budgeted_model(pb, ds, cent, lib.irwls_par(ds, cent, pb))
    pp = abi.make_predict_properties(labels=0)
    want = [lib.soft_test(td, mm, pp) for mm in (m, mb)]
    want_many = lib.soft_test_many(td, [m, mb], pp)
    td.pin(with_means=False)
    try:
        got = [lib.soft_test(td, mm, pp) for mm in (m, mb)]
        got_many = lib.soft_test_many(td, [m, mb], pp)
    finally:
        td.unpin()
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    assert np.array_equal(got_many, want_many)
