"""CPU pins against the reference's own kernel functions and predictors
(src/kernels.c, src/LIBIRWLS-predict.c), compiled from its sources into
oracle/_ref/libref_predict.so by oracle/Makefile: the oracle's orc_kernel,
orc_kernel_test, orc_softTest and orc_test, and the library's exported
scalar kernelFunction and kernelTest, equal them bit for bit on dense and
CSR datasets, RBF and linear kernels, the class-mean rows included.
Skipped where the reference tree was not present to build the library."""
import numpy as np
import pytest

from libirwls_amd import abi, lib, synth
from oracle import pyoracle as po
from oracle import refio

pytestmark = pytest.mark.skipif(not refio.predict_available(),
                                reason="oracle/_ref not built (needs the reference tree)")


def _datasets():
    X, y = synth.two_gaussians(120, 9, seed=3)
    rng = np.random.default_rng(4)
    rp, col, val = synth.sparse_uniform(150, 40, 0.15, rng)
    ys = np.where(rng.random(150) < 0.5, 1.0, -1.0)
    return {"dense": (po.Dataset.dense(X, y), lib.Dataset.dense(X, y)),
            "csr": (po.Dataset.csr(rp, col, val, ys), lib.Dataset.csr(rp, col, val, ys))}


@pytest.mark.parametrize("kind", ["dense", "csr"])
@pytest.mark.parametrize("kernel", [1, 0])
def test_kernel_function_equals_reference(kind, kernel):
    R = refio.predict_lib()
    L = lib.load()
    ods, gds = _datasets()[kind]
    props = abi.make_properties(gamma=0.37, kernel=kernel)
    n = ods.ds.l
    rng = np.random.default_rng(5)
    pairs = [(int(i), int(j)) for i, j in rng.integers(0, n + 2, size=(300, 2))] + [(n, n + 1), (n + 1, 3), (0, 0)]
    for i, j in pairs:
        want = R.kernelFunction(ods.ds, i, j, props)
        assert po.kernel(ods, i, j, props) == want, (i, j)
        assert L.kernelFunction(gds.ds, i, j, props) == want, (i, j)


@pytest.mark.parametrize("kind", ["dense", "csr"])
@pytest.mark.parametrize("kernel", [1, 0])
def test_kernel_test_and_predictors_equal_reference(kind, kernel):
    """kernelTest of test rows against a model's support vectors (the norm
    form with stored norms, also for dense data), and softTest / test over
    the whole test set."""
    R = refio.predict_lib()
    L = lib.load()
    ods, gds = _datasets()[kind]
    props = abi.make_properties(gamma=0.37, C_=10.0, W=60, kernel=kernel)
    rng = np.random.default_rng(6)
    n = ods.ds.l
    beta = np.zeros(n + 1)
    sv = rng.choice(n, 40, replace=False)
    beta[sv] = rng.standard_normal(40)
    beta[-1] = 0.3
    om = po.full_model(props, ods, beta)
    gm = lib.full_model(props, gds, beta)
    assert om.bytes() == gm.bytes()
    for i in range(0, n, 7):
        for j in range(om.m.nSVs):
            want = R.kernelTest(ods.ds, i, om.m, j)
            assert po.lib().orc_kernel_test(ods.ds, i, om.m, j) == want
            assert L.kernelTest(gds.ds, i, gm.m, j) == want
    pp = abi.make_predict_properties(labels=0)
    soft = po.take_doubles(R.softTest(ods.ds, om.m, pp), n)
    hard = po.take_doubles(R.test(ods.ds, om.m, pp), n)
    assert np.array_equal(po.soft_test(ods, om), soft)
    assert np.array_equal(po.hard_test(ods, om), hard)
    assert np.array_equal(hard, np.where(soft >= 0.0, 1.0, -1.0))


def test_budgeted_model_predictor_equals_reference():
    """A budgeted model (centroids l and l+1 among its vectors, so the stored
    norm of the negative-class mean row enters) predicted by the reference's
    softTest and by the oracle's."""
    R = refio.predict_lib()
    ods, _ = _datasets()["csr"]
    props = abi.make_properties(gamma=0.37, C_=10.0, size=12)
    n = ods.ds.l
    cent = np.array([n, n + 1] + list(range(0, 50, 5)), dtype=np.int32)
    beta = np.random.default_rng(7).standard_normal(12)
    om = po.budgeted_model(props, ods, cent, beta)
    pp = abi.make_predict_properties(labels=0)
    assert np.array_equal(po.soft_test(ods, om), po.take_doubles(R.softTest(ods.ds, om.m, pp), n))
