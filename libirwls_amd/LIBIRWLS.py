"""Python 3 module with the interface of the reference's Python extension
LIBIRWLS (python-package/pythonmodule.c): full_train, budgeted_train,
predict, save and load on numpy arrays, over this library's C ABI (the GPU
trainers and predictor behind the reference's function names).

The reference extension builds only for CPython 2 (Py_InitModule,
python-package/pythonmodule.c:519).  This module keeps its function names,
keyword names, defaults and call sequence, so a script written for it runs
unchanged apart from the import:

    from libirwls_amd import LIBIRWLS
    clf = LIBIRWLS.full_train(X, y, gamma=0.1, C=10.0, workingSet=500)
    yhat = LIBIRWLS.predict(clf, Xt)

Datasets are converted as the extension converts them
(irwls_dataset_from_numpy): zeros dropped, `sparse` set as soon as one value
is nonzero, and for budgeted training the extension's two mean rows.  Like
the extension, nothing seeds the C library's rand() here: call
`lib.srand(seed)` first for a reproducible working-set or centroid draw.
"""
import ctypes

import numpy as np

from . import abi, lib

__all__ = ["Classifier", "full_train", "budgeted_train", "predict", "save", "load"]


class Classifier:
    """A trained model; the extension's "CLASSIFIER" capsule
    (python-package/pythonmodule.c:490-504).  The model's memory is released
    with freeModel when the object is collected."""

    def __init__(self, owner):
        self._model = owner   # a lib.Model, which frees the struct

    @property
    def model(self):
        """The C `model` struct (include/IOStructures.h)."""
        return self._model.m

    @property
    def nSVs(self):
        return self._model.m.nSVs

    @property
    def bias(self):
        return self._model.m.bias


def _matrix(data):
    X = np.ascontiguousarray(data, dtype=np.float64)
    if X.ndim != 2:
        raise ValueError("data must be a two-dimensional array (rows x features)")
    return X


def _labels(labels, n):
    y = np.ascontiguousarray(labels, dtype=np.float64).reshape(-1)
    if y.shape[0] != n:
        raise ValueError("labels must have one entry per row of data")
    return y


def _properties(gamma, C_, threads, kernel, verbose, MaxSize, Eta, size, algorithm):
    p = abi.properties()
    p.Kgamma = float(gamma)
    p.C = float(C_)
    p.Threads = int(threads)
    p.MaxSize = int(MaxSize)
    p.Eta = float(Eta)
    p.size = int(size)
    p.algorithm = int(algorithm)
    p.kernelType = int(kernel)
    p.verbose = int(verbose)
    p.file = 0
    p.separator = b","
    return p


def full_train(data, labels, gamma=1.0, C=1.0, threads=1, workingSet=500, eta=0.001, kernel=1, verbose=1):
    """Full SVM training by working-set IRWLS (trainFULL), returned as a
    Classifier.  Reference: python-package/pythonmodule.c:450-505 (defaults
    gamma 1, C 1, threads 1, workingSet 500, eta 0.001, kernel 1 = RBF,
    verbose 1; initMemory, trainFULL, calculateFULLModel)."""
    L = lib.load()
    X = _matrix(data)
    y = _labels(labels, X.shape[0])
    props = _properties(gamma, C, threads, kernel, verbose, MaxSize=workingSet, Eta=eta, size=10, algorithm=0)
    ds = lib.Dataset.numpy(X, y)
    try:
        L.initMemory(props.Threads, props.MaxSize + 1)
        W = L.trainFULL(ds.ds, props)
        m = L.calculateFULLModel(props, ds.ds, W)
        lib._libc.free(ctypes.cast(W, ctypes.c_void_p))
    finally:
        ds.free()
    return Classifier(lib.Model(m))


def budgeted_train(data, labels, gamma=1.0, C=1.0, threads=1, size=10, algorithm=0, kernel=1, verbose=1):
    """Budgeted SVM training: `size` centroids drawn at random (algorithm 0,
    randomCentroids) or by SGMA (algorithm 1), then IRWLSpar, returned as a
    Classifier.  Reference: python-package/pythonmodule.c:369-438 (defaults
    gamma 1, C 1, threads 1, size 10, algorithm 0, kernel 1, verbose 1)."""
    L = lib.load()
    X = _matrix(data)
    y = _labels(labels, X.shape[0])
    props = _properties(gamma, C, threads, kernel, verbose, MaxSize=0, Eta=0.001, size=size, algorithm=algorithm)
    ds = lib.Dataset.numpy(X, y, with_average=True)
    try:
        L.initMemory(props.Threads, props.size)
        cent = L.randomCentroids(ds.ds, props) if props.algorithm == 0 else L.SGMA(ds.ds, props)
        W = L.IRWLSpar(ds.ds, cent, props)
        m = L.calculateBudgetedModel(props, ds.ds, cent, W)
        lib._libc.free(ctypes.cast(W, ctypes.c_void_p))
        lib._libc.free(ctypes.cast(cent, ctypes.c_void_p))
    finally:
        ds.free()
    return Classifier(lib.Model(m))


def predict(classifier, data, labels=None, threads=1, Soft=0, verbose=1):
    """Classifier output for every row of `data`: the labels +1/-1 (test),
    or the decision values with Soft=1 (softTest).  With `labels` the
    accuracy is reported as the predictor reports it.  Reference:
    python-package/pythonmodule.c:72-128."""
    L = lib.load()
    X = _matrix(data)
    has_labels = labels is not None
    y = _labels(labels, X.shape[0]) if has_labels else np.zeros(X.shape[0])
    pp = abi.make_predict_properties(labels=1 if has_labels else 0, soft=int(Soft), threads=int(threads),
                                     verbose=int(verbose))
    ds = lib.Dataset.numpy(X, y)
    try:
        fn = L.softTest if pp.Soft else L.test
        out = lib.take_doubles(fn(ds.ds, classifier.model, pp), X.shape[0])
    finally:
        ds.free()
    return out


def save(classifier, file):
    """Writes the model file (storeModel's raw format).  Reference:
    python-package/pythonmodule.c:138-160."""
    classifier._model.store(str(file))


def load(file):
    """Reads a model file written by save or by the command-line trainers.
    Reference: python-package/pythonmodule.c:173-198."""
    return Classifier(lib.Model.read(str(file)))
