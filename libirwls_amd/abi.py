"""ctypes mirror of the C boundary structs (include/IOStructures.h).

Member order and types follow the reference layout
(reference: include/IOStructures.h:44-122); the sizes are asserted below so a
mismatch fails at import time rather than inside a C call.
"""
import ctypes as C

import numpy as np


class svm_sample(C.Structure):
    _fields_ = [("index", C.c_int), ("value", C.c_double)]


class svm_dataset(C.Structure):
    _fields_ = [
        ("l", C.c_int),
        ("sparse", C.c_int),
        ("maxdim", C.c_int),
        ("y", C.POINTER(C.c_double)),
        ("x", C.POINTER(C.POINTER(svm_sample))),
        ("quadratic_value", C.POINTER(C.c_double)),
        ("features", C.POINTER(svm_sample)),
    ]


class properties(C.Structure):
    _fields_ = [
        ("Kgamma", C.c_double),
        ("kernelType", C.c_int),
        ("C", C.c_double),
        ("Threads", C.c_int),
        ("MaxSize", C.c_int),
        ("size", C.c_int),
        ("algorithm", C.c_int),
        ("Eta", C.c_double),
        ("file", C.c_int),
        ("separator", C.c_char_p),
        ("verbose", C.c_int),
    ]


class predictProperties(C.Structure):
    _fields_ = [
        ("Labels", C.c_int),
        ("Threads", C.c_int),
        ("Soft", C.c_int),
        ("file", C.c_int),
        ("separator", C.c_char_p),
        ("verbose", C.c_int),
    ]


class model(C.Structure):
    _fields_ = [
        ("Kgamma", C.c_double),
        ("kernelType", C.c_int),
        ("sparse", C.c_int),
        ("nSVs", C.c_int),
        ("nElem", C.c_int),
        ("weights", C.POINTER(C.c_double)),
        ("x", C.POINTER(C.POINTER(svm_sample))),
        ("quadratic_value", C.POINTER(C.c_double)),
        ("maxdim", C.c_int),
        ("bias", C.c_double),
        ("features", C.POINTER(svm_sample)),
    ]


assert C.sizeof(svm_sample) == 16
assert C.sizeof(svm_dataset) == 48
assert C.sizeof(properties) == 72
assert C.sizeof(predictProperties) == 32
assert C.sizeof(model) == 72

SAMPLE_DTYPE = np.dtype({"names": ["index", "value"], "formats": ["<i4", "<f8"],
                         "offsets": [0, 8], "itemsize": 16})


def make_properties(gamma=1.0, C_=1.0, kernel=1, W=500, eta=0.001, size=10,
                    algorithm=1, threads=1, verbose=0):
    """Defaults of the reference's parsers (src/full-train.c:710-720,
    src/budgeted-train.c:583-594)."""
    p = properties()
    p.Kgamma = gamma
    p.kernelType = kernel
    p.C = C_
    p.Threads = threads
    p.MaxSize = W
    p.size = size
    p.algorithm = algorithm
    p.Eta = eta
    p.file = 1
    p.separator = b","
    p.verbose = verbose
    return p


def make_predict_properties(labels=0, soft=0, threads=1, verbose=0):
    p = predictProperties()
    p.Labels = labels
    p.Threads = threads
    p.Soft = soft
    p.file = 1
    p.separator = b","
    p.verbose = verbose
    return p


def dataset_arrays(ds, with_means=True):
    """Copies of (y, quadratic_value, features) of a dataset as numpy arrays."""
    rows = ds.l + (2 if with_means else 0)
    y = np.ctypeslib.as_array(ds.y, shape=(rows,)).copy()
    q = np.ctypeslib.as_array(ds.quadratic_value, shape=(rows,)).copy()
    # walk to the last terminator to find the feature count
    last = ds.x[rows - 1]
    base = C.addressof(ds.features.contents)
    k = 0
    while last[k].index != -1:
        k += 1
    end = (C.addressof(last.contents) - base) // 16 + k + 1
    buf = (C.c_char * (end * 16)).from_address(base)
    feats = np.frombuffer(bytes(buf), dtype=SAMPLE_DTYPE)
    offs = np.array([(C.addressof(ds.x[i].contents) - base) // 16 for i in range(rows)],
                    dtype=np.int64)
    return y, q, feats, offs


def model_arrays(m):
    w = np.ctypeslib.as_array(m.weights, shape=(m.nSVs,)).copy() if m.nSVs else np.zeros(0)
    q = np.ctypeslib.as_array(m.quadratic_value, shape=(m.nSVs,)).copy() if m.nSVs else np.zeros(0)
    return w, q
