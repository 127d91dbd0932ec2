"""ORACLE -- test infrastructure only.

ctypes wrappers over the reference's own code compiled from its sources by
oracle/Makefile where the reference tree is present:
oracle/_ref/libref_io.so (dataset readers and model file,
src/IOStructures.c) and oracle/_ref/libref_predict.so (kernel functions and
predictors, src/kernels.c and src/LIBIRWLS-predict.c).  The tests compare
this library and the oracle with them; they skip when they have not been
built.
"""
import ctypes as C
import os

from libirwls_amd.abi import model, predictProperties, properties, svm_dataset

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libref_io.so")
PREDICT_PATH = os.path.join(HERE, "_ref", "libref_predict.so")

_lib = None
_plib = None


def available():
    return os.path.exists(LIB_PATH)


def predict_available():
    return os.path.exists(PREDICT_PATH)


def predict_lib():
    """kernelFunction, kernelTest, softTest, test of the reference."""
    global _plib
    if _plib is None:
        L = C.CDLL(PREDICT_PATH)
        L.kernelFunction.restype = C.c_double
        L.kernelFunction.argtypes = [svm_dataset, C.c_int, C.c_int, properties]
        L.kernelTest.restype = C.c_double
        L.kernelTest.argtypes = [svm_dataset, C.c_int, model, C.c_int]
        for name in ("softTest", "test"):
            getattr(L, name).restype = C.POINTER(C.c_double)
            getattr(L, name).argtypes = [svm_dataset, model, predictProperties]
        _plib = L
    return _plib


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        for name in ("readTrainFile", "readUnlabeledFile"):
            getattr(L, name).restype = svm_dataset
            getattr(L, name).argtypes = [C.c_char_p]
        for name in ("readTrainFileCSV", "readUnlabeledFileCSV"):
            getattr(L, name).restype = svm_dataset
            getattr(L, name).argtypes = [C.c_char_p, C.c_char_p]
        L.freeDataset.argtypes = [svm_dataset]
        L.storeModel.argtypes = [C.POINTER(model), C.c_void_p]
        L.readModel.argtypes = [C.POINTER(model), C.c_void_p]
        _lib = L
    return _lib
