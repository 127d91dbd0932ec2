"""ORACLE -- test infrastructure only.

ctypes wrapper over oracle/_ref/libref_io.so: the reference's own dataset
readers and model-file functions (src/IOStructures.c), compiled from its
sources by oracle/Makefile where the reference tree is present.  The tests
compare this library's readers and model file with it; they skip when it
has not been built.
"""
import ctypes as C
import os

from libirwls_amd.abi import model, svm_dataset

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libref_io.so")

_lib = None


def available():
    return os.path.exists(LIB_PATH)


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
