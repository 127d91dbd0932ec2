"""The C-ABI library loads and exports every function include/*.h declares
(no GPU call is made)."""
import ctypes
import os
import re

import pytest

from libirwls_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["IOStructures.h", "full-train.h", "budgeted-train.h", "LIBIRWLS-predict.h", "kernels.h",
           "ParallelAlgorithms.h", "irwls_gpu.h", "irwls_device.h"]


def declared_functions(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?[\s\*]([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
        name = m.group(1)
        if name not in ("if", "while", "for", "return", "sizeof"):
            names.add(name)
    return names


@pytest.mark.parametrize("header", HEADERS)
def test_exports(header):
    L = lib.load()
    names = declared_functions(header)
    assert names, header
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, "%s: not exported: %s" % (header, missing)


def test_reference_entry_points_present():
    names = set()
    for h in HEADERS[:6]:
        names |= declared_functions(h)
    for n in ["trainFULL", "subIRWLS", "rpermute", "calculateFULLModel", "SGMA", "randomCentroids",
              "IRWLSpar", "calculateBudgetedModel", "test", "softTest", "kernelFunction", "kernelTest",
              "initMemory", "freeMemory", "ParallelLinearSystem", "ParallelVectorMatrixT",
              "ParallelVectorMatrix", "readTrainFile", "storeModel", "readModel"]:
        assert n in names


def test_struct_sizes():
    assert ctypes.sizeof(abi.svm_sample) == 16
    assert ctypes.sizeof(abi.svm_dataset) == 48
    assert ctypes.sizeof(abi.properties) == 72
    assert ctypes.sizeof(abi.model) == 72


def test_c_caller_compiles_and_links(tmp_path):
    """A C program written against include/*.h only (tests/c_caller/caller.c)
    compiles with gcc and links against libirwls.so: every function it calls
    resolves (no GPU call is made here)."""
    import subprocess
    exe = str(tmp_path / "caller")
    subprocess.check_call(["gcc", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c_caller", "caller.c"),
                           "-L", os.path.join(ROOT, "libirwls_amd"), "-lirwls",
                           "-Wl,-rpath," + os.path.join(ROOT, "libirwls_amd"), "-Wl,--no-undefined", "-lm",
                           "-o", exe])
    assert os.path.exists(exe)
    out = subprocess.run(["nm", "-u", exe], stdout=subprocess.PIPE, text=True).stdout
    for name in ["trainFULL", "SGMA", "IRWLSpar", "calculateFULLModel", "calculateBudgetedModel", "test",
                 "readTrainFile", "storeModel", "readModel", "initMemory", "freeMemory",
                 "parseTrainFULLParameters", "parseTrainParameters"]:
        assert name in out.split(), name
