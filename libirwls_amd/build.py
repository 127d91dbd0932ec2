"""Builds libirwls.so in-tree: HIP kernels for gfx950 with hipcc, the C host
code with gcc, linked by hipcc.

    python -m libirwls_amd.build          # compile and link
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libirwls.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IRWLS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC,
             # packed FP32 ops stay off on this target
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-pthread", "-I" + INC, "-Wall", "-Wno-unused-result",
           "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def sources():
    hip = sorted(glob.glob(os.path.join(HERE, "csrc", "device", "*.hip")))
    c = sorted(glob.glob(os.path.join(HERE, "csrc", "host", "*.c")))
    return hip, c


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hip, c = sources()
    jobs = []
    for s in hip:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o], o))
    for s in c:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        jobs.append((["gcc"] + C_FLAGS + ["-c", s, "-o", o], o))
    workers = min(8, len(jobs))
    with cf.ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [j[1] for j in jobs] + ["-lm", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
