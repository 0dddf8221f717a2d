"""Builds the native libraries in-tree.

* cpanel_amd/libcpgpu.so - the product: host C++ + gfx950 HIP kernels behind
  the C ABI of include/cpgpu.h (hipcc, offload arch gfx950 only).
* oracle/liboracle.so    - the FP64 CPU oracle used by tests and the CPU
  baseline (g++; test infrastructure, not linked by the product).

Usage: python -m cpanel_amd.build   (or __graft_entry__.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcpgpu.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 VALU forms are disabled for every HIP translation unit.
HIP_FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
             "-Wno-unknown-warning-option", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def build_product(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()

    def comp(src):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        log = _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        if verbose and log.strip():
            print(log)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    srcs = [os.path.join(ORACLE_DIR, f) for f in sorted(os.listdir(ORACLE_DIR)) if f.endswith(".cpp")]
    tmp = ORACLE_LIB + ".tmp"
    # -ffp-contract=off: no FMA contraction, like the reference's SSE2 build
    _run(["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
          "-std=c++17", "-o", tmp] + srcs)
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_all(verbose=False):
    with ThreadPoolExecutor(max_workers=2) as ex:
        f1 = ex.submit(build_product, 8, verbose)
        f2 = ex.submit(build_oracle)
        return f1.result(), f2.result()


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
