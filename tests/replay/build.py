"""Builds the CPU replay of libcpgpu (test infrastructure): the library's host
code and its non-matrix-core kernels compiled with g++ against the HIP
replay in tests/replay/hip, host translation units under AddressSanitizer +
UndefinedBehaviorSanitizer (mode "asan") or ThreadSanitizer (mode "tsan").

    python tests/replay/build.py [asan|tsan]   -> tests/replay/_build/libcpgpu_<mode>.so

Run a program against it with the sanitizer runtime preloaded (see
preload_env()).  The batched mixed-precision path (batched.hip,
blockinv.hip) is replaced by stubs.cpp.
"""
import hashlib
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "cpanel_amd", "csrc")
OUT = os.path.join(HERE, "_build")

HOST_TU = ["capi.cpp", "vp_case.cpp", "serve.cpp", "geometry.cpp", "post_host.cpp", "precond.cpp", "vp_tree.cpp",
           "timer.cpp"]
KERNEL_TU = ["krylov.hip", "post.hip", "vp.hip", "assemble.hip", "noalloc.hip", "nearcsr.hip"]


def available():
    gxx = shutil.which("g++")
    if not gxx:
        return False
    for lib in ("libasan.so", "libtsan.so", "libstdc++.so.6"):
        r = subprocess.run([gxx, "-print-file-name=" + lib], stdout=subprocess.PIPE, text=True)
        if not os.path.isabs(r.stdout.strip()):
            return False
    return True


def preload_env(mode="asan"):
    """Environment for a Python child process that loads the replay library:
    the sanitizer runtime and libstdc++ preloaded (Python itself is not
    instrumented), leak checking off."""
    def lib(name):
        return subprocess.run(["g++", "-print-file-name=" + name], stdout=subprocess.PIPE, text=True).stdout.strip()
    env = dict(os.environ)
    san = lib("libasan.so" if mode == "asan" else "libtsan.so")
    env["LD_PRELOAD"] = san + " " + lib("libstdc++.so.6")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:report_signal_unsafe=0"
    return env


def _sources():
    return ([os.path.join(CSRC, f) for f in HOST_TU + KERNEL_TU] +
            [os.path.join(HERE, f) for f in ("rt.cpp", "stubs.cpp", os.path.join("hip", "hip_runtime.h"))] +
            [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".h")] +
            [os.path.join(ROOT, "include", "cpgpu.h")])


def build(mode="asan", jobs=8):
    """Compiles (when any source changed) and returns the library path."""
    lib = os.path.join(OUT, "libcpgpu_%s.so" % mode)
    h = hashlib.sha256(mode.encode())
    for f in _sources():
        with open(f, "rb") as fh:
            h.update(fh.read())
    stamp = lib + ".sha256"
    if os.path.exists(lib) and os.path.exists(stamp) and open(stamp).read() == h.hexdigest():
        return lib
    obj = os.path.join(OUT, "obj_" + mode)
    os.makedirs(obj, exist_ok=True)
    inc = ["-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    san = (["-fsanitize=address,undefined"] if mode == "asan" else ["-fsanitize=thread"]) + \
        ["-fno-omit-frame-pointer"]
    ksan = ["-fsanitize=undefined"] if mode == "asan" else []
    base = ["g++", "-std=c++17", "-g", "-fPIC", "-Wno-unknown-pragmas"]
    cmds = []
    for f in HOST_TU:
        cmds.append(base + ["-O1"] + san + inc + ["-c", os.path.join(CSRC, f), "-o", os.path.join(obj, f + ".o")])
    for f in KERNEL_TU:
        cmds.append(base + ["-O2", "-x", "c++"] + ksan + inc +
                    ["-c", os.path.join(CSRC, f), "-o", os.path.join(obj, f + ".o")])
    # the fiber switch runs uninstrumented
    cmds.append(base + ["-O2"] + inc + ["-c", os.path.join(HERE, "rt.cpp"), "-o", os.path.join(obj, "rt.o")])
    cmds.append(base + ["-O1"] + san + inc + ["-c", os.path.join(HERE, "stubs.cpp"),
                                              "-o", os.path.join(obj, "stubs.o")])

    def run(c):
        r = subprocess.run(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("replay build failed: %s\n%s" % (" ".join(c), r.stdout))
        return c[-1]

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(run, cmds))
    run(["g++", "-shared", "-fPIC"] + san + ["-o", lib] + objs + ["-lpthread"])
    with open(stamp, "w") as fh:
        fh.write(h.hexdigest())
    return lib


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else "asan"))
