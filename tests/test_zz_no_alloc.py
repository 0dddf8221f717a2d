"""cp_matvec_f64_noalloc (the allocation-free FP64 product y = A x through
caller-given device buffers, split-K partials in a static device array)
against cp_matvec kernel 0 (the same kernels with a workspace): bit-identical
on two calls in a row, the second on a smaller geometry with fewer columns.
Device buffers come from the HIP runtime the library itself uses (loaded
through ctypes), so no second runtime shares the process."""
import ctypes as C

import numpy as np
import pytest

from cpanel_amd import CP_FP64, abi, meshgen
from cpanel_amd.case import Context

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DevBuf:
    """Device memory from the library's HIP runtime."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def hip_runtime():
    abi.lib()  # the library loads its runtime; this handle is the same one
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def test_noalloc_matvec_equals_cp_matvec():
    hip = hip_runtime()
    ctx = Context(0)
    rng = np.random.default_rng(7)
    for mesh, k in ((meshgen.config_c1(), 3), (meshgen.small_wing("4412", n_chord=20, n_span=6), 2)):
        g = ctx.geometry(mesh)
        g.assemble(CP_FP64)
        n, lda = g.nb, g.lda
        X = rng.normal(size=(k, n))
        ref = g.matvec(X.T, kernel=0).T  # k x n
        xd, yd = DevBuf(hip, X.nbytes), DevBuf(hip, k * lda * 8)
        xd.put(X)
        yd.put(np.full((k, lda), np.nan))
        st = abi.lib().cp_matvec_f64_noalloc(g.h, k, xd.p, n, yd.p, lda)
        assert st == abi.CP_OK, abi.lib().cp_last_error(ctx.h).decode()
        Y = yd.get(np.zeros((k, lda)))
        print("[gpu] noalloc matvec n=%d k=%d max |dy| %.1e" % (n, k, np.abs(Y[:, :n] - ref).max()), flush=True)
        assert np.array_equal(Y[:, :n], ref)
        assert np.all(Y[:, n:] == 0)
        xd.free()
        yd.free()
        g.release()
    ctx.close()


def test_noalloc_matvec_rejects_short_strides():
    """Strides shorter than a column are refused before any launch (the
    buffers are real allocations all the same)."""
    hip = hip_runtime()
    ctx = Context(0)
    g = ctx.geometry(meshgen.small_wing("4412", n_chord=12, n_span=4))
    g.assemble(CP_FP64)
    xd, yd = DevBuf(hip, 2 * g.lda * 8), DevBuf(hip, 2 * g.lda * 8)
    assert abi.lib().cp_matvec_f64_noalloc(g.h, 1, xd.p, g.nb - 1, yd.p, g.lda) == abi.CP_BAD_ARG
    assert abi.lib().cp_matvec_f64_noalloc(g.h, 1, xd.p, g.nb, yd.p, g.lda - 1) == abi.CP_BAD_ARG
    xd.free()
    yd.free()
    g.release()
    ctx.close()
