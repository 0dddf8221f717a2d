"""The C-ABI library loads, exports every symbol include/cpgpu.h declares, and
refuses device work on a host-only context.  CPU only (no compute calls)."""
import ctypes as C

import numpy as np

from cpanel_amd import abi, meshgen
from cpanel_amd.case import Context, CPanelError


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = abi.declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
        assert n in abi.SIGNATURES, n


def test_status_strings():
    L = abi.lib()
    for s in range(8):
        assert L.cp_status_string(s)


def test_host_only_context_refuses_device_work():
    ctx = Context(device=-1)
    g = ctx.geometry(meshgen.small_wing())
    assert g.nb > 0
    L = abi.lib()
    assert L.cp_assemble(g.h, abi.CP_FP64) == abi.CP_NO_DEVICE
    A = np.zeros((g.nb, g.nb), order="F")
    assert L.cp_read_matrices(g.h, abi.CP_FP64, A.ctypes.data_as(C.POINTER(C.c_double)), None, g.nb) == \
        abi.CP_NO_DEVICE
    try:
        g.solve([(1, 5, 0)])
        raise AssertionError("expected CPanelError")
    except CPanelError as e:
        assert e.status == abi.CP_NO_DEVICE


def test_bad_args():
    L = abi.lib()
    assert L.cp_create(0, 0, None) == abi.CP_BAD_ARG
    assert L.cp_geometry_info(None, None) == abi.CP_BAD_ARG


def test_vp_solve_refused_without_device():
    ctx = Context(device=-1)
    g = ctx.geometry(meshgen.small_wing())
    L = abi.lib()
    sig = np.zeros(g.nb)
    o = abi.SolveOpts()
    st = L.cp_solve_rhs(g.h, abi.CP_FP64, 1, sig.ctypes.data_as(C.POINTER(C.c_double)), None, None, C.byref(o),
                        None, None)
    assert st == abi.CP_NO_DEVICE
    assert L.cp_set_buffer_wake(g.h, 1, None, None) == abi.CP_BAD_ARG
