"""ctypes binding of libcpgpu.so (include/cpgpu.h).

The library is the product: host C++ topology and post-processing plus the
gfx950 HIP kernels.  Loading fails loudly if it has not been built; there is
no Python or CPU fallback for any compute entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpgpu.so")

CP_OK, CP_NOT_CONVERGED, CP_OUT_OF_BUDGET, CP_BAD_MESH, CP_HIP_ERROR, CP_BAD_ARG, CP_NO_DEVICE, \
    CP_NOT_ASSEMBLED = range(8)
CP_FP64, CP_FP32_REFINE, CP_FP32_REFINE64 = 0, 1, 2

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("xyz", _d), ("n_tris", C.c_int32), ("tri", _i32), ("ids", _i32)]


class GeomInfo(C.Structure):
    _fields_ = [("n_body", C.c_int32), ("n_wake", C.c_int32), ("n_wakelines", C.c_int32),
                ("n_nodes", C.c_int32), ("lda", C.c_int64), ("n_wakes", C.c_int32), ("n_stations", C.c_int32)]


class CaseFields(C.Structure):
    _fields_ = [("sigma", C.POINTER(C.c_double)), ("velocity", C.POINTER(C.c_double)),
                ("potential", C.POINTER(C.c_double)),
                ("cp", C.POINTER(C.c_double)), ("wake_mu", C.POINTER(C.c_double)),
                ("wake_potential", C.POINTER(C.c_double)), ("spanwise", C.POINTER(C.c_double))]


class Freestream(C.Structure):
    _fields_ = [("V", C.c_double), ("alpha_deg", C.c_double), ("beta_deg", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("mach", C.c_double), ("Sref", C.c_double), ("bref", C.c_double), ("cref", C.c_double),
                ("cg", C.c_double * 3), ("tol", C.c_double), ("max_iter", C.c_int32),
                ("max_refine", C.c_int32), ("inner_tol", C.c_double), ("precond", C.c_int32),
                ("pad", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("CL", C.c_double), ("CDi", C.c_double), ("Fbody", C.c_double * 3), ("Fwind", C.c_double * 3),
                ("CM", C.c_double * 3), ("rel_residual", C.c_double), ("iters", C.c_int32),
                ("refinements", C.c_int32), ("converged", C.c_int32), ("pad", C.c_int32)]


class RequestStats(C.Structure):
    _fields_ = [("queue_s", C.c_double), ("prepare_s", C.c_double), ("solve_s", C.c_double),
                ("latency_s", C.c_double), ("cache_hit", C.c_int32), ("batch_conditions", C.c_int32),
                ("n_body", C.c_int32), ("group_geometries", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("conditions", C.c_int64),
                ("batches", C.c_int64), ("cache_hits", C.c_int64), ("cache_misses", C.c_int64),
                ("evictions", C.c_int64), ("resident_geometries", C.c_int64),
                ("resident_bytes_est", C.c_uint64), ("device_bytes", C.c_uint64),
                ("grouped_solves", C.c_int64)]


class VpOpts(C.Structure):
    _fields_ = [("fs", Freestream), ("mach", C.c_double), ("Sref", C.c_double), ("bref", C.c_double),
                ("cref", C.c_double), ("cg", C.c_double * 3), ("prec", C.c_int32), ("accelerate", C.c_int32),
                ("warm_start", C.c_int32), ("max_members", C.c_int32), ("theta", C.c_double), ("tol", C.c_double),
                ("kernel", C.c_int32), ("per_edge", C.c_int32), ("core", C.c_double), ("fp32", C.c_int32),
                ("pad", C.c_int32)]


class VpStepInfo(C.Structure):
    _fields_ = [("step", C.c_int32), ("n_particles", C.c_int32), ("converged", C.c_int32),
                ("solve_iters", C.c_int32), ("CL_trefftz", C.c_double), ("CDi_trefftz", C.c_double),
                ("rel_residual", C.c_double), ("ms", C.c_double)]


class BhOpts(C.Structure):
    _fields_ = [("max_members", C.c_int32), ("kernel", C.c_int32), ("theta", C.c_double), ("core", C.c_double),
                ("fp32", C.c_int32), ("host_build", C.c_int32)]


CP_BH_VEL_POINTS, CP_BH_VEL, CP_BH_STRETCH, CP_BH_DIFF = 0, 1, 2, 3
CP_VP_REFERENCE, CP_VP_ALGEBRAIC = 0, 1

# Every symbol declared in include/cpgpu.h with (restype, argtypes).
SIGNATURES = {
    "cp_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "cp_destroy": (None, [C.c_void_p]),
    "cp_last_error": (C.c_char_p, [C.c_void_p]),
    "cp_last_hip_error": (C.c_int, [C.c_void_p]),
    "cp_status_string": (C.c_char_p, [C.c_int]),
    "cp_bytes_in_use": (C.c_uint64, [C.c_void_p]),
    "cp_geometry_upload": (C.c_int, [C.c_void_p, C.POINTER(Mesh), C.c_uint64, C.POINTER(C.c_void_p)]),
    "cp_geometry_from_tri": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "cp_geometry_release": (None, [C.c_void_p]),
    "cp_geometry_info": (C.c_int, [C.c_void_p, C.POINTER(GeomInfo)]),
    "cp_geometry_body": (C.c_int, [C.c_void_p, _i32, _d, _d, _d]),
    "cp_geometry_wake": (C.c_int, [C.c_void_p, _i32, _i32, _d]),
    "cp_geometry_clusters": (C.c_int, [C.c_void_p, _i32, _i32, _i32]),
    "cp_assemble": (C.c_int, [C.c_void_p, C.c_int]),
    "cp_read_matrices": (C.c_int, [C.c_void_p, C.c_int, _d, _d, C.c_int64]),
    "cp_load_matrices": (C.c_int, [C.c_void_p, _d, _d, C.c_int64]),
    "cp_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_int32, C.POINTER(Freestream), C.POINTER(SolveOpts), _d, _d,
                           C.POINTER(SolveInfo)]),
    "cp_solve_group": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_int32),
                                 C.POINTER(C.POINTER(Freestream)), C.POINTER(SolveOpts), C.POINTER(_d),
                                 C.POINTER(_d), C.POINTER(C.POINTER(SolveInfo))]),
    "cp_matvec": (C.c_int, [C.c_void_p, C.c_int, C.c_int32, _d, _d]),
    "cp_matvec_f64_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]),
    "cp_solver_report": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "cp_geometry_upload_vp": (C.c_int, [C.c_void_p, C.POINTER(Mesh), C.c_double, C.c_double, C.c_uint64,
                                        C.POINTER(C.c_void_p)]),
    "cp_vp_buffer_rows": (C.c_int, [C.c_void_p, _i32, _d, _d, _d]),
    "cp_vp_create": (C.c_int, [C.c_void_p, C.POINTER(VpOpts), C.POINTER(C.c_void_p)]),
    "cp_vp_step": (C.c_int, [C.c_void_p, C.POINTER(VpStepInfo)]),
    "cp_vp_finish": (C.c_int, [C.c_void_p, C.POINTER(SolveInfo), _d]),
    "cp_vp_particles": (C.c_int, [C.c_void_p, _i32, _d, _d, _d, _i32]),
    "cp_vp_state": (C.c_int, [C.c_void_p, _d, _d, _d, _d]),
    "cp_vp_release": (None, [C.c_void_p]),
    "cp_case_fields_eval": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, _d, C.POINTER(CaseFields)]),
    "cp_block_jacobi_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "cp_bh_build": (C.c_int, [C.c_void_p, C.c_int32, _d, _d, _d, C.POINTER(BhOpts), C.POINTER(C.c_void_p)]),
    "cp_bh_read": (C.c_int, [C.c_void_p, _d, _i32, _i32]),
    "cp_bh_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _d, _d]),
    "cp_bh_eval": (C.c_int, [C.c_void_p, C.c_int, C.c_int32, _d, _d]),
    "cp_bh_release": (None, [C.c_void_p]),
    "cp_panel_to_points": (C.c_int, [C.c_void_p, _d, _d, _d, C.c_int32, _d, _d, C.c_int32]),
    "cp_panel_stretching": (C.c_int, [C.c_void_p, _d, _d, _d, C.c_int32, _d, _d, _d, C.c_int32]),
    "cp_set_buffer_wake": (C.c_int, [C.c_void_p, C.c_int32, _i32, _d]),
    "cp_solve_rhs": (C.c_int, [C.c_void_p, C.c_int, C.c_int32, _d, _d, _d, C.POINTER(SolveOpts), _d,
                               C.POINTER(SolveInfo)]),
    "cp_server_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "cp_server_submit": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(Mesh), C.c_int32, C.POINTER(Freestream),
                                   C.POINTER(SolveOpts), C.c_int32, C.POINTER(C.c_uint64)]),
    "cp_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(SolveInfo), _d, C.POINTER(RequestStats)]),
    "cp_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "cp_server_destroy": (None, [C.c_void_p]),
    "cp_timing_reset": (C.c_int, [C.c_void_p]),
    "cp_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "cp_timing_only": (C.c_int, [C.c_void_p, C.c_char_p]),
    "cp_timing_get": (C.c_int, [C.c_void_p, C.c_char_p, _d, _i64]),
}

_lib = None


def lib():
    """Loads libcpgpu.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcpgpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names of the functions declared in include/cpgpu.h (parsed from the header)."""
    import re
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "cpgpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(cp_[a-z0-9_]+)\s*\(", hdr)))
