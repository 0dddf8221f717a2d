"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)


class CaseParams(C.Structure):
    _fields_ = [("V", C.c_double), ("alpha_deg", C.c_double), ("beta_deg", C.c_double),
                ("mach", C.c_double), ("Sref", C.c_double), ("bref", C.c_double),
                ("cref", C.c_double), ("cg", C.c_double * 3)]


class CaseResult(C.Structure):
    _fields_ = [("CL", C.c_double), ("CDi", C.c_double), ("Fbody", C.c_double * 3),
                ("Fwind", C.c_double * 3), ("CM", C.c_double * 3), ("solve_err", C.c_double),
                ("solve_iters", C.c_int)]


_lib = None


def use_library(path):
    """Loads the oracle from `path` (a build of the same sources with other
    flags, e.g. build_native's) in place of liboracle.so."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None
    return lib()


def build_native(out_dir):
    """Builds the oracle with -march=native for this host's CPU into out_dir
    (the CPU baseline is timed with the host's own instruction set; the
    in-tree liboracle.so is built for x86-64-v3 so that it loads anywhere).
    Returns the library path or None when no compiler is available."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if not gxx:
        return None
    srcs = [os.path.join(_HERE, f) for f in sorted(os.listdir(_HERE)) if f.endswith(".cpp")]
    out = os.path.join(out_dir, "liboracle_native.so")
    r = subprocess.run([gxx, "-O3", "-march=native", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                        "-std=c++17", "-o", out] + srcs, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return out if r.returncode == 0 else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: %s (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.or_geom_create.restype = C.c_void_p
        L.or_geom_create.argtypes = [C.c_int, _d, C.c_int, _i, _i]
        L.or_geom_read_tri.restype = C.c_void_p
        L.or_geom_read_tri.argtypes = [C.c_char_p]
        L.or_geom_destroy.argtypes = [C.c_void_p]
        for f in ("or_geom_n_body", "or_geom_n_wake", "or_geom_n_wakelines"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        L.or_geom_body_info.argtypes = [C.c_void_p, _i, _d, _d, _d, _d, _i]
        L.or_geom_wake_info.argtypes = [C.c_void_p, _i, _i, _d]
        L.or_geom_clusters.argtypes = [C.c_void_p, _i, _i]
        L.or_assemble.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.c_int]
        L.or_assemble_cols.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.or_bicgstab.argtypes = [C.c_int, _d, C.c_int64, _d, _d, C.c_double, C.c_int, C.c_int, _d]
        L.or_bicgstab.restype = C.c_int
        L.or_case_run.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.POINTER(CaseParams), C.c_double,
                                  C.c_int, C.c_int, C.POINTER(CaseResult), _d, _d, _d, _d, _d, _d,
                                  _d, _d, _d]
        L.or_case_run.restype = C.c_int
        L.or_case_post.argtypes = [C.c_void_p, C.POINTER(CaseParams), _d, C.POINTER(CaseResult),
                                   _d, _d, _d, _d]
        L.or_panel_phi_inf.argtypes = [C.c_int, _d, _d, _d, _d]
        L.or_dub_phi_inf.argtypes = [C.c_int, _d, _d]
        L.or_dub_phi_inf.restype = C.c_double
        L.or_panel_v_inf.argtypes = [C.c_int, _d, _d, _d, _d]
        L.or_dub_v_inf.argtypes = [C.c_int, _d, _d, _d]
        L.or_wake_potential.argtypes = [C.c_void_p, _d]
        L.or_wake_potential.restype = C.c_double
        _lib = L
        try:
            from . import vp_oracle  # noqa: F401  (registers VP symbols when present)
        except Exception:
            pass
    return _lib


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def panel_phi_inf(verts, poi):
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(poi, dtype=np.float64)
    s, d = C.c_double(), C.c_double()
    lib().or_panel_phi_inf(len(v) // 3, _p(v), _p(q), C.byref(s), C.byref(d))
    return s.value, d.value


def dub_phi_inf(verts, poi):
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(poi, dtype=np.float64)
    return lib().or_dub_phi_inf(len(v) // 3, _p(v), _p(q))


def dub_v_inf(verts, poi):
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(poi, dtype=np.float64)
    d = np.zeros(3)
    lib().or_dub_v_inf(len(v) // 3, _p(v), _p(q), _p(d))
    return d


def panel_v_inf(verts, poi):
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(poi, dtype=np.float64)
    s = np.zeros(3)
    d = np.zeros(3)
    lib().or_panel_v_inf(len(v) // 3, _p(v), _p(q), _p(s), _p(d))
    return s, d


class Geom:
    """Reference topology + panel geometry built by the oracle."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("oracle failed to build geometry")
        self.h = C.c_void_p(handle)
        L = lib()
        self.nb = L.or_geom_n_body(self.h)
        self.nw = L.or_geom_n_wake(self.h)

    @classmethod
    def from_tri(cls, path):
        return cls(lib().or_geom_read_tri(path.encode()))

    @classmethod
    def from_arrays(cls, xyz, tri0, ids):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tri0 = np.ascontiguousarray(tri0, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        return cls(lib().or_geom_create(len(xyz), _p(xyz), len(tri0), _p(tri0, _i), _p(ids, _i)))

    def __del__(self):
        try:
            lib().or_geom_destroy(self.h)
        except Exception:
            pass

    def body_info(self):
        n = self.nb
        ti = np.zeros(n, np.int32)
        c = np.zeros((n, 3))
        nr = np.zeros((n, 3))
        a = np.zeros(n)
        ls = np.zeros(n)
        tip = np.zeros(n, np.int32)
        lib().or_geom_body_info(self.h, _p(ti, _i), _p(c), _p(nr), _p(a), _p(ls), _p(tip, _i))
        return dict(tri_index=ti, center=c, normal=nr, area=a, long_side=ls, tip=tip)

    def wake_info(self):
        n = self.nw
        ti = np.zeros(n, np.int32)
        idx = np.zeros((n, 4), np.int32)
        cf = np.zeros(n)
        lib().or_geom_wake_info(self.h, _p(ti, _i), _p(idx, _i), _p(cf))
        return dict(tri_index=ti, idx4=idx, coef=cf)

    def clusters(self):
        out = np.zeros((self.nb, 16), np.int32)
        ts = np.zeros(self.nb, np.int32)
        lib().or_geom_clusters(self.h, _p(out, _i), _p(ts, _i))
        return out, ts

    def assemble(self, nthreads=0):
        n = self.nb
        A = np.zeros((n, n), order="F")
        B = np.zeros((n, n), order="F")
        lib().or_assemble(self.h, _p(A), _p(B), n, nthreads)
        return A, B

    def assemble_cols(self, j0, j1, nthreads=0):
        """Columns j0..j1-1 of the body-body block (no wake folding)."""
        n = self.nb
        A = np.zeros((n, j1 - j0), order="F")
        B = np.zeros((n, j1 - j0), order="F")
        lib().or_assemble_cols(self.h, _p(A), _p(B), n, j0, j1, nthreads)
        return A, B

    def run_case(self, A, B, V=1.0, alpha=0.0, beta=0.0, mach=0.1, Sref=1.0, bref=1.0, cref=1.0,
                 cg=(0.0, 0.0, 0.0), tol=1e-13, maxit=0, nthreads=0):
        n = self.nb
        p = CaseParams(V, alpha, beta, mach, Sref, bref, cref, (C.c_double * 3)(*cg))
        r = CaseResult()
        out = {k: np.zeros(n) for k in ("mu", "sigma", "phi", "cp")}
        out["vel"] = np.zeros((n, 3))
        out["wake_mu"] = np.zeros(max(self.nw, 1))
        for k in ("yloc", "span_cl", "span_cd"):
            out[k] = np.zeros(165)
        A = np.asfortranarray(A)
        B = np.asfortranarray(B)
        rc = lib().or_case_run(self.h, _p(A), _p(B), n, C.byref(p), tol, maxit, nthreads, C.byref(r),
                               _p(out["mu"]), _p(out["sigma"]), _p(out["phi"]), _p(out["vel"]),
                               _p(out["cp"]), _p(out["wake_mu"]), _p(out["yloc"]),
                               _p(out["span_cl"]), _p(out["span_cd"]))
        out.update(CL=r.CL, CDi=r.CDi, Fbody=np.array(r.Fbody[:]), Fwind=np.array(r.Fwind[:]),
                   CM=np.array(r.CM[:]), err=r.solve_err, iters=r.solve_iters, converged=(rc == 0))
        return out

    def post(self, mu, V=1.0, alpha=0.0, beta=0.0, mach=0.1, Sref=1.0, bref=1.0, cref=1.0,
             cg=(0.0, 0.0, 0.0)):
        n = self.nb
        p = CaseParams(V, alpha, beta, mach, Sref, bref, cref, (C.c_double * 3)(*cg))
        r = CaseResult()
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        phi = np.zeros(n)
        vel = np.zeros((n, 3))
        cp = np.zeros(n)
        wmu = np.zeros(max(self.nw, 1))
        lib().or_case_post(self.h, C.byref(p), _p(mu), C.byref(r), _p(phi), _p(vel), _p(cp), _p(wmu))
        return dict(CL=r.CL, CDi=r.CDi, Fbody=np.array(r.Fbody[:]), Fwind=np.array(r.Fwind[:]),
                    CM=np.array(r.CM[:]), phi=phi, vel=vel, cp=cp, wake_mu=wmu)


def bicgstab(A, b, tol=0.0, maxit=0, nthreads=0):
    A = np.asfortranarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    err = C.c_double()
    it = lib().or_bicgstab(len(b), _p(A), A.shape[0], _p(b), _p(x), tol, maxit, nthreads, C.byref(err))
    return x, it, err.value
