"""ctypes binding of the vortex-particle part of the FP64 oracle.

TEST INFRASTRUCTURE ONLY (see oracle.py).
"""
import ctypes as C

import numpy as np

from . import oracle as _o

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)

MODE_VEL_POINTS, MODE_VEL, MODE_STRETCH, MODE_DIFF, MODE_GAUSS_POINTS = 0, 1, 2, 3, 4


def _lib():
    L = _o.lib()
    if not getattr(L, "_vp_ready", False):
        L.or_bh_build.restype = C.c_void_p
        L.or_bh_build.argtypes = [C.c_int, _d, _d, _d, C.c_int, C.c_double]
        L.or_bh_destroy.argtypes = [C.c_void_p]
        L.or_bh_info.argtypes = [C.c_void_p, _i, _i, _d, _d]
        L.or_bh_eval.argtypes = [C.c_void_p, C.c_int, C.c_int, _d, _d]
        L.or_vp_direct.argtypes = [C.c_int, C.c_int, _d, _d, _d, C.c_int, _d, _d]
        L.or_vp_pair.argtypes = [C.c_int, _d, _d, C.c_double, _d, _d, C.c_double, _d]
        L.or_vp_direct_core.argtypes = [C.c_int, C.c_double, C.c_int, _d, _d, C.c_int, _d, _d]
        L.or_vp_pair_core.argtypes = [C.c_int, C.c_double, _d, _d, _d, _d, _d]
        L.or_panels_to_points.argtypes = [C.c_int, _i, _d, _d, _d, C.c_int, _i, _d, _d, C.c_int, _d, _d]
        L.or_velgrad.argtypes = [C.c_int, _d, C.c_int, C.c_int, C.c_double, C.c_double, _d, _d]
        L.or_panels_stretching.argtypes = [C.c_int, _i, _d, _d, _d, C.c_int, _i, _d, _d, C.c_int, _d, _d, _d]
        L._vp_ready = True
    return L


def _a(x, dt=np.float64):
    return np.ascontiguousarray(x, dtype=dt)


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


class BHTree:
    """Particle octree + pseudo-particles with the reference's semantics."""

    def __init__(self, xyz, alpha, radius, max_members=10, theta=0.5):
        self.xyz, self.alpha, self.radius = _a(xyz).reshape(-1, 3), _a(alpha).reshape(-1, 3), _a(radius).reshape(-1)
        self.n = len(self.radius)
        self.h = C.c_void_p(_lib().or_bh_build(self.n, _p(self.xyz), _p(self.alpha), _p(self.radius),
                                               max_members, theta))

    def __del__(self):
        if getattr(self, "h", None):
            _lib().or_bh_destroy(self.h)
            self.h = None

    def info(self):
        nn, lv = C.c_int32(), C.c_int32()
        o, h = np.zeros(3), np.zeros(3)
        _lib().or_bh_info(self.h, C.byref(nn), C.byref(lv), _p(o), _p(h))
        return dict(n_nodes=nn.value, levels=lv.value, origin=o, half=h)

    def eval(self, mode, pts=None):
        if mode == MODE_VEL_POINTS:
            q = _a(pts).reshape(-1, 3)
            out = np.zeros_like(q)
            _lib().or_bh_eval(self.h, mode, len(q), _p(q), _p(out))
        else:
            out = np.zeros((self.n, 3))
            _lib().or_bh_eval(self.h, mode, 0, None, _p(out))
        return out


def direct(mode, xyz, alpha, radius, pts=None):
    xyz, alpha, radius = _a(xyz).reshape(-1, 3), _a(alpha).reshape(-1, 3), _a(radius).reshape(-1)
    if mode in (MODE_VEL_POINTS, MODE_GAUSS_POINTS):
        q = _a(pts).reshape(-1, 3)
    else:
        q = np.zeros((len(radius), 3))
    out = np.zeros_like(q)
    _lib().or_vp_direct(mode, len(radius), _p(xyz), _p(alpha), _p(radius), len(q), _p(q), _p(out))
    return out


CORE_VEL_POINTS, CORE_VEL, CORE_STRETCH = 0, 1, 2


def direct_core(mode, core, xyz, alpha, pts=None):
    """Direct sums of the configured case's kernel (algebraic, one core
    radius): CORE_VEL_POINTS at pts, CORE_VEL / CORE_STRETCH on the particles."""
    xyz, alpha = _a(xyz).reshape(-1, 3), _a(alpha).reshape(-1, 3)
    q = _a(pts).reshape(-1, 3) if mode == CORE_VEL_POINTS else np.zeros((len(xyz), 3))
    out = np.zeros_like(q)
    _lib().or_vp_direct_core(mode, float(core), len(xyz), _p(xyz), _p(alpha), len(q), _p(q), _p(out))
    return out


def pair_core(mode, core, xs, as_, xt, at=(0.0, 0.0, 0.0)):
    out = np.zeros(3)
    _lib().or_vp_pair_core(mode, float(core), _p(_a(xs)), _p(_a(as_)), _p(_a(xt)), _p(_a(at)), _p(out))
    return out


def pair(mode, xs, as_, rs, xt, at=(0.0, 0.0, 0.0), rt=0.0):
    out = np.zeros(3)
    _lib().or_vp_pair(mode, _p(_a(xs)), _p(_a(as_)), float(rs), _p(_a(xt)), _p(_a(at)), float(rt), _p(out))
    return out


def panels_to_points(body_verts, sigma, mu, wake_verts, mu_w, pts):
    """body_verts/wake_verts: (k, nv, 3) arrays (nv = 3 or 4)."""
    def pack(V):
        V = _a(V)
        k = V.shape[0]
        nv = np.full(k, V.shape[1], dtype=np.int32)
        buf = np.zeros((k, 12))
        buf[:, :3 * V.shape[1]] = V.reshape(k, -1)
        return k, nv, buf
    nb, nvb, vb = pack(body_verts)
    nw, nvw, vw = pack(wake_verts) if wake_verts is not None and len(wake_verts) else (0, np.zeros(1, np.int32), np.zeros((1, 12)))
    q = _a(pts).reshape(-1, 3)
    out = np.zeros_like(q)
    mw = _a(mu_w) if nw else np.zeros(1)
    _lib().or_panels_to_points(nb, _p(nvb, _i), _p(vb), _p(_a(sigma)), _p(_a(mu)), nw, _p(nvw, _i), _p(vw),
                               _p(mw), len(q), _p(q), _p(out))
    return out


VG_PART, VG_POINT_SOURCE, VG_PANEL_SOURCE, VG_POINT_DOUBLET, VG_PANEL_DOUBLET, VG_EDGE = 0, 1, 2, 3, 4, 5


def velgrad(verts, pos, kind=VG_PART, body=True, sigma=1.0, mu=1.0):
    """3x3 M(i, j) = d u_j / d x_i of one panel at pos (see or_velgrad)."""
    v = _a(verts).reshape(-1)
    out = np.zeros(9)
    _lib().or_velgrad(len(v) // 3, _p(v), 1 if body else 0, kind, float(sigma), float(mu), _p(_a(pos)), _p(out))
    return out.reshape(3, 3)


def panels_stretching(body_verts, sigma, mu, wake_verts, mu_w, pos, alpha):
    def pack(V):
        V = _a(V)
        k = V.shape[0]
        nv = np.full(k, V.shape[1], dtype=np.int32)
        buf = np.zeros((k, 12))
        buf[:, :3 * V.shape[1]] = V.reshape(k, -1)
        return k, nv, buf
    nb, nvb, vb = pack(body_verts)
    if wake_verts is not None and len(wake_verts):
        nw, nvw, vw = pack(wake_verts)
        mw = _a(mu_w)
    else:
        nw, nvw, vw, mw = 0, np.zeros(1, np.int32), np.zeros((1, 12)), np.zeros(1)
    q = _a(pos).reshape(-1, 3)
    al = _a(alpha).reshape(-1, 3)
    out = np.zeros_like(q)
    _lib().or_panels_stretching(nb, _p(nvb, _i), _p(vb), _p(_a(sigma)), _p(_a(mu)), nw, _p(nvw, _i), _p(vw), _p(mw),
                                len(q), _p(q), _p(al), _p(out))
    return out
