"""Host-side mirror of CPanel's geometry / cpCase interface over the C ABI.

    ctx  = Context(device=0)
    geom = ctx.geometry(mesh)          # geometry(inputParams*) + readTri
    geom.assemble(CP_FP64)             # geometry::setInfCoeff
    res  = geom.solve([(V, alpha, beta), ...], prec)   # cpCase::run per condition

Results carry mu, Cp, CL, CDi, Fbody, Fwind, CM as cpCase reports them
(cpCase.h:129-133).  Errors raise CPanelError with the library's message.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import CP_FP32_REFINE, CP_FP64  # noqa: F401


class CPanelError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (abi.lib().cp_status_string(status).decode(), msg))
        self.status = status


# cp_solve_info as a NumPy record (the layout of abi.SolveInfo)
_SOLVE_INFO = np.dtype([("CL", "<f8"), ("CDi", "<f8"), ("Fbody", "<f8", (3,)), ("Fwind", "<f8", (3,)),
                        ("CM", "<f8", (3,)), ("rel_residual", "<f8"), ("iters", "<i4"), ("refinements", "<i4"),
                        ("converged", "<i4"), ("pad", "<i4")])
assert _SOLVE_INFO.itemsize == C.sizeof(abi.SolveInfo)


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class Context:
    def __init__(self, device=0, hbm_budget_bytes=0):
        L = abi.lib()
        h = C.c_void_p()
        st = L.cp_create(device, hbm_budget_bytes, C.byref(h))
        if st != abi.CP_OK:
            raise CPanelError(st, "cp_create(device=%d)" % device)
        self.h = h

    def check(self, st, what=""):
        if st not in (abi.CP_OK,):
            raise CPanelError(st, "%s %s" % (what, abi.lib().cp_last_error(self.h).decode()))
        return st

    def close(self):
        if self.h:
            abi.lib().cp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self, mesh, key=0):
        return Geometry(self, mesh, key)

    def solve_group(self, jobs, prec=CP_FP32_REFINE, mach=0.1, Sref=1.0, bref=1.0, cref=1.0, cg=(0, 0, 0),
                    tol=1e-10, max_iter=0, max_refine=0, inner_tol=0.0, want_mu=True, want_cp=False,
                    allow_unconverged=False, precond=0):
        """Several geometries solved in one pass (cp_solve_group): jobs is a
        list of (geometry, conditions) or (geometry, conditions, options)
        with distinct geometries and at most 64 conditions each; options (a
        dict of this method's keywords) override the shared ones for that
        geometry.  Returns one list of per-condition results per job, as
        Geometry.solve does."""
        shared = dict(mach=mach, Sref=Sref, bref=bref, cref=cref, cg=cg, tol=tol, max_iter=max_iter,
                      max_refine=max_refine, inner_tol=inner_tol, precond=precond)
        opts = []
        for j in jobs:
            o = dict(shared, **(j[2] if len(j) > 2 else {}))
            opts.append(abi.SolveOpts(o["mach"], o["Sref"], o["bref"], o["cref"], (C.c_double * 3)(*o["cg"]),
                                      o["tol"], o["max_iter"], o["max_refine"], o["inner_tol"], o["precond"], 0))
        jobs = [(j[0], j[1]) for j in jobs]
        G = len(jobs)
        geoms = (C.c_void_p * G)(*[g.h for g, _ in jobs])
        ks = (C.c_int32 * G)(*[len(c) for _, c in jobs])
        fs_keep = [(abi.Freestream * len(c)).from_buffer_copy(np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3))
                   for _, c in jobs]
        fs = (C.POINTER(abi.Freestream) * G)(*[C.cast(f, C.POINTER(abi.Freestream)) for f in fs_keep])
        o = (abi.SolveOpts * G)(*opts)
        mus = [np.zeros((len(c), g.nb)) if want_mu else None for g, c in jobs]
        cps = [np.zeros((len(c), g.nb)) if want_cp else None for g, c in jobs]
        infos = [(abi.SolveInfo * len(c))() for _, c in jobs]
        dp = C.POINTER(C.c_double)
        mu_p = (dp * G)(*[_p(m) if m is not None else dp() for m in mus])
        cp_p = (dp * G)(*[_p(q) if q is not None else dp() for q in cps])
        in_p = (C.POINTER(abi.SolveInfo) * G)(*[C.cast(i, C.POINTER(abi.SolveInfo)) for i in infos])
        st = abi.lib().cp_solve_group(G, geoms, prec, ks, fs, o, mu_p, cp_p, in_p)
        if not (st == abi.CP_NOT_CONVERGED and allow_unconverged):
            self.check(st, "solve_group")
        out = []
        for (g, c), m, q, info in zip(jobs, mus, cps, infos):
            rec = np.frombuffer(info, dtype=_SOLVE_INFO)
            out.append([dict(mu=None if m is None else m[i], cp=None if q is None else q[i], CL=float(rec["CL"][i]),
                             CDi=float(rec["CDi"][i]), Fbody=rec["Fbody"][i].copy(), Fwind=rec["Fwind"][i].copy(),
                             CM=rec["CM"][i].copy(), rel_residual=float(rec["rel_residual"][i]),
                             iters=int(rec["iters"][i]), refinements=int(rec["refinements"][i]),
                             converged=bool(rec["converged"][i])) for i in range(len(c))])
        return out

    def geometry_vp(self, mesh, V=1.0, dt=0.0, key=0):
        """Vortex-particle geometry: no prescribed wake, two buffer rows of
        quads behind the trailing edge (time step dt; <= 0: computed)."""
        return Geometry(self, mesh, key, vp=(V, dt))

    def bh_tree(self, xyz, alpha, radius, max_members=10, theta=0.5, kernel=0, core=0.0, fp32=False,
                host_build=False):
        return BHTree(self, xyz, alpha, radius, max_members, theta, kernel, core, fp32, host_build)

    def bytes_in_use(self):
        return abi.lib().cp_bytes_in_use(self.h)

    def timing(self, name):
        ms = C.c_double()
        n = C.c_int64()
        self.check(abi.lib().cp_timing_get(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timing_enable(self, on=True):
        abi.lib().cp_timing_enable(self.h, 1 if on else 0)

    def timing_only(self, name=None):
        """Record only `name`'s intervals while timing is on (None: all)."""
        self.check(abi.lib().cp_timing_only(self.h, None if name is None else name.encode()))

    def timing_reset(self):
        self.check(abi.lib().cp_timing_reset(self.h))


class Geometry:
    def __init__(self, ctx, mesh, key=0, vp=None):
        self.ctx = ctx
        L = abi.lib()
        h = C.c_void_p()
        if vp is not None:
            if isinstance(mesh, str):
                from . import meshgen
                mesh = meshgen.read_tri(mesh)
            self._xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float64)
            self._tri = np.ascontiguousarray(mesh["tri"], dtype=np.int32)
            self._ids = np.ascontiguousarray(mesh["ids"], dtype=np.int32)
            m = abi.Mesh(len(self._xyz), _p(self._xyz), len(self._tri), _p(self._tri, C.c_int32),
                         _p(self._ids, C.c_int32))
            st = L.cp_geometry_upload_vp(ctx.h, C.byref(m), float(vp[0]), float(vp[1]), key, C.byref(h))
        elif isinstance(mesh, str):
            st = L.cp_geometry_from_tri(ctx.h, mesh.encode(), key, C.byref(h))
        else:
            self._xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float64)
            self._tri = np.ascontiguousarray(mesh["tri"], dtype=np.int32)
            self._ids = np.ascontiguousarray(mesh["ids"], dtype=np.int32)
            m = abi.Mesh(len(self._xyz), _p(self._xyz), len(self._tri), _p(self._tri, C.c_int32),
                         _p(self._ids, C.c_int32))
            st = L.cp_geometry_upload(ctx.h, C.byref(m), key, C.byref(h))
        ctx.check(st, "geometry upload")
        self.h = h
        info = abi.GeomInfo()
        ctx.check(L.cp_geometry_info(h, C.byref(info)))
        self.nb, self.nw, self.lda = info.n_body, info.n_wake, info.lda
        self.n_wakelines = info.n_wakelines

    def release(self):
        if self.h:
            abi.lib().cp_geometry_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def body(self):
        n = self.nb
        ti = np.zeros(n, np.int32)
        c = np.zeros((n, 3))
        nr = np.zeros((n, 3))
        a = np.zeros(n)
        self.ctx.check(abi.lib().cp_geometry_body(self.h, _p(ti, C.c_int32), _p(c), _p(nr), _p(a)))
        return dict(tri_index=ti, center=c, normal=nr, area=a)

    def wake(self):
        n = self.nw
        ti = np.zeros(max(n, 1), np.int32)
        idx = np.zeros((max(n, 1), 4), np.int32)
        cf = np.zeros(max(n, 1))
        self.ctx.check(abi.lib().cp_geometry_wake(self.h, _p(ti, C.c_int32), _p(idx, C.c_int32), _p(cf)))
        return dict(tri_index=ti[:n], idx4=idx[:n], coef=cf[:n])

    def clusters(self):
        n = self.nb
        cl = np.zeros((n, 16), np.int32)
        ts = np.zeros(n, np.int32)
        tip = np.zeros(n, np.int32)
        self.ctx.check(abi.lib().cp_geometry_clusters(self.h, _p(cl, C.c_int32), _p(ts, C.c_int32),
                                                      _p(tip, C.c_int32)))
        return cl, ts, tip

    def assemble(self, prec=CP_FP64):
        self.ctx.check(abi.lib().cp_assemble(self.h, prec), "assemble")

    def matrices(self, prec=CP_FP64):
        n = self.nb
        A = np.zeros((n, n), order="F")
        B = np.zeros((n, n), order="F")
        self.ctx.check(abi.lib().cp_read_matrices(self.h, prec, _p(A), _p(B), n), "read matrices")
        return A, B

    def load_matrices(self, A, B):
        """Set the FP64 A and B (n x n) in place of assemble(CP_FP64), as the
        reference does when it reads a .infCoeff file (geometry::readInfCoeff,
        geometry.cpp:1125-1166; cpanel_amd.infcoeff reads the file)."""
        n = self.nb
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        if A.shape != (n, n) or B.shape != (n, n):
            raise ValueError("A and B must be %d x %d" % (n, n))
        self.ctx.check(abi.lib().cp_load_matrices(self.h, _p(A), _p(B), n), "load matrices")

    def matvec(self, X, kernel=0):
        """Y = A X through one device kernel (0: FP64 GEMV, 1: FP32 MFMA, 2: FP64-acc MFMA,
        3: FP64-acc MFMA + FP64 near-field correction)."""
        X = np.asfortranarray(np.atleast_2d(np.asarray(X, dtype=np.float64).T).T)
        Y = np.zeros_like(X, order="F")
        self.ctx.check(abi.lib().cp_matvec(self.h, kernel, X.shape[1], _p(X), _p(Y)), "matvec")
        return Y

    def panel_to_points(self, sigma, mu, pts, mu_wake=None, fp32=False):
        """Velocity induced at pts (m x 3) by the body panels (source sigma,
        doublet mu, solve order) and the wake panels (doublet mu_wake);
        fp32: far-field pairs in FP32 arithmetic."""
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        m_ = np.ascontiguousarray(mu, dtype=np.float64)
        if sg.shape != (self.nb,) or m_.shape != (self.nb,):
            raise ValueError("sigma and mu need one value per body panel")
        mw = None if mu_wake is None else np.ascontiguousarray(mu_wake, dtype=np.float64)
        if mw is not None and mw.shape != (self.nw,):
            raise ValueError("mu_wake needs one value per wake panel")
        q = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(q)
        self.ctx.check(abi.lib().cp_panel_to_points(self.h, _p(sg), _p(m_), _p(mw), len(q), _p(q), _p(out),
                                                    int(fp32)), "cp_panel_to_points")
        return out

    def solve_with_derivatives(self, condition, prec=CP_FP64, delta_deg=0.5, **kw):
        """cpCase::run with stability derivatives (cpCase.cpp:635-665): the
        condition and its alpha + delta, beta + delta neighbours solved as one
        batch; returns (result, derivatives)."""
        from .cpout import stability_derivatives
        V, a, b = condition
        r = self.solve([(V, a, b), (V, a + delta_deg, b), (V, a, b + delta_deg)], prec, **kw)
        return r[0], stability_derivatives(r[0], r[1], r[2], delta_deg)

    def set_buffer_wake(self, verts):
        """verts: (m, nv, 3) panels of the second buffer-wake row (nv = 3 or 4)."""
        V = np.asarray(verts, dtype=np.float64)
        m = V.shape[0]
        nv = np.full(m, V.shape[1] if m else 3, dtype=np.int32)
        buf = np.zeros((max(m, 1), 12))
        if m:
            buf[:m, :3 * V.shape[1]] = V.reshape(m, -1)
        self.ctx.check(abi.lib().cp_set_buffer_wake(self.h, m, _p(nv, C.c_int32), _p(buf)), "cp_set_buffer_wake")
        self.nw2 = m

    def solve_rhs(self, sigma, prec=CP_FP64, mu_w2=None, x0=None, tol=1e-10, max_iter=0, max_refine=0,
                  inner_tol=0.0, precond=0, allow_unconverged=False):
        """Solves A mu = -B sigma - C mu_w2 (cpCaseVP::solveVPmatrixEq) for the
        columns of sigma (n_body x k or n_body); returns (mu, infos)."""
        sg = np.asarray(sigma, dtype=np.float64)
        one = sg.ndim == 1
        sg = np.ascontiguousarray(sg.reshape(self.nb, -1).T)  # k x n (column c contiguous)
        k = sg.shape[0]
        mw = None if mu_w2 is None else np.ascontiguousarray(np.asarray(mu_w2, np.float64).reshape(-1, k).T)
        x = None if x0 is None else np.ascontiguousarray(np.asarray(x0, np.float64).reshape(self.nb, -1).T)
        o = abi.SolveOpts(0.1, 1.0, 1.0, 1.0, (C.c_double * 3)(0, 0, 0), tol, max_iter, max_refine, inner_tol,
                          precond, 0)
        mu = np.zeros((k, self.nb))
        info = (abi.SolveInfo * k)()
        st = abi.lib().cp_solve_rhs(self.h, prec, k, _p(sg), _p(mw), _p(x), C.byref(o), _p(mu), info)
        if st != abi.CP_OK and not (st == abi.CP_NOT_CONVERGED and allow_unconverged):
            self.ctx.check(st, "cp_solve_rhs")
        infos = [dict(rel_residual=i.rel_residual, iters=i.iters, refinements=i.refinements,
                      converged=bool(i.converged)) for i in info]
        return (mu[0] if one else mu.T), infos

    def panel_stretching(self, sigma, mu, pos, alpha, mu_wake=None, fp32=False):
        """Stretching of particles (pos, alpha: m x 3) by the body panels
        (sigma, mu) and wake panels (mu_wake): sum of partStretching;
        fp32: far-field pairs in FP32 arithmetic."""
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        m_ = np.ascontiguousarray(mu, dtype=np.float64)
        if sg.shape != (self.nb,) or m_.shape != (self.nb,):
            raise ValueError("sigma and mu need one value per body panel")
        mw = None if mu_wake is None else np.ascontiguousarray(mu_wake, dtype=np.float64)
        q = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        al = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, 3)
        if al.shape != q.shape:
            raise ValueError("pos and alpha must describe the same particles")
        out = np.zeros_like(q)
        self.ctx.check(abi.lib().cp_panel_stretching(self.h, _p(sg), _p(m_), _p(mw), len(q), _p(q), _p(al), _p(out),
                                                     int(fp32)), "cp_panel_stretching")
        return out

    def case_fields(self, conditions, mu, mach=0.1, Sref=1.0, bref=1.0, cref=1.0, cg=(0, 0, 0)):
        """Per-panel and per-station outputs of solved conditions, from their
        body doublet strengths mu (k x n_body): the arrays the reference writes
        to surfaceData.vtu, wakeData.vtu and spanwiseData_Wake<i>.csv.  Returns
        one dict per condition."""
        k = len(conditions)
        mu = np.ascontiguousarray(np.asarray(mu, dtype=np.float64).reshape(k, self.nb))
        info = abi.GeomInfo()
        self.ctx.check(abi.lib().cp_geometry_info(self.h, C.byref(info)))
        nwk, nst = info.n_wakes, info.n_stations
        fs = (abi.Freestream * k)(*[abi.Freestream(*map(float, c)) for c in conditions])
        o = abi.SolveOpts(mach, Sref, bref, cref, (C.c_double * 3)(*cg), 0.0, 0, 0, 0.0, 0, 0)
        sig = np.zeros((k, self.nb))
        vel = np.zeros((k, self.nb, 3))
        pot = np.zeros((k, self.nb))
        cp = np.zeros((k, self.nb))
        wmu = np.zeros((k, max(self.nw, 1)))
        wpot = np.zeros((k, max(self.nw, 1)))
        span = np.zeros((k, max(nwk, 1), nst, 3))
        f = abi.CaseFields(_p(sig), _p(vel), _p(pot), _p(cp), _p(wmu), _p(wpot), _p(span))
        self.ctx.check(abi.lib().cp_case_fields_eval(self.h, k, fs, C.byref(o), _p(mu), C.byref(f)), "case fields")
        return [dict(mu=mu[q], sigma=sig[q], potential=pot[q], velocity=vel[q], cp=cp[q],
                     wake_mu=wmu[q, :self.nw], wake_potential=wpot[q, :self.nw], spanwise=span[q, :nwk])
                for q in range(k)]

    def vp_rows(self):
        """(dt, first-row quads, second-row quads) of a vortex-particle
        geometry; quads as (m, 4, 3) vertex arrays in node order."""
        m = C.c_int32()
        dt = C.c_double()
        L = abi.lib()
        self.ctx.check(L.cp_vp_buffer_rows(self.h, C.byref(m), C.byref(dt), None, None), "buffer rows")
        r1 = np.zeros((m.value, 4, 3))
        r2 = np.zeros((m.value, 4, 3))
        self.ctx.check(L.cp_vp_buffer_rows(self.h, None, None, _p(r1), _p(r2)), "buffer rows")
        return dt.value, r1, r2

    def vp_case(self, condition, prec=CP_FP64, mach=0.1, Sref=1.0, bref=1.0, cref=1.0, cg=(0, 0, 0), accelerate=True,
                warm_start=False, max_members=10, theta=0.5, tol=1e-12, kernel=0, per_edge=1, core=0.0, fp32=False):
        """kernel=abi.CP_VP_ALGEBRAIC, per_edge, core, fp32: the configured
        form (algebraic fixed-core particle kernel without diffusion, q
        particles per second-row panel and step, FP32 particle and far-field
        panel sums); the defaults are the reference's case."""
        return VpCase(self, condition, prec, mach, Sref, bref, cref, cg, accelerate, warm_start, max_members, theta,
                      tol, kernel, per_edge, core, fp32)

    def block_jacobi(self):
        """(ptr, perm, [inverse of each block as a b x b FP32 array]) of the
        mixed-precision solve's strip block-Jacobi preconditioner, or None
        when the point-Jacobi fallback is in use."""
        sz = (C.c_int64 * 3)()
        lib = abi.lib()
        self.ctx.check(lib.cp_block_jacobi_read(self.h, sz, None, None, None), "block jacobi")
        nb, npan, tot = sz[0], sz[1], sz[2]
        if nb == 0:
            return None
        ptr = np.zeros(nb + 1, np.int32)
        perm = np.zeros(npan, np.int32)
        inv = np.zeros(tot, np.float32)
        self.ctx.check(lib.cp_block_jacobi_read(self.h, sz, _p(ptr, C.c_int32), _p(perm, C.c_int32),
                                                _p(inv, C.c_float)), "block jacobi")
        blocks, o = [], 0
        for q in range(nb):
            b = int(ptr[q + 1] - ptr[q])
            blocks.append(inv[o:o + b * b].reshape(b, b).T)  # stored column-major
            o += b * b
        return ptr, perm, blocks

    def solver_report(self):
        import json
        buf = C.create_string_buffer(1 << 16)
        self.ctx.check(abi.lib().cp_solver_report(self.h, buf, len(buf)), "solver report")
        return json.loads(buf.value.decode())

    def solve(self, conditions, prec=CP_FP64, mach=0.1, Sref=1.0, bref=1.0, cref=1.0, cg=(0, 0, 0), tol=1e-10,
              max_iter=0, max_refine=0, inner_tol=0.0, want_cp=True, allow_unconverged=False, precond=0,
              want_mu=True):
        """Solves k freestream conditions (V, alpha_deg, beta_deg) on the
        assembled geometry (cpCase::run per condition).  want_mu / want_cp
        False leave the per-panel arrays on the device (the forces, CL and
        CDi still come back)."""
        k = len(conditions)
        # the k cp_freestream records as one (k, 3) FP64 buffer
        fs = (abi.Freestream * k).from_buffer_copy(np.ascontiguousarray(conditions, dtype=np.float64).reshape(k, 3))
        o = abi.SolveOpts(mach, Sref, bref, cref, (C.c_double * 3)(*cg), tol, max_iter, max_refine, inner_tol,
                          precond, 0)
        mu = np.zeros((k, self.nb)) if want_mu else None
        cp = np.zeros((k, self.nb)) if want_cp else None
        info = (abi.SolveInfo * k)()
        st = abi.lib().cp_solve(self.h, prec, k, fs, C.byref(o), _p(mu), _p(cp), info)
        if st == abi.CP_NOT_CONVERGED and allow_unconverged:
            pass
        else:
            self.ctx.check(st, "solve")
        # the k cp_solve_info records read through one structured view
        rec = np.frombuffer(info, dtype=_SOLVE_INFO)
        CL, CDi, rel = rec["CL"].tolist(), rec["CDi"].tolist(), rec["rel_residual"].tolist()
        iters, refs, conv = rec["iters"].tolist(), rec["refinements"].tolist(), rec["converged"].tolist()
        Fb, Fw, CM = rec["Fbody"].copy(), rec["Fwind"].copy(), rec["CM"].copy()
        return [dict(mu=None if mu is None else mu[q], cp=None if cp is None else cp[q], CL=CL[q], CDi=CDi[q],
                     Fbody=Fb[q], Fwind=Fw[q], CM=CM[q], rel_residual=rel[q], iters=iters[q], refinements=refs[q],
                     converged=bool(conv[q])) for q in range(k)]


class VpCase:
    """The vortex-particle case of cpCaseVP (cpCaseVP.cpp:26-158) on a
    vortex-particle geometry: step() advances one time step, run() loops to
    the step count or to the Trefftz convergence and returns finish()'s
    forces."""

    def __init__(self, geom, condition, prec, mach, Sref, bref, cref, cg, accelerate, warm_start, max_members,
                 theta, tol, kernel=0, per_edge=1, core=0.0, fp32=False):
        self.g = geom
        o = abi.VpOpts()
        o.fs = abi.Freestream(*map(float, condition))
        o.mach, o.Sref, o.bref, o.cref = mach, Sref, bref, cref
        o.cg = (C.c_double * 3)(*cg)
        o.prec, o.accelerate, o.warm_start, o.max_members = prec, int(accelerate), int(warm_start), max_members
        o.theta, o.tol = theta, tol
        o.kernel, o.per_edge, o.core, o.fp32 = kernel, per_edge, core, int(fp32)
        h = C.c_void_p()
        geom.ctx.check(abi.lib().cp_vp_create(geom.h, C.byref(o), C.byref(h)), "vp create")
        self.h = h

    def step(self):
        i = abi.VpStepInfo()
        st = abi.lib().cp_vp_step(self.h, C.byref(i))
        if st not in (abi.CP_OK, abi.CP_NOT_CONVERGED):
            self.g.ctx.check(st, "vp step")
        return dict(step=i.step, n_particles=i.n_particles, converged=bool(i.converged), iters=i.solve_iters,
                    CL=i.CL_trefftz, CDi=i.CDi_trefftz, rel_residual=i.rel_residual, ms=i.ms,
                    solved=st == abi.CP_OK)

    def run(self, steps, on_step=None):
        """cpCaseVP::run's loop: steps time steps, stopping early when the
        Trefftz values converge; then finish()."""
        for _ in range(steps):
            s = self.step()
            if on_step:
                on_step(s)
            if s["converged"]:
                break
        return self.finish()

    def finish(self):
        r = abi.SolveInfo()
        cp = np.zeros(self.g.nb)
        self.g.ctx.check(abi.lib().cp_vp_finish(self.h, C.byref(r), _p(cp)), "vp finish")
        return dict(CL=r.CL, CDi=r.CDi, Fbody=np.array(r.Fbody[:]), Fwind=np.array(r.Fwind[:]),
                    CM=np.array(r.CM[:]), cp=cp)

    def particles(self):
        n = C.c_int32()
        L = abi.lib()
        L.cp_vp_particles(self.h, C.byref(n), None, None, None, None)
        k = n.value
        pos, st, rad = np.zeros((k, 3)), np.zeros((k, 3)), np.zeros(k)
        shed = np.zeros(k, np.int32)
        L.cp_vp_particles(self.h, None, _p(pos), _p(st), _p(rad), _p(shed, C.c_int32))
        return dict(pos=pos, strength=st, radius=rad, shed=shed)

    def state(self):
        m = self.g.nw
        mu, sig = np.zeros(self.g.nb), np.zeros(self.g.nb)
        m1, m2 = np.zeros(max(m, 1)), np.zeros(max(m, 1))
        abi.lib().cp_vp_state(self.h, _p(mu), _p(sig), _p(m1), _p(m2))
        return dict(mu=mu, sigma=sig, mu_row1=m1[:m], mu_row2=m2[:m])

    def close(self):
        if self.h:
            abi.lib().cp_vp_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BHTree:
    """Particle octree with pseudo-particles on the device
    (particleFMM::build, particleFMM.cpp:22-27); eval() runs the Barnes-Hut
    sums of particleFMM::barnesHutVel/Stretch/Diff (particleFMM.cpp:65-116)."""

    def __init__(self, ctx, xyz, alpha, radius, max_members=10, theta=0.5, kernel=0, core=0.0, fp32=False,
                 host_build=False):
        """kernel=abi.CP_VP_ALGEBRAIC: the algebraic kernel of one core
        radius `core` for every mode (no CP_BH_DIFF), in FP32 with fp32.
        host_build: build the tree on the host (the device build's check)."""
        self.ctx = ctx
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.alpha = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, 3)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64).reshape(-1)
        self.n = len(self.radius)
        if self.xyz.shape[0] != self.n or self.alpha.shape[0] != self.n:
            raise ValueError("xyz, alpha and radius must describe the same particles")
        h = C.c_void_p()
        o = abi.BhOpts(max_members, kernel, theta, core, int(fp32), int(host_build))
        ctx.check(abi.lib().cp_bh_build(ctx.h, self.n, _p(self.xyz), _p(self.alpha), _p(self.radius), C.byref(o),
                                        C.byref(h)), "cp_bh_build")
        self.h = h

    def info(self):
        nn, lv = C.c_int32(), C.c_int32()
        o, h = np.zeros(3), C.c_double()
        abi.lib().cp_bh_info(self.h, C.byref(nn), C.byref(lv), _p(o), C.byref(h))
        return dict(n_nodes=nn.value, levels=lv.value, origin=o, half=np.full(3, h.value))

    def read(self):
        """The flattened tree: nodes (n_nodes x 8: pseudo-particle position,
        cell width, strength, radius), links (n_nodes x 3: skip, first
        member slot, member count or -1) and order (slot -> particle)."""
        nn = self.info()["n_nodes"]
        nodes = np.zeros((max(nn, 1), 8))
        links = np.zeros((max(nn, 1), 3), np.int32)
        order = np.zeros(max(self.n, 1), np.int32)
        self.ctx.check(abi.lib().cp_bh_read(self.h, _p(nodes), _p(links, C.c_int32), _p(order, C.c_int32)),
                       "cp_bh_read")
        return nodes[:nn], links[:nn], order[:self.n]

    def eval(self, mode, pts=None):
        if mode == abi.CP_BH_VEL_POINTS:
            q = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
            out = np.zeros_like(q)
            self.ctx.check(abi.lib().cp_bh_eval(self.h, mode, len(q), _p(q), _p(out)), "cp_bh_eval")
        else:
            out = np.zeros((self.n, 3))
            self.ctx.check(abi.lib().cp_bh_eval(self.h, mode, 0, None, _p(out)), "cp_bh_eval")
        return out

    def close(self):
        if getattr(self, "h", None):
            abi.lib().cp_bh_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

