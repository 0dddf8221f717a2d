"""Scenarios run against the CPU replay of libcpgpu in a child process with
the sanitizer runtime preloaded (tests/test_replay.py).

    python tests/replay/run_case.py LIB assembly
    python tests/replay/run_case.py LIB vp STEPS
    python tests/replay/run_case.py LIB vpcfg STEPS
    python tests/replay/run_case.py LIB panels
    python tests/replay/run_case.py LIB teardown

Each prints "ok ..." lines; any sanitizer report aborts the process.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cpanel_amd import abi  # noqa: E402

abi.LIB_PATH = sys.argv[1]

from cpanel_amd import CP_FP32_REFINE, CP_FP64, meshgen  # noqa: E402
from cpanel_amd.case import Context  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def rel(a, b):
    scale = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (scale if scale > 0 else 1.0)


def assembly():
    """The far and near passes of the assembly (lane-permute compaction
    included) against the FP64 oracle on a small wing."""
    from oracle import oracle as O
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    og = O.Geom.from_arrays(m["xyz"], m["tri"], m["ids"])
    A, B = og.assemble()
    ctx = Context(0)
    g = ctx.geometry(m)
    g.assemble(CP_FP64)
    Ag, Bg = g.matrices(CP_FP64)
    ea, eb = np.abs(Ag - A).max() / np.abs(A).max(), np.abs(Bg - B).max() / np.abs(B).max()
    assert ea <= 1e-12 and eb <= 1e-12, (ea, eb)
    g.assemble(CP_FP32_REFINE)
    A32, B32 = g.matrices(CP_FP32_REFINE)
    ea32, eb32 = np.abs(A32 - A).max() / np.abs(A).max(), np.abs(B32 - B).max() / np.abs(B).max()
    assert ea32 <= 4e-7 and eb32 <= 4e-7, (ea32, eb32)
    print("ok assembly n=%d fp64 %.1e %.1e fp32 %.1e %.1e" % (g.nb, ea, eb, ea32, eb32), flush=True)
    g.release()
    ctx.close()


def vp(steps):
    """The vortex-particle driver on the reference's steadyTest case against
    its stored steps; the handles are left to the interpreter, as a script
    that never closes them would (this module's functions hold its globals,
    so at exit the handles are finalized together, in no fixed order)."""
    global ctx, geom, case
    f = np.load(os.path.join(GOLDEN, "steadyTest_vp.npz"))
    ctx = Context(0)
    geom = ctx.geometry_vp(meshgen.read_tri(os.path.join(GOLDEN, "steadyTest.tri")), V=1.0, dt=0.0)
    case = geom.vp_case((1.0, 10.0, 0.0), prec=CP_FP64, mach=0.1, Sref=6, bref=6, cref=1, cg=(0.416, 0, 0.03),
                        accelerate=True, tol=1e-12)
    for k in range(1, steps + 1):
        case.step()
        if ("mu_%d" % k) in f:
            st = case.state()
            d = rel(st["mu"], f["mu_%d" % k])
            assert d < 2e-5, (k, d)
            if k >= 3:
                p = case.particles()
                assert np.max(np.abs(p["pos"] - f["ppos_%d" % k])) < 2e-5
            print("ok vp step %d dmu %.2e" % (k, d), flush=True)
    fin = case.finish()
    assert np.isfinite(fin["CL"])
    print("ok vp finish", flush=True)


def vp_config(steps):
    """The configured form of the vortex-particle case (algebraic fixed-core
    kernel, two particles per second-row panel, FP64 and FP32 sums) on a
    small swept, tapered wing; after the steps, the Barnes-Hut sums of a
    theta = 0 tree on the case's particles against the oracle's direct
    sums."""
    from oracle import vp_oracle as VO
    m = meshgen.wing("4412", 1.0, 4.0, 16, 6, "uniform", 1.0, taper=0.5, sweep_deg=20.0, with_wake=False)
    c = Context(0)
    g = c.geometry_vp(m, V=1.0, dt=0.05)
    out = {}
    for f32 in (False, True):
        case = g.vp_case((1.0, 6.0, 0.0), prec=CP_FP64, mach=0.1, Sref=3, bref=4, cref=0.75, accelerate=True,
                         warm_start=True, max_members=32, theta=0.5, tol=1e-12, kernel=abi.CP_VP_ALGEBRAIC,
                         per_edge=2, fp32=f32)
        for k in range(1, steps + 1):
            s = case.step()
            assert s["rel_residual"] <= 1e-12, (k, s)
        p = case.particles()
        assert len(p["radius"]) == 2 * 6 * (steps - 2), len(p["radius"])
        out[f32] = (case.state()["mu"], p)
        case.close()
    mu64, p64 = out[False]
    mu32, p32 = out[True]
    dmu = rel(mu32, mu64)
    dpos = float(np.max(np.abs(p32["pos"] - p64["pos"])))
    assert dmu < 1e-4 and dpos < 1e-4, (dmu, dpos)
    core = float(p64["radius"][0])
    x, a = p64["pos"], p64["strength"]
    for f32, tol in ((False, 1e-12), (True, 2e-5)):
        t = c.bh_tree(x, a, p64["radius"], max_members=32, theta=0.0, kernel=abi.CP_VP_ALGEBRAIC, core=core,
                      fp32=f32)
        for mode, omode in ((abi.CP_BH_VEL, VO.CORE_VEL), (abi.CP_BH_STRETCH, VO.CORE_STRETCH)):
            d = rel(t.eval(mode), VO.direct_core(omode, core, x, a))
            assert d <= tol, (f32, mode, d)
        pts = x[:7] + 0.01
        d = rel(t.eval(abi.CP_BH_VEL_POINTS, pts), VO.direct_core(VO.CORE_VEL_POINTS, core, x, a, pts))
        assert d <= tol, (f32, d)
        t.close()
    print("ok vp config n=%d dmu %.1e dpos %.1e" % (len(x), dmu, dpos), flush=True)
    g.release()
    c.close()


def panels():
    """The panel -> point velocity and panel -> particle stretching sums
    (far pass over LDS-staged panels, near pass over queued pairs) against
    the oracle on a small wing, points near the surface and far away, FP64
    and with the far field in FP32."""
    from oracle import vp_oracle as VO
    m = meshgen.wing("4412", 1.0, 4.0, 24, 8, "uniform", 2.0)
    c = Context(0)
    g = c.geometry(m)
    tri, xyz = np.asarray(m["tri"]), np.asarray(m["xyz"])
    bv, wv = xyz[tri[g.body()["tri_index"]]], xyz[tri[g.wake()["tri_index"]]]
    rng = np.random.default_rng(4)
    sigma, mu, muw = rng.normal(size=g.nb), rng.normal(size=g.nb), rng.normal(size=g.nw)
    ctr, nrm = g.body()["center"], g.body()["normal"]
    pts = np.vstack([ctr[::5] + 0.03 * nrm[::5], rng.normal(size=(150, 3)) * 3.0,
                     np.column_stack([rng.uniform(1.0, 3.0, 70), rng.uniform(-2, 2, 70), rng.normal(0, 0.05, 70)])])
    alpha = rng.normal(size=pts.shape)
    ref = VO.panels_to_points(bv, sigma, mu, wv, muw, pts)
    sref = VO.panels_stretching(bv, sigma, mu, wv, muw, pts, alpha)
    ok = np.isfinite(sref).all(axis=1)
    for f32, tol in ((False, 1e-10), (True, 1e-5)):
        got = g.panel_to_points(sigma, mu, pts, mu_wake=muw, fp32=f32)
        e = np.abs(got - ref).max() / np.abs(ref).max()
        sg = g.panel_stretching(sigma, mu, pts, alpha, mu_wake=muw, fp32=f32)
        assert np.array_equal(np.isnan(sg), np.isnan(sref))
        es = np.abs(sg[ok] - sref[ok]).max() / np.abs(sref[ok]).max()
        assert e <= tol and es <= (1e-8 if not f32 else 1e-5), (f32, e, es)
        print("ok panels fp32=%d velocity %.1e stretching %.1e (m=%d, %d panels)" %
              (f32, e, es, len(pts), g.nb + g.nw), flush=True)
    g.release()
    c.close()


def teardown():
    """Parents closed before their children: the context before a geometry,
    a tree and a vortex-particle case, the geometry before its case."""
    m = meshgen.small_wing("4412", n_chord=12, n_span=4)
    c = Context(0)
    g = c.geometry(m)
    g.assemble(CP_FP64)
    t = c.bh_tree(np.random.default_rng(0).normal(size=(40, 3)), np.ones((40, 3)), np.full(40, 0.1))
    c.close()
    r = g.solve([(1.0, 5.0, 0.0)], CP_FP64, tol=1e-12, mach=0.1, Sref=4, bref=4, cref=1)[0]
    assert np.isfinite(r["CL"])
    assert np.all(np.isfinite(t.eval(abi.CP_BH_VEL)))
    g.release()
    t.close()
    c2 = Context(0)
    gv = c2.geometry_vp(meshgen.read_tri(os.path.join(GOLDEN, "steadyTest.tri")), V=1.0, dt=0.0)
    case = gv.vp_case((1.0, 10.0, 0.0), prec=CP_FP64, mach=0.1, Sref=6, bref=6, cref=1, cg=(0.416, 0, 0.03),
                      accelerate=True, tol=1e-10)
    gv.release()
    c2.close()
    case.step()
    case.close()
    print("ok teardown", flush=True)


if __name__ == "__main__":
    what = sys.argv[2]
    if what == "assembly":
        assembly()
    elif what == "vp":
        vp(int(sys.argv[3]))
    elif what == "vpcfg":
        vp_config(int(sys.argv[3]))
    elif what == "panels":
        panels()
    elif what == "teardown":
        teardown()
