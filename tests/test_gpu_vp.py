"""GPU parity of the vortex-particle kernels through the C ABI.

Tolerances (stated per quantity): FP64 throughout, so the Barnes-Hut sums on
the same tree and the panel -> point sums agree with the oracle to rounding:
|out - oracle| <= 1e-11 * max|oracle| (different summation association and
last-bit differences of erf/exp/atan2/log only).  With theta = 0 the
traversal must equal the oracle's direct sum to the same bound.
"""
import sys

import numpy as np
import pytest

from cpanel_amd import abi, meshgen
from cpanel_amd.case import Context
from oracle import vp_oracle as V

pytestmark = pytest.mark.gpu

TOL = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu-vp] creating context on device 0")
    c = Context(0)
    yield c
    # the geometries of this module's fixtures keep it alive until they go
    c.close()


def cloud(n, seed, spread=0.5):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3)) * spread
    x[:, 0] *= 3.0  # a wake-like elongated cloud
    a = rng.normal(size=(n, 3)) * 0.05
    r = 0.02 + 0.03 * rng.random(n)
    return x, a, r


def close(got, ref, tol=TOL):
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    return err <= tol * scale, err / max(scale, 1e-300)


def tree_cases():
    rng = np.random.default_rng(11)
    X, A, R = meshgen.c4_particles(meshgen.config_c4())
    grid = np.stack(np.meshgrid(np.arange(12.), np.arange(12.), np.arange(12.)), -1).reshape(-1, 3)
    Y = rng.normal(size=(3000, 3))
    Y[:100] = Y[0] + 1e-9 * rng.normal(size=(100, 3))
    return [("c4 80k, 10 per leaf", X, A, R, 10), ("c4 80k, 32 per leaf", X, A, R, 32),
            ("random 5000", rng.normal(size=(5000, 3)), rng.normal(size=(5000, 3)), np.full(5000, 0.1), 10),
            ("grid 1728 (particles on the split planes)", grid, rng.normal(size=grid.shape), np.full(len(grid), 0.2),
             10),
            ("7 particles (the root is a leaf)", rng.normal(size=(7, 3)), rng.normal(size=(7, 3)), np.full(7, 0.1), 10),
            ("clustered within 1e-9 (beyond the keys: host build)", Y, rng.normal(size=Y.shape), np.full(3000, 0.1),
             10)]


def test_device_tree_equals_host_tree(ctx):
    """The device build (octant keys by the host's own centre recurrence,
    radix sort, nodes level by level, pseudo-particles bottom-up) gives the
    host build's tree: the same nodes in the same pre-order, skip links,
    member ranges and member order, and the same pseudo-particles (FP64 with
    contraction off on both sides, correctly rounded sqrt and division)."""
    for name, X, A, R, maxm in tree_cases():
        say("[gpu-vp] device tree: %s" % name)
        dev = ctx.bh_tree(X, A, R, max_members=maxm)
        host = ctx.bh_tree(X, A, R, max_members=maxm, host_build=True)
        dn, dl, do = dev.read()
        hn, hl, ho = host.read()
        assert dev.info()["n_nodes"] == host.info()["n_nodes"] and dev.info()["levels"] == host.info()["levels"]
        assert np.array_equal(dl, hl) and np.array_equal(do, ho), name
        scale = np.maximum(np.abs(hn), 1e-300)
        rel = float(np.nanmax(np.abs(dn - hn) / scale))
        bitwise = np.array_equal(np.isnan(dn), np.isnan(hn)) and np.array_equal(np.nan_to_num(dn), np.nan_to_num(hn))
        say("[gpu-vp]   %d nodes, bitwise %s, max rel diff %.1e" % (len(dn), bitwise, rel))
        assert rel <= 1e-15, (name, rel)
        for mode in (abi.CP_BH_VEL, abi.CP_BH_STRETCH):
            assert np.array_equal(dev.eval(mode), host.eval(mode)) or rel > 0
        dev.close()
        host.close()


@pytest.mark.parametrize("theta", [0.5, 0.0])
@pytest.mark.parametrize("mode", [abi.CP_BH_VEL_POINTS, abi.CP_BH_VEL, abi.CP_BH_STRETCH, abi.CP_BH_DIFF])
def test_bh_matches_oracle(ctx, mode, theta):
    x, a, r = cloud(4000, seed=21)
    pts = np.random.default_rng(22).normal(size=(700, 3)) * 0.5
    say("[gpu-vp] bh mode %d theta %.2f n %d" % (mode, theta, len(r)))
    t = ctx.bh_tree(x, a, r, theta=theta)
    got = t.eval(mode, pts)
    ref = V.BHTree(x, a, r, theta=theta).eval(mode, pts)
    ok, rel = close(got, ref)
    say("[gpu-vp]   max rel err %.2e" % rel)
    assert ok, rel
    if theta == 0.0:
        d = V.direct(mode, x, a, r, pts)
        ok, rel = close(got, d)
        assert ok, rel


def test_bh_root_leaf_and_empty(ctx):
    x, a, r = cloud(10, seed=3)
    t = ctx.bh_tree(x, a, r)
    say("[gpu-vp] root leaf")
    assert np.all(t.eval(abi.CP_BH_VEL) == 0)
    assert np.all(t.eval(abi.CP_BH_VEL_POINTS, x) == 0)
    t0 = ctx.bh_tree(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    assert t0.eval(abi.CP_BH_VEL).shape == (0, 3)


def test_bh_large_tree_sample(ctx):
    """80k particles (the C4 upper size): the traversal against the oracle on
    the same tree for a sample of targets (particle velocity)."""
    x, a, r = cloud(80000, seed=5, spread=1.0)
    say("[gpu-vp] bh 80k")
    t = ctx.bh_tree(x, a, r)
    got = t.eval(abi.CP_BH_VEL)
    ot = V.BHTree(x, a, r)
    ref = ot.eval(abi.CP_BH_VEL)
    ok, rel = close(got, ref)
    assert ok, rel


def panels_of(g, mesh):
    b = g.body()
    w = g.wake()
    tri = np.asarray(mesh["tri"])
    xyz = np.asarray(mesh["xyz"])
    return xyz[tri[b["tri_index"]]], xyz[tri[w["tri_index"]]]


def test_panel_to_points_matches_oracle(ctx):
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    g = ctx.geometry(m)
    bv, wv = panels_of(g, m)
    rng = np.random.default_rng(4)
    sigma = rng.normal(size=g.nb)
    mu = rng.normal(size=g.nb)
    muw = rng.normal(size=g.nw)
    c = g.body()["center"]
    nrm = g.body()["normal"]
    # points off the surface (near field of many panels) and far away; over
    # 4,096 points, so the near pass tiles them in their Morton order
    pts = np.vstack([c[::7] + 0.02 * nrm[::7], rng.normal(size=(4200, 3)) * 4.0])
    say("[gpu-vp] panel to points m=%d panels=%d" % (len(pts), g.nb + g.nw))
    got = g.panel_to_points(sigma, mu, pts, mu_wake=muw)
    ref = V.panels_to_points(bv, sigma, mu, wv, muw, pts)
    # per point: tolerance scaled by that point's velocity magnitude over the
    # whole set (sums of ~1,100 terms of both signs)
    ok, rel = close(got, ref, 1e-10)
    say("[gpu-vp]   max rel err %.2e" % rel)
    assert ok, rel
    # no wake strengths = zero wake
    got0 = g.panel_to_points(sigma, mu, pts)
    ref0 = V.panels_to_points(bv, sigma, mu, None, None, pts)
    assert close(got0, ref0, 1e-10)[0]
    # far-field pairs in FP32 (near field FP64): a pair's point source and
    # doublet carry a few FP32 roundings (~1e-6 relative), summed over up to
    # ~1,100 far panels of both signs
    ok, rel32 = close(g.panel_to_points(sigma, mu, pts, mu_wake=muw, fp32=True), ref, 1e-5)
    say("[gpu-vp]   fp32 far field: max rel err %.2e" % rel32)
    assert ok, rel32


@pytest.fixture(scope="module")
def vp_case(ctx):
    from cpanel_amd import CP_FP32_REFINE, CP_FP64
    from oracle import oracle as O
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 1.0)
    g = ctx.geometry(m)
    g.assemble(CP_FP64)
    g.assemble(CP_FP32_REFINE)
    og = O.Geom.from_arrays(m["xyz"], m["tri"], m["ids"])
    A, B = og.assemble()
    # second buffer-wake row: the first row moved one chord downstream
    _, wv = panels_of(g, m)
    w2 = wv + np.array([1.0, 0.0, 0.0])
    g.set_buffer_wake(w2)
    c = g.body()["center"]
    C = np.array([[O.dub_phi_inf(w2[j], c[i]) for j in range(len(w2))] for i in range(g.nb)])
    rng = np.random.default_rng(17)
    nrm = g.body()["normal"]
    vinf = np.array([np.cos(np.radians(6.0)), 0.0, np.sin(np.radians(6.0))])
    sig = np.stack([-(nrm @ (vinf + 0.05 * rng.normal(size=3))) for _ in range(3)], axis=1)
    muw2 = rng.normal(size=(len(w2), 3)) * 0.1
    rhs = -B @ sig - C @ muw2
    ref = np.linalg.solve(A, rhs)
    return g, sig, muw2, ref


@pytest.mark.parametrize("prec_name", ["fp64", "fp32_refine"])
def test_vp_solve_matches_dense(vp_case, prec_name):
    """cpCaseVP::solveVPmatrixEq: A mu = -B sigma - C mu_w2 against LAPACK on
    the oracle's A, B and C (FP64: 1e-8 relative; mixed: 1e-7)."""
    from cpanel_amd import CP_FP32_REFINE, CP_FP64
    g, sig, muw2, ref = vp_case
    prec = CP_FP64 if prec_name == "fp64" else CP_FP32_REFINE
    say("[gpu-vp] vp solve %s" % prec_name)
    mu, info = g.solve_rhs(sig, prec, mu_w2=muw2, tol=1e-12)
    for c in range(sig.shape[1]):
        rel = np.linalg.norm(mu[:, c] - ref[:, c]) / np.linalg.norm(ref[:, c])
        assert rel <= (1e-8 if prec == CP_FP64 else 1e-7), (c, rel)
        assert info[c]["rel_residual"] <= 1e-12


@pytest.mark.parametrize("prec_name", ["fp64", "fp32_refine"])
def test_vp_warm_start(vp_case, prec_name):
    """A warm start (the previous step's mu) gives the same solution in fewer
    iterations than the reference's zero start."""
    from cpanel_amd import CP_FP32_REFINE, CP_FP64
    g, sig, muw2, ref = vp_case
    prec = CP_FP64 if prec_name == "fp64" else CP_FP32_REFINE
    say("[gpu-vp] warm start %s" % prec_name)
    cold, ic = g.solve_rhs(sig[:, 0], prec, mu_w2=muw2[:, 0], tol=1e-12)
    x0 = ref[:, 0] * (1 + 1e-4 * np.random.default_rng(2).normal(size=g.nb))
    warm, iw = g.solve_rhs(sig[:, 0], prec, mu_w2=muw2[:, 0], x0=x0, tol=1e-12)
    say("[gpu-vp]   iterations cold %d warm %d" % (ic[0]["iters"], iw[0]["iters"]))
    # both solve the same system to 1e-12: they agree to cond(A) * 1e-12;
    # against the FP64 LAPACK solution the mixed path carries its 3e-8
    assert np.linalg.norm(warm - cold) / np.linalg.norm(cold) <= 1e-9
    assert np.linalg.norm(warm - ref[:, 0]) / np.linalg.norm(ref[:, 0]) <= (1e-8 if prec == CP_FP64 else 1e-7)
    assert iw[0]["iters"] < ic[0]["iters"]


def test_panel_stretching_matches_oracle(ctx):
    """Panel -> particle stretching (bodyPanel/wakePanel::partStretching summed
    over all panels) against the oracle: FP64, |out - oracle| <= 1e-8 *
    max|oracle| (near-field source gradients divide by (z r)^2 and amplify
    last-bit differences of sqrt/pow; NaNs of edges parallel to the local y
    axis, which the reference also produces, must coincide)."""
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    g = ctx.geometry(m)
    bv, wv = panels_of(g, m)
    rng = np.random.default_rng(23)
    sigma, mu, muw = rng.normal(size=g.nb), rng.normal(size=g.nb), rng.normal(size=g.nw)
    c = g.body()["center"]
    nrm = g.body()["normal"]
    pos = np.vstack([c[::9] + 0.05 * nrm[::9], rng.normal(size=(150, 3)) * 3.0,
                     np.column_stack([rng.uniform(1.0, 6.0, 100), rng.uniform(-3, 3, 100), rng.normal(0, 0.05, 100)])])
    alpha = rng.normal(size=pos.shape)
    say("[gpu-vp] panel stretching m=%d panels=%d" % (len(pos), g.nb + g.nw))
    got = g.panel_stretching(sigma, mu, pos, alpha, mu_wake=muw)
    ref = V.panels_stretching(bv, sigma, mu, wv, muw, pos, alpha)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = np.isfinite(ref).all(axis=1)
    say("[gpu-vp]   finite rows %d of %d" % (ok.sum(), len(ok)))
    assert ok.sum() > 0.5 * len(ok)
    scale = np.abs(ref[ok]).max()
    err = np.abs(got[ok] - ref[ok]).max()
    say("[gpu-vp]   max rel err %.2e" % (err / scale))
    assert err <= 1e-8 * scale
    # far-field gradients in FP32 (bound as for the FP32 panel velocity)
    got32 = g.panel_stretching(sigma, mu, pos, alpha, mu_wake=muw, fp32=True)
    assert np.array_equal(np.isnan(got32), np.isnan(ref))
    err32 = np.abs(got32[ok] - ref[ok]).max()
    say("[gpu-vp]   fp32 far field: max rel err %.2e" % (err32 / scale))
    assert err32 <= 1e-5 * scale
