"""Serving on the GPU: a mixed-size request stream through the native
continuous-batching server, then every request again alone.

Tolerances: a request's results in the mix equal its results alone exactly
(each column of a batched solve follows its own recurrence with fixed-order
reductions; only the batch composition differs), and its CL is within 1e-5 of
the plain FP64 GPU solve of the same condition (checked on a sample).
"""
import sys

import numpy as np
import pytest

from cpanel_amd import CP_FP64, abi, serve
from cpanel_amd.case import Context

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu-serve] creating context on device 0")
    c = Context(0)
    yield c
    # the geometries of this module's fixtures keep it alive until they go
    c.close()


def test_mix_equals_alone_and_fp64(ctx):
    stream, geoms = serve.make_stream(24, 200.0, seed=11, n_lo=400, n_hi=3000, reuse=0.4, max_cond=12)
    say("[gpu-serve] stream of %d requests, %d geometries" % (len(stream), len(geoms)))
    S = serve.Server(ctx, slots=3)
    mix, wall = serve.run_stream(S, stream, geoms, want_mu=True)
    st = S.stats()
    S.close()
    say("[gpu-serve] mix done in %.2fs: %s" % (wall, serve.summarize(mix, wall)))
    assert all(r["status"] == abi.CP_OK for r in mix)
    assert st["cache_hits"] > 0
    merged = sum(1 for r in mix if r["stats"]["batch_conditions"] > len(r["info"]))
    say("[gpu-serve] requests solved in a merged batch: %d" % merged)
    # each request alone, one at a time, on a fresh server
    A = serve.Server(ctx, slots=1)
    for i, r in enumerate(stream):
        mesh, prm = geoms[r["key"]]
        nb = int(np.sum(np.asarray(mesh["ids"]) < 1000))
        opts = serve.solve_opts(Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"])
        rid = A.submit(r["key"], mesh, r["conds"], opts, want_mu=True)
        s, info, mu, rs = A.wait(rid, nb)
        assert s == abi.CP_OK
        assert rs["batch_conditions"] == len(r["conds"])
        for a, b in zip(info, mix[i]["info"]):
            assert a["CL"] == b["CL"] and a["CDi"] == b["CDi"], (i, a["CL"], b["CL"])
        assert np.array_equal(mu, mix[i]["mu"]), i
    A.close()
    # sample against the plain FP64 path
    for i in (0, len(stream) // 2, len(stream) - 1):
        r = stream[i]
        mesh, prm = geoms[r["key"]]
        g = ctx.geometry(mesh)
        g.assemble(CP_FP64)
        ref = g.solve(r["conds"][:4], CP_FP64, mach=0.1, Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"],
                      cg=(0.25, 0.0, 0.0), tol=1e-12)
        for a, b in zip(ref, mix[i]["info"][:4]):
            assert abs(a["CL"] - b["CL"]) <= 1e-5


def test_burst_groups_equal_alone(ctx, monkeypatch):
    """A burst of requests over resident small geometries: one worker solves
    several geometries' requests together (cp_solve_group), and every result
    equals the same request solved alone with grouping off."""
    stream, geoms = serve.make_stream(40, 1e6, seed=21, n_lo=400, n_hi=2500, reuse=0.8, max_cond=6)
    S = serve.Server(ctx, slots=1)
    serve.run_stream(S, stream, geoms)  # builds every geometry
    burst, wall = serve.run_stream(S, stream, geoms, want_mu=True)  # all resident, all queued at once
    st = S.stats()
    S.close()
    grouped = sum(1 for r in burst if r["stats"]["group_geometries"] > 1)
    say("[gpu-serve] burst: %d grouped solves, %d of %d requests in a group" % (st["grouped_solves"], grouped,
                                                                                 len(burst)))
    assert all(r["status"] == abi.CP_OK for r in burst)
    assert st["grouped_solves"] > 0 and grouped > 0
    monkeypatch.setenv("CPGPU_SERVE_GROUP", "1")
    A = serve.Server(ctx, slots=1)
    for i, r in enumerate(stream):
        mesh, prm = geoms[r["key"]]
        nb = int(np.sum(np.asarray(mesh["ids"]) < 1000))
        opts = serve.solve_opts(Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"])
        rid = A.submit(r["key"], mesh, r["conds"], opts, want_mu=True)
        s, info, mu, rs = A.wait(rid, nb)
        assert s == abi.CP_OK and rs["group_geometries"] == 1
        for a, b in zip(info, burst[i]["info"]):
            assert a["CL"] == b["CL"] and a["iters"] == b["iters"], (i, a["CL"], b["CL"])
        assert np.array_equal(mu, burst[i]["mu"]), i
    A.close()


def test_budget_evicts_and_results_hold(ctx):
    stream, geoms = serve.make_stream(12, 1000.0, seed=5, n_lo=800, n_hi=1500, reuse=0.3, max_cond=6)
    say("[gpu-serve] budget stream")
    S = serve.Server(ctx, slots=2, hbm_budget=int(5e8))
    res, wall = serve.run_stream(S, stream, geoms)
    st = S.stats()
    S.close()
    say("[gpu-serve] budget stats %s" % st)
    assert all(r["status"] == abi.CP_OK for r in res)
    assert st["evictions"] > 0
    assert st["resident_bytes_est"] <= int(5e8)


@pytest.mark.timeout(600)
def test_block_jacobi_escalates_on_stalled_sweeps():
    """A geometry of the serving stream (NACA 0009, AR 5, 126 x 78, 19,904
    panels) on which GMRES with the bf16 hi plane of the strip block-Jacobi
    inverses stagnates (the bf16 rounding of the ill-conditioned tip block's
    inverse spoils the preconditioner): the refinement detects the stalled
    sweep, reads the inverses' lo plane (then falls back to point Jacobi if
    that stalls too), and every condition converges to the FP64 solution."""
    from cpanel_amd import CP_FP32_REFINE, CP_FP64, serve
    from cpanel_amd.case import Context
    stream, geoms = serve.make_stream(1132, 20.0, seed=0, n_lo=500, n_hi=20000, reuse=0.3, max_cond=32)
    r = stream[1131]
    mesh, prm = geoms[r["key"]]
    ctx = Context(0)
    g = ctx.geometry(mesh)
    g.assemble(CP_FP32_REFINE)
    kw = dict(mach=0.1, Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"], cg=(0.25, 0.0, 0.0), tol=1e-10)
    res = g.solve(r["conds"], CP_FP32_REFINE, want_cp=False, **kw)
    rep = g.solver_report()
    print("[gpu] n=%d refinements %s, block-Jacobi level %s" % (g.nb, sorted(set(x["refinements"] for x in res)),
                                                               rep.get("block_jacobi_level")), flush=True)
    assert all(x["converged"] and x["rel_residual"] <= 1e-10 for x in res)
    assert rep.get("block_jacobi_level", 0) >= 1
    g.assemble(CP_FP64)
    r64 = g.solve(r["conds"][:3], CP_FP64, want_cp=False, **kw)
    for a, b in zip(res, r64):
        assert abs(a["CL"] - b["CL"]) <= 1e-6
    g.release()
    ctx.close()
