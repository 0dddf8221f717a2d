"""Grouped solves on the GPU (cp_solve_group): several geometries of uneven
panel count solved in one pass, every Krylov iteration stage one launch over
all of them, against each geometry's own cp_solve.

Tolerances: a geometry's mu, Cp, CL and iteration counts in the group equal
its own solve exactly (its launches inside the group are the ones its own
solve makes, the group only interleaves them with the other geometries');
CL within 1e-5 of the plain FP64 GPU solve on a sample condition.
"""
import sys

import numpy as np
import pytest

from cpanel_amd import CP_FP32_REFINE, CP_FP32_REFINE64, CP_FP64, abi, serve
from cpanel_amd.case import CPanelError, Context

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    say("[gpu-group] creating context and 6 geometries")
    ctx = Context(0)
    rng = np.random.default_rng(5)
    geoms = []
    for i in range(6):
        mesh, prm = serve.request_geometry(rng, 400, 4000)
        g = ctx.geometry(mesh, key=100 + i)
        g.assemble(CP_FP32_REFINE)
        geoms.append((g, prm, mesh))
    say("[gpu-group] body panels %s" % [g.nb for g, _, _ in geoms])
    yield ctx, geoms, rng
    ctx.close()


def conds_for(rng, k):
    return [(1.0, float(rng.uniform(-4, 10)), float(rng.uniform(-5, 5))) for _ in range(k)]


def kw_of(prm):
    return dict(mach=0.1, Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"], cg=(0.25, 0.0, 0.0))


def test_group_equals_alone(setup):
    ctx, geoms, rng = setup
    ks = [7, 1, 16, 3, 32, 12]
    conds = [conds_for(rng, k) for k in ks]
    # each geometry with its own reference lengths
    alone = []
    for (g, prm, _), c in zip(geoms, conds):
        alone.append(g.solve(c, CP_FP32_REFINE, tol=1e-10, want_cp=True, **kw_of(prm)))
    say("[gpu-group] alone done; iterations %s" % [max(r["iters"] for r in a) for a in alone])
    grp = ctx.solve_group([(g, c, kw_of(prm)) for (g, prm, _), c in zip(geoms, conds)], CP_FP32_REFINE, tol=1e-10,
                          want_cp=True)
    say("[gpu-group] group done")
    for i, (a, b) in enumerate(zip(alone, grp)):
        for q, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x["mu"], y["mu"]), (i, q)
            assert np.array_equal(x["cp"], y["cp"]), (i, q)
            assert x["CL"] == y["CL"] and x["CDi"] == y["CDi"], (i, q)
            assert x["iters"] == y["iters"] and x["refinements"] == y["refinements"], (i, q, x["iters"], y["iters"])
            assert y["converged"] and y["rel_residual"] <= 1e-10


def test_group_subsets_and_order(setup):
    """The same geometry's result does not depend on which other geometries
    share its group, nor on its slot in the group."""
    ctx, geoms, rng = setup
    c = [conds_for(rng, 5) for _ in geoms]
    kw = kw_of(geoms[1][1])
    full = ctx.solve_group([(g, cc) for (g, _, _), cc in zip(geoms, c)], CP_FP32_REFINE, **kw)
    rev = ctx.solve_group([(g, cc) for (g, _, _), cc in reversed(list(zip(geoms, c)))], CP_FP32_REFINE, **kw)[::-1]
    pair = ctx.solve_group([(geoms[4][0], c[4]), (geoms[1][0], c[1])], CP_FP32_REFINE, **kw)
    for q in range(5):
        for i in range(len(geoms)):
            assert np.array_equal(full[i][q]["mu"], rev[i][q]["mu"]), (i, q)
        assert np.array_equal(full[4][q]["mu"], pair[0][q]["mu"])
        assert np.array_equal(full[1][q]["mu"], pair[1][q]["mu"])


def test_group_refine64_and_fp64_sample(setup):
    ctx, geoms, rng = setup
    sel = [geoms[0], geoms[2], geoms[5]]
    for g, _, _ in sel:
        g.assemble(CP_FP64)
    kw = kw_of(sel[0][1])
    c = [conds_for(rng, 9) for _ in sel]
    grp = ctx.solve_group([(g, cc) for (g, _, _), cc in zip(sel, c)], CP_FP32_REFINE64, tol=1e-12, **kw)
    for (g, _, _), cc, r in zip(sel, c, grp):
        alone = g.solve(cc, CP_FP32_REFINE64, tol=1e-12, want_cp=False, **kw)
        for x, y in zip(alone, r):
            assert np.array_equal(x["mu"], y["mu"])
        ref = g.solve(cc[:2], CP_FP64, tol=1e-12, want_cp=False, **kw)
        for a, b in zip(ref, r[:2]):
            assert abs(a["CL"] - b["CL"]) <= 1e-5
            rel = np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"])
            assert rel <= 1e-8, rel


def test_group_rejects_bad_calls(setup):
    ctx, geoms, rng = setup
    g = geoms[0][0]
    c = conds_for(rng, 2)
    with pytest.raises(CPanelError):  # the same geometry twice
        ctx.solve_group([(g, c), (g, c)], CP_FP32_REFINE)
    with pytest.raises(CPanelError):  # more than 64 conditions
        ctx.solve_group([(g, conds_for(rng, 65))], CP_FP32_REFINE)
    with pytest.raises(CPanelError):  # the plain FP64 path is not grouped
        ctx.solve_group([(g, c)], CP_FP64)
    st = abi.lib().cp_solve_group(0, None, CP_FP32_REFINE, None, None, None, None, None, None)
    assert st == abi.CP_BAD_ARG
