"""GPU parity of the steady path through the C ABI against the FP64 oracle.

Tolerances (stated per quantity):
  * FP64 assembly: |A - A_oracle| <= 1e-12 * max|A| entry-wise (same formulas,
    FMA contraction off in the near field; differences come from atan2/log/pow
    last-bit differences only);
  * FP32 assembly: near-field entries are the FP64 value rounded once, far
    field within 4e-7 relative;
  * FP64 solve: ||mu - mu_oracle|| / ||mu_oracle|| <= 1e-8;
  * mixed FP32 + FP64 refinement: |dCL| <= 1e-5, |dCp| <= 1e-4 against FP64.
"""
import os
import sys

import numpy as np
import pytest

from cpanel_amd import CP_FP32_REFINE, CP_FP64, meshgen
from cpanel_amd.case import Context
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P4412 = dict(mach=0.1, Sref=6.0, bref=6.0, cref=1.0, cg=(0.25, 0.0, 0.0))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = Context(0)
    yield c
    # the geometries of this module's fixtures keep it alive until they go
    c.close()


def build_pair(ctx, mesh):
    og = O.Geom.from_arrays(mesh["xyz"], mesh["tri"], mesh["ids"])
    pg = ctx.geometry(mesh)
    return og, pg


@pytest.fixture(scope="module")
def c1s(ctx):
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    og, pg = build_pair(ctx, m)
    A, B = og.assemble()
    return og, pg, A, B


def test_assembly_fp64_matches_oracle(c1s):
    og, pg, A, B = c1s
    say("[gpu] assemble fp64 n=%d" % pg.nb)
    pg.assemble(CP_FP64)
    Ag, Bg = pg.matrices(CP_FP64)
    sa, sb = np.abs(A).max(), np.abs(B).max()
    assert np.abs(Ag - A).max() <= 1e-12 * sa
    assert np.abs(Bg - B).max() <= 1e-12 * sb


def test_assembly_fp32(c1s):
    og, pg, A, B = c1s
    say("[gpu] assemble fp32 n=%d" % pg.nb)
    pg.assemble(CP_FP32_REFINE)
    Ag, Bg = pg.matrices(CP_FP32_REFINE)
    assert np.abs(Ag - A).max() <= 4e-7 * np.abs(A).max()
    assert np.abs(Bg - B).max() <= 4e-7 * np.abs(B).max()
    # entry-wise relative error of the far field (FP32 arithmetic)
    rel = np.abs(Bg - B) / np.maximum(np.abs(B), 1e-30)
    assert np.median(rel) < 1e-7


def test_matvec_kernels(c1s):
    """Each matrix-vector kernel against numpy on the same matrix:
    FP64 GEMV (A64 x), FP64-accumulated MFMA (A32 x, widened exactly), FP32
    MFMA (A32 x32 with FP32 accumulation: error ~ sqrt(n) eps32 ||A|| ||x||)."""
    og, pg, A, B = c1s
    pg.assemble(CP_FP64)
    pg.assemble(CP_FP32_REFINE)
    A32, _ = pg.matrices(CP_FP32_REFINE)
    rng = np.random.default_rng(5)
    for k in (1, 7, 64):
        say("[gpu] matvec kernels k=%d" % k)
        X = rng.normal(size=(pg.nb, k))
        Y0 = pg.matvec(X, 0)
        ref = A @ X
        assert np.abs(Y0 - ref).max() <= 1e-13 * np.abs(ref).max()
        Y2 = pg.matvec(X, 2)
        ref2 = A32 @ X
        assert np.abs(Y2 - ref2).max() <= 1e-13 * np.abs(ref2).max()
        # near-field correction: entries within one panel diameter carry
        # their FP64 value, so the product moves toward A64 x
        Y3 = pg.matvec(X, 3)
        e2 = np.abs(Y2 - ref).max()
        e3 = np.abs(Y3 - ref).max()
        assert e3 <= 0.3 * e2, (e2, e3)
        Y1 = pg.matvec(X, 1)
        ref1 = A32 @ X.astype(np.float32).astype(np.float64)
        scale = np.abs(A32) @ np.abs(X)
        assert np.all(np.abs(Y1 - ref1) <= 4e-6 * scale + 1e-30)
        # FP64 MFMA product on the FP64 matrix (refinement residual of
        # CP_FP32_REFINE64)
        Y4 = pg.matvec(X, 4)
        assert np.abs(Y4 - ref).max() <= 1e-13 * np.abs(ref).max()


def test_solve_fp64_matches_oracle(c1s):
    og, pg, A, B = c1s
    pg.assemble(CP_FP64)
    conds = [(1.0, 5.0, 0.0), (1.0, -2.0, 3.0), (2.0, 8.0, -4.0)]
    say("[gpu] solve fp64 k=%d" % len(conds))
    res = pg.solve(conds, CP_FP64, tol=1e-12, **P4412)
    for (V, a, b), r in zip(conds, res):
        ref = og.run_case(A, B, V=V, alpha=a, beta=b, tol=1e-13, **P4412)
        rel = np.linalg.norm(r["mu"] - ref["mu"]) / np.linalg.norm(ref["mu"])
        assert rel <= 1e-8, rel
        assert r["rel_residual"] <= 1e-12
        # outputs on the same mu: post-processing parity
        post = og.post(r["mu"], V=V, alpha=a, beta=b, **P4412)
        assert abs(r["CL"] - post["CL"]) <= 1e-10
        assert abs(r["CDi"] - post["CDi"]) <= 1e-10
        # forces sum Cp over panels whose CHTLS fits can be ill conditioned
        # (thin tip-cap triangles): compare relative to the output scale
        fs = max(1.0, np.abs(post["Fbody"]).max())
        assert np.abs(r["Fbody"] - post["Fbody"]).max() <= 1e-9 * fs
        assert np.abs(r["Fwind"] - post["Fwind"]).max() <= 1e-9 * fs
        assert np.abs(r["CM"] - post["CM"]).max() <= 1e-9 * max(1.0, np.abs(post["CM"]).max())
        assert np.all(np.abs(r["cp"] - post["cp"]) <= 1e-9 * np.maximum(1.0, np.abs(post["cp"])))
        # and end to end
        assert abs(r["CL"] - ref["CL"]) <= 1e-8
        assert abs(r["CDi"] - ref["CDi"]) <= 1e-8


def test_4412_fixture_on_gpu(ctx, golden, fixtures_json):
    say("[gpu] 4412 fixture")
    g = ctx.geometry(os.path.join(golden, "4412.tri"))
    g.assemble(CP_FP64)
    fx = fixtures_json["4412"]
    r = g.solve([(1.0, 5.0, 0.0)], CP_FP64, tol=1e-12, **P4412)[0]
    tol = 0.5e-5 + 1e-12
    assert abs(r["CL"] - fx["CL"]) <= tol
    assert abs(r["CDi"] - fx["CDi"]) <= tol
    assert abs(r["Fwind"][2] - fx["CL_surface"]) <= tol
    assert abs(r["Fwind"][0] - fx["CD_surface"]) <= tol
    assert abs(r["CM"][1] - fx["Cm"]) <= tol
    s = np.load(os.path.join(golden, "4412_surface.npz"))
    assert np.all(np.abs(r["mu"] - s["mu"]) <= 5e-6 * np.maximum(np.abs(s["mu"]), 0.1))
    assert np.all(np.abs(r["cp"] - s["cp"]) <= 5e-6 * np.maximum(np.abs(s["cp"]), 0.1))


def close6(a, b):
    """Agreement to the 6 significant digits the reference's files carry."""
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= 5e-6 * np.maximum(np.abs(b), 1e-1) + 1e-12))


def test_4412_case_files_on_gpu(ctx, golden, tmp_path):
    """surfaceData.vtu, wakeData.vtu and spanwiseData_Wake1.csv of the 4412
    case, evaluated on the device and written in the reference's formats,
    against the arrays of the reference's own files for that case (6
    significant digits)."""
    from cpanel_amd import casefiles, meshgen
    say("[gpu] 4412 case files")
    mesh = meshgen.read_tri(os.path.join(golden, "4412.tri"))
    g = ctx.geometry(mesh)
    g.assemble(CP_FP64)
    cond = (1.0, 5.0, 0.0)
    r = g.solve([cond], CP_FP64, tol=1e-12, **P4412)[0]
    f = g.case_fields([cond], r["mu"][None, :], **P4412)[0]
    s = np.load(os.path.join(golden, "4412_surface.npz"))
    assert close6(f["mu"], s["mu"])
    assert close6(f["sigma"], s["sigma"])
    assert close6(f["potential"], s["phi"])
    assert close6(f["velocity"], s["vel"])
    assert close6(f["cp"], s["cp"])
    assert np.array_equal(f["cp"], r["cp"])  # the same post kernel as the solve
    assert close6(f["wake_mu"], np.load(os.path.join(golden, "4412_wake_mu.npy")))
    sp = np.load(os.path.join(golden, "4412_spanwise.npy"))
    assert f["spanwise"].shape == (1, 165, 3)
    assert np.all(np.abs(2 * f["spanwise"][0, :, 0] / P4412["bref"] - sp[:, 0]) <= 5e-6)
    assert close6(f["spanwise"][0, :, 1], sp[:, 1])
    assert close6(f["spanwise"][0, :, 2], sp[:, 2])
    # the Simpson sum of the station values is the case's CL
    PG = np.sqrt(1 - P4412["mach"] ** 2)
    y = f["spanwise"][0, :, 0]
    h = y[1] - y[0]
    cl = f["spanwise"][0, :, 1]
    assert abs(h / 3 * (cl[0] + 4 * cl[1:-1:2].sum() + 2 * cl[2:-1:2].sum() + cl[-1]) - r["CL"]) <= 1e-9
    # the files, read back
    d = casefiles.write_case(str(tmp_path), g, mesh, cond, P4412["mach"], f, P4412["bref"])
    assert os.path.basename(d) == "V1_Mach0.1_alpha5_beta0"
    v = casefiles.read_vtu_arrays(os.path.join(d, "surfaceData.vtu"))
    assert list(v)[:9] == ["Doublet Strengths", "Source Strengths", "Velocity Potential", "Velocity", "Cp",
                           "bezNormals", "xPosition", "yPosition", "zPostition"]
    assert close6(v["Doublet Strengths"][:, 0], s["mu"])
    assert close6(v["Velocity"], s["vel"])
    assert close6(v["Cp"][:, 0], s["cp"])
    assert close6(np.column_stack([v["xPosition"][:, 0], v["yPosition"][:, 0], v["zPostition"][:, 0]]), s["center"])
    w = casefiles.read_vtu_arrays(os.path.join(d, "wakeData.vtu"))
    assert close6(w["Doublet Strengths"][:, 0], np.load(os.path.join(golden, "4412_wake_mu.npy")))
    csv = np.loadtxt(os.path.join(d, "spanwiseData_Wake1.csv"), delimiter=",", skiprows=1)
    assert close6(csv[:, 1], sp[:, 1]) and close6(csv[:, 2], sp[:, 2])
    g.release()


def test_mixed_precision_against_fp64(c1s):
    og, pg, A, B = c1s
    pg.assemble(CP_FP64)
    pg.assemble(CP_FP32_REFINE)
    conds = [(1.0, a, b) for a in (-4.0, 2.0, 10.0) for b in (-5.0, 0.0, 5.0)]
    say("[gpu] mixed k=%d" % len(conds))
    r64 = pg.solve(conds, CP_FP64, tol=1e-12, **P4412)
    r32 = pg.solve(conds, CP_FP32_REFINE, tol=1e-10, allow_unconverged=True, **P4412)
    say("[gpu] mixed: rel_residual %s iters %s refinements %s" % (
        [float("%.2e" % b["rel_residual"]) for b in r32], [b["iters"] for b in r32], [b["refinements"] for b in r32]))
    say("[gpu] fp64 iters %s" % [a["iters"] for a in r64])
    dcl = [abs(a["CL"] - b["CL"]) for a, b in zip(r64, r32)]
    dmu = [np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"]) for a, b in zip(r64, r32)]
    dcp = [np.abs(a["cp"] - b["cp"]) for a, b in zip(r64, r32)]
    say("[gpu] mixed vs fp64: max dCL %.2e, max rel dmu %.2e, max dCp %.2e, max dCp/max(1,|Cp|) %.2e" % (
        max(dcl), max(dmu), max(d.max() for d in dcp),
        max((d / np.maximum(1.0, np.abs(a["cp"]))).max() for d, a in zip(dcp, r64))))
    for a, b, d in zip(r64, r32, dcp):
        assert b["rel_residual"] <= 1e-10
        assert abs(a["CL"] - b["CL"]) <= 1e-5
        assert abs(a["CDi"] - b["CDi"]) <= 1e-5
        # Cp: 1e-4 absolute where |Cp| <= 1, relative to |Cp| beyond (panels
        # with |Cp| >> 1 are thin tip-cap triangles whose CHTLS fits are ill
        # conditioned and amplify any change of mu)
        assert np.all(d <= 1e-4 * np.maximum(1.0, np.abs(a["cp"])))
        assert np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"]) <= 1e-7


def test_batched_column_independent_of_batch(c1s):
    og, pg, A, B = c1s
    pg.assemble(CP_FP32_REFINE)
    target = (1.0, 3.0, 1.5)
    say("[gpu] batch independence")
    alone = pg.solve([target], CP_FP32_REFINE, **P4412)[0]
    mix = pg.solve([(1.0, -3.0, 2.0), target, (1.5, 9.0, -4.0), (0.7, 0.0, 0.0)], CP_FP32_REFINE, **P4412)[1]
    assert np.array_equal(alone["mu"], mix["mu"])
    assert alone["CL"] == mix["CL"]
    again = pg.solve([target], CP_FP32_REFINE, **P4412)[0]
    assert np.array_equal(alone["mu"], again["mu"])


def test_block_jacobi_same_solution_fewer_iterations(c1s):
    """Strip block-Jacobi changes only the inner iteration count, not the
    refined solution (both satisfy ||b - A x|| / ||b|| <= 1e-10 on the same
    corrected FP32 system)."""
    og, pg, A, B = c1s
    pg.assemble(CP_FP32_REFINE)
    conds = [(1.0, a, b) for a in (-4.0, 4.0, 10.0) for b in (-5.0, 5.0)]
    say("[gpu] block jacobi vs jacobi k=%d" % len(conds))
    rj = pg.solve(conds, CP_FP32_REFINE, precond=1, **P4412)
    rb = pg.solve(conds, CP_FP32_REFINE, precond=0, **P4412)
    say("[gpu] inner iterations jacobi %s block-jacobi %s" % ([r["iters"] for r in rj], [r["iters"] for r in rb]))
    for a, b in zip(rj, rb):
        assert b["rel_residual"] <= 1e-10
        # both stop at ||r||/||b|| <= 1e-10: solutions agree to cond(A) * 1e-10
        assert np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"]) <= 1e-7
        assert abs(a["CL"] - b["CL"]) <= 1e-8
    assert sum(r["iters"] for r in rb) < sum(r["iters"] for r in rj)


def check_block_inverses(pg):
    """Every strip block-Jacobi inverse (device Gauss-Jordan in FP64, stored
    FP32) against LAPACK's FP64 inverse of the same FP32 block, rounded to
    FP32: |inv - fl32(inv_ref)| <= 4 ulp of the block's largest entry (the two
    FP64 inverses differ by rounding only; the FP32 store can then differ by
    one rounding step)."""
    A32, _ = pg.matrices(CP_FP32_REFINE)
    bj = pg.block_jacobi()
    assert bj is not None
    ptr, perm, blocks = bj
    assert sorted(perm.tolist()) == list(range(pg.nb))
    worst = 0.0
    for q, inv in enumerate(blocks):
        idx = perm[ptr[q]:ptr[q + 1]]
        ref = np.linalg.inv(A32[np.ix_(idx, idx)]).astype(np.float32)
        scale = np.abs(ref).max()
        err = np.abs(inv.astype(np.float64) - ref).max() / scale
        worst = max(worst, err)
        assert err <= 4 * 2.0 ** -23, (q, len(idx), err)
    return len(blocks), worst


def test_block_jacobi_inverses_match_lapack(c1s, c1):
    for name, pg in (("c1 small", c1s[1]), ("c1", c1[1])):
        pg.assemble(CP_FP32_REFINE)
        say("[gpu] block jacobi inverses %s n=%d" % (name, pg.nb))
        nb, worst = check_block_inverses(pg)
        say("[gpu]   %d blocks, max rel err %.2e" % (nb, worst))


@pytest.mark.parametrize("target,rows", [(2048, (1024, 2048)), (3072, (2048, 3072))])
def test_block_jacobi_blocks_above_1024_rows(ctx, monkeypatch, target, rows):
    """Blocks of more than 1024 rows (the panel kernel's two- and
    three-rows-per-thread forms): inverses against LAPACK as above, and the
    solve refined against the FP64 matrices with them equals the FP64 solve."""
    from cpanel_amd import CP_FP32_REFINE64
    monkeypatch.setenv("CPGPU_BJ_TARGET", str(target))
    m = meshgen.config_c1()
    g = ctx.geometry(m)
    g.assemble(CP_FP32_REFINE)
    g.assemble(CP_FP64)
    ptr = g.block_jacobi()[0]
    say("[gpu] block jacobi with target %d: block sizes %s" % (target, sorted(set(np.diff(ptr).tolist()))))
    assert rows[0] < np.diff(ptr).max() <= rows[1]
    nb, worst = check_block_inverses(g)
    say("[gpu]   %d blocks, max rel err %.2e" % (nb, worst))
    conds = [(1.0, 4.0, 0.0), (1.0, -3.0, 5.0)]
    rm = g.solve(conds, CP_FP32_REFINE64, tol=1e-12, **P4412)
    r64 = g.solve(conds, CP_FP64, tol=1e-12, **P4412)
    for a, b in zip(rm, r64):
        assert a["rel_residual"] <= 1e-12
        assert abs(a["CL"] - b["CL"]) <= 1e-9
    g.release()


def bf16_round(x):
    """Round-to-nearest-even of FP32 values to bf16 (kept as float64)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def near_field_mask(g, m):
    """|c_i - c_j| <= longSide_j: the entries the mixed path keeps beside its
    FP32 matrices (longSide = the panel's longest edge)."""
    b = g.body()
    V = m["xyz"][m["tri"][b["tri_index"]]]
    ls = np.max(np.linalg.norm(V - np.roll(V, 1, axis=1), axis=2), axis=1)
    c = b["center"]
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    return d2 <= ls[None, :] ** 2


def test_near_pairs_match_brute_force(ctx):
    """The near-field pairs the device lists (|c_i - c_j| <= longSide_j, the
    entries whose FP64 value the mixed path keeps) against all n^2 pairs
    tested in numpy with the same arithmetic: the same count, and the
    corrected product equals the product of A32 with A64's values on exactly
    those entries."""
    rng = np.random.default_rng(13)
    for m in (meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0), meshgen.config_c1()):
        g = ctx.geometry(m)
        g.assemble(CP_FP64)
        g.assemble(CP_FP32_REFINE)
        b = g.body()
        V = m["xyz"][m["tri"][b["tri_index"]]]
        e = V - np.roll(V, 1, axis=1)
        ls = np.sqrt((e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]) + e[..., 2] * e[..., 2]).max(axis=1)
        c = b["center"]
        d = c[:, None, :] - c[None, :, :]
        near = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]) <= ls[None, :]
        del d
        rep = g.solver_report()
        say("[gpu] near pairs n=%d: device %d, numpy %d" % (g.nb, rep["near_nnz"], near.sum()))
        assert rep["near_nnz"] == int(near.sum())
        A64, _ = g.matrices(CP_FP64)
        A32, _ = g.matrices(CP_FP32_REFINE)
        X = rng.normal(size=(g.nb, 8))
        Y3 = g.matvec(X, 3)
        ref = np.where(near, A64, A32) @ X
        assert np.abs(Y3 - ref).max() <= 1e-12 * np.abs(ref).max()
        g.release()


@pytest.mark.parametrize("shape", ["0,1,8", "0,2,16", "0,4,4", "1,4,8", "2,4,32", "3,4,16", "4,4,32", "3,4,4",
                                   "5,4,4", "5,4,32", "6,4,8", "7,4,4", "8,4,16", "9,4,32",
                                   "10,4,8", "11,4,16", "12,4,4", "10,4,32", "13,2,8", "14,2,4"])
def test_gemm_shapes(ctx, shape, monkeypatch):
    """Every GEMM variant/shape the autotuner can pick gives its product
    (FP32 accumulation bound as in test_matvec_kernels): the FP32 product for
    variants 0-9, and for the bf16x2 variants 10-12 the product of A32
    rounded to bf16 outside the packed tiles holding near-field entries and
    kept to 2^-16 inside them."""
    monkeypatch.setenv("CPGPU_GEMM_SHAPE", shape)
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    g = ctx.geometry(m)
    g.assemble(CP_FP32_REFINE)
    A32, _ = g.matrices(CP_FP32_REFINE)
    say("[gpu] gemm shape %s" % shape)
    X = np.random.default_rng(9).normal(size=(g.nb, 64))
    Y = g.matvec(X, 1)
    if int(shape.split(",")[0]) >= 10:
        near = near_field_mask(g, m)
        # the restoration works per packed 32 x 16 tile: every tile holding a
        # near-field entry carries A32 (to 2^-16), the rest bf16(A32)
        ti, tj = np.nonzero(near)
        T = np.zeros(((g.nb + 31) // 32, (g.nb + 15) // 16), bool)
        T[ti // 32, tj // 16] = True
        tiles = T[np.arange(g.nb)[:, None] // 32, np.arange(g.nb)[None, :] // 16]
        Aq = np.where(tiles, A32, bf16_round(A32))
        ref = Aq @ X.astype(np.float32).astype(np.float64)
        scale = np.abs(Aq) @ np.abs(X)
        # the multivector is split hi + lo (16 significant bits), FP32 sums
        err = np.abs(Y - ref) / (scale + 1e-300)
        say("[gpu]   max error / (|Aq||x|) = %.2e (near-field entries %d, tiles %d)" % (err.max(), near.sum(),
                                                                                  T.sum()))
        assert np.all(np.abs(Y - ref) <= 4e-5 * scale + 1e-30)
        # and it is not the FP32 product: far-field entries carry bf16 rounding
        full = A32 @ X.astype(np.float32).astype(np.float64)
        assert np.max(np.abs(Y - full) / (np.abs(A32) @ np.abs(X))) > 1e-4
        rep = g.solver_report()
        assert rep["gemm_shape"] == dict(zip(("variant", "rt", "kslices"), map(int, shape.split(","))))
        return
    ref = A32 @ X.astype(np.float32).astype(np.float64)
    scale = np.abs(A32) @ np.abs(X)
    # FP32 accumulation; the bf16x3 form also rounds each operand to 16
    # significant bits (split hi + lo) and drops lo*lo: <= 2^-15 |a x| per term
    bf16 = int(shape.split(",")[0]) >= 5
    bound = 4e-5 if bf16 else 4e-6
    err = np.abs(Y - ref) / (scale + 1e-300)
    say("[gpu]   max error / (|A||x|) = %.2e" % err.max())
    assert np.all(np.abs(Y - ref) <= bound * scale + 1e-30)
    if bf16:
        assert err.max() > 1e-9  # it really is the bf16 path
    rep = g.solver_report()
    assert rep["gemm_shape"] == dict(zip(("variant", "rt", "kslices"), map(int, shape.split(","))))


def test_bf16x2_product_same_refined_solution(ctx, monkeypatch):
    """The bf16x2 inner product (A as bf16 outside the near field) changes
    only the inner iterations: refined against the FP64 matrices, its
    solutions equal the bf16x3 product's to the refinement tolerance."""
    from cpanel_amd import CP_FP32_REFINE64
    m = meshgen.config_c1()
    conds = [(1.0, a, b) for a in (-4.0, 3.0, 10.0) for b in (-5.0, 0.0, 5.0)]
    kw = dict(mach=0.1, Sref=6.0, bref=6.0, cref=1.0, cg=(0.25, 0.0, 0.0), tol=1e-12)
    out = {}
    for shape in ("9,4,8", "10,4,8"):
        monkeypatch.setenv("CPGPU_GEMM_SHAPE", shape)
        g = ctx.geometry(m)
        g.assemble(CP_FP32_REFINE64)
        out[shape] = g.solve(conds, CP_FP32_REFINE64, **kw)
        say("[gpu] shape %s: inner iterations %s refinements %s" % (
            shape, [r["iters"] for r in out[shape]], [r["refinements"] for r in out[shape]]))
        g.release()
    for a, b in zip(out["9,4,8"], out["10,4,8"]):
        assert a["rel_residual"] <= 1e-12 and b["rel_residual"] <= 1e-12
        assert np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"]) <= 1e-10
        assert abs(a["CL"] - b["CL"]) <= 1e-10
        # bf16 rounding of the far field costs at most a modest number of
        # extra inner iterations
        assert b["iters"] <= 2 * a["iters"] + 10


@pytest.mark.parametrize("krylov_dim", ["4", "40"])
def test_gmres_inner_same_refined_solution(ctx, monkeypatch, krylov_dim):
    """The GMRES inner solver (restarted every `krylov_dim` iterations, one
    restart per refinement sweep) and the BiCGSTAB one give the same refined
    solutions: both are refined against the FP64 matrices to 1e-12."""
    from cpanel_amd import CP_FP32_REFINE64
    m = meshgen.config_c1()
    conds = [(1.0, a, b) for a in (-4.0, 3.0, 10.0) for b in (-5.0, 0.0, 5.0)]
    kw = dict(mach=0.1, Sref=6.0, bref=6.0, cref=1.0, cg=(0.25, 0.0, 0.0), tol=1e-12)
    out = {}
    for inner in ("bicgstab", "gmres"):
        monkeypatch.setenv("CPGPU_INNER", inner)
        monkeypatch.setenv("CPGPU_GMRES_M", krylov_dim)
        g = ctx.geometry(m)
        g.assemble(CP_FP32_REFINE64)
        out[inner] = g.solve(conds, CP_FP32_REFINE64, **kw)
        say("[gpu] inner %s m=%s: inner iterations %s refinements %s" % (
            inner, krylov_dim, [r["iters"] for r in out[inner]], [r["refinements"] for r in out[inner]]))
        g.release()
    for a, b in zip(out["bicgstab"], out["gmres"]):
        assert a["rel_residual"] <= 1e-12 and b["rel_residual"] <= 1e-12
        assert np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"]) <= 1e-10
        assert abs(a["CL"] - b["CL"]) <= 1e-10
        assert abs(a["CDi"] - b["CDi"]) <= 1e-10


@pytest.mark.parametrize("bj_prec", ["bf16x3", "bf16x2", "bf16x2-rt1", "fp32"])
def test_block_jacobi_mfma_matches_valu(ctx, monkeypatch, bj_prec):
    """The matrix-core block-Jacobi products (bf16 hi/lo pairs of the
    inverses, their hi plane alone, or FP32 operands) and the vector-ALU FP32
    one give the same refined solutions: the preconditioner changes only the
    inner iterations."""
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    conds = [(1.0, a, b) for a in (-4.0, 6.0) for b in (-5.0, 2.5)]
    kw = dict(mach=0.1, Sref=6.0, bref=6.0, cref=1.0, cg=(0.25, 0.0, 0.0), tol=1e-12)
    say("[gpu] block jacobi mfma (%s) vs valu" % bj_prec)
    if bj_prec.endswith("-rt1"):  # one 32-row tile per workgroup
        monkeypatch.setenv("CPGPU_BJ_RT", "1")
    monkeypatch.setenv("CPGPU_BJ_PREC", bj_prec.split("-")[0])
    g1 = ctx.geometry(m)
    g1.assemble(CP_FP32_REFINE)
    r1 = g1.solve(conds, CP_FP32_REFINE, **kw)
    say("[gpu]   inner iterations %s" % [r["iters"] for r in r1])
    monkeypatch.setenv("CPGPU_BJ_VALU", "1")
    g2 = ctx.geometry(m)
    g2.assemble(CP_FP32_REFINE)
    r2 = g2.solve(conds, CP_FP32_REFINE, **kw)
    for a, b in zip(r1, r2):
        assert a["rel_residual"] <= 1e-12 and b["rel_residual"] <= 1e-12
        assert np.linalg.norm(a["mu"] - b["mu"]) / np.linalg.norm(a["mu"]) <= 1e-9
        assert abs(a["CL"] - b["CL"]) <= 1e-10


@pytest.mark.parametrize("prec", [CP_FP64, CP_FP32_REFINE])
def test_4412ar6_stability_derivatives_on_gpu(ctx, golden, fixtures_json, prec):
    """The 4412AR6 case with its stability derivatives (3 conditions in one
    batch) against every number 4412AR6.CPout prints (5 decimals; CDi carries
    the recorded 2.7e-7 drift past its print boundary)."""
    from cpanel_amd import cpout
    fx = fixtures_json["4412AR6"]
    p = fx["params"]
    say("[gpu] 4412AR6 stability derivatives prec %d" % prec)
    g = ctx.geometry(os.path.join(golden, "4412AR6.tri"))
    g.assemble(prec)
    kw = dict(mach=p["mach"], Sref=p["Sref"], bref=p["bref"], cref=p["cref"], cg=tuple(p["cg"]), tol=1e-12)
    r, d = g.solve_with_derivatives((p["V"], p["alpha"], p["beta"]), prec, **kw)
    tol = 0.5e-5 + 1e-9
    assert abs(r["CL"] - fx["CL"]) <= tol
    assert abs(r["CDi"] - fx["CDi"]) <= 1e-5
    assert abs(r["Fwind"][2] - fx["CL_surface"]) <= tol and abs(r["Fwind"][0] - fx["CD_surface"]) <= tol
    assert abs(r["Fbody"][2] - fx["CN"]) <= tol and abs(r["Fbody"][0] - fx["CA"]) <= tol
    assert abs(r["CM"][1] - fx["Cm"]) <= tol and abs(r["CM"][0] - fx["Cl"]) <= tol
    assert abs(r["CM"][2] - fx["Cn"]) <= tol
    # derivatives: differences of 5-decimal outputs over 0.5 deg amplify the
    # print rounding by 1/0.0087: compare to 1e-4
    got = dict(CL_alpha=d["dF_dAlpha"][2], CY_alpha=d["dF_dAlpha"][1], Cm_alpha=d["dM_dAlpha"][1],
               Cl_alpha=d["dM_dAlpha"][0], Cn_alpha=d["dM_dAlpha"][2], CL_beta=d["dF_dBeta"][2],
               CY_beta=d["dF_dBeta"][1], Cm_beta=d["dM_dBeta"][1], Cl_beta=d["dM_dBeta"][0],
               Cn_beta=d["dM_dBeta"][2])
    for k, v in got.items():
        assert abs(v - fx[k]) <= 1e-4, (k, v, fx[k])
    txt = cpout.case_block(1, p["V"], p["mach"], p["alpha"], p["beta"], r, d)
    assert "CL = 0.65503" in txt


@pytest.fixture(scope="module")
def c1(ctx):
    """The plain steady case at its configured size: NACA 4412, AR 6, 80 x 30,
    flat wake 20 c (4,956 body panels)."""
    m = meshgen.config_c1()
    og, pg = build_pair(ctx, m)
    return og, pg, m


def test_c1_fp64_matches_oracle(c1):
    """configs' plain case (alpha 5, beta 0) at full size: GPU FP64 mu within
    1e-8 of the FP64 oracle (both to ||r||/||b|| <= 1e-12), CL/CDi to 1e-8."""
    og, pg, m = c1
    say("[gpu] C1 fp64 vs oracle n=%d" % pg.nb)
    A, B = og.assemble()
    pg.assemble(CP_FP64)
    r = pg.solve([(1.0, 5.0, 0.0)], CP_FP64, tol=1e-12, **P4412)[0]
    ref = og.run_case(A, B, V=1.0, alpha=5.0, beta=0.0, tol=1e-12, **P4412)
    rel = np.linalg.norm(r["mu"] - ref["mu"]) / np.linalg.norm(ref["mu"])
    say("[gpu]   rel dmu %.2e, dCL %.2e" % (rel, abs(r["CL"] - ref["CL"])))
    assert rel <= 1e-8
    assert abs(r["CL"] - ref["CL"]) <= 1e-8 and abs(r["CDi"] - ref["CDi"]) <= 1e-8


@pytest.mark.parametrize("prec_name,tol", [("refine", 1e-10), ("refine64", 1e-10), ("refine64", 1e-11)])
def test_c1_mixed_64_conditions(c1, prec_name, tol):
    """The 64-condition grid on the full-size plain case against the plain
    FP64 GPU solve (1e-12): every column within |dCL| <= 1e-5.  Cp within
    1e-4 * max(1, |Cp|): with CP_FP32_REFINE64 (refined against the FP64
    matrices) on every panel; with CP_FP32_REFINE (FP32 far field) on every
    panel outside the tip patches, whose 2-D CHTLS fits amplify mu
    perturbations of the FP32 far field (3e-8) by ~1e6."""
    from cpanel_amd import CP_FP32_REFINE64
    og, pg, m = c1
    prec = CP_FP32_REFINE64 if prec_name == "refine64" else CP_FP32_REFINE
    conds = [(1.0, float(a), float(b)) for b in (-5.0, 0.0, 2.5, 5.0) for a in np.linspace(-4, 10, 16)]
    say("[gpu] C1 mixed 64 conditions %s tol %.0e" % (prec_name, tol))
    pg.assemble(CP_FP32_REFINE64)
    r64 = pg.solve(conds, CP_FP64, tol=1e-12, **P4412)
    r32 = pg.solve(conds, prec, tol=tol, **P4412)
    tip = np.asarray(pg.clusters()[2]).astype(bool)
    keep = np.ones(pg.nb, bool) if prec == CP_FP32_REFINE64 else ~tip
    dcl = max(abs(a["CL"] - b["CL"]) for a, b in zip(r64, r32))
    rel = [np.abs(a["cp"] - b["cp"]) / np.maximum(1.0, np.abs(a["cp"])) for a, b in zip(r64, r32)]
    dcp = max(r[keep].max() for r in rel)
    dcp_all = max(r.max() for r in rel)
    say("[gpu]   max dCL %.2e, max dCp/max(1,|Cp|) %.2e (all panels %.2e), refinements %s" % (
        dcl, dcp, dcp_all, sorted(set(b["refinements"] for b in r32))))
    assert dcl <= 1e-5 and dcp <= 1e-4
    assert all(b["rel_residual"] <= tol for b in r32)


@pytest.mark.timeout(600)
def test_c3_64_conditions_full_size(ctx):
    """The multi-condition batch at full size: the 64-condition grid (alpha in
    linspace(-4, 10, 16) x beta in {-5, 0, 2.5, 5}) on the large steady mesh
    (NACA 0012 AR 10, 25,916 panels).  CP_FP32_REFINE64 (bf16 Krylov
    products, FP64 refinement) and the plain FP64 path, both to
    ||r||/||b|| <= 1e-12: every column within |dCL| <= 1e-5 and every panel's
    Cp within 1e-4 * max(1, |Cp|), tip patches included."""
    from cpanel_amd import CP_FP32_REFINE64
    m = meshgen.config_c2()
    say("[gpu] C3 full size, building geometry")
    g = ctx.geometry(m)
    g.assemble(CP_FP32_REFINE64)
    kw = dict(mach=0.1, Sref=10.0, bref=10.0, cref=1.0, cg=(0.25, 0.0, 0.0), tol=1e-12)
    conds = [(1.0, float(a), float(b)) for b in (-5.0, 0.0, 2.5, 5.0) for a in np.linspace(-4, 10, 16)]
    say("[gpu] C3 64 plain fp64 solves (n=%d)" % g.nb)
    r64 = g.solve(conds, CP_FP64, **kw)
    say("[gpu] C3 refine64, 64 conditions in one batch")
    r32 = g.solve(conds, CP_FP32_REFINE64, **kw)
    dcl = np.array([abs(a["CL"] - b["CL"]) for a, b in zip(r64, r32)])
    dcp = np.array([(np.abs(a["cp"] - b["cp"]) / np.maximum(1.0, np.abs(a["cp"]))).max() for a, b in zip(r64, r32)])
    say("[gpu]   C3 max dCL %.2e, max dCp/max(1,|Cp|) %.2e, refinements %s" % (
        dcl.max(), dcp.max(), sorted(set(b["refinements"] for b in r32))))
    g.release()
    assert np.all(dcl <= 1e-5) and np.all(dcp <= 1e-4)
    assert all(b["rel_residual"] <= 1e-12 for b in r32)


def test_load_matrices_from_infcoeff(ctx, tmp_path):
    """A and B written to a .infCoeff file and loaded back (cp_load_matrices,
    the counterpart of geometry::readInfCoeff) solve exactly as the assembled
    ones; loading 2 B doubles every solution (the solver's cached -B n
    columns follow the new B)."""
    from cpanel_amd import CP_FP32_REFINE64, infcoeff
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    conds = [(1.0, 4.0, 0.0), (1.0, -2.0, 3.0)]
    kw = dict(tol=1e-12, **P4412)
    g1 = ctx.geometry(m)
    g1.assemble(CP_FP32_REFINE64)
    A, B = g1.matrices(CP_FP64)
    r64 = g1.solve(conds, CP_FP64, **kw)
    rmx = g1.solve(conds, CP_FP32_REFINE64, **kw)
    path = tmp_path / "wing.infCoeff"
    infcoeff.write_infcoeff(path, A, B, precision=17)
    assert infcoeff.matches(path, g1.nb)
    A2, B2, _ = infcoeff.read_infcoeff(path)
    say("[gpu] load matrices n=%d" % g1.nb)
    g2 = ctx.geometry(m)
    g2.load_matrices(A2, B2)
    g2.assemble(CP_FP32_REFINE)
    for a, b in zip(g2.solve(conds, CP_FP64, **kw), r64):
        assert np.array_equal(a["mu"], b["mu"])
    for a, b in zip(g2.solve(conds, CP_FP32_REFINE64, **kw), rmx):
        assert np.array_equal(a["mu"], b["mu"])
    g2.load_matrices(A2, 2.0 * B2)
    for a, b in zip(g2.solve(conds, CP_FP32_REFINE64, **kw), rmx):
        assert np.linalg.norm(a["mu"] - 2.0 * b["mu"]) <= 1e-10 * np.linalg.norm(b["mu"])
    g1.release()
    g2.release()


def test_timing_only_records_the_named_kernel(ctx):
    """cp_timing_only: with a name set only that kernel's launches are timed;
    cleared, every timed kernel is recorded again."""
    from cpanel_amd import CP_FP32_REFINE64
    m = meshgen.wing("4412", 1.0, 6.0, 40, 12, "uniform", 20.0)
    g = ctx.geometry(m)
    g.assemble(CP_FP32_REFINE64)
    conds = [(1.0, 4.0, 0.0), (1.0, -2.0, 3.0)]
    ctx.timing_enable(True)
    try:
        ctx.timing_reset()
        ctx.timing_only("gemm_f32")
        g.solve(conds, CP_FP32_REFINE64, tol=1e-10, **P4412)
        assert ctx.timing("gemm_f32")[1] > 0
        assert ctx.timing("bj_apply")[1] == 0 and ctx.timing("resid_f64")[1] == 0
        ctx.timing_only(None)
        ctx.timing_reset()
        g.solve(conds, CP_FP32_REFINE64, tol=1e-10, **P4412)
        assert ctx.timing("gemm_f32")[1] > 0 and ctx.timing("bj_apply")[1] > 0 and ctx.timing("resid_f64")[1] > 0
    finally:
        ctx.timing_only(None)
        ctx.timing_enable(False)
        g.release()


def test_handles_close_in_any_order():
    """A geometry, a tree and a vortex-particle case keep their context (and
    the case its geometry) alive: closing the parents first must leave the
    children usable and their release clean."""
    m = meshgen.small_wing("4412", n_chord=20, n_span=6)
    c = Context(0)
    g = c.geometry(m)
    g.assemble(CP_FP64)
    t = c.bh_tree(np.random.default_rng(0).normal(size=(64, 3)), np.ones((64, 3)), np.full(64, 0.1))
    c.close()
    say("[gpu] solve after the context handle was closed")
    r = g.solve([(1.0, 5.0, 0.0)], CP_FP64, tol=1e-12, **P4412)[0]
    assert np.isfinite(r["CL"]) and r["converged"]
    assert np.all(np.isfinite(t.eval(1)))
    g.release()
    t.close()


def test_converged_flag_follows_requested_tol(c1s):
    """cp_solve reports convergence against the caller's tol (the reference's
    1e-10 flag is the default), on both paths."""
    og, pg, A, B = c1s
    pg.assemble(CP_FP64)
    pg.assemble(CP_FP32_REFINE)
    for prec in (CP_FP64, CP_FP32_REFINE):
        r = pg.solve([(1.0, 5.0, 0.0)], prec, tol=1e-6, **P4412)[0]
        assert r["converged"] and r["rel_residual"] <= 1e-6
        r = pg.solve([(1.0, 5.0, 0.0)], prec, tol=1e-12, **P4412)[0]
        assert r["converged"] and r["rel_residual"] <= 1e-12
