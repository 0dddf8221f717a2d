"""Benchmark of the CPanel solve hot path on one MI355X.

Workload (the large steady case with its multi-condition batch): NACA 0012
rectangular wing, aspect ratio 10, 160 cosine chordwise x 80 spanwise panels
(25,916 body triangles), 30-chord prescribed wake, 64 freestream conditions
alpha in linspace(-4, 10, 16) x beta in {-5, 0, 2.5, 5} deg, |V| = 1, M = 0.1.

One step = solve all 64 conditions on the cached geometry (A, B resident in
HBM) with FP32-class Krylov iterations + FP64 refinement to ||r||/||b|| <= tol
(1e-12 by default, against the FP64 matrices), plus per-condition
post-processing (Cp, forces, Trefftz CL/CDi).  A solution counts only if its
Trefftz CL is within 1e-5 of the plain FP64 GPU solution of the same
condition AND every panel's Cp is within 1e-4 * max(1, |Cp|) of it; the
comparator is the plain FP64 path at tol 1e-12 (checked against the FP64 CPU
oracle by the test suite and, for one condition, in the cpu_baseline leg),
solved once outside the timed loop and timed as the plain FP64 leg.

Prints one JSON line (rank 0).  --gpus N > 1 is refused: this build runs on
one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBS = 8000.0
MI355X_FP32_MFMA_TFLOPS = 157.3
MI355X_FP64_TFLOPS = 78.6


def conditions():
    return [(1.0, float(a), float(b)) for b in (-5.0, 0.0, 2.5, 5.0) for a in np.linspace(-4, 10, 16)]


def gemm_kernel_name(shape):
    """Kernel symbol of the autotuned FP32 GEMM shape (as the profiler names it)."""
    if shape is None:
        return None
    if shape["variant"] == 1:
        return "gemm_f32_x4_kernel<8>"
    if shape["variant"] == 2:
        return "gemm_f32_x4_kernel<4>"
    if shape["variant"] == 3:
        return "gemm_f32_x4s3_kernel<4>"
    if shape["variant"] == 4:
        return "gemm_f32_x4s3_kernel<2>"
    if shape["variant"] == 5:
        return "gemm_bf16x3_kernel"
    if shape["variant"] in (6, 7, 8, 9):
        return "gemm_bf16x3_s_kernel<%d, %d>" % {6: (2, 3), 7: (2, 4), 8: (4, 3), 9: (4, 2)}[shape["variant"]]
    if shape["variant"] in (10, 11, 12, 13, 14):
        return "gemm_bf16x2_s_kernel<%d, %d>" % {10: (4, 2), 11: (4, 3), 12: (4, 4), 13: (2, 2),
                                                 14: (2, 3)}[shape["variant"]]
    return "gemm_f32_kernel<%d>" % shape["rt"]


def profiled_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", workload + ".traffic.json")
    try:
        with open(path) as f:
            ks = json.load(f)["kernels"]
    except (OSError, ValueError):
        return None
    k = ks.get(kernel)
    for fam in ("gemm_bf16x3", "gemm_bf16x2"):
        if k is None and kernel and kernel.startswith(fam):
            # the tilings of one family differ only in pipeline depth and
            # row tiles: the same bytes
            k = next((v for name, v in ks.items() if name.startswith(fam)), None)
    return None if k is None else k["read_bytes"] + k["write_bytes"]


def pair_flops():
    """Per-pair flop counts (constexpr int kName = N) of the kernels'
    flop-count header."""
    import re
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpanel_amd", "csrc", "flopcount.h")) as f:
        txt = f.read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"constexpr int (k\w+) = (\d+);", txt)}


def vp_case_bench(ctx, log, steps, faithful=False):
    """The unsteady vortex-particle case on the swept tapered 4412 wing
    (6,116 tris, 50 trailing-edge edges) from the impulsive start at alpha
    6 deg, dt 0.02: the configured form (algebraic core 0.2, 4 particles per
    edge per step, 32 per leaf, FP32 particle and far-field panel sums) or the reference's (Gaussian kernels with
    diffusion, one particle per edge per step, 10 per leaf, FP64).  Warm
    starts, solves refined to 1e-12 against the FP64 matrices."""
    from cpanel_amd import CP_FP32_REFINE64, abi, meshgen
    g = ctx.geometry_vp(meshgen.config_c4(with_wake=False), V=1.0, dt=0.02)
    # core 0.2: the specified 0.1 x mean panel size (0.0043) leaves the
    # particles without overlap and the run diverges (FP64 and FP32) within
    # 40-70 steps; 0.2 is the smallest core probed that stays bounded over
    # 400 steps
    kw = dict(max_members=10) if faithful else dict(kernel=abi.CP_VP_ALGEBRAIC, per_edge=4, fp32=True,
                                                    max_members=32, core=0.2)
    case = g.vp_case((1.0, 6.0, 0.0), prec=CP_FP32_REFINE64, mach=0.1, Sref=6.0, bref=8.0, cref=0.75,
                     accelerate=True, warm_start=True, theta=0.5, tol=1e-12, **kw)
    names = ("bh_eval", "panel_to_points", "panel_stretching")
    ctx.timing_reset()
    infos = []
    t0 = time.perf_counter()
    for k in range(steps):
        infos.append(case.step())
        if (k + 1) % 50 == 0:
            i = infos[-1]
            log("vp case (%s): step %d, %d particles, %.1f ms, CL_trefftz %.5f" %
                ("reference" if faithful else "configured", i["step"], i["n_particles"], i["ms"], i["CL"]))
    wall = time.perf_counter() - t0
    fin = case.finish()
    phases = {nm: ctx.timing(nm)[0] for nm in names}
    p = case.particles()
    case.close()
    g.release()
    last = infos[-10:]
    return {"form": "reference" if faithful else "configured", "core": None if faithful else 0.2,
            "steps": steps, "wall_s": wall,
            "ms_per_step_mean": 1e3 * wall / steps, "ms_per_step_last10": float(np.mean([i["ms"] for i in last])),
            "particles": infos[-1]["n_particles"], "CL_trefftz": infos[-1]["CL"], "CDi_trefftz": infos[-1]["CDi"],
            "CL": fin["CL"], "CDi": fin["CDi"], "max_rel_residual": max(i["rel_residual"] for i in infos),
            "solve_iters_last": infos[-1]["iters"], "finite": bool(np.isfinite(p["pos"]).all() and
                                                                     np.isfinite(p["strength"]).all()),
            "phase_ms_total": phases}


def vp_bench(ctx, log, case_steps=0, faithful_steps=0):
    """Vortex-particle kernels on the swept tapered 4412 case: 80,000
    particles (4 per trailing-edge edge per step over 400 steps) and 6,116
    body panels.  Barnes-Hut (theta 0.5, 10 per leaf) in each mode, the
    configured form's algebraic-core tree (32 per leaf) in FP64 and FP32,
    and the panel -> particle velocity and stretching of every particle
    against every panel in FP64 and with the far field in FP32; then the
    case itself grown over case_steps (configured) and faithful_steps
    (reference form) steps."""
    from cpanel_amd import abi, meshgen
    mesh = meshgen.config_c4()
    X, A, R = meshgen.c4_particles(mesh)
    out = {"particles": int(len(R))}
    ctx.bh_tree(X, A, R).close()  # first build: the pool's allocations
    ctx.timing_reset()
    t0 = time.perf_counter()
    tree = ctx.bh_tree(X, A, R)
    out["bh_build_wall_ms"] = 1e3 * (time.perf_counter() - t0)  # with the upload of the particles
    out["bh_build_ms"] = ctx.timing("bh_build")[0]  # the device build (events on its stream)
    out["bh_nodes"] = tree.info()["n_nodes"]
    names = {abi.CP_BH_VEL: "vel", abi.CP_BH_STRETCH: "stretch", abi.CP_BH_DIFF: "diff"}
    for mode, nm in names.items():
        tree.eval(mode)  # warm
        ctx.timing_reset()
        for _ in range(3):
            tree.eval(mode)
        ms, n = ctx.timing("bh_eval")
        out["bh_%s_ms" % nm] = ms / max(n, 1)
    tree.close()
    for f32 in (False, True):
        tree = ctx.bh_tree(X, A, R, max_members=32, theta=0.5, kernel=abi.CP_VP_ALGEBRAIC, core=float(R[0]),
                           fp32=f32)
        for mode, nm in ((abi.CP_BH_VEL, "vel"), (abi.CP_BH_STRETCH, "stretch")):
            tree.eval(mode)
            ctx.timing_reset()
            for _ in range(3):
                tree.eval(mode)
            ms, n = ctx.timing("bh_eval")
            out["alg_core_%s_%s_ms" % (nm, "fp32" if f32 else "fp64")] = ms / max(n, 1)
        tree.close()
    fl = pair_flops()
    g = ctx.geometry(mesh)
    rng = np.random.default_rng(1)
    sig, mu = rng.normal(size=g.nb), rng.normal(size=g.nb)
    g.panel_to_points(sig, mu, X[:1000])
    ctx.timing_reset()
    g.panel_to_points(sig, mu, X, mu_wake=rng.normal(size=g.nw))
    ms, n = ctx.timing("panel_to_points")
    out["panel_to_points_ms"] = ms
    out["panel_to_points_gpairs_per_s"] = len(R) * (g.nb + g.nw) / (ms * 1e-3) / 1e9
    g.panel_stretching(sig, mu, X[:1000], A[:1000])
    ctx.timing_reset()
    g.panel_stretching(sig, mu, X, A, mu_wake=rng.normal(size=g.nw))
    ms, n = ctx.timing("panel_stretching")
    out["panel_stretching_ms"] = ms
    out["panel_stretching_gpairs_per_s"] = len(R) * (g.nb + g.nw) / (ms * 1e-3) / 1e9
    # far field in FP32 (near field FP64); the flop rate counts every pair
    # at the far-field count, a lower bound (near pairs cost ~13x more)
    muw = rng.normal(size=g.nw)
    g.panel_to_points(sig, mu, X[:1000], fp32=True)
    ctx.timing_reset()
    g.panel_to_points(sig, mu, X, mu_wake=muw, fp32=True)
    ms, n = ctx.timing("panel_to_points")
    pairs = len(R) * (g.nb + g.nw)
    out["panel_to_points_fp32_ms"] = ms
    out["panel_to_points_fp32_gpairs_per_s"] = pairs / (ms * 1e-3) / 1e9
    tf = pairs * fl["kP2PFarPairF32"] / (ms * 1e-3) / 1e12
    out["panel_to_points_fp32_tflops_lower_bound"] = tf
    out["panel_to_points_fp32_frac_of_fp32_valu_peak"] = tf / MI355X_FP32_MFMA_TFLOPS
    tf64 = pairs * fl["kP2PFarPair"] / (out["panel_to_points_ms"] * 1e-3) / 1e12
    out["panel_to_points_fp64_tflops_lower_bound"] = tf64
    out["panel_to_points_fp64_frac_of_fp64_valu_peak"] = tf64 / MI355X_FP64_TFLOPS
    g.panel_stretching(sig, mu, X[:1000], A[:1000], fp32=True)
    ctx.timing_reset()
    g.panel_stretching(sig, mu, X, A, mu_wake=muw, fp32=True)
    ms, n = ctx.timing("panel_stretching")
    out["panel_stretching_fp32_ms"] = ms
    out["panel_stretching_fp32_gpairs_per_s"] = pairs / (ms * 1e-3) / 1e9
    g.release()
    log("vp: %s" % json.dumps(out))
    if case_steps > 0:
        out["case_configured"] = vp_case_bench(ctx, log, case_steps)
        log("vp case: %s" % json.dumps(out["case_configured"]))
    if faithful_steps > 0:
        out["case_reference_form"] = vp_case_bench(ctx, log, faithful_steps, faithful=True)
        log("vp case: %s" % json.dumps(out["case_reference_form"]))
    return out


def serve_bench(ctx, n_req, qps, slots, seed, log):
    """Mixed-size request stream through the continuous-batching server:
    Poisson arrivals, NACA wings of 500..20,000 body panels (log-uniform),
    1..32 conditions each, 30 % geometry reuse, 200 GB HBM budget."""
    from cpanel_amd import serve
    stream, geoms = serve.make_stream(n_req, qps, seed=seed, n_lo=500, n_hi=20000, reuse=0.3, max_cond=32)
    S = serve.Server(ctx, slots=slots, hbm_budget=200 << 30)
    res, wall = serve.run_stream(S, stream, geoms)
    st = S.stats()
    S.close()
    summ = serve.summarize(res, wall)
    summ.update(dict(qps=qps, slots=slots, seed=seed, geometries=len(geoms), evictions=st["evictions"],
                     batches=st["batches"],
                     mean_prepare_ms=1e3 * float(np.mean([r["stats"]["prepare_s"] for r in res
                                                          if not r["stats"]["cache_hit"]] or [0])),
                     mean_solve_ms=1e3 * float(np.mean([r["stats"]["solve_s"] for r in res]))))
    log("serving: %s" % json.dumps(summ))
    return summ


def group_bench(ctx, log, seed=3, reps=3):
    """Windows of small geometries (uneven panel counts, the serving stream's
    generator) solved one after another with cp_solve and together with
    cp_solve_group: wall ms per window, geometries cached, CP_FP32_REFINE
    at tol 1e-10 as in serving; the grouped results equal the single ones
    bit for bit (checked)."""
    from cpanel_amd import CP_FP32_REFINE, serve
    rng = np.random.default_rng(seed)
    out = []
    for G, lo, hi, k in ((8, 500, 3000, 16), (16, 500, 1500, 8), (4, 4000, 8000, 32)):
        geoms = []
        for _ in range(G):
            mesh, prm = serve.request_geometry(rng, lo, hi)
            g = ctx.geometry(mesh)
            g.assemble(CP_FP32_REFINE)
            geoms.append((g, dict(mach=0.1, Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"], cg=(0.25, 0, 0))))
        conds = [[(1.0, float(rng.uniform(-4, 10)), float(rng.uniform(-5, 5))) for _ in range(k)] for _ in range(G)]
        jobs = [(g, c, kw) for (g, kw), c in zip(geoms, conds)]
        ref = [g.solve(c, CP_FP32_REFINE, want_cp=False, **kw) for (g, kw), c in zip(geoms, conds)]  # warm
        grp = ctx.solve_group(jobs, CP_FP32_REFINE, want_cp=False)
        same = all(np.array_equal(a["mu"], b["mu"]) for ra, rb in zip(ref, grp) for a, b in zip(ra, rb))
        ts, tg = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            for (g, kw), c in zip(geoms, conds):
                g.solve(c, CP_FP32_REFINE, want_mu=False, want_cp=False, **kw)
            t1 = time.perf_counter()
            ctx.solve_group(jobs, CP_FP32_REFINE, want_mu=False, want_cp=False)
            t2 = time.perf_counter()
            ts.append(t1 - t0)
            tg.append(t2 - t1)
        r = {"geometries": G, "n_body": [g.nb for g, _ in geoms], "conditions_each": k,
             "sequential_ms": 1e3 * min(ts), "grouped_ms": 1e3 * min(tg), "speedup": min(ts) / min(tg),
             "solves_per_s_grouped": G * k / min(tg), "bit_identical": bool(same)}
        log("grouped solves: %s" % json.dumps(r))
        out.append(r)
        for g, _ in geoms:
            g.release()
    return out


def serve_hot_bench(ctx, log, n_req=2000, n_geoms=100, qps=1000.0, seed=4):
    """A hot-cache burst through the server: 2,000 requests at 1,000/s over
    100 resident geometries of 500..4,000 body panels (log-uniform), 1..16
    conditions each, four slots; with grouping of small geometries (the
    default) and with it turned off (CPGPU_SERVE_GROUP=1)."""
    from cpanel_amd import serve
    rng = np.random.default_rng(seed)
    geoms = [serve.request_geometry(rng, 500, 4000) for _ in range(n_geoms)]
    t = 0.0
    stream = []
    for _ in range(n_req):
        t += rng.exponential(1.0 / qps)
        key = int(rng.integers(0, n_geoms))
        k = int(rng.integers(1, 17))
        stream.append(dict(t=t, key=key, conds=[(1.0, float(rng.uniform(-4, 10)), float(rng.uniform(-5, 5)))
                                                for _ in range(k)]))
    warm = [dict(t=0.0, key=i, conds=[(1.0, 2.0, 0.0)]) for i in range(n_geoms)]
    out = {}
    for label, grp in (("grouped", None), ("alone", "1")):
        if grp:
            os.environ["CPGPU_SERVE_GROUP"] = grp
        try:
            S = serve.Server(ctx, slots=4)
            serve.run_stream(S, warm, geoms)  # every geometry resident
            st0 = S.stats()
            res, wall = serve.run_stream(S, stream, geoms)
            st = S.stats()
            S.close()
        finally:
            os.environ.pop("CPGPU_SERVE_GROUP", None)
        summ = serve.summarize(res, wall)
        summ.update(dict(grouped_solves=st["grouped_solves"] - st0["grouped_solves"],
                         requests_in_groups=sum(1 for r in res if r["stats"]["group_geometries"] > 1)))
        out[label] = summ
        log("serving, hot cache (%s): %s" % (label, json.dumps(summ)))
    out["workload"] = ("%d requests at %.0f/s over %d resident geometries of 500..4,000 panels, 1..16 conditions, "
                       "4 slots, seed %d" % (n_req, qps, n_geoms, seed))
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Host CPU description: model string, logical CPUs, the process's CPU set."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity}


def cpu_baseline(mesh, kw, cond, nthreads, tol, log):
    """The FP64 oracle (the reference algorithm restated in C++, built for this
    host with -march=native when a compiler is present) on the host cores.

    * multi-threaded (nthreads): full assembly, then one condition solved to
      tol on the full matrix (the BiCGSTAB GEMVs, the RHS product and the
      post-processing run on nthreads);
    * one thread, the reference's own configuration (it is single-threaded):
      a 512-column block of the assembly and BiCGSTAB runs capped at 2 and 6
      iterations, extrapolated linearly to the full assembly and to the
      multi-threaded run's iteration count."""
    import tempfile
    from oracle import oracle as O
    native = O.build_native(tempfile.mkdtemp(prefix="cpanel_oracle_"))
    if native:
        O.use_library(native)
    og = O.Geom.from_arrays(mesh["xyz"], mesh["tri"], mesh["ids"])
    n = og.nb
    t0 = time.perf_counter()
    A, B = og.assemble(nthreads=nthreads)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = og.run_case(A, B, V=cond[0], alpha=cond[1], beta=cond[2], tol=tol, nthreads=nthreads, **kw)
    t_solve = time.perf_counter() - t0
    log("cpu baseline: %d threads, assembly %.2fs, solve %.2fs (%d iterations)" % (nthreads, t_asm, t_solve,
                                                                                   r["iters"]))
    cols = min(512, n)
    t0 = time.perf_counter()
    og.assemble_cols(0, cols, nthreads=1)
    t_cols = time.perf_counter() - t0
    t0 = time.perf_counter()
    og.run_case(A, B, V=cond[0], alpha=cond[1], beta=cond[2], tol=tol, maxit=2, nthreads=1, **kw)
    t2 = time.perf_counter() - t0
    t0 = time.perf_counter()
    og.run_case(A, B, V=cond[0], alpha=cond[1], beta=cond[2], tol=tol, maxit=6, nthreads=1, **kw)
    t6 = time.perf_counter() - t0
    per_it = (t6 - t2) / 4
    t1_solve = t2 + per_it * max(r["iters"] - 2, 0)
    t1_asm = t_cols * n / cols
    log("cpu baseline, one thread: %.2fs per iteration -> %.1fs per solve (extrapolated), assembly %.1fs "
        "(extrapolated from %d columns)" % (per_it, t1_solve, t1_asm, cols))
    return dict(t_assemble=t_asm, t_solve=t_solve, CL=r["CL"], mu=r["mu"], cp=r["cp"], iters=r["iters"],
                native=bool(native), t1_solve=t1_solve, t1_assemble=t1_asm, t1_per_iter=per_it, cols=cols)


def validity(ref, res, tip):
    """Per condition: |dCL| <= 1e-5 and |dCp| <= 1e-4 * max(1, |Cp|) on every
    panel, against the comparator `ref` (the plain FP64 solutions)."""
    dcl = np.array([abs(a["CL"] - b["CL"]) for a, b in zip(ref, res)])
    rel = [np.abs(a["cp"] - b["cp"]) / np.maximum(1.0, np.abs(a["cp"])) for a, b in zip(ref, res)]
    dcp = np.array([r.max() for r in rel])
    dcp_body = np.array([r[~tip].max() for r in rel])
    ok = (dcl <= 1e-5) & (dcp <= 1e-4)
    return ok, dcl, dcp, dcp_body


def mixed_steps(g, conds, mprec, tol, inner_tol, kw, warmup, steps):
    for _ in range(warmup):
        g.solve(conds, mprec, tol=tol, inner_tol=inner_tol, want_mu=False, want_cp=False, **kw)
    times = []
    res = None
    for _ in range(steps):
        t0 = time.perf_counter()
        res = g.solve(conds, mprec, tol=tol, inner_tol=inner_tol, want_mu=False, want_cp=False, **kw)
        times.append(time.perf_counter() - t0)
    return float(np.mean(times)), res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="C1-size mesh (quick check)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison path")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0: 16, the host cores one GPU of the box is given)")
    ap.add_argument("--tol", type=float, default=1e-12, help="refinement target ||r||/||b|| of the mixed path")
    ap.add_argument("--cmp-tol", type=float, default=1e-12, help="tol of the plain FP64 comparator leg")
    ap.add_argument("--inner-tol", type=float, default=0.0, help="inner FP32 solve reduction per sweep (0: default)")
    ap.add_argument("--prec", choices=("refine64", "refine"), default="refine64",
                    help="refine64: refinement against the FP64 matrices; refine: FP32 far field + FP64 near field")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests in the serving stream (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=20.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-seed", type=int, default=0)
    ap.add_argument("--no-vp", action="store_true", help="skip the vortex-particle kernels")
    ap.add_argument("--vp-steps", type=int, default=400,
                    help="steps of the configured vortex-particle case (0: skip)")
    ap.add_argument("--vp-reference-steps", type=int, default=100,
                    help="steps of the vortex-particle case in the reference's form (0: skip)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the amortised-setup and FP32-operand measurements")
    ap.add_argument("--group-only", action="store_true",
                    help="run only the grouped-solve windows (profiling of the grouped kernels)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one elementwise copy of 256 MiB first (HBM counter calibration under a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from cpanel_amd import CP_FP32_REFINE, CP_FP32_REFINE64, CP_FP64, meshgen
    from cpanel_amd.case import Context

    mesh = meshgen.config_c1() if args.small else meshgen.config_c2()
    span = 6.0 if args.small else 10.0
    kw = dict(mach=0.1, Sref=span * 1.0, bref=span, cref=1.0, cg=(0.25, 0.0, 0.0))
    conds = conditions()
    k = len(conds)

    if args.copy_calib:
        # reads 256 MiB and writes 256 MiB: the FETCH_SIZE / WRITE_SIZE scale of a plain stream
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        torch.neg(x, out=y)
        torch.cuda.synchronize()
        del x, y

    ctx = Context(0)
    if args.group_only:
        print(json.dumps({"grouped_solves": group_bench(ctx, log)}), flush=True)
        return
    t0 = time.perf_counter()
    g = ctx.geometry(mesh)
    t_geom = time.perf_counter() - t0
    n = g.nb
    log("geometry: %d body panels, %d wake panels (%.2fs host topology+upload)" % (n, g.nw, t_geom))

    ctx.timing_enable(True)
    ctx.timing_reset()
    g.assemble(CP_FP64)
    g.assemble(CP_FP32_REFINE)
    ctx.timing_reset()
    g.assemble(CP_FP64)
    g.assemble(CP_FP32_REFINE)
    t_a64, _ = ctx.timing("assemble_f64")
    t_a32, _ = ctx.timing("assemble_f32")
    asm_parts = {nm: ctx.timing(nm)[0] for nm in ("assemble_far_f64", "assemble_near_f64", "assemble_far_f32",
                                                  "assemble_near_f32", "near_f64")}
    pairs = float(n) * n
    log("assembly: fp64 %.2f ms (%.1f Ginfl/s), fp32 %.2f ms (%.1f Ginfl/s); passes (ms) %s" %
        (t_a64, pairs / t_a64 / 1e6, t_a32, pairs / t_a32 / 1e6, json.dumps(asm_parts)))
    tip = g.clusters()[2].astype(bool)

    # plain FP64 GPU path, every condition solved alone (reference
    # semantics): the comparator of the validity check and the leg the
    # speed-up is quoted against, both at cmp_tol
    fp64 = None
    t_fp64 = None
    if not args.no_fp64:
        ctx.timing_reset()
        t0 = time.perf_counter()
        fp64 = g.solve(conds, CP_FP64, tol=args.cmp_tol, **kw)
        t_fp64 = time.perf_counter() - t0
        gm_ms, gm_n = ctx.timing("gemv_f64")
        log("plain fp64 (tol %.0e): %d solves in %.3fs (%.1f ms/solve), gemv avg %.3f ms x %d = %.0f GB/s" %
            (args.cmp_tol, k, t_fp64, 1e3 * t_fp64 / k, gm_ms / max(gm_n, 1), gm_n,
             8.0 * n * n / (gm_ms / max(gm_n, 1)) / 1e6))

    # mixed-precision batched path; timed steps leave mu and Cp in HBM (CL,
    # CDi and the forces of every condition come back) and carry HIP events
    # on the dominant kernel's launches only (the roofline's duration)
    mprec = CP_FP32_REFINE64 if args.prec == "refine64" else CP_FP32_REFINE
    g.solve(conds, mprec, tol=args.tol, inner_tol=args.inner_tol, want_mu=False, want_cp=False, **kw)
    ctx.timing_reset()
    ctx.timing_only("gemm_f32")
    t_step, res = mixed_steps(g, conds, mprec, args.tol, args.inner_tol, kw, max(args.warmup - 1, 0), args.steps)
    mm_ms, mm_n = ctx.timing("gemm_f32")
    # the same step with mu and Cp of all 64 conditions copied to host memory
    # (PCIe-inclusive; reported beside value, never as value): its arrays are
    # the ones the validity check compares (every column's result is
    # bit-identical run to run)
    t0 = time.perf_counter()
    res_full = g.solve(conds, mprec, tol=args.tol, inner_tol=args.inner_tol, **kw)
    t_step_host = time.perf_counter() - t0
    # one more step outside both measurements with every kernel timed (log only)
    ctx.timing_only(None)
    ctx.timing_reset()
    g.solve(conds, mprec, tol=args.tol, inner_tol=args.inner_tol, want_mu=False, want_cp=False, **kw)
    rs_ms, rs_n = ctx.timing("resid_f64")
    bj_ms, bj_n = ctx.timing("bj_apply")
    gemm_avg = mm_ms / max(mm_n, 1)
    gemm_flops = 2.0 * n * n * 64
    achieved = gemm_flops / (gemm_avg * 1e-3) / 1e12
    log("mixed batched (tol %.0e): %.1f ms/step (%d conditions); gemm_f32 avg %.3f ms x %d (%.1f TFLOP/s), "
        "resid_f64 avg %.3f ms x %d, bj_apply avg %.1f us x %d; refinements max %d, inner iterations max %d" % (
            args.tol, 1e3 * t_step, k, gemm_avg, mm_n, achieved, rs_ms / max(rs_n, 1), rs_n,
            1e3 * bj_ms / max(bj_n, 1), bj_n, max(r["refinements"] for r in res), max(r["iters"] for r in res)))

    rep = g.solver_report()
    log("solver: %s" % json.dumps(rep))
    kname = gemm_kernel_name(rep.get("gemm_shape"))
    traffic = profiled_traffic("mixed_solve_naca0012_ar10_64cond", kname) if not args.small else None
    variant = rep.get("gemm_shape", {}).get("variant", 0)
    if variant >= 10:
        # bf16x2: the product streams the bf16 hi plane of A (2 bytes per
        # entry) plus the packed multivector and the output; the same launch
        # runs the near-field restoration on the workgroup slots the product
        # leaves free (its gathers hit L2/MALL)
        # (near-field CSR: 4-byte column + 4-byte value per entry; its
        # partial slice lda x 64 FP32)
        nnz = float(rep.get("near_nnz", 0))
        alg_bytes = 2.0 * n * n + 2 * 4.0 * n * 64 + 8.0 * nnz + 4.0 * n * 64
        gbs = alg_bytes / (gemm_avg * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": gbs, "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": gbs / MI355X_HBM_GBS,
                "traffic": traffic,
                "kernel": "%s (A32 as bf16 tiles x 64-column multivector + near-field FP32 restoration of %d "
                          "entries in the same launch, %.0f MB per launch)" % (kname, int(nnz), alg_bytes / 1e6)}
    elif variant >= 5:
        # bf16x3: the product needs a third of the FP32 matrix-core time and is
        # bound by streaming the packed matrix (4 bytes per entry: bf16 hi + lo)
        # plus the packed multivector and the output
        alg_bytes = 4.0 * n * n + 2 * 4.0 * n * 64
        gbs = alg_bytes / (gemm_avg * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": gbs, "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": gbs / MI355X_HBM_GBS,
                "traffic": traffic,
                "kernel": "%s (A32 as bf16 hi/lo tiles x 64-column multivector, %.0f MB per launch; %.1f "
                          "FP32-equivalent TFLOP/s)" % (kname, alg_bytes / 1e6, achieved)}
    else:
        roof = {"bound": "mfma", "achieved": achieved, "peak": MI355X_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / MI355X_FP32_MFMA_TFLOPS, "traffic": traffic,
                "kernel": "%s (A32 x 64-column multivector, %.0f MB of A per launch)" % (kname, 4.0 * n * n / 1e6)}
    valid = k
    vinfo = {}
    if fp64 is not None:
        ok, dcl, dcp, dcp_body = validity(fp64, res_full, tip)
        valid = int(ok.sum())
        vinfo = {"valid": valid, "max_abs_dCL": float(dcl.max()), "max_rel_dCp": float(dcp.max()),
                 "max_rel_dCp_outside_tips": float(dcp_body.max()), "failing_CL": int(np.sum(dcl > 1e-5)),
                 "failing_Cp": int(np.sum(dcp > 1e-4))}
        log("validity vs plain fp64 (tol %.0e): %d/%d conditions meet |dCL|<=1e-5 and |dCp|<=1e-4*max(1,|Cp|); "
            "max |dCL| = %.3e, max |dCp|/max(1,|Cp|) = %.3e (%.3e outside tip patches); speedup vs plain fp64 = "
            "%.1fx" % (args.cmp_tol, valid, k, dcl.max(), dcp.max(), dcp_body.max(), t_fp64 / t_step))
    value = valid / t_step

    # (b) the geometry's setup amortised over one 64-condition request: a
    # fresh geometry of the same mesh, from host topology to the solved
    # batch (both assemblies, near-field correction, bf16 packing, block
    # inversion, B n columns, post tables), the autotuner's table for this
    # leading dimension already filled by the geometry above
    amort = None
    fp32op = None
    if not args.no_variants:
        ctx.timing_reset()
        t0 = time.perf_counter()
        g2 = ctx.geometry(mesh)
        if mprec == CP_FP32_REFINE64:
            g2.assemble(CP_FP64)
        g2.assemble(CP_FP32_REFINE)
        r2 = g2.solve(conds, mprec, tol=args.tol, inner_tol=args.inner_tol, want_mu=False, want_cp=False, **kw)
        t_amort = time.perf_counter() - t0
        bi_ms, _ = ctx.timing("bj_invert")
        g2.release()
        amort = {"setup_plus_solve_s": t_amort, "solves_per_s": k / t_amort, "block_inversion_ms": bi_ms,
                 "same_CL_as_cached": bool(all(a["CL"] == b["CL"] for a, b in zip(r2, res)))}
        log("amortised: new geometry + setup + one %d-condition solve = %.3fs (%.1f solves/s; block inversion %.1f"
            " ms)" % (k, t_amort, k / t_amort, bi_ms))
        # the Krylov products on FP32 operands (the tuner without the bf16
        # products): C3's stated operator, for comparison with the bf16 one
        os.environ["CPGPU_NO_BF16X2"] = "1"
        os.environ["CPGPU_NO_BF16X3"] = "1"
        try:
            g3 = ctx.geometry(mesh)
            if mprec == CP_FP32_REFINE64:
                g3.assemble(CP_FP64)
            g3.assemble(CP_FP32_REFINE)
            t3, r3 = mixed_steps(g3, conds, mprec, args.tol, args.inner_tol, kw, 1, 3)
            res3 = g3.solve(conds, mprec, tol=args.tol, inner_tol=args.inner_tol, **kw)
            ok3 = validity(fp64, res3, tip)[0] if fp64 is not None else np.ones(k, bool)
            fp32op = {"ms_per_step": 1e3 * t3, "valid": int(ok3.sum()), "solves_per_s": float(ok3.sum()) / t3,
                      "gemm_shape": g3.solver_report().get("gemm_shape"),
                      "inner_iterations_max": max(r["iters"] for r in r3)}
            g3.release()
            log("fp32-operand products: %.1f ms/step, %d/%d valid, shape %s" % (1e3 * t3, fp32op["valid"], k,
                                                                                fp32op["gemm_shape"]))
        finally:
            del os.environ["CPGPU_NO_BF16X2"]
            del os.environ["CPGPU_NO_BF16X3"]

    groups = None if args.no_variants else group_bench(ctx, log)
    hot = serve_hot_bench(ctx, log) if args.serve_requests > 0 else None
    vp = None if args.no_vp else vp_bench(ctx, log, args.vp_steps, args.vp_reference_steps)
    srv = serve_bench(ctx, args.serve_requests, args.serve_qps, args.serve_slots, args.serve_seed, log) \
        if args.serve_requests > 0 else None

    cpu = None
    oracle_check = None
    hc = host_cpu()
    if not args.no_cpu:
        # 16 host cores: the share of the box's CPUs that one GPU is given
        nth = args.cpu_threads or 16
        c = cpu_baseline(mesh, kw, conds[0], nth, args.cmp_tol, log)
        cpu = {"value": 1.0 / c["t_solve"], "unit": "solves/s", "cores": nth, "kind": "port",
               "sample": "FP64 oracle (%s), 1 of the 64 conditions (alpha=-4, beta=-5) on the full %d-panel "
                         "matrix to tol %.0e (%d BiCGSTAB iterations), geometry cached (assembly %.1fs on %d "
                         "threads excluded); host %s, nproc %s, %s CPUs in this process's set" % (
                             "-march=native" if c["native"] else "x86-64-v3", n, args.cmp_tol, c["iters"],
                             c["t_assemble"], nth, hc["model"], hc["nproc"], hc["affinity"]),
               "single_thread_s_per_solve": c["t1_solve"],
               "single_thread_sample": "one thread (the reference's configuration): BiCGSTAB capped at 2 and 6 "
                                       "iterations, %.2fs per iteration, extrapolated to %d iterations; assembly "
                                       "%.1fs extrapolated from a %d-column block" % (
                                           c["t1_per_iter"], c["iters"], c["t1_assemble"], c["cols"])}
        if fp64 is not None:
            oracle_check = {"condition": list(conds[0]), "CL_oracle": c["CL"],
                            "dCL_fp64_gpu": abs(fp64[0]["CL"] - c["CL"]),
                            "dCL_mixed_gpu": abs(res_full[0]["CL"] - c["CL"]),
                            "mu_rel_fp64_gpu": float(np.linalg.norm(fp64[0]["mu"] - c["mu"]) /
                                                     np.linalg.norm(c["mu"])),
                            "mu_rel_mixed_gpu": float(np.linalg.norm(res_full[0]["mu"] - c["mu"]) /
                                                      np.linalg.norm(c["mu"]))}
            log("oracle check (condition 0): %s" % json.dumps(oracle_check))

    out = {
        "metric": "steady solves/sec at |dCL|<=1e-5 vs FP64",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("bf16(hi plane)+fp32 near tiles" if variant >= 10 else
                  ("bf16x3" if variant >= 5 else "fp32")) +
                 ("+fp64-refine(fp64 A,B)" if args.prec == "refine64" else "+fp64-refine"),
        "data": "synthetic (generated NACA wing mesh)",
        "config": {"workload": ("naca4412_ar6_4956tri_64cond" if args.small else "naca0012_ar10_25916tri_64cond"),
                   "n_body": n, "conditions": k,
                   "precision": "Krylov products on %s; FP64 refinement against %s" % (
                       "bf16 A (hi plane) outside near-field tiles, FP32 A on the near-field tiles "
                       "(32x16 tiles holding an entry within one panel diameter)" if variant >= 10 else
                       ("bf16x3 A (hi+lo planes)" if variant >= 5 else "FP32 A"),
                       "the FP64 A, B" if args.prec == "refine64" else "FP32 A, B + FP64 near field"),
                   "tol": args.tol, "comparator_tol": args.cmp_tol,
                   "valid_if": "|dCL|<=1e-5 and every panel |dCp|<=1e-4*max(1,|Cp|) vs the plain FP64 path"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "latency_p50_ms": srv["p50_ms"] if srv else None,
        "latency_p99_ms": srv["p99_ms"] if srv else None,
        "extra": {
            "validity": vinfo,
            "assembly_ginfl_per_s_fp64": pairs / t_a64 / 1e6, "assembly_ms_fp64": t_a64,
            "assembly_ginfl_per_s_fp32": pairs / t_a32 / 1e6, "assembly_ms_fp32": t_a32,
            "assembly_passes_ms": asm_parts,
            "plain_fp64_solves_per_s": (k / t_fp64) if t_fp64 else None,
            "plain_fp64_tol": args.cmp_tol,
            "solves_per_s_with_host_copy_of_mu_cp": valid / t_step_host,
            "ms_per_step_with_host_copy_of_mu_cp": 1e3 * t_step_host,
            "speedup_vs_plain_fp64": (t_fp64 / t_step) if t_fp64 else None,
            "amortised_setup": amort,
            "fp32_operand_products": fp32op,
            "oracle_check": oracle_check,
            "host_cpu": hc,
            "resid_f64_avg_ms": rs_ms / max(rs_n, 1),
            "gemm_f32_avg_ms": gemm_avg,
            "vortex_particles": vp,
            "serving": srv,
            "grouped_solves": groups,
            "serving_hot_cache": hot,
            "gemm_shape": rep.get("gemm_shape"),
            "block_jacobi_blocks": rep.get("block_jacobi_blocks"),
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
