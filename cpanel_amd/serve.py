"""Serving loop on one GPU: continuous batching of (geometry, conditions)
requests through the native server (cp_server_* in include/cpgpu.h), plus
the seeded mixed-size request stream used to measure it.

The stream: Poisson arrivals at a given rate; NACA 4-digit wings (camber
0..6 %, its position 0.3..0.5, thickness 8..18 %, aspect ratio 4..12,
rectangular or tapered) with body-panel counts log-uniform over a range;
1..32 conditions uniform in alpha [-4, 10] x beta [-5, 5] deg; a fraction of
the requests reuse an earlier geometry.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import abi, meshgen
from .case import CPanelError, _p


RequestStats = abi.RequestStats
ServerStats = abi.ServerStats


def _sig():
    return abi.lib()


def solve_opts(mach=0.1, Sref=1.0, bref=1.0, cref=1.0, cg=(0.25, 0.0, 0.0), tol=1e-10, max_iter=0, max_refine=0,
               inner_tol=0.0, precond=0):
    o = abi.SolveOpts()
    o.mach, o.Sref, o.bref, o.cref = mach, Sref, bref, cref
    for a in range(3):
        o.cg[a] = cg[a]
    o.tol, o.max_iter, o.max_refine, o.inner_tol, o.precond, o.pad = tol, max_iter, max_refine, inner_tol, precond, 0
    return o


class Server:
    """Native continuous-batching server over one context."""

    def __init__(self, ctx, slots=4, hbm_budget=0):
        L = _sig()
        h = C.c_void_p()
        ctx.check(L.cp_server_create(ctx.h, slots, hbm_budget, C.byref(h)), "cp_server_create")
        self.ctx, self.h, self._meta = ctx, h, {}

    def submit(self, key, mesh, conditions, opts, want_mu=False):
        L = _sig()
        xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float64)
        tri = np.ascontiguousarray(mesh["tri"], dtype=np.int32)
        ids = np.ascontiguousarray(mesh["ids"], dtype=np.int32)
        m = abi.Mesh(len(xyz), _p(xyz), len(tri), _p(tri, C.c_int32), _p(ids, C.c_int32))
        k = len(conditions)
        fs = (abi.Freestream * k)(*[abi.Freestream(*c) for c in conditions])
        rid = C.c_uint64()
        st = L.cp_server_submit(self.h, key, C.byref(m), k, fs, C.byref(opts), 1 if want_mu else 0, C.byref(rid))
        if st != abi.CP_OK:
            raise CPanelError(st, "cp_server_submit")
        self._meta[rid.value] = k
        return rid.value

    def wait(self, rid, n_body=None):
        """Blocks until request rid is done: (status, info list, mu or None, stats)."""
        L = _sig()
        k = self._meta.pop(rid)
        info = (abi.SolveInfo * k)()
        mu = np.zeros((k, n_body)) if n_body else None
        rs = RequestStats()
        st = L.cp_server_wait(self.h, rid, info, None if mu is None else _p(mu), C.byref(rs))
        out = [dict(CL=i.CL, CDi=i.CDi, Fwind=np.array(i.Fwind), rel_residual=i.rel_residual, iters=i.iters,
                    refinements=i.refinements, converged=bool(i.converged)) for i in info]
        stats = {f: getattr(rs, f) for f, _ in RequestStats._fields_ if f != "pad"}
        return st, out, mu, stats

    def stats(self):
        s = ServerStats()
        _sig().cp_server_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            _sig().cp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- workload
def request_geometry(rng, n_lo=500, n_hi=20000):
    """Seeded NACA 4-digit wing with a body-panel count log-uniform in
    [n_lo, n_hi]; returns (mesh, params)."""
    m = int(rng.integers(0, 7))
    p = int(rng.integers(3, 6))
    t = int(rng.integers(8, 19))
    code = "%d%d%02d" % (m, p if m > 0 else 0, t)
    ar = float(rng.uniform(4, 12))
    tapered = bool(rng.random() < 0.5)
    taper = float(rng.uniform(0.4, 0.8)) if tapered else 1.0
    target = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
    quads = target / 2.0
    n_span = max(4, int(round(np.sqrt(quads * ar / 8.0))))
    n_chord = max(12, 2 * int(round(quads / n_span / 2.0)))
    chord = 1.0
    span = ar * chord * (1 + taper) / 2.0
    mesh = meshgen.wing(code, chord, span, n_chord, n_span, "uniform", 10.0 * chord, taper=taper,
                        seed=int(rng.integers(1 << 30)))
    mean_chord = chord * (1 + taper) / 2.0
    prm = dict(code=code, ar=ar, taper=taper, n_chord=n_chord, n_span=n_span, span=span,
               Sref=span * mean_chord, bref=span, cref=mean_chord)
    return mesh, prm


def make_stream(n_requests, qps, seed=0, n_lo=500, n_hi=20000, reuse=0.3, max_cond=32):
    """Seeded request list: dicts with arrival time, geometry key, mesh,
    conditions and reference lengths."""
    rng = np.random.default_rng(seed)
    t = 0.0
    geoms = []
    out = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / qps)
        if geoms and rng.random() < reuse:
            key = int(rng.integers(0, len(geoms)))
        else:
            geoms.append(request_geometry(rng, n_lo, n_hi))
            key = len(geoms) - 1
        k = int(rng.integers(1, max_cond + 1))
        conds = [(1.0, float(rng.uniform(-4, 10)), float(rng.uniform(-5, 5))) for _ in range(k)]
        out.append(dict(t=t, key=key, conds=conds))
    return out, geoms


def run_stream(server, stream, geoms, want_mu=False, mach=0.1):
    """Submits every request at its arrival time (relative to now) and
    collects the results; returns (results in request order, wall seconds)."""
    results = [None] * len(stream)
    ids = [None] * len(stream)
    nb = {}
    ready = threading.Semaphore(0)

    def collector():
        for i in range(len(stream)):
            ready.acquire()
            st, info, mu, rs = server.wait(ids[i], nb.get(stream[i]["key"]) if want_mu else None)
            results[i] = dict(status=st, info=info, mu=mu, stats=rs)

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    for i, r in enumerate(stream):
        mesh, prm = geoms[r["key"]]
        if want_mu and r["key"] not in nb:
            nb[r["key"]] = int(np.sum(np.asarray(mesh["ids"]) < 1000))
        opts = solve_opts(mach=mach, Sref=prm["Sref"], bref=prm["bref"], cref=prm["cref"])
        delay = r["t"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids[i] = server.submit(r["key"], mesh, r["conds"], opts, want_mu)
        ready.release()
    th.join()
    return results, time.perf_counter() - t0


def summarize(results, wall):
    lat = np.array([r["stats"]["latency_s"] for r in results])
    ok = [r for r in results if r["status"] == abi.CP_OK]
    conds = sum(len(r["info"]) for r in ok)
    failed = {}
    failed_detail = []
    for r in results:
        if r["status"] != abi.CP_OK:
            name = abi.lib().cp_status_string(r["status"]).decode()
            failed[name] = failed.get(name, 0) + 1
            if len(failed_detail) < 8:
                failed_detail.append(dict(status=name, n_body=r["stats"]["n_body"], conditions=len(r["info"]),
                                          max_rel_residual=max((i["rel_residual"] for i in r["info"]), default=0),
                                          refinements=max((i["refinements"] for i in r["info"]), default=0)))
    return dict(requests=len(results), ok=len(ok), failed=failed, failed_detail=failed_detail, conditions=conds,
                wall_s=wall,
                solves_per_s=conds / wall,
                p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                mean_queue_ms=1e3 * float(np.mean([r["stats"]["queue_s"] for r in results])),
                cache_hit_frac=float(np.mean([r["stats"]["cache_hit"] for r in results])))


def main(argv=None):
    """python -m cpanel_amd.serve --requests N --qps Q [--slots S] [--seed K]
    [--budget-gb G]: runs the seeded mixed-size stream on device 0 and prints
    one JSON summary line (p50/p99 latency, solves/s, cache behaviour)."""
    import argparse
    import json
    import sys

    from .case import Context
    ap = argparse.ArgumentParser(prog="python -m cpanel_amd.serve")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--qps", type=float, default=10.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--min-panels", type=int, default=500)
    ap.add_argument("--max-panels", type=int, default=20000)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    stream, geoms = make_stream(a.requests, a.qps, seed=a.seed, n_lo=a.min_panels, n_hi=a.max_panels)
    print("stream: %d requests, %d geometries generated in %.1fs" % (len(stream), len(geoms),
                                                                     time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    ctx = Context(0)
    S = Server(ctx, slots=a.slots, hbm_budget=int(a.budget_gb * (1 << 30)))
    res, wall = run_stream(S, stream, geoms)
    st = S.stats()
    S.close()
    out = summarize(res, wall)
    out.update(dict(qps=a.qps, slots=a.slots, seed=a.seed, geometries=len(geoms), evictions=st["evictions"],
                    batches=st["batches"], grouped_solves=st["grouped_solves"],
                    requests_in_groups=sum(1 for r in res if r["stats"]["group_geometries"] > 1),
                    budget_gb=a.budget_gb,
                    p90_ms=1e3 * float(np.percentile([r["stats"]["latency_s"] for r in res], 90)),
                    mean_prepare_ms_new_geometry=1e3 * float(np.mean(
                        [r["stats"]["prepare_s"] for r in res if not r["stats"]["cache_hit"]] or [0])),
                    mean_solve_ms=1e3 * float(np.mean([r["stats"]["solve_s"] for r in res]))))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
