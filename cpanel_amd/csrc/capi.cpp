// C ABI of the MI355X CPanel solve path (include/cpgpu.h).
//
// Objects:
//   cp_ctx  - one GPU: device id, stream, HBM budget accounting, timers.
//   cp_geom - one mesh: host topology, device panel tables, resident A/B in
//             FP64 and/or FP32, post-processing tables, solver workspaces.
#include "capi_internal.h"

#include <cstdlib>

using namespace cpgpu;
using namespace cpgpu::capi;

namespace cpgpu {
namespace capi {

// the bytes the context may hold: the caller's budget within the device cap
static uint64_t limit(const cp_ctx* c) {
    if (c->budget && c->cap) return std::min(c->budget, c->cap);
    return c->budget ? c->budget : c->cap;
}

// frees pooled blocks (largest first) until `need` more bytes fit the limit
void trim_pool(cp_ctx* c, uint64_t need) {
    const uint64_t lim = limit(c);
    while (!c->pool.empty() && lim && c->in_use + need > lim) {
        auto it = std::prev(c->pool.end());
        (void)hipFree(it->second);
        c->in_use -= it->first;
        c->pooled -= it->first;
        c->pool.erase(it);
    }
}

void dalloc(cp_ctx* c, DevBuf& b, size_t bytes) {
    b.release();
    bytes = (bytes + 255) / 256 * 256;
    std::lock_guard<std::mutex> lk(c->alloc_mu);
    b.ctx = c;
    // reuse a pooled block of at most 1.25x the request
    auto it = c->pool.lower_bound(bytes);
    if (it != c->pool.end() && it->first <= bytes + bytes / 4) {
        b.p = it->second;
        b.bytes = it->first;
        c->pooled -= it->first;
        c->pool.erase(it);
        return;
    }
    trim_pool(c, bytes);
    if (c->budget && c->in_use + bytes > c->budget) throw BudgetError{bytes};
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {  // device full: give the pool back and retry once
        (void)hipGetLastError();
        while (!c->pool.empty()) {
            auto q = std::prev(c->pool.end());
            (void)hipFree(q->second);
            c->in_use -= q->first;
            c->pooled -= q->first;
            c->pool.erase(q);
        }
        CP_HIP(hipMalloc(&b.p, bytes));
    }
    b.bytes = bytes;
    c->in_use += bytes;
}

// device bytes allocated outside DevBuf (solver workspaces)
void account(cp_ctx* c, int64_t delta) {
    std::lock_guard<std::mutex> lk(c->alloc_mu);
    c->in_use = (uint64_t)((int64_t)c->in_use + delta);
}

// would `bytes` more fit the budget (counting pooled blocks as free)?
bool fits(cp_ctx* c, uint64_t bytes) {
    std::lock_guard<std::mutex> lk(c->alloc_mu);
    return !c->budget || c->in_use - c->pooled + bytes <= c->budget;
}



cp_status fail(cp_ctx* c, cp_status s, const std::string& msg) {
    if (c) {
        std::lock_guard<std::mutex> lk(c->err_mu);
        c->err = msg;
    }
    return s;
}

void set_err(cp_ctx* c, const std::string& msg) {
    std::lock_guard<std::mutex> lk(c->err_mu);
    c->err = msg;
}

void ctx_ref(cp_ctx* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

// the last reference frees the context: its streams, pinned blocks and pooled
// device blocks (every child has returned its buffers by then)
void ctx_unref(cp_ctx* c) {
    if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    if (c->device >= 0) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        for (hipStream_t st : c->streams) (void)hipStreamDestroy(st);
        for (void* p : c->pinned) (void)hipHostFree(p);
        for (auto& e : c->pool) (void)hipFree(e.second);
    }
    delete c;
}

void mark_served(cp_ctx* c) { c->served.store(true); }

void geom_ref(cp_geom* g) { g->refs.fetch_add(1, std::memory_order_relaxed); }

// the last reference frees the geometry; its buffers, pinned block and
// stream go back to the context, which it then releases
void geom_unref(cp_geom* g) {
    if (g->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    cp_ctx* c = g->ctx;
    hipStream_t st = nullptr;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        if (c->device >= 0) {
            (void)hipSetDevice(c->device);
            if (g->stream) (void)hipStreamSynchronize(g->stream);
        }
        if (g->solver_bytes) account(c, -g->solver_bytes);
        st = g->stream;
    }
    void* host = g->solver_host;
    g->batched.reset();
    delete g;
    {
        std::lock_guard<std::mutex> lk(c->alloc_mu);
        if (host) c->pinned.push_back(host);
        if (st) c->streams.push_back(st);
    }
    ctx_unref(c);
}


// cpCase::windToBody (cpCase.cpp:142-166): body-axis freestream and the
// wind<-body transform T (row major)
void wind_to_body(const cp_freestream& f, double T[9], double Vb[3]) {
    double a = f.alpha_deg * kPI / 180, b = f.beta_deg * kPI / 180;
    T[0] = std::cos(a) * std::cos(b);
    T[1] = std::cos(a) * std::sin(b);
    T[2] = -std::sin(a);
    T[3] = -std::sin(b);
    T[4] = std::cos(b);
    T[5] = 0;
    T[6] = std::sin(a) * std::cos(b);
    T[7] = std::sin(a) * std::sin(b);
    T[8] = std::cos(a);
    Vb[0] = T[0] * f.V;
    Vb[1] = T[3] * f.V;
    Vb[2] = T[6] * f.V;
}

cp_status build_geom(cp_ctx* c, const cp_mesh* m, uint64_t key, cp_geom** out, const VpParams* vp) {
    std::unique_ptr<cp_geom> g(new cp_geom());
    g->ctx = c;
    g->key = key;
    std::string err;
    if (!build_geometry(m->n_nodes, m->xyz, m->n_tris, m->tri, m->ids, g->host, err, vp))
        return fail(c, CP_BAD_MESH, err);
    const Geometry& G = g->host;
    const int n = (int)G.body.size(), nw = (int)G.wake.size();
    DevGeom& d = g->dev;
    d.n = n;
    d.nw = nw;
    d.lda = round_up(n, 512);
    if (c->device < 0) {  // host-only context: topology tables only
        ctx_ref(c);
        *out = g.release();
        return CP_OK;
    }
    {
        std::lock_guard<std::mutex> lk(c->alloc_mu);
        if (!c->streams.empty()) {
            g->stream = c->streams.back();
            c->streams.pop_back();
        }
    }
    if (!g->stream) CP_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    std::vector<PanelRec> br(n), wr(std::max(nw, 1));
    for (int i = 0; i < n; i++) br[i] = make_panel_rec(G.body[i]);
    for (int j = 0; j < nw; j++) wr[j] = make_panel_rec(G.wake[j]);
    upload(c, g->body, br);
    upload(c, g->wake, wr);
    // collocation points padded with a far point (padding rows are zeroed)
    std::vector<double> cxyz(3 * (size_t)d.lda, 1.0e6), nrm(3 * (size_t)d.lda, 0.0);
    for (int i = 0; i < n; i++) {
        cxyz[i] = G.body[i].center.x;
        cxyz[d.lda + i] = G.body[i].center.y;
        cxyz[2 * d.lda + i] = G.body[i].center.z;
        nrm[i] = G.body[i].normal.x;
        nrm[d.lda + i] = G.body[i].normal.y;
        nrm[2 * d.lda + i] = G.body[i].normal.z;
    }
    upload(c, g->cxyz, cxyz);
    upload(c, g->normals, nrm);
    // bounding sphere of each 64-row tile of collocation points (box centre,
    // largest distance to it, widened by a relative 1e-12): the near-field
    // pass of the assembly skips the tiles it proves far
    d.ntiles = (n + 63) / 64;
    std::vector<double> ts(4 * (size_t)d.ntiles);
    for (int t = 0; t < d.ntiles; t++) {
        const int i0 = 64 * t, i1 = std::min(n, i0 + 64);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = i0; i < i1; i++) {
            const Vec3& p = G.body[i].center;
            const double v[3] = {p.x, p.y, p.z};
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], v[a]);
                hi[a] = std::max(hi[a], v[a]);
            }
        }
        double ctr[3], r2 = 0;
        for (int a = 0; a < 3; a++) ctr[a] = 0.5 * (lo[a] + hi[a]);
        for (int i = i0; i < i1; i++) {
            const Vec3& p = G.body[i].center;
            const double dx = p.x - ctr[0], dy = p.y - ctr[1], dz = p.z - ctr[2];
            r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
        }
        ts[4 * t] = ctr[0];
        ts[4 * t + 1] = ctr[1];
        ts[4 * t + 2] = ctr[2];
        ts[4 * t + 3] = std::sqrt(r2) * (1 + 1e-12) + 1e-300;
    }
    upload(c, g->tsph, ts);
    // fold lists per body column in (wake panel, slot) order
    std::vector<std::vector<std::pair<int, double>>> cols(n);
    for (int j = 0; j < nw; j++) {
        double t = G.fold_coef[j];
        double cf[4] = {1 - t, t - 1, t, -t};
        for (int s = 0; s < 4; s++) cols[G.fold_idx[4 * j + s]].push_back({j, cf[s]});
    }
    std::vector<int> fp(n + 1, 0), fw;
    std::vector<double> fc;
    for (int i = 0; i < n; i++) {
        fp[i] = (int)fw.size();
        for (auto& e : cols[i]) {
            fw.push_back(e.first);
            fc.push_back(e.second);
        }
    }
    fp[n] = (int)fw.size();
    upload(c, g->fold_ptr, fp);
    upload(c, g->fold_w, fw);
    upload(c, g->fold_c, fc);
    d.body = g->body.as<PanelRec>();
    d.wake = g->wake.as<PanelRec>();
    d.cx = g->cxyz.as<double>();
    d.cy = d.cx + d.lda;
    d.cz = d.cx + 2 * d.lda;
    d.fold_ptr = g->fold_ptr.as<int>();
    d.fold_w = g->fold_w.as<int>();
    d.fold_c = g->fold_c.as<double>();
    d.tsph = g->tsph.as<double>();
    // the post tables depend on the host geometry only (the CHTLS fits are the
    // longest host step of a new geometry)
    cp_geom* gp = g.get();
    gp->post_job = std::async(std::launch::async, [gp] { build_post_tables(gp->host, gp->post_host); });
    ctx_ref(c);
    *out = g.release();
    return CP_OK;
}

void ensure_wake_influence(cp_geom* g) {
    if (g->W.p || g->dev.nw == 0) return;
    cp_ctx* c = g->ctx;
    dalloc(c, g->W, (size_t)g->dev.lda * g->dev.nw * sizeof(double));
    launch_wake_influence(g->dev, g->W.as<double>(), g->dev.lda, g->stream);
}

void ensure_post(cp_geom* g) {
    if (g->post_ready) return;
    cp_ctx* c = g->ctx;
    if (g->post_job.valid())
        g->post_job.get();
    else
        build_post_tables(g->host, g->post_host);
    const PostHost& H = g->post_host;
    const int n = H.n;
    upload(c, g->p_nbr, H.nbr);
    upload(c, g->p_w1, H.w1);
    upload(c, g->p_w2, H.w2);
    upload(c, g->p_mode, H.mode);
    upload(c, g->p_te, H.te_idx);
    upload(c, g->p_rphi, H.rphi);
    upload(c, g->p_rv, H.rv);
    upload(c, g->p_rw, H.rw);
    upload(c, g->p_simp, H.simpson);
    std::vector<double> ctr(3 * (size_t)n), bez(3 * (size_t)n), area(n);
    for (int i = 0; i < n; i++) {
        const PanelGeom& p = g->host.body[i];
        ctr[i] = p.center.x;
        ctr[n + i] = p.center.y;
        ctr[2 * (size_t)n + i] = p.center.z;
        bez[i] = p.bez.x;
        bez[n + i] = p.bez.y;
        bez[2 * (size_t)n + i] = p.bez.z;
        area[i] = p.area;
    }
    upload(c, g->p_ctr, ctr);
    upload(c, g->p_bez, bez);
    upload(c, g->p_area, area);
    PostDev& P = g->post;
    P.n = n;
    P.n_te = H.n_te;
    P.n_st = H.n_st;
    P.n_wakes = H.n_wakes;
    P.nbr = g->p_nbr.as<int32_t>();
    P.w1 = g->p_w1.as<double>();
    P.w2 = g->p_w2.as<double>();
    P.mode = g->p_mode.as<int32_t>();
    P.te_idx = g->p_te.as<int32_t>();
    P.rphi = g->p_rphi.as<double>();
    P.rv = g->p_rv.as<double>();
    P.rw = g->p_rw.as<double>();
    P.simpson = g->p_simp.as<double>();
    P.ctr = g->p_ctr.as<double>();
    P.bez = g->p_bez.as<double>();
    P.area = g->p_area.as<double>();
    g->post_ready = true;
}

// The batched mixed-precision solver of a geometry, built on first use: its
// workspace comes from the context's device pool, its pinned host block from
// the context's pinned pool.
void ensure_batched(cp_geom* g) {
    if (g->batched) return;
    cp_ctx* c = g->ctx;
    const int64_t lda = g->dev.lda;
    dalloc(c, g->solver_ws, BatchedSolver::workspace_bytes(lda));
    if (!g->solver_host) {
        {
            std::lock_guard<std::mutex> lk(c->alloc_mu);
            if (!c->pinned.empty()) {
                g->solver_host = c->pinned.back();
                c->pinned.pop_back();
            }
        }
        if (!g->solver_host) CP_HIP(hipHostMalloc(&g->solver_host, BatchedSolver::host_bytes(), hipHostMallocDefault));
    }
    g->batched.reset(new BatchedSolver(g->dev.n, lda, kRhsTile, g->solver_ws.p, g->solver_host, g->stream));
}

// Strip block-Jacobi inverses of the FP32 matrix: plan on the host, invert
// the diagonal blocks on the device (FP64 Gauss-Jordan, FP32 result).
void ensure_block_jacobi(cp_geom* g) {
    if (g->bj_ready) return;
    cp_ctx* c = g->ctx;
    BlockPlan& P = g->bj_plan;
    if (P.perm.empty()) {
        build_strip_blocks(g->host, bj_block_target(g->dev.n), kBjHardMax, P);
        upload(c, g->bj_perm, P.perm);
        std::vector<int> ip(g->dev.n, 0);
        for (size_t q = 0; q < P.perm.size(); q++) ip[P.perm[q]] = (int)q;
        upload(c, g->bj_iperm, ip);
        upload(c, g->bj_ptr, P.ptr);
        upload(c, g->bj_off, P.off);
        dalloc(c, g->bj_inv, std::max<int64_t>(P.off.back(), 1) * sizeof(float));
        std::vector<int64_t> poff(P.ptr.size(), 0);
        for (int q = 0; q + 1 < (int)P.ptr.size(); q++)
            poff[q + 1] = poff[q] + bj_packed_floats(P.ptr[q + 1] - P.ptr[q]);
        upload(c, g->bj_poff, poff);
        dalloc(c, g->bj_pk, std::max<int64_t>(poff.back(), 1) * sizeof(float));
    }
    const int nb = P.nblocks();
    const int np = (int)P.perm.size();
    DevBuf X, flags, st, Y;
    dalloc(c, X, (size_t)std::max<int64_t>(P.off.back(), 1) * sizeof(double));
    dalloc(c, Y, (size_t)bj_invert_rows_doubles(nb, P.max_b) * sizeof(double));
    dalloc(c, flags, 2 * (size_t)std::max(np, 1) * sizeof(int));
    dalloc(c, st, (size_t)std::max(nb, 1) * sizeof(int));
    {
        TimedScope ts("bj_invert", g->stream);
        launch_bj_invert(g->A32.as<float>(), g->dev.lda, g->bj_perm.as<int>(), g->bj_ptr.as<int>(),
                         g->bj_off.as<int64_t>(), nb, P.max_b, X.as<double>(), flags.as<int>(), flags.as<int>() + np,
                         st.as<int>(), Y.as<double>(), g->bj_inv.as<float>(), g->stream);
        // CPGPU_BJ_PREC=fp32: FP32 matrix-core product (default: bf16 operands)
        const char* e = std::getenv("CPGPU_BJ_PREC");
        g->bj_pk_kind = (e && std::string(e) == "fp32") ? kBjPackFp32 : kBjPackBf16;
        launch_bj_pack(g->bj_inv.as<float>(), g->bj_ptr.as<int>(), g->bj_off.as<int64_t>(), g->bj_poff.as<int64_t>(),
                       nb, P.max_b, g->bj_pk.as<float>(), g->bj_pk_kind, g->stream);
    }
    std::vector<int> sv(std::max(nb, 1), 0);
    if (nb > 0)
        CP_HIP(hipMemcpyAsync(sv.data(), st.p, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost, g->stream));
    CP_HIP(hipStreamSynchronize(g->stream));
    g->bj_ready = true;
    // a singular block: fall back to Jacobi
    if (std::any_of(sv.begin(), sv.end(), [](int v) { return v != 0; })) g->bj_plan.perm.clear();
}

void fill_info(const double* r, const double T[9], cp_solve_info& inf) {
    inf.CL = r[0];
    inf.CDi = r[1];
    for (int a = 0; a < 3; a++) inf.Fbody[a] = r[2 + a];
    for (int a = 0; a < 3; a++) inf.CM[a] = r[8 + a];
    // bodyToWind: T^T F (cpCase.cpp:168-172)
    for (int a = 0; a < 3; a++) inf.Fwind[a] = T[0 * 3 + a] * r[2] + T[1 * 3 + a] * r[3] + T[2 * 3 + a] * r[4];
}

// The geometry's mixed-precision solver with its matrices, near-field
// correction, packed tiles and preconditioner attached for `prec`.
void prepare_mixed(cp_geom* g, cp_prec prec, const cp_solve_opts* opts) {
    const int n = g->dev.n;
    ensure_batched(g);
    g->batched->set_near(g->near_rowptr.as<int>(), g->near_cols.as<int>(), g->near_dA.as<double>(),
                         g->near_dB.as<double>());
    g->batched->set_packed(g->A16.p, packed_kblocks(n));
    g->batched->set_near_tiles(g->near_tptr.as<int>(), g->near_tkb.as<int>());
    if (prec == CP_FP32_REFINE64)
        g->batched->set_f64(g->A64.as<double>(), g->B64.as<double>());
    else
        g->batched->set_f64(nullptr, nullptr);
    if (opts->precond != 1) ensure_block_jacobi(g);
    if (opts->precond != 1 && !g->bj_plan.perm.empty())
        g->batched->set_block_jacobi(g->bj_perm.as<int>(), g->bj_ptr.as<int>(), g->bj_off.as<int64_t>(),
                                     g->bj_inv.as<float>(), g->bj_plan.nblocks(), g->bj_plan.max_b,
                                     g->bj_iperm.as<int>(), g->bj_pk.as<float>(), g->bj_poff.as<int64_t>(),
                                     g->bj_pk_kind);
    else
        g->batched->set_block_jacobi(nullptr, nullptr, nullptr, nullptr, 0, 0);
    if (g->ctx->served.load(std::memory_order_relaxed)) g->batched->allow_graphs(false);
}

BatchedOpts mixed_opts(const cp_solve_opts* opts, int n) {
    BatchedOpts bo;
    bo.tol = opts->tol > 0 ? opts->tol : 1e-10;
    bo.max_iter = opts->max_iter > 0 ? opts->max_iter : 2 * n;
    bo.max_refine = opts->max_refine > 0 ? opts->max_refine : 20;
    bo.inner_tol = opts->inner_tol > 0 ? opts->inner_tol : 0;  // 0: adaptive
    return bo;
}

// Cp, forces and Trefftz CL/CDi of k solutions mu (device, leading dimension
// lda) on stream s; the per-condition results land in res (host, 16 doubles
// per condition) after the caller synchronises s.  dpart, dres, dcp: the
// launch's device buffers (dcp only when Cp is wanted).
void launch_post_k(cp_geom* g, const cp_solve_opts* opts, int k, const double* dmu, const double* dv4, DevBuf& dpart,
                   DevBuf& dres, DevBuf* dcp, double* res, hipStream_t s) {
    cp_ctx* c = g->ctx;
    const int n = g->dev.n;
    const int64_t lda = g->dev.lda;
    int nblk = post_partial_blocks(n);
    dalloc(c, dpart, (size_t)nblk * k * 6 * sizeof(double));
    dalloc(c, dres, (size_t)16 * k * sizeof(double));
    if (dcp) dalloc(c, *dcp, (size_t)lda * k * sizeof(double));
    PostParams pp;
    pp.PG = std::sqrt(1 - opts->mach * opts->mach);
    pp.Sref = opts->Sref;
    pp.bref = opts->bref;
    pp.cref = opts->cref;
    for (int a = 0; a < 3; a++) pp.cg[a] = opts->cg[a];
    launch_post(g->post, pp, k, dmu, lda, dv4, dcp ? dcp->as<double>() : nullptr, lda, dpart.as<double>(), nblk,
                dres.as<double>(), s);
    CP_HIP(hipMemcpyAsync(res, dres.p, (size_t)16 * k * sizeof(double), hipMemcpyDeviceToHost, s));
}

}  // namespace capi
}  // namespace cpgpu


extern "C" {

const char* cp_status_string(cp_status s) {
    switch (s) {
        case CP_OK: return "ok";
        case CP_NOT_CONVERGED: return "not converged";
        case CP_OUT_OF_BUDGET: return "out of HBM budget";
        case CP_BAD_MESH: return "bad mesh";
        case CP_HIP_ERROR: return "HIP error";
        case CP_BAD_ARG: return "bad argument";
        case CP_NO_DEVICE: return "no device";
        case CP_NOT_ASSEMBLED: return "matrices not assembled";
    }
    return "unknown";
}

cp_status cp_create(int device, uint64_t budget, cp_ctx** out) {
    if (!out) return CP_BAD_ARG;
    *out = nullptr;
    if (device == -1) {  // host-only: geometry/topology queries, no device work
        cp_ctx* c = new cp_ctx();
        c->device = -1;
        c->budget = budget;
        *out = c;
        return CP_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return CP_NO_DEVICE;
    cp_ctx* c = new cp_ctx();
    c->device = device;
    c->budget = budget;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CP_HIP_ERROR;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0) {
        const uint64_t reserve = std::max<uint64_t>((uint64_t)8 << 30, total_b / 20);
        c->cap = free_b > 2 * reserve ? free_b - reserve : free_b / 2;
    }
    *out = c;
    return CP_OK;
}

void cp_destroy(cp_ctx* c) {
    if (c) ctx_unref(c);
}

const char* cp_last_error(const cp_ctx* c) {
    if (!c) return "no context";
    // a thread-local copy: the context's string may be replaced by another
    // thread's error while the caller reads it
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(const_cast<cp_ctx*>(c)->err_mu);
    copy = c->err;
    return copy.c_str();
}
int cp_last_hip_error(const cp_ctx* c) { return c ? c->hip_err : 0; }
uint64_t cp_bytes_in_use(const cp_ctx* c) { return c ? c->in_use : 0; }

cp_status cp_geometry_upload(cp_ctx* c, const cp_mesh* m, uint64_t key, cp_geom** out) {
    if (!m || !out || !m->xyz || !m->tri || !m->ids || m->n_nodes <= 0 || m->n_tris <= 0) return CP_BAD_ARG;
    if (!c) return CP_BAD_ARG;
    return guarded_impl(c, nullptr, [&] { return build_geom(c, m, key, out); });
}

cp_status cp_geometry_from_tri(cp_ctx* c, const char* path, uint64_t key, cp_geom** out) {
    if (!path || !out) return CP_BAD_ARG;
    std::vector<double> xyz;
    std::vector<int32_t> tri, ids;
    std::string err;
    if (!read_tri_file(path, xyz, tri, ids, err)) return fail(c, CP_BAD_MESH, err);
    cp_mesh m{(int32_t)(xyz.size() / 3), xyz.data(), (int32_t)ids.size(), tri.data(), ids.data()};
    return cp_geometry_upload(c, &m, key, out);
}

void cp_geometry_release(cp_geom* g) {
    if (g) geom_unref(g);
}

cp_status cp_geometry_info(const cp_geom* g, cp_geom_info* info) {
    if (!g || !info) return CP_BAD_ARG;
    info->n_body = g->dev.n;
    info->n_wake = g->dev.nw;
    int nl = 0;
    for (auto& w : g->host.wakes) nl += (int)w.lines.size();
    info->n_wakelines = nl;
    info->n_nodes = g->host.n_nodes;
    info->lda = g->dev.lda;
    info->n_wakes = (int32_t)g->host.wakes.size();
    info->n_stations = 165;  // wake::trefftzPlane: 164 intervals (wake.cpp:148-156)
    return CP_OK;
}

cp_status cp_geometry_body(const cp_geom* g, int32_t* tri_index, double* center, double* normal, double* area) {
    if (!g) return CP_BAD_ARG;
    const Geometry& G = g->host;
    for (size_t i = 0; i < G.body.size(); i++) {
        if (tri_index) tri_index[i] = G.body_tri[i];
        if (center) {
            center[3 * i] = G.body[i].center.x;
            center[3 * i + 1] = G.body[i].center.y;
            center[3 * i + 2] = G.body[i].center.z;
        }
        if (normal) {
            normal[3 * i] = G.body[i].normal.x;
            normal[3 * i + 1] = G.body[i].normal.y;
            normal[3 * i + 2] = G.body[i].normal.z;
        }
        if (area) area[i] = G.body[i].area;
    }
    return CP_OK;
}

cp_status cp_geometry_wake(const cp_geom* g, int32_t* tri_index, int32_t* idx4, double* coef) {
    if (!g) return CP_BAD_ARG;
    const Geometry& G = g->host;
    for (size_t j = 0; j < G.wake.size(); j++) {
        if (tri_index) tri_index[j] = G.wake_tri[j];
        if (idx4)
            for (int s = 0; s < 4; s++) idx4[4 * j + s] = G.fold_idx[4 * j + s];
        if (coef) coef[j] = G.fold_coef[j];
    }
    return CP_OK;
}

cp_status cp_geometry_clusters(const cp_geom* g, int32_t* cl16, int32_t* ts_order, int32_t* tip) {
    if (!g) return CP_BAD_ARG;
    const Geometry& G = g->host;
    for (size_t i = 0; i < G.body.size(); i++) {
        if (cl16)
            for (int k = 0; k < 16; k++) cl16[16 * i + k] = G.cluster[16 * i + k];
        if (ts_order) ts_order[i] = G.ts_order[i];
        if (tip) tip[i] = G.tip[i];
    }
    return CP_OK;
}

// The near-field pairs |c_i - c_j| <= longSide_j (CSR by row, columns
// ascending, with each entry's row) and the packed 32 x 16 tiles of A that
// hold one, built on the geometry's stream; two small reads of the counts
// size the lists.
static void build_near_lists(cp_geom* g) {
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    const int n = g->dev.n;
    const int nrt = (int)(g->dev.lda / 32), nkb = packed_kblocks(n);
    DevBuf cnt, qcnt, mark, reach;
    dalloc(c, cnt, (size_t)(std::max(n, nrt) + 1) * sizeof(int));
    dalloc(c, qcnt, (size_t)std::max(near_chunks(g->dev), 1) * std::max(n, 1) * sizeof(int));
    dalloc(c, reach, (size_t)std::max(g->dev.ntiles, 1) * sizeof(double));
    dalloc(c, g->near_rowptr, (size_t)(n + 1) * sizeof(int));
    launch_near_reach(g->dev, 1.0, reach.as<double>(), s);
    launch_near_count(g->dev, 1.0, reach.as<double>(), qcnt.as<int>(), cnt.as<int>(), s);
    launch_exclusive_scan(cnt.as<int>(), n, g->near_rowptr.as<int>(), s);
    int nnz = 0;
    CP_HIP(hipMemcpyAsync(&nnz, g->near_rowptr.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, s));
    CP_HIP(hipStreamSynchronize(s));
    g->near_nnz = nnz;
    dalloc(c, g->near_cols, (size_t)std::max(nnz, 1) * sizeof(int));
    dalloc(c, g->near_rows, (size_t)std::max(nnz, 1) * sizeof(int));
    launch_near_fill(g->dev, 1.0, reach.as<double>(), qcnt.as<int>(), g->near_rowptr.as<int>(),
                     g->near_cols.as<int>(), g->near_rows.as<int>(), s);
    dalloc(c, mark, std::max<size_t>((size_t)nrt * nkb, 1));
    CP_HIP(hipMemsetAsync(mark.p, 0, std::max<size_t>((size_t)nrt * nkb, 1), s));
    launch_near_tile_count(nnz, g->near_rows.as<int>(), g->near_cols.as<int>(), nrt, nkb, mark.as<unsigned char>(),
                           cnt.as<int>(), s);
    dalloc(c, g->near_tptr, (size_t)(nrt + 1) * sizeof(int));
    launch_exclusive_scan(cnt.as<int>(), nrt, g->near_tptr.as<int>(), s);
    int ntk = 0;
    CP_HIP(hipMemcpyAsync(&ntk, g->near_tptr.as<int>() + nrt, sizeof(int), hipMemcpyDeviceToHost, s));
    CP_HIP(hipStreamSynchronize(s));
    dalloc(c, g->near_tkb, (size_t)std::max(ntk, 1) * sizeof(int));
    launch_near_tile_fill(mark.as<unsigned char>(), nrt, nkb, g->near_tptr.as<int>(), g->near_tkb.as<int>(), s);
    dalloc(c, g->near_dA, (size_t)std::max(nnz, 1) * sizeof(double));
    dalloc(c, g->near_dB, (size_t)std::max(nnz, 1) * sizeof(double));
    CP_HIP(hipStreamSynchronize(s));  // the work buffers go back to the pool
}

cp_status cp_assemble(cp_geom* g, cp_prec prec) {
    if (!g || (prec != CP_FP64 && prec != CP_FP32_REFINE && prec != CP_FP32_REFINE64)) return CP_BAD_ARG;
    if (prec == CP_FP32_REFINE64) {  // both sets: FP64 for refinement, FP32 for the Krylov iterations
        cp_status st = cp_assemble(g, CP_FP64);
        return st == CP_OK ? cp_assemble(g, CP_FP32_REFINE) : st;
    }
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&] {
        ensure_wake_influence(g);
        const int64_t lda = g->dev.lda;
        const size_t elems = (size_t)lda * g->dev.n;
        const double* W = g->W.as<double>();
        if (prec == CP_FP64) {
            if (!g->have64) {
                dalloc(c, g->A64, elems * sizeof(double));
                dalloc(c, g->B64, elems * sizeof(double));
            }
            {
                TimedScope ts("assemble_f64", g->stream);
                launch_assemble_f64(g->dev, W, lda, g->A64.as<double>(), g->B64.as<double>(), g->stream);
            }
            g->have64 = true;
        } else {
            if (!g->have32) {
                dalloc(c, g->A32, elems * sizeof(float));
                dalloc(c, g->B32, elems * sizeof(float));
            }
            {
                TimedScope ts("assemble_f32", g->stream);
                launch_assemble_f32(g->dev, W, lda, g->A32.as<float>(), g->B32.as<float>(), g->stream);
            }
            if (!g->near_rowptr.p) build_near_lists(g);
            {
                TimedScope ts("near_f64", g->stream);
                launch_near_correction(g->dev, W, lda, g->A32.as<float>(), g->B32.as<float>(), g->near_nnz,
                                       g->near_rows.as<int>(), g->near_cols.as<int>(), g->near_dA.as<double>(),
                                       g->near_dB.as<double>(), g->stream);
            }
            if (!g->A16.p) dalloc(c, g->A16, packed_a_bytes(lda, g->dev.n));
            launch_pack_a(g->A32.as<float>(), lda, g->dev.n, g->A16.p, g->stream);
            g->have32 = true;
            g->bj_ready = false;  // inverses follow the new A32
        }
        CP_HIP(hipStreamSynchronize(g->stream));
        return CP_OK;
    });
}

cp_status cp_load_matrices(cp_geom* g, const double* A, const double* B, int64_t ld) {
    if (!g || !A || !B || ld < g->dev.n) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&] {
        const int n = g->dev.n;
        const int64_t lda = g->dev.lda;
        const size_t elems = (size_t)lda * n;
        if (!g->have64) {
            dalloc(c, g->A64, elems * sizeof(double));
            dalloc(c, g->B64, elems * sizeof(double));
        }
        // rows n..lda of each column stay zero, as the assembly leaves them
        CP_HIP(hipMemsetAsync(g->A64.p, 0, elems * sizeof(double), g->stream));
        CP_HIP(hipMemsetAsync(g->B64.p, 0, elems * sizeof(double), g->stream));
        CP_HIP(hipMemcpy2DAsync(g->A64.p, lda * sizeof(double), A, ld * sizeof(double), n * sizeof(double), n,
                                hipMemcpyHostToDevice, g->stream));
        CP_HIP(hipMemcpy2DAsync(g->B64.p, lda * sizeof(double), B, ld * sizeof(double), n * sizeof(double), n,
                                hipMemcpyHostToDevice, g->stream));
        CP_HIP(hipStreamSynchronize(g->stream));
        g->have64 = true;
        if (g->batched) g->batched->invalidate_rhs_cache();  // its -B n columns came from the old B
        return CP_OK;
    });
}

cp_status cp_read_matrices(cp_geom* g, cp_prec prec, double* A, double* B, int64_t ld) {
    if (!g || ld < g->dev.n) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&] {
        const int n = g->dev.n;
        const int64_t lda = g->dev.lda;
        if (prec == CP_FP64) {
            if (!g->have64) return fail(c, CP_NOT_ASSEMBLED, "FP64 matrices not assembled");
            if (A)
                CP_HIP(hipMemcpy2DAsync(A, ld * sizeof(double), g->A64.p, lda * sizeof(double), n * sizeof(double), n,
                                        hipMemcpyDeviceToHost, g->stream));
            if (B)
                CP_HIP(hipMemcpy2DAsync(B, ld * sizeof(double), g->B64.p, lda * sizeof(double), n * sizeof(double), n,
                                        hipMemcpyDeviceToHost, g->stream));
            CP_HIP(hipStreamSynchronize(g->stream));
        } else {
            if (!g->have32) return fail(c, CP_NOT_ASSEMBLED, "FP32 matrices not assembled");
            std::vector<float> tmp((size_t)n * n);
            const void* src[2] = {g->A32.p, g->B32.p};
            double* dst[2] = {A, B};
            for (int q = 0; q < 2; q++) {
                if (!dst[q]) continue;
                CP_HIP(hipMemcpy2DAsync(tmp.data(), n * sizeof(float), src[q], lda * sizeof(float), n * sizeof(float),
                                        n, hipMemcpyDeviceToHost, g->stream));
                CP_HIP(hipStreamSynchronize(g->stream));
                for (int j = 0; j < n; j++)
                    for (int i = 0; i < n; i++) dst[q][i + (int64_t)j * ld] = tmp[i + (size_t)j * n];
            }
        }
        return CP_OK;
    });
}

cp_status cp_solve(cp_geom* g, cp_prec prec, int32_t k, const cp_freestream* fs, const cp_solve_opts* opts,
                   double* mu, double* cpout, cp_solve_info* info) {
    if (!g || !fs || !opts || k <= 0) return CP_BAD_ARG;
    if (prec != CP_FP64 && prec != CP_FP32_REFINE && prec != CP_FP32_REFINE64) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&]() -> cp_status {
        if (prec == CP_FP64 && !g->have64) return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP64) first");
        if (prec == CP_FP32_REFINE && !g->have32)
            return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP32_REFINE) first");
        if (prec == CP_FP32_REFINE64 && !(g->have32 && g->have64))
            return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP32_REFINE64) first");
        const int n = g->dev.n;
        const int64_t lda = g->dev.lda;
        hipStream_t s = g->stream;
        const double tol = opts->tol > 0 ? opts->tol : 1e-10;
        std::vector<double> T(9 * (size_t)k), vinf4(4 * (size_t)k);
        for (int q = 0; q < k; q++) {
            double Vb[3];
            wind_to_body(fs[q], &T[9 * (size_t)q], Vb);
            vinf4[4 * q] = Vb[0];
            vinf4[4 * q + 1] = Vb[1];
            vinf4[4 * q + 2] = Vb[2];
            vinf4[4 * q + 3] = fs[q].V;
        }
        DevBuf dv4, dmu, dcp, dpart, dres;
        upload(c, dv4, vinf4);
        dalloc(c, dmu, (size_t)lda * k * sizeof(double));
        std::vector<SolveStats> stats(k);
        if (prec == CP_FP64) {
            if (!g->bicg) g->bicg.reset(new Bicgstab64(n, lda));
            DevBuf dsig, drhs;
            dalloc(c, dsig, (size_t)lda * sizeof(double));
            dalloc(c, drhs, (size_t)lda * sizeof(double));
            for (int q = 0; q < k; q++) {
                // sigma_j = -Vinf . n_j  (bodyPanel::setSigma, bodyPanel.cpp:97-100)
                launch_sigma(n, lda, g->normals.as<double>(), dv4.as<double>() + 4 * q, dsig.as<double>(), s);
                // RHS = -B sigma
                g->bicg->gemv().apply(g->B64.as<double>(), dsig.as<double>(), drhs.as<double>(), -1.0,
                                      g->bicg->gemv_scratch(), s);
                stats[q] = g->bicg->solve(g->A64.as<double>(), drhs.as<double>(), dmu.as<double>() + lda * q, tol,
                                          opts->max_iter, s);
            }
        } else {
            prepare_mixed(g, prec, opts);
            const BatchedOpts bo = mixed_opts(opts, n);
            g->batched->solve(g->A32.as<float>(), g->B32.as<float>(), g->normals.as<double>(), dv4.as<double>(), k,
                              bo, dmu.as<double>(), lda, stats.data(), s);
        }
        // the post-processing tables (built on a host thread since the
        // geometry's upload) are joined only now, so that a new geometry's
        // table build overlaps its block inversion and first solve
        ensure_post(g);
        int nblk = post_partial_blocks(n);
        dalloc(c, dpart, (size_t)nblk * k * 6 * sizeof(double));
        dalloc(c, dres, (size_t)16 * k * sizeof(double));
        if (cpout) dalloc(c, dcp, (size_t)lda * k * sizeof(double));
        PostParams pp;
        pp.PG = std::sqrt(1 - opts->mach * opts->mach);
        pp.Sref = opts->Sref;
        pp.bref = opts->bref;
        pp.cref = opts->cref;
        for (int a = 0; a < 3; a++) pp.cg[a] = opts->cg[a];
        launch_post(g->post, pp, k, dmu.as<double>(), lda, dv4.as<double>(), cpout ? dcp.as<double>() : nullptr, lda,
                    dpart.as<double>(), nblk, dres.as<double>(), s);
        std::vector<double> res(16 * (size_t)k);
        CP_HIP(hipMemcpyAsync(res.data(), dres.p, res.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        if (mu)
            CP_HIP(hipMemcpy2DAsync(mu, n * sizeof(double), dmu.p, lda * sizeof(double), n * sizeof(double), k,
                                    hipMemcpyDeviceToHost, s));
        if (cpout)
            CP_HIP(hipMemcpy2DAsync(cpout, n * sizeof(double), dcp.p, lda * sizeof(double), n * sizeof(double), k,
                                    hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        bool all_conv = true;
        for (int q = 0; q < k; q++) {
            if (info) {
                std::memset(&info[q], 0, sizeof(cp_solve_info));
                fill_info(&res[16 * (size_t)q], &T[9 * (size_t)q], info[q]);
                info[q].rel_residual = stats[q].rel_residual;
                info[q].iters = stats[q].iters;
                info[q].refinements = stats[q].refinements;
                info[q].converged = stats[q].rel_residual <= tol;
            }
            all_conv = all_conv && stats[q].rel_residual <= tol;
        }
        return all_conv ? CP_OK : CP_NOT_CONVERGED;
    });
}

cp_status cp_solve_group(int32_t G, cp_geom* const* geoms, cp_prec prec, const int32_t* k,
                         const cp_freestream* const* fs, const cp_solve_opts* opts, double* const* mu,
                         double* const* cpout, cp_solve_info* const* info) {
    if (G <= 0 || G > kMaxGroup || !geoms || !k || !fs || !opts) return CP_BAD_ARG;
    if (prec != CP_FP32_REFINE && prec != CP_FP32_REFINE64) return CP_BAD_ARG;
    for (int i = 0; i < G; i++) {
        if (!geoms[i] || !fs[i] || k[i] <= 0 || k[i] > kRhsTile) return CP_BAD_ARG;
        if (geoms[i]->ctx != geoms[0]->ctx) return CP_BAD_ARG;
        for (int j = 0; j < i; j++)
            if (geoms[j] == geoms[i]) return CP_BAD_ARG;
    }
    cp_ctx* c = geoms[0]->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    // every geometry's lock, taken in address order (no deadlock with other
    // groups or single calls)
    std::vector<cp_geom*> order(geoms, geoms + G);
    std::sort(order.begin(), order.end());
    std::vector<std::unique_lock<std::mutex>> locks;
    for (cp_geom* g : order) locks.emplace_back(g->mu);
    return guarded_impl(c, nullptr, [&]() -> cp_status {
        for (int i = 0; i < G; i++) {
            cp_geom* g = geoms[i];
            if (!g->have32 || (prec == CP_FP32_REFINE64 && !g->have64))
                return fail(c, CP_NOT_ASSEMBLED, "cp_solve_group: a geometry is not assembled for this precision");
        }
        // the group runs on the first geometry's stream; the others' set-up
        // work (ensure_*) is on their own streams, finished before the group
        hipStream_t s = geoms[0]->stream;
        std::vector<std::vector<double>> T(G), res(G);
        std::vector<DevBuf> dv4(G), dmu(G), dpart(G), dres(G), dcp(G);
        std::vector<std::vector<SolveStats>> stats(G);
        std::vector<GroupItem> items(G);
        std::vector<BatchedOpts> bo(G);
        for (int i = 0; i < G; i++) {
            cp_geom* g = geoms[i];
            prepare_mixed(g, prec, &opts[i]);
            CP_HIP(hipStreamSynchronize(g->stream));
            bo[i] = mixed_opts(&opts[i], g->dev.n);
            std::vector<double> vinf4(4 * (size_t)k[i]);
            T[i].resize(9 * (size_t)k[i]);
            for (int q = 0; q < k[i]; q++) {
                double Vb[3];
                wind_to_body(fs[i][q], &T[i][9 * (size_t)q], Vb);
                vinf4[4 * q] = Vb[0];
                vinf4[4 * q + 1] = Vb[1];
                vinf4[4 * q + 2] = Vb[2];
                vinf4[4 * q + 3] = fs[i][q].V;
            }
            upload(c, dv4[i], vinf4);
            dalloc(c, dmu[i], (size_t)g->dev.lda * k[i] * sizeof(double));
            stats[i].resize(k[i]);
            items[i] = GroupItem{g->batched.get(), g->A32.as<float>(), g->B32.as<float>(), g->normals.as<double>(),
                                 dv4[i].as<double>(), k[i], dmu[i].as<double>(), g->dev.lda, stats[i].data(),
                                 &bo[i]};
        }
        DevBuf gbuf;
        dalloc(c, gbuf, BatchedSolver::group_bytes(G));
        BatchedSolver::solve_group(items, gbuf.p, s);
        for (int i = 0; i < G; i++) ensure_post(geoms[i]);  // after the solve, as in cp_solve
        for (int i = 0; i < G; i++) {
            cp_geom* g = geoms[i];
            const int n = g->dev.n;
            const bool want_cp = cpout && cpout[i];
            res[i].resize(16 * (size_t)k[i]);
            launch_post_k(g, &opts[i], k[i], dmu[i].as<double>(), dv4[i].as<double>(), dpart[i], dres[i],
                          want_cp ? &dcp[i] : nullptr, res[i].data(), s);
            if (mu && mu[i])
                CP_HIP(hipMemcpy2DAsync(mu[i], n * sizeof(double), dmu[i].p, g->dev.lda * sizeof(double),
                                        n * sizeof(double), k[i], hipMemcpyDeviceToHost, s));
            if (want_cp)
                CP_HIP(hipMemcpy2DAsync(cpout[i], n * sizeof(double), dcp[i].p, g->dev.lda * sizeof(double),
                                        n * sizeof(double), k[i], hipMemcpyDeviceToHost, s));
        }
        CP_HIP(hipStreamSynchronize(s));
        bool all_conv = true;
        for (int i = 0; i < G; i++) {
            const double tol = opts[i].tol > 0 ? opts[i].tol : 1e-10;
            for (int q = 0; q < k[i]; q++) {
                const SolveStats& st = stats[i][q];
                if (info && info[i]) {
                    cp_solve_info& inf = info[i][q];
                    std::memset(&inf, 0, sizeof(cp_solve_info));
                    fill_info(&res[i][16 * (size_t)q], &T[i][9 * (size_t)q], inf);
                    inf.rel_residual = st.rel_residual;
                    inf.iters = st.iters;
                    inf.refinements = st.refinements;
                    inf.converged = st.rel_residual <= tol;
                }
                all_conv = all_conv && st.rel_residual <= tol;
            }
        }
        return all_conv ? CP_OK : CP_NOT_CONVERGED;
    });
}

cp_status cp_matvec(cp_geom* g, int kernel, int32_t k, const double* X, double* Y) {
    if (!g || !X || !Y || k <= 0 || kernel < 0 || kernel > 4) return CP_BAD_ARG;
    if (kernel > 0 && k > kRhsTile) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&]() -> cp_status {
        const int n = g->dev.n;
        const int64_t lda = g->dev.lda;
        hipStream_t s = g->stream;
        if (kernel == 0) {
            if (!g->have64) return fail(c, CP_NOT_ASSEMBLED, "FP64 matrices not assembled");
            if (!g->bicg) g->bicg.reset(new Bicgstab64(n, lda));
            DevBuf dx, dy;
            dalloc(c, dx, (size_t)lda * sizeof(double));
            dalloc(c, dy, (size_t)lda * sizeof(double));
            CP_HIP(hipMemsetAsync(dx.p, 0, (size_t)lda * sizeof(double), s));
            for (int q = 0; q < k; q++) {
                CP_HIP(hipMemcpyAsync(dx.p, X + (size_t)q * n, n * sizeof(double), hipMemcpyHostToDevice, s));
                g->bicg->gemv().apply(g->A64.as<double>(), dx.as<double>(), dy.as<double>(), 1.0,
                                      g->bicg->gemv_scratch(), s);
                CP_HIP(hipMemcpyAsync(Y + (size_t)q * n, dy.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
            }
            CP_HIP(hipStreamSynchronize(s));
            return CP_OK;
        }
        if (!g->have32) return fail(c, CP_NOT_ASSEMBLED, "FP32 matrices not assembled");
        if (kernel == 4 && !g->have64) return fail(c, CP_NOT_ASSEMBLED, "FP64 matrices not assembled");
        ensure_batched(g);
        g->batched->set_near(g->near_rowptr.as<int>(), g->near_cols.as<int>(), g->near_dA.as<double>(),
                             g->near_dB.as<double>());
        g->batched->set_packed(g->A16.p, packed_kblocks(n));
        g->batched->set_near_tiles(g->near_tptr.as<int>(), g->near_tkb.as<int>());
        // row-major lda x 64 multivectors
        const size_t elems = (size_t)lda * kRhsTile;
        if (kernel == 1) {
            std::vector<float> xr(elems, 0.f), yr(elems);
            for (int q = 0; q < k; q++)
                for (int i = 0; i < n; i++) xr[(size_t)i * kRhsTile + q] = (float)X[(size_t)q * n + i];
            DevBuf dx, dy;
            upload(c, dx, xr);
            dalloc(c, dy, elems * sizeof(float));
            g->batched->matvec32(g->A32.as<float>(), dx.as<float>(), dy.as<float>(), s);
            CP_HIP(hipMemcpyAsync(yr.data(), dy.p, elems * sizeof(float), hipMemcpyDeviceToHost, s));
            CP_HIP(hipStreamSynchronize(s));
            for (int q = 0; q < k; q++)
                for (int i = 0; i < n; i++) Y[(size_t)q * n + i] = yr[(size_t)i * kRhsTile + q];
        } else {
            std::vector<double> xr(elems, 0.0), yr(elems);
            for (int q = 0; q < k; q++)
                for (int i = 0; i < n; i++) xr[(size_t)i * kRhsTile + q] = X[(size_t)q * n + i];
            DevBuf dx, dy;
            upload(c, dx, xr);
            dalloc(c, dy, elems * sizeof(double));
            g->batched->matvec64acc(g->A32.as<float>(), dx.as<double>(), dy.as<double>(), nullptr, 1.0, s,
                                    kernel == 3 ? g->near_dA.as<double>() : nullptr,
                                    kernel == 4 ? g->A64.as<double>() : nullptr);
            CP_HIP(hipMemcpyAsync(yr.data(), dy.p, elems * sizeof(double), hipMemcpyDeviceToHost, s));
            CP_HIP(hipStreamSynchronize(s));
            for (int q = 0; q < k; q++)
                for (int i = 0; i < n; i++) Y[(size_t)q * n + i] = yr[(size_t)i * kRhsTile + q];
        }
        return CP_OK;
    });
}

cp_status cp_solver_report(cp_geom* g, char* buf, int32_t len) {
    if (!g || !buf || len <= 0) return CP_BAD_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    std::string r = "{";
    if (g->batched) {
        const GemmShape& sh = g->batched->shape();
        r += "\"gemm_shape\": {\"variant\": " + std::to_string(sh.variant) + ", \"rt\": " + std::to_string(sh.rt) +
             ", \"kslices\": " + std::to_string(sh.kslices) + "}, \"tune\": [";
        bool first = true;
        for (const TuneRecord& t : g->batched->tune_log()) {
            if (!first) r += ", ";
            first = false;
            r += "[" + std::to_string(t.shape.variant) + ", " + std::to_string(t.shape.rt) + ", " +
                 std::to_string(t.shape.kslices) + ", " + std::to_string(t.ms) + "]";
        }
        r += "], ";
    }
    if (g->batched) r += "\"block_jacobi_level\": " + std::to_string(g->batched->block_jacobi_level()) + ", ";
    r += "\"block_jacobi_blocks\": " + std::to_string(g->bj_plan.perm.empty() ? 0 : g->bj_plan.nblocks()) +
         ", \"block_jacobi_max\": " + std::to_string(g->bj_plan.max_b) + ", \"near_nnz\": " +
         std::to_string(g->near_nnz) + "}";
    std::snprintf(buf, (size_t)len, "%s", r.c_str());
    return (int32_t)r.size() < len ? CP_OK : CP_BAD_ARG;
}

cp_status cp_case_fields_eval(cp_geom* g, int32_t k, const cp_freestream* fs, const cp_solve_opts* opts,
                              const double* mu, cp_case_fields* out) {
    if (!g || !fs || !opts || !mu || !out || k <= 0) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&]() -> cp_status {
        ensure_post(g);
        const int n = g->dev.n, nw = g->dev.nw;
        const int64_t lda = g->dev.lda;
        hipStream_t s = g->stream;
        std::vector<double> vinf4(4 * (size_t)k);
        for (int q = 0; q < k; q++) {
            double T[9], Vb[3];
            wind_to_body(fs[q], T, Vb);
            for (int a = 0; a < 3; a++) vinf4[4 * q + a] = Vb[a];
            vinf4[4 * q + 3] = fs[q].V;
        }
        DevBuf dv4, dmu, dpart, dres, dvel, dphi, dcp, dspan, dwmu, dwphi;
        upload(c, dv4, vinf4);
        dalloc(c, dmu, (size_t)lda * k * sizeof(double));
        CP_HIP(hipMemcpy2DAsync(dmu.p, lda * sizeof(double), mu, n * sizeof(double), n * sizeof(double), k,
                                hipMemcpyHostToDevice, s));
        const int nblk = post_partial_blocks(n);
        dalloc(c, dpart, (size_t)nblk * k * 6 * sizeof(double));
        dalloc(c, dres, (size_t)16 * k * sizeof(double));
        const PostHost& H = g->post_host;
        const size_t nspan = (size_t)H.n_wakes * H.n_st;
        PostFields F;
        if (out->velocity) {
            dalloc(c, dvel, (size_t)3 * n * k * sizeof(double));
            F.vel = dvel.as<double>();
            F.ldvel = 3 * (int64_t)n;
        }
        if (out->potential) {
            dalloc(c, dphi, (size_t)n * k * sizeof(double));
            F.phi = dphi.as<double>();
            F.ldphi = n;
        }
        if (out->cp) dalloc(c, dcp, (size_t)n * k * sizeof(double));
        DevBuf dsig;
        if (out->sigma) {
            dalloc(c, dsig, (size_t)lda * k * sizeof(double));
            for (int q = 0; q < k; q++)
                launch_sigma(n, lda, g->normals.as<double>(), dv4.as<double>() + 4 * q, dsig.as<double>() + lda * q, s);
        }
        if (out->spanwise && nspan) {
            dalloc(c, dspan, nspan * 2 * k * sizeof(double));
            F.span = dspan.as<double>();
        }
        PostParams pp;
        pp.PG = std::sqrt(1 - opts->mach * opts->mach);
        pp.Sref = opts->Sref;
        pp.bref = opts->bref;
        pp.cref = opts->cref;
        for (int a = 0; a < 3; a++) pp.cg[a] = opts->cg[a];
        launch_post(g->post, pp, k, dmu.as<double>(), lda, dv4.as<double>(), out->cp ? dcp.as<double>() : nullptr, n,
                    dpart.as<double>(), nblk, dres.as<double>(), s, &F);
        if ((out->wake_mu || out->wake_potential) && nw > 0) {
            if (!g->wf_idx.p) {
                const Geometry& G = g->host;
                std::vector<int32_t> idx(G.fold_idx.begin(), G.fold_idx.end());
                std::vector<double> ctr(3 * (size_t)nw);
                for (int j = 0; j < nw; j++) {
                    ctr[j] = G.wake[j].center.x;
                    ctr[nw + j] = G.wake[j].center.y;
                    ctr[2 * (size_t)nw + j] = G.wake[j].center.z;
                }
                upload(c, g->wf_idx, idx);
                upload(c, g->wf_t, G.fold_coef);
                upload(c, g->wf_ctr, ctr);
            }
            if (out->wake_mu) dalloc(c, dwmu, (size_t)nw * k * sizeof(double));
            if (out->wake_potential) dalloc(c, dwphi, (size_t)nw * k * sizeof(double));
            launch_post_wake(nw, g->wf_idx.as<int32_t>(), g->wf_t.as<double>(), g->wf_ctr.as<double>(), k,
                             dmu.as<double>(), lda, dv4.as<double>(), out->wake_mu ? dwmu.as<double>() : nullptr,
                             out->wake_potential ? dwphi.as<double>() : nullptr, s);
        }
        auto d2h = [&](double* dst, const DevBuf& src, size_t count) {
            if (dst && src.p) CP_HIP(hipMemcpyAsync(dst, src.p, count * sizeof(double), hipMemcpyDeviceToHost, s));
        };
        if (out->sigma)
            CP_HIP(hipMemcpy2DAsync(out->sigma, n * sizeof(double), dsig.p, lda * sizeof(double), n * sizeof(double), k,
                                    hipMemcpyDeviceToHost, s));
        d2h(out->velocity, dvel, (size_t)3 * n * k);
        d2h(out->potential, dphi, (size_t)n * k);
        d2h(out->cp, dcp, (size_t)n * k);
        d2h(out->wake_mu, dwmu, (size_t)nw * k);
        d2h(out->wake_potential, dwphi, (size_t)nw * k);
        std::vector<double> span(nspan * 2 * k);
        if (out->spanwise && nspan) d2h(span.data(), dspan, span.size());
        CP_HIP(hipStreamSynchronize(s));
        if (out->spanwise)
            for (int q = 0; q < k; q++)
                for (size_t e = 0; e < nspan; e++) {
                    double* o = out->spanwise + ((size_t)q * nspan + e) * 3;
                    o[0] = H.yloc[e];
                    o[1] = span[((size_t)q * nspan + e) * 2];
                    o[2] = span[((size_t)q * nspan + e) * 2 + 1];
                }
        return CP_OK;
    });
}

cp_status cp_block_jacobi_read(cp_geom* g, int64_t* sizes, int32_t* ptr, int32_t* perm, float* inv) {
    if (!g || !sizes) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded_g(g, [&] {
        if (!g->have32) return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP32_REFINE) first");
        ensure_block_jacobi(g);
        const BlockPlan& P = g->bj_plan;
        const bool on = !P.perm.empty();
        sizes[0] = on ? P.nblocks() : 0;
        sizes[1] = on ? (int64_t)P.perm.size() : 0;
        sizes[2] = on ? P.off.back() : 0;
        if (!on) return CP_OK;
        if (ptr) std::copy(P.ptr.begin(), P.ptr.end(), ptr);
        if (perm) std::copy(P.perm.begin(), P.perm.end(), perm);
        if (inv) {
            CP_HIP(hipMemcpyAsync(inv, g->bj_inv.p, (size_t)P.off.back() * sizeof(float), hipMemcpyDeviceToHost,
                                  g->stream));
            CP_HIP(hipStreamSynchronize(g->stream));
        }
        return CP_OK;
    });
}

cp_status cp_timing_reset(cp_ctx* c) {
    if (!c) return CP_BAD_ARG;
    return guarded(c, [&] {
        c->timer.reset();
        return CP_OK;
    });
}

cp_status cp_timing_enable(cp_ctx* c, int on) {
    if (!c) return CP_BAD_ARG;
    c->timer.enabled = on != 0;
    return CP_OK;
}

cp_status cp_timing_only(cp_ctx* c, const char* kernel) {
    if (!c) return CP_BAD_ARG;
    c->timer.set_only(kernel);
    return CP_OK;
}

cp_status cp_timing_get(cp_ctx* c, const char* kernel, double* total_ms, int64_t* launches) {
    if (!c || !kernel) return CP_BAD_ARG;
    return guarded(c, [&] {
        double ms = 0;
        long long nl = 0;
        c->timer.collect();
        c->timer.get(kernel, ms, nl);
        if (total_ms) *total_ms = ms;
        if (launches) *launches = nl;
        return CP_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- VP kernels

namespace cpgpu {
namespace capi {

void build_tree_host(cp_ctx* c, int n, const double* xyz, const double* alpha, const double* radius,
                     int max_members, cp_tree* t) {
    build_bh_tree(n, xyz, alpha, radius, max_members, t->host);
    t->dev.n = n;
    t->dev.n_nodes = (int)t->host.nodes.size();
    if (c->device < 0) return;  // host-only context: tree queries only
    upload(c, t->nodes, t->host.nodes);
    upload(c, t->mpos, t->host.mpos);
    upload(c, t->mstr, t->host.mstr);
    upload(c, t->mrad, t->host.mrad);
    upload(c, t->order, t->host.order);
    upload(c, t->slot, t->host.slot);
    t->dev.nodes = t->nodes.as<BhNode>();
    t->dev.mpos = t->mpos.as<double>();
    t->dev.mstr = t->mstr.as<double>();
    t->dev.mrad = t->mrad.as<double>();
    t->dev.order = t->order.as<int>();
}

}  // namespace capi
}  // namespace cpgpu

cp_status cp_bh_build(cp_ctx* c, int32_t n, const double* xyz, const double* alpha, const double* radius,
                      const cp_bh_opts* opts, cp_tree** out) {
    if (!c || !out || n < 0 || (n > 0 && (!xyz || !alpha || !radius))) return CP_BAD_ARG;
    *out = nullptr;
    int maxm = (opts && opts->max_members > 0) ? opts->max_members : 10;
    double theta = (opts && opts->theta >= 0) ? opts->theta : 0.5;
    const int kernel = opts ? opts->kernel : CP_VP_REFERENCE;
    if (kernel != CP_VP_REFERENCE && kernel != CP_VP_ALGEBRAIC) return CP_BAD_ARG;
    if (kernel == CP_VP_ALGEBRAIC && !(opts->core > 0)) return CP_BAD_ARG;
    const bool host_build = opts && opts->host_build != 0;
    std::unique_ptr<cp_tree> t(new cp_tree);
    t->ctx = c;
    t->theta = theta;
    t->kernel = kernel;
    if (kernel == CP_VP_ALGEBRAIC) {
        t->core = opts->core;
        t->fp32 = opts->fp32 != 0;
    }
    cp_status st = guarded(c, [&]() -> cp_status {
        if (c->device >= 0 && !host_build && n > 0) {
            DevBuf dx, da, dr;
            dalloc(c, dx, (size_t)3 * n * sizeof(double));
            dalloc(c, da, (size_t)3 * n * sizeof(double));
            dalloc(c, dr, (size_t)n * sizeof(double));
            CP_HIP(hipMemcpyAsync(dx.p, xyz, (size_t)3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            CP_HIP(hipMemcpyAsync(da.p, alpha, (size_t)3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            CP_HIP(hipMemcpyAsync(dr.p, radius, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            TimedScope ts("bh_build", c->stream);
            const bool ok = build_tree_device(c, n, dx.as<double>(), da.as<double>(), dr.as<double>(), maxm,
                                              c->stream, t.get());
            CP_HIP(hipStreamSynchronize(c->stream));
            if (ok) return CP_OK;
        }
        build_tree_host(c, n, xyz, alpha, radius, maxm, t.get());
        return CP_OK;
    });
    if (st == CP_OK) {
        ctx_ref(c);
        *out = t.release();
    }
    return st;
}

cp_status cp_bh_info(const cp_tree* t, int32_t* n_nodes, int32_t* levels, double* origin, double* half) {
    if (!t) return CP_BAD_ARG;
    if (n_nodes) *n_nodes = (int32_t)t->dev.n_nodes;
    if (levels) *levels = t->host.levels;
    if (origin)
        for (int d = 0; d < 3; d++) origin[d] = t->host.origin[d];
    if (half) *half = t->host.half;
    return CP_OK;
}

cp_status cp_bh_read(const cp_tree* t, double* nodes, int32_t* links, int32_t* order) {
    if (!t) return CP_BAD_ARG;
    cp_ctx* c = t->ctx;
    const int nn = t->dev.n_nodes, n = t->dev.n;
    return guarded(c, [&]() -> cp_status {
        std::vector<BhNode> nd;
        std::vector<int32_t> od;
        if (!t->host.nodes.empty() || nn == 0) {
            nd = t->host.nodes;
            od = t->host.order;
        } else {
            nd.resize(nn);
            od.resize(n);
            CP_HIP(hipMemcpyAsync(nd.data(), t->nodes.p, (size_t)nn * sizeof(BhNode), hipMemcpyDeviceToHost,
                                  c->stream));
            CP_HIP(hipMemcpyAsync(od.data(), t->order.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  c->stream));
            CP_HIP(hipStreamSynchronize(c->stream));
        }
        for (int i = 0; i < nn; i++) {
            const BhNode& b = nd[i];
            if (nodes) {
                const double v[8] = {b.ex, b.ey, b.ez, b.width, b.sx, b.sy, b.sz, b.radius};
                std::memcpy(nodes + 8 * (size_t)i, v, sizeof(v));
            }
            if (links) {
                links[3 * i] = b.skip;
                links[3 * i + 1] = b.first;
                links[3 * i + 2] = b.count;
            }
        }
        if (order) std::copy(od.begin(), od.end(), order);
        return CP_OK;
    });
}

cp_status cp_bh_eval(cp_tree* t, cp_bh_mode mode, int32_t m, const double* pts, double* out) {
    if (!t || !out || mode < CP_BH_VEL_POINTS || mode > CP_BH_DIFF) return CP_BAD_ARG;
    if (mode == CP_BH_VEL_POINTS && (m < 0 || (m > 0 && !pts))) return CP_BAD_ARG;
    const bool alg = t->kernel == CP_VP_ALGEBRAIC;
    if (alg && mode == CP_BH_DIFF) return fail(t->ctx, CP_BAD_ARG, "the algebraic kernel has no diffusion term");
    cp_ctx* c = t->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    return guarded(c, [&]() -> cp_status {
        hipStream_t s = c->stream;
        const int nt = mode == CP_BH_VEL_POINTS ? m : t->dev.n;
        if (nt == 0) return CP_OK;
        DevBuf dout, dpts;
        dalloc(c, dout, (size_t)3 * nt * sizeof(double));
        if (t->dev.n_nodes <= 1) {
            // root alone (a leaf, or no particles): the sums start at the
            // root's children, so every result is zero
            std::memset(out, 0, (size_t)3 * nt * sizeof(double));
            return CP_OK;
        }
        if (mode == CP_BH_VEL_POINTS) {
            dalloc(c, dpts, (size_t)3 * m * sizeof(double));
            CP_HIP(hipMemcpyAsync(dpts.p, pts, (size_t)3 * m * sizeof(double), hipMemcpyHostToDevice, s));
            TimedScope ts("bh_eval", s);
            if (alg)
                launch_core_bh(t->dev, CP_BH_VEL, t->theta, t->core, t->fp32, m, dpts.as<double>(), nullptr, nullptr,
                               nullptr, dout.as<double>(), s);
            else
                launch_bh_points(t->dev, t->theta, m, dpts.as<double>(), dout.as<double>(), s);
        } else if (alg) {
            TimedScope ts("bh_eval", s);
            launch_core_bh(t->dev, (int)mode, t->theta, t->core, t->fp32, t->dev.n, t->dev.mpos, t->dev.mstr,
                           nullptr, t->dev.order, dout.as<double>(), s);
        } else {
            TimedScope ts("bh_eval", s);
            launch_bh_particles(t->dev, (int)mode, t->theta, dout.as<double>(), s);
        }
        CP_HIP(hipMemcpyAsync(out, dout.p, (size_t)3 * nt * sizeof(double), hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        return CP_OK;
    });
}

void cp_bh_release(cp_tree* t) {
    if (!t) return;
    cp_ctx* c = t->ctx;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->device >= 0) (void)hipSetDevice(c->device);
        delete t;
    }
    ctx_unref(c);
}

cp_status cp_panel_to_points(cp_geom* g, const double* sigma, const double* mu, const double* mu_wake, int32_t m,
                             const double* pts, double* vel, int32_t fp32) {
    if (!g || !sigma || !mu || m < 0 || (m > 0 && (!pts || !vel))) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    if (m == 0) return CP_OK;
    return guarded_g(g, [&]() -> cp_status {
        hipStream_t s = g->stream;
        const int nb = g->dev.n, nw = g->dev.nw;
        std::vector<double> str((size_t)nb * 2 + std::max(nw, 1), 0.0);
        std::memcpy(str.data(), sigma, nb * sizeof(double));
        std::memcpy(str.data() + nb, mu, nb * sizeof(double));
        if (mu_wake && nw > 0) std::memcpy(str.data() + 2 * nb, mu_wake, nw * sizeof(double));
        DevBuf dstr, dpts, dpart, dout;
        upload(c, dstr, str);
        dalloc(c, dpts, (size_t)3 * m * sizeof(double));
        dalloc(c, dpart, p2p_part_doubles(m, nb + nw) * sizeof(double));
        dalloc(c, dout, (size_t)3 * m * sizeof(double));
        CP_HIP(hipMemcpyAsync(dpts.p, pts, (size_t)3 * m * sizeof(double), hipMemcpyHostToDevice, s));
        {
            TimedScope ts("panel_to_points", s);
            launch_panel_to_points(g->dev, dstr.as<double>(), dstr.as<double>() + nb, dstr.as<double>() + 2 * nb, m,
                                   dpts.as<double>(), dpart.as<double>(), dout.as<double>(), s, fp32 != 0);
        }
        CP_HIP(hipMemcpyAsync(vel, dout.p, (size_t)3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        return CP_OK;
    });
}

// ---------------------------------------------------------------- VP solve
cp_status cp_set_buffer_wake(cp_geom* g, int32_t m, const int32_t* nv, const double* verts) {
    if (!g || m < 0 || (m > 0 && (!nv || !verts))) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    for (int j = 0; j < m; j++)
        if (nv[j] != 3 && nv[j] != 4) return fail(c, CP_BAD_ARG, "buffer wake panels need 3 or 4 vertices");
    return guarded_g(g, [&]() -> cp_status {
        g->nw2 = m;
        if (m == 0) {
            g->Cw.release();
            g->w2.release();
            return CP_OK;
        }
        std::vector<PanelRec> rec(m);
        for (int j = 0; j < m; j++) {
            PanelGeom P;
            P.nv = nv[j];
            for (int q = 0; q < nv[j]; q++)
                P.v[q] = Vec3{verts[12 * j + 3 * q], verts[12 * j + 3 * q + 1], verts[12 * j + 3 * q + 2]};
            panel_setup(P);
            rec[j] = make_panel_rec(P);
        }
        upload(c, g->w2, rec);
        dalloc(c, g->Cw, (size_t)g->dev.lda * m * sizeof(double));
        // C(i, j) = dubPhiInf of buffer panel j at collocation point i
        // (geometry.cpp:876-882): the wake-influence kernel on these panels
        DevGeom d = g->dev;
        d.wake = g->w2.as<PanelRec>();
        d.nw = m;
        launch_wake_influence(d, g->Cw.as<double>(), g->dev.lda, g->stream);
        CP_HIP(hipStreamSynchronize(g->stream));
        return CP_OK;
    });
}

cp_status cp_solve_rhs(cp_geom* g, cp_prec prec, int32_t k, const double* sigma, const double* mu_w2,
                       const double* x0, const cp_solve_opts* opts, double* mu, cp_solve_info* info) {
    if (!g || !sigma || !opts || k <= 0 || (prec != CP_FP64 && prec != CP_FP32_REFINE && prec != CP_FP32_REFINE64))
        return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    if (mu_w2 && g->nw2 == 0) return fail(c, CP_BAD_ARG, "mu_w2 given but no buffer wake set");
    return guarded_g(g, [&]() -> cp_status {
        if (prec == CP_FP64 && !g->have64) return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP64) first");
        if (prec == CP_FP32_REFINE && !g->have32)
            return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP32_REFINE) first");
        if (prec == CP_FP32_REFINE64 && !(g->have32 && g->have64))
            return fail(c, CP_NOT_ASSEMBLED, "call cp_assemble(CP_FP32_REFINE64) first");
        const int n = g->dev.n;
        const int64_t lda = g->dev.lda;
        hipStream_t s = g->stream;
        const double tol = opts->tol > 0 ? opts->tol : 1e-10;
        // caller columns, padded to lda
        std::vector<double> hs((size_t)lda * k, 0.0), hx;
        for (int q = 0; q < k; q++) std::memcpy(&hs[(size_t)q * lda], sigma + (size_t)q * n, n * sizeof(double));
        DevBuf dsig, dadd, dx0, dmu, dmw;
        upload(c, dsig, hs);
        if (x0) {
            hx.assign((size_t)lda * k, 0.0);
            for (int q = 0; q < k; q++) std::memcpy(&hx[(size_t)q * lda], x0 + (size_t)q * n, n * sizeof(double));
            upload(c, dx0, hx);
        }
        if (mu_w2) {
            std::vector<double> mw(mu_w2, mu_w2 + (size_t)g->nw2 * k);
            upload(c, dmw, mw);
            dalloc(c, dadd, (size_t)lda * k * sizeof(double));
            CP_HIP(hipMemsetAsync(dadd.p, 0, (size_t)lda * k * sizeof(double), s));
            launch_buffer_rhs(n, lda, g->nw2, g->Cw.as<double>(), dmw.as<double>(), k, dadd.as<double>(), lda, s);
        }
        dalloc(c, dmu, (size_t)lda * k * sizeof(double));
        std::vector<SolveStats> stats(k);
        if (prec == CP_FP64) {
            if (!g->bicg) g->bicg.reset(new Bicgstab64(n, lda));
            DevBuf drhs;
            dalloc(c, drhs, (size_t)lda * sizeof(double));
            for (int q = 0; q < k; q++) {
                // RHS = -B sigma - C mu_w2 (cpCaseVP.cpp:272)
                g->bicg->gemv().apply(g->B64.as<double>(), dsig.as<double>() + lda * q, drhs.as<double>(), -1.0,
                                      g->bicg->gemv_scratch(), s);
                if (mu_w2) launch_add_vec(n, dadd.as<double>() + lda * q, drhs.as<double>(), s);
                stats[q] = g->bicg->solve(g->A64.as<double>(), drhs.as<double>(), dmu.as<double>() + lda * q, tol,
                                          opts->max_iter, s, nullptr, x0 ? dx0.as<double>() + lda * q : nullptr);
            }
        } else {
            ensure_batched(g);
            g->batched->set_near(g->near_rowptr.as<int>(), g->near_cols.as<int>(), g->near_dA.as<double>(),
                                 g->near_dB.as<double>());
            g->batched->set_packed(g->A16.p, packed_kblocks(n));
            g->batched->set_near_tiles(g->near_tptr.as<int>(), g->near_tkb.as<int>());
            if (prec == CP_FP32_REFINE64)
                g->batched->set_f64(g->A64.as<double>(), g->B64.as<double>());
            else
                g->batched->set_f64(nullptr, nullptr);
            if (opts->precond != 1) ensure_block_jacobi(g);
            if (opts->precond != 1 && !g->bj_plan.perm.empty())
                g->batched->set_block_jacobi(g->bj_perm.as<int>(), g->bj_ptr.as<int>(), g->bj_off.as<int64_t>(),
                                             g->bj_inv.as<float>(), g->bj_plan.nblocks(), g->bj_plan.max_b,
                                             g->bj_iperm.as<int>(), g->bj_pk.as<float>(),
                                             g->bj_poff.as<int64_t>(), g->bj_pk_kind);
            else
                g->batched->set_block_jacobi(nullptr, nullptr, nullptr, nullptr, 0, 0);
            BatchedOpts bo;
            bo.tol = tol;
            bo.max_iter = opts->max_iter > 0 ? opts->max_iter : 2 * n;
            bo.max_refine = opts->max_refine > 0 ? opts->max_refine : 20;
            bo.inner_tol = opts->inner_tol > 0 ? opts->inner_tol : 0;  // 0: adaptive
            g->batched->solve_rhs(g->A32.as<float>(), g->B32.as<float>(), dsig.as<double>(),
                                  mu_w2 ? dadd.as<double>() : nullptr, x0 ? dx0.as<double>() : nullptr, k, bo,
                                  dmu.as<double>(), lda, stats.data(), s);
        }
        if (mu)
            CP_HIP(hipMemcpy2DAsync(mu, n * sizeof(double), dmu.p, lda * sizeof(double), n * sizeof(double), k,
                                    hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        bool all_conv = true;
        for (int q = 0; q < k; q++) {
            if (info) {
                std::memset(&info[q], 0, sizeof(cp_solve_info));
                info[q].rel_residual = stats[q].rel_residual;
                info[q].iters = stats[q].iters;
                info[q].refinements = stats[q].refinements;
                info[q].converged = stats[q].rel_residual <= tol;
            }
            all_conv = all_conv && stats[q].rel_residual <= tol;
        }
        return all_conv ? CP_OK : CP_NOT_CONVERGED;
    });
}

cp_status cp_panel_stretching(cp_geom* g, const double* sigma, const double* mu, const double* mu_wake, int32_t m,
                              const double* pos, const double* alpha, double* out, int32_t fp32) {
    if (!g || !sigma || !mu || m < 0 || (m > 0 && (!pos || !alpha || !out))) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    if (m == 0) return CP_OK;
    return guarded_g(g, [&]() -> cp_status {
        hipStream_t s = g->stream;
        const int nb = g->dev.n, nw = g->dev.nw;
        std::vector<double> str((size_t)nb * 2 + std::max(nw, 1), 0.0);
        std::memcpy(str.data(), sigma, nb * sizeof(double));
        std::memcpy(str.data() + nb, mu, nb * sizeof(double));
        if (mu_wake && nw > 0) std::memcpy(str.data() + 2 * nb, mu_wake, nw * sizeof(double));
        std::vector<double> pa(6 * (size_t)m);
        std::memcpy(pa.data(), pos, 3 * (size_t)m * sizeof(double));
        std::memcpy(pa.data() + 3 * (size_t)m, alpha, 3 * (size_t)m * sizeof(double));
        DevBuf dstr, dpa, dpart, dout;
        upload(c, dstr, str);
        upload(c, dpa, pa);
        dalloc(c, dpart, p2p_part_doubles(m, nb + nw) * sizeof(double));
        dalloc(c, dout, (size_t)3 * m * sizeof(double));
        {
            TimedScope ts("panel_stretching", s);
            launch_panel_stretching(g->dev, dstr.as<double>(), dstr.as<double>() + nb, dstr.as<double>() + 2 * nb, m,
                                    dpa.as<double>(), dpa.as<double>() + 3 * (size_t)m, dpart.as<double>(),
                                    dout.as<double>(), s, fp32 != 0);
        }
        CP_HIP(hipMemcpyAsync(out, dout.p, (size_t)3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        return CP_OK;
    });
}

