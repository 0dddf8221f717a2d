// Internal objects of the C ABI (include/cpgpu.h), shared by its translation
// units: the context and geometry objects, pooled device buffers and the
// error-guard wrappers every entry point runs its body in.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/cpgpu.h"
#include "batched.h"
#include "device.h"
#include "geometry.h"
#include "krylov.h"
#include "panelrec.h"
#include "post.h"
#include "precond.h"
#include "timer.h"
#include "vp.h"
#include "vp_tree.h"

// the ABI objects are global-namespace types (include/cpgpu.h) built from
// the solver's types
using namespace cpgpu;

struct cp_ctx {
    // one reference for the caller's handle and one per live geometry, tree
    // and server: cp_destroy drops the handle's, and the context is freed
    // when the last child is released, whatever order a caller tears down in
    std::atomic<int> refs{1};
    int device = 0;
    uint64_t budget = 0;
    // device bytes this context lets itself hold (live + pooled): the free
    // memory at creation less a reserve for the runtime's own allocations
    // (kernel scratch, queues) and other users of the device.  Pooled blocks
    // are freed to stay under it even without a caller budget: a pool that
    // grew to the physical limit left the runtime unable to allocate a
    // queue's scratch, which aborts the queue.
    uint64_t cap = 0;
    uint64_t in_use = 0;          // device bytes held (live + pooled), under alloc_mu
    hipStream_t stream = nullptr; // context-level calls (Barnes-Hut trees)
    std::string err;              // last error, under err_mu
    int hip_err = 0;
    KernelTimer timer;
    std::mutex mu;                // serialises context-level calls
    std::mutex alloc_mu;
    std::mutex err_mu;
    // freed device blocks kept for reuse: a hipFree waits for the whole
    // device, which would stall the other geometries' streams
    std::multimap<size_t, void*> pool;
    uint64_t pooled = 0;
    // pinned host blocks of BatchedSolver::host_bytes() and idle
    // non-blocking streams of released geometries, reused for the same
    // reason (hipHostFree and stream teardown synchronise)
    std::vector<void*> pinned;
    std::vector<hipStream_t> streams;
    // a server runs on this context: its solvers are used from several host
    // threads at once and never capture graphs
    std::atomic<bool> served{false};
};

namespace cpgpu {
namespace capi {


struct DevBuf {
    cp_ctx* ctx = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    // returns the block to the context's pool; callers release only buffers
    // whose last use on the device has completed (every API call
    // synchronises its stream before its temporaries go)
    void release() {
        if (p) {
            std::lock_guard<std::mutex> lk(ctx->alloc_mu);
            ctx->pool.emplace(bytes, p);
            ctx->pooled += bytes;
            p = nullptr;
            bytes = 0;
        }
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct BudgetError {
    uint64_t need;
};

void trim_pool(cp_ctx* c, uint64_t need);
void dalloc(cp_ctx* c, DevBuf& b, size_t bytes);
void account(cp_ctx* c, int64_t delta);
bool fits(cp_ctx* c, uint64_t bytes);

template <class T>
void upload(cp_ctx* c, DevBuf& b, const std::vector<T>& v) {
    dalloc(c, b, std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty()) CP_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace capi
}  // namespace cpgpu

using cpgpu::capi::DevBuf;

struct cp_geom {
    cp_ctx* ctx = nullptr;
    // the caller's handle plus one per vortex-particle case built on it;
    // the geometry is freed when the last goes
    std::atomic<int> refs{1};
    uint64_t key = 0;
    hipStream_t stream = nullptr;  // every device call on this geometry
    std::mutex mu;                 // serialises calls on this geometry
    int64_t solver_bytes = 0;      // workspaces counted in ctx->in_use
    DevBuf solver_ws;              // BatchedSolver workspace
    void* solver_host = nullptr;   // BatchedSolver pinned host block (ctx pool)
    Geometry host;
    DevGeom dev;
    DevBuf body, wake, cxyz, fold_ptr, fold_w, fold_c, normals;
    DevBuf tsph;       // bounding spheres of the 64-row collocation tiles
    DevBuf W;          // wake influence N x Nw (FP64)
    DevBuf A64, B64, A32, B32;
    DevBuf A16;        // bf16 hi and lo tile planes of A32 (bf16x3 / bf16x2 inner products)
    // second buffer-wake row of the vortex-particle case: panels and C
    DevBuf w2, Cw;
    int nw2 = 0;
    bool have64 = false, have32 = false;
    // FP64 near-field corrections of A32/B32 (entries within one panel
    // diameter: |c_i - c_j| <= longSide_j)
    DevBuf near_rowptr, near_rows, near_cols, near_dA, near_dB;
    DevBuf near_tptr, near_tkb;  // the packed 32 x 16 tiles holding them (bf16x2 product)
    int near_nnz = 0;
    // strip block-Jacobi preconditioner of the batched FP32 inner solve
    BlockPlan bj_plan;
    DevBuf bj_perm, bj_ptr, bj_off, bj_inv, bj_iperm, bj_poff, bj_pk;
    bool bj_ready = false;
    int bj_pk_kind = kBjPackBf16;  // layout of bj_pk (launch_bj_pack)
    // post tables
    bool post_ready = false;
    PostHost post_host;
    PostDev post;
    DevBuf p_nbr, p_w1, p_w2, p_mode, p_te, p_rphi, p_rv, p_rw, p_simp, p_ctr, p_bez, p_area;
    std::unique_ptr<Bicgstab64> bicg;
    std::unique_ptr<BatchedSolver> batched;
    // wake fold tables for the wake output fields: U1, L1, U2, L2 and t per
    // wake panel, wake centres (SoA)
    DevBuf wf_idx, wf_t, wf_ctr;
    // host build of post_host, started at upload so that it runs beside the
    // assembly and the solve; declared last so that it is joined before the
    // members it writes are destroyed
    std::future<void> post_job;
};

struct cp_tree {
    cp_ctx* ctx = nullptr;
    BhTreeHost host;
    DevBhTree dev;
    DevBuf nodes, mpos, mstr, mrad, order;
    DevBuf slot;                   // particle index -> member slot (device)
    double theta = 0.5;
    int kernel = CP_VP_REFERENCE;  // cp_vp_kernel
    double core = 0;               // CP_VP_ALGEBRAIC core radius
    bool fp32 = false;
};

namespace cpgpu {
namespace capi {

cp_status fail(cp_ctx* c, cp_status s, const std::string& msg);
// Builds t's octree on the device from n particles in device memory (the
// tree build_bh_tree makes on the host, node for node, in t->nodes, mpos,
// mstr, mrad, order, slot, t->dev and t->host's levels / origin / half).
// false: a node would split below the deepest level its keys hold (or the
// node capacity ran out); nothing is built and the caller builds on the host.
bool build_tree_device(cp_ctx* c, int n, const double* dxyz, const double* dalpha, const double* drad,
                       int max_members, hipStream_t s, cp_tree* t);
// the host build of t (build_bh_tree) uploaded; the device copy included
void build_tree_host(cp_ctx* c, int n, const double* xyz, const double* alpha, const double* radius,
                     int max_members, cp_tree* t);
// context and geometry lifetimes (see cp_ctx::refs, cp_geom::refs)
void ctx_ref(cp_ctx* c);
void ctx_unref(cp_ctx* c);
void geom_ref(cp_geom* g);
void geom_unref(cp_geom* g);
void set_err(cp_ctx* c, const std::string& msg);

// runs f with HIP and budget errors turned into cp_status; `lock` (if any)
// is held for the call
template <class F>
cp_status guarded_impl(cp_ctx* c, std::mutex* lock, F&& f) {
    if (!c) return CP_BAD_ARG;
    std::unique_lock<std::mutex> lk;
    if (lock) lk = std::unique_lock<std::mutex>(*lock);
    set_current_timer(&c->timer);
    try {
        if (c->device >= 0) CP_HIP(hipSetDevice(c->device));
        cp_status s = f();
        set_current_timer(nullptr);
        return s;
    } catch (const HipError& e) {
        set_current_timer(nullptr);
        c->hip_err = (int)e.code;
        set_err(c, std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return CP_HIP_ERROR;
    } catch (const BudgetError& e) {
        set_current_timer(nullptr);
        set_err(c, "HBM budget exceeded: need " + std::to_string(e.need) + " more bytes, " +
                       std::to_string(c->in_use) + " of " + std::to_string(c->budget) + " in use");
        return CP_OUT_OF_BUDGET;
    } catch (const std::bad_alloc&) {
        set_current_timer(nullptr);
        set_err(c, "host allocation failed");
        return CP_OUT_OF_BUDGET;
    }
}

template <class F>
cp_status guarded(cp_ctx* c, F&& f) {
    return guarded_impl(c, c ? &c->mu : nullptr, std::forward<F>(f));
}

template <class F>
cp_status guarded_g(cp_geom* g, F&& f) {
    return guarded_impl(g->ctx, &g->mu, std::forward<F>(f));
}

void wind_to_body(const cp_freestream& f, double T[9], double Vb[3]);
cp_status build_geom(cp_ctx* c, const cp_mesh* m, uint64_t key, cp_geom** out, const VpParams* vp = nullptr);
void ensure_wake_influence(cp_geom* g);
void ensure_post(cp_geom* g);
void ensure_batched(cp_geom* g);
void ensure_block_jacobi(cp_geom* g);
void fill_info(const double* r, const double T[9], cp_solve_info& inf);
void prepare_mixed(cp_geom* g, cp_prec prec, const cp_solve_opts* opts);
BatchedOpts mixed_opts(const cp_solve_opts* opts, int n);

}  // namespace capi
}  // namespace cpgpu
