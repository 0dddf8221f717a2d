// K3: batched multi-RHS solve in FP32 with FP64 iterative refinement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "krylov.h"

namespace cpgpu {

constexpr int kRhsTile = 64;  // right-hand sides per batched launch

// Default inner reduction per refinement sweep.  How far one sweep lowers
// the FP64 residual is limited by the bf16x2 inner operator: on the
// 25,916-panel wing a 2e-4 target reaches tol = 1e-10 in three sweeps, a 5e-4
// target needs a fourth and 1e-4 or 3e-5 only add inner iterations.
constexpr double kInnerTolTarget = 2e-4;

// block-Jacobi inverses in matrix-core operand order: FP32 (FP32 MFMA) or
// bf16 hi/lo pairs (bf16 MFMA)
enum { kBjPackFp32 = 0, kBjPackBf16 = 1 };

// geometries per grouped solve (BatchedSolver::solve_group)
constexpr int kMaxGroup = 16;

struct GmProb;  // one geometry's arrays and sizes in a grouped launch (batched.hip)
class BatchedSolver;

// one geometry's part of a grouped solve: its solver (set up as for solve())
// and k <= 64 steady conditions
struct GroupItem {
    BatchedSolver* sv;
    const float* A32;
    const float* B32;
    const double* normals;
    const double* vinf4;
    int k;
    double* x;
    int64_t ldx;
    SolveStats* stats;
    const struct BatchedOpts* o;  // this geometry's options (as for its own solve())
};

struct BatchedOpts {
    double tol = 1e-10;       // outer: ||b - A x|| / ||b|| in FP64
    int max_iter = 0;         // inner FP32 iterations per sweep
    int max_refine = 20;      // FP64 refinement sweeps
    // inner residual reduction per sweep; <= 0: kInnerTolTarget, relaxed on
    // a sweep that can finish to what the outer tolerance still needs
    double inner_tol = 0;
};

// GEMM tile shape chosen by the autotuner: rt = 32-row tiles per wavefront,
// kslices = split of the K range over workgroups
struct GemmShape {
    int variant = 0;  // 0: 4-byte A loads, rt tiles per wave; 1/2: 16-byte A loads, 4 interleaved
                      // tiles, 16/8 K values per chunk, two register stages; 3/4: the same with
                      // 8/4 K values per chunk and three stages; 5: bf16x3 on packed tiles
                      // (4 row tiles, 2 stages); 6/7: bf16x3, 2 row tiles, 3/4 stages;
                      // 8/9: bf16x3, 4 row tiles, 3/2 stages (generic form);
                      // 10/11/12: bf16x2 (A as bf16, near field restored in
                      // FP32), 4 row tiles, 2/3/4 stages; 13/14: bf16x2, 2 row
                      // tiles, 2/3 stages
    int rt = 2;
    int kslices = 8;
};

struct TuneRecord {
    GemmShape shape;
    float ms;
};

class BatchedSolver {
public:
    // ws: device workspace of workspace_bytes(lda) bytes; host: pinned host
    // memory of host_bytes() bytes.  Both stay owned by the caller (pooled
    // there: freeing device or pinned memory waits for the whole device).
    // The workspace's zero-initialised parts are cleared on stream s.
    BatchedSolver(int n, int64_t lda, int capacity, void* ws, void* host, hipStream_t s);
    // device bytes of the workspace of a solver with leading dimension lda
    static size_t workspace_bytes(int64_t lda);
    static size_t host_bytes();
    ~BatchedSolver();
    BatchedSolver(const BatchedSolver&) = delete;
    BatchedSolver& operator=(const BatchedSolver&) = delete;
    int capacity() const { return cap_; }
    // Solves A32 x_c = -B32 sigma_c (sigma_c = -Vinf_c . n) for k columns.
    // normals: device SoA (3*lda); vinf4: device (Vx,Vy,Vz,|V|) per column;
    // x: device FP64 output, column c at x + c*ldx.
    void solve(const float* A32, const float* B32, const double* normals, const double* vinf4, int k,
               const BatchedOpts& o, double* x, int64_t ldx, SolveStats* stats, hipStream_t s);
    // Same with caller-given columns (device, column-major, leading dimension
    // ldx): sigma (replaces -Vinf . n), rhs_add (added to -B sigma, or null)
    // and x0 (initial guess, or null = zero)
    void solve_rhs(const float* A32, const float* B32, const double* sigma, const double* rhs_add, const double* x0,
                   int k, const BatchedOpts& o, double* x, int64_t ldx, SolveStats* stats, hipStream_t s);
    // Several geometries' steady solves at once (k <= 64 conditions each, on
    // one stream s): every GMRES iteration stage is one launch covering all
    // geometries still sweeping (grouped kernels), the rest of the
    // refinement loop runs per geometry.  A geometry's launches are exactly
    // those of its own solve(), so its results are bit-identical to it.
    // dev: group_bytes(items.size()) bytes of device memory.
    static void solve_group(std::vector<GroupItem>& items, void* dev, hipStream_t s);
    static size_t group_bytes(int G);
    // the grouped iteration applies to this solver now: GMRES with the bf16
    // block-Jacobi product and a bf16x2 product shape (valid after tuning)
    bool group_capable() const;
    void set_shape(const GemmShape& g) {
        shape_ = g;
        tuned_ = true;
    }
    void set_autotune(bool on) { autotune_ = on; }
    // off: never capture graphs (solvers used from several host threads at
    // once: a capture there failed other threads' HIP calls)
    void allow_graphs(bool on) {
        if (!on) graphs_ = false;
    }
    // times every candidate shape on A32 and keeps the fastest (once per
    // leading dimension and process)
    void tune(const float* A32, hipStream_t s);
    const std::vector<TuneRecord>& tune_log() const { return tune_log_; }
    // 0, 1, 2: see bj_level_
    int block_jacobi_level() const { return bj_level_; }
    const GemmShape& shape() const { return shape_; }
    // Y = A32 * P for one 64-column tile (exposed for tests / tuning)
    void matvec32(const float* A32, const float* P, float* Y, hipStream_t s);
    // Y = Badd + sign * (A32 + D) X64 with FP64 accumulation; D = the sparse
    // near-field correction given by corr (values on the set_near pattern)
    void matvec64acc(const float* A32, const double* X, double* Y, const double* Badd, double sign, hipStream_t s,
                     const double* corr = nullptr, const double* A64 = nullptr);
    // FP64 matrices for the refinement residual and the right-hand side (the
    // refined solution is then the FP64 system's); null = FP32 + near-field
    // correction
    // forget the cached -B n columns (the matrices were replaced)
    void invalidate_rhs_cache() { bn_mode_ = -1; }
    void set_f64(const double* A64, const double* B64) {
        a64_ = A64;
        b64_ = B64;
    }
    // strip block-Jacobi preconditioner for the inner solve (device arrays;
    // null perm = Jacobi).  inv: FP32 inverses, column-major per block; pk /
    // poff: the same in matrix-core operand order (launch_bj_pack).
    void set_block_jacobi(const int* perm, const int* ptr, const int64_t* off, const float* inv, int nblocks,
                          int max_b, const int* iperm = nullptr, const float* pk = nullptr,
                          const int64_t* poff = nullptr, int pk_kind = kBjPackBf16) {
        bj_ = perm != nullptr && bj_level_ < 2;
        bj_pk_ = pk;
        bj_poff_ = poff;
        bj_kind_ = pk_kind;
        bj_iperm_ = pk && poff ? iperm : nullptr;  // block-ordered products need the packed inverses
        bj_perm_ = perm;
        bj_ptr_ = ptr;
        bj_off_ = off;
        bj_inv_ = inv;
        bj_nblocks_ = nblocks;
        bj_maxb_ = max_b;
    }
    // bf16 hi/lo tile planes of A32 (pack_a_bf16x2); enables GEMM variants 5-9
    void set_packed(const void* apk, int nkb) {
        apk_ = apk;
        nkb_ = nkb;
        plane_ = (lda_ / 32) * (int64_t)nkb * 64;
    }
    // FP64 near-field corrections of A32 and B32 (device CSR by row)
    void set_near(const int* rowptr, const int* cols, const double* dA, const double* dB) {
        near_rowptr_ = rowptr;
        near_cols_ = cols;
        near_dA_ = dA;
        near_dB_ = dB;
    }
    // near-field tiles of the packed A (CSR by 32-row tile: the 16-wide K
    // blocks holding an entry |c_i - c_j| <= longSide_j); enables the bf16x2
    // GEMM variants 10-14
    void set_near_tiles(const int* tptr, const int* tkb) {
        tptr_ = tptr;
        tkb_ = tkb;
    }
    // partial slices a product of shape g leaves for the reduction
    static int slices(const GemmShape& g) { return g.kslices + (g.variant >= 10 ? 1 : 0); }

private:
    void tile_solve(const float* A32, const float* B32, const double* normals, const double* vinf4, int c0,
                    int kc, const BatchedOpts& o, double* x, int64_t ldx, SolveStats* stats, hipStream_t s,
                    const double* sig_cm = nullptr, const double* add_cm = nullptr, const double* x0_cm = nullptr);
    // parts: 1 packs P into bf16 tiles (variants >= 5), 2 runs the product, 3 both
    void launch_gemm32(const GemmShape& sh, const float* A32, const float* P, hipStream_t s, int parts = 3);
    void apply_block_jacobi(const float* P, float* Y, hipStream_t s);
    // Y = B (rows >= n zeroed) with the column partials of ||Y_c||^2 in dpart_
    void copy64(const double* B, double* Y, hipStream_t s);
    // inner solve of one refinement sweep by right-preconditioned GMRES
    void bj_apply_pk(const float* PP, float* Y, const double* cscale, int cstride, hipStream_t s);
    // one 64-column tile of a solve: inputs and the refinement loop's host state
    struct Tile {
        const float* A32 = nullptr;
        const float* B32 = nullptr;
        const double* normals = nullptr;
        const double* vinf4 = nullptr;
        int kc = 0;
        BatchedOpts o;
        double* x = nullptr;
        int64_t ldx = 0;
        SolveStats* stats = nullptr;
        const double* sig_cm = nullptr;
        const double* add_cm = nullptr;
        const double* x0_cm = nullptr;
        std::vector<double> bb, rel, prev;
        std::vector<int> done;
        int sweep = 0;
    };
    void tile_begin(Tile& t, hipStream_t s);
    void tile_begin_host(Tile& t);
    void tile_residual(Tile& t, hipStream_t s);
    bool tile_check(Tile& t);
    void tile_sweep_open(Tile& t, hipStream_t s);
    void tile_sweep_close(Tile& t, hipStream_t s);
    void tile_end(Tile& t, hipStream_t s);
    void bicgstab_sweep(Tile& t, hipStream_t s);
    // column sums of nq quantities' block partials (nblk blocks; 0: nrb_)
    void colsum_q(const double* dp, int nq, int q0, int q1, int op, hipStream_t s, int nblk = 0);
    bool gmres_ready() const;
    int gm_bound(const BatchedOpts& o) const;
    void gm_begin(hipStream_t s);
    void gm_iterations(const float* A32, const BatchedOpts& o, hipStream_t s);
    void gm_end(hipStream_t s);
    bool gm_counted_ = false;  // a GMRES sweep left its iteration counts in sc_[Q_GMNJ]
    void fill_prob(GmProb& P, const BatchedOpts& o, int* flag) const;
    static void group_iterations(const std::vector<BatchedSolver*>& sv, const GmProb* probs,
                                 const std::vector<GmProb>& hp, int* flags, hipStream_t s);
    // GMRES iteration j's launches captured once as HIP graphs and replayed
    // in every later sweep.  The iteration is four segments (block-Jacobi
    // product, P packing, product, Gram-Schmidt + Givens); a segment whose
    // kernel is being timed is launched on the stream between graphs (events
    // recorded inside a graph carry no timestamps here), the others are
    // captured, consecutive ones into one graph.  The cache is keyed by
    // everything the captured launches depend on.
    struct IterPlan {
        std::vector<std::pair<hipGraphExec_t, int>> items;  // (graph, -1) or (null, segment launched directly)
        ~IterPlan();
    };
    bool graphs_ = false;  // CPGPU_GRAPH=1: replay captured iteration graphs
    std::vector<std::unique_ptr<IterPlan>> gm_plans_;
    std::vector<uintptr_t> gm_plan_key_;
    void gm_plan_clear();
    const double* a64_ = nullptr;
    const double* b64_ = nullptr;
    const void* apk_ = nullptr;
    int nkb_ = 0;
    void* ppk_ = nullptr;  // bf16 hi/lo tiles of the multivector being multiplied
    bool bj_ = false;
    bool bj_mfma_ = true;  // matrix-core block product (else vector ALU)
    bool bf16x3_ = true;   // bf16x3 GEMM is a tuner candidate
    bool bf16x2_ = true;   // bf16x2 GEMM is a tuner candidate
    int64_t plane_ = 0;    // uint4 elements from the hi plane to the lo plane
    const int* tptr_ = nullptr;  // near-field tile lists (set_near_tiles)
    const int* tkb_ = nullptr;
    bool f64_pipelined_ = true;  // pipelined FP64-accumulated product
    int f64_depth_ = 2;          // K steps per stage of the FP64-matrix product
    int f64_ks_ = 0;             // pinned K split of the FP64-accumulated products (0: by fill)
    const int* bj_perm_ = nullptr;
    const int* bj_iperm_ = nullptr;  // body index -> position in block order
    float* pp_ = nullptr;            // multivector in block order
    const int* bj_ptr_ = nullptr;
    const int64_t* bj_off_ = nullptr;
    const float* bj_inv_ = nullptr;
    const float* bj_pk_ = nullptr;
    const int64_t* bj_poff_ = nullptr;
    int cus_ = 256;  // compute units of the device
    int bj_kind_ = kBjPackBf16;  // layout of bj_pk_
    bool bj_lo_ = false;         // bf16 product: read the lo plane of the inverses too
    // preconditioner strength after stalled refinement sweeps (sticky for
    // this solver): 0 as configured, 1 bf16 product with the lo plane,
    // 2 point Jacobi
    int bj_level_ = 0;
    int bj_escalations_ = 0;
    void escalate_block_jacobi();
    int bj_rt_ = 2;              // bf16 product: 32-row tiles per workgroup
    int bj_u_ = 1;  // K chunks per load batch of the packed block-Jacobi product
    int bj_nblocks_ = 0, bj_maxb_ = 0;
    bool tuned_ = false, autotune_ = true;
    std::vector<TuneRecord> tune_log_;
    int n_, cap_;
    int64_t lda_;
    int nrb_;      // row blocks of the reduction kernels
    GemmShape shape_;
    void* ws_ = nullptr;
    // FP64 multivectors (lda x 64 row-major)
    double *X_, *Bv_, *R_, *S_;
    // FP32 multivectors
    float *r_, *r0_, *p_, *v_, *sv_, *t_, *y_, *z_, *d_;
    float* invd_;
    float* part32_;     // kslices x lda x 64
    double* part64_;    // kslices x lda x 64
    double* dpart_;     // nrb x 64 x 4 partial column dots
    double* bn_;        // lda x 4: -B n_x, -B n_y, -B n_z (steady right-hand sides)
    int bn_mode_ = -1;  // matrix bn_ was formed with: 1 FP64 B, 0 B32 + near field, -1 none
    // GMRES inner solve (with the block-Jacobi preconditioner): basis in
    // block order, per-column state, dot partials; Krylov dimension gm_m_
    bool gmres_ = true;
    int gm_m_ = 40;
    float* gm_u_ = nullptr;
    double* gm_st_ = nullptr;
    double* gm_dpart_ = nullptr;
    // [t][column] h_{t,j} / s_t of the current iteration; after a sweep, y_q / s_q
    // (rows q) and the iteration count (row kGmMax)
    float* gm_coef_ = nullptr;
    double* sc_;        // 64 x 16 per-column scalars
    int* act_;          // 64 active flags
    const int* near_rowptr_ = nullptr;
    const int* near_cols_ = nullptr;
    const double* near_dA_ = nullptr;
    const double* near_dB_ = nullptr;
    double* host_sc_ = nullptr;
    int* host_act_ = nullptr;
    hipEvent_t act_ev_[2] = {nullptr, nullptr};  // active-flag read-backs of the inner loop
    // active flags of the inner iteration being queued (null outside it): the
    // heavy kernels of an iteration with no active column return at once
    const int* gate_ = nullptr;
};

void launch_sigma(int n, int64_t lda, const double* normals, const double* vinf4, double* sigma, hipStream_t s);
// bf16 hi/lo tiles of A32 for the bf16x3 GEMM: K blocks of 16 and bytes
int packed_kblocks(int n);
size_t packed_a_bytes(int64_t lda, int n);
// floats reserved for one b x b inverse in operand order (either kind)
int64_t bj_packed_floats(int b);
// block q's inverse (column-major at inv + off[q]) to operand order of the
// given kind at pk + poff[q] (poff[q + 1] - poff[q] = bj_packed_floats(b_q))
void launch_bj_pack(const float* inv, const int* ptr, const int64_t* off, const int64_t* poff, int nblocks, int max_b,
                    float* pk, int kind, hipStream_t s);
void launch_pack_a(const float* A32, int64_t lda, int n, void* out, hipStream_t s);
// Inverts the diagonal blocks of A32 named by the block plan (perm, ptr,
// off; rows per block <= 3072) in FP64 on the device and stores the inverses
// in FP32, packed column-major, in out.  Work buffers: X (off[nblocks]
// doubles), used and piv (n ints each, n = ptr[nblocks]), status (nblocks
// ints; status[q] != 0 marks a singular block, whose inverse is not written),
// Y (bj_invert_rows_doubles(nblocks, max_b) doubles).
int64_t bj_invert_rows_doubles(int nblocks, int max_b);
void launch_bj_invert(const float* A32, int64_t lda, const int* perm, const int* ptr, const int64_t* off, int nblocks,
                      int max_b, double* X, int* used, int* piv, int* status, double* Y, float* out, hipStream_t s);

}  // namespace cpgpu
