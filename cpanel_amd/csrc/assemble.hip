// K1: dense source/doublet influence assembly for gfx950.
//
// Replaces geometry::setInfCoeff (geometry.cpp:809-891).  A and B are
// column-major (A[i + j*lda], column j = panel j, row i = collocation point
// i), the layout the reference writes.  Two passes per matrix pair:
//
//   * far field (every pair): the point source / point doublet of
//     bodyPanel::panelPhiInf (bodyPanel.cpp:132-190) with no branch.  One
//     lane per row i (the collocation point in registers, loaded once,
//     coalesced); a wavefront walks a chunk of columns j whose panel record
//     is wave-uniform (scalar loads); each column store is 64 consecutive
//     rows per wavefront.  This pass is bound by the matrix writes.
//   * near field (pairs with |c_i - c_j| <= 5 longSide_j, 7.5 % of the pairs
//     on the 25,916-panel wing): the VSAERO edge sums overwrite the far-field
//     values.  One wavefront per column: a bounding-sphere test of each
//     64-row tile of collocation points against 5 longSide_j (64 tiles per
//     test, one per lane) keeps only the tiles that can hold a near row; the
//     near rows of those tiles are compacted with lane permutes into a queue
//     held one row per lane, and every full queue is evaluated by the whole
//     wavefront, so the transcendental-heavy edge formulas run with all 64
//     lanes busy.  The far/near decision is the reference's |pjk|/longSide > 5
//     (is_far), the same test the far pass would have made.
//
// Kutta folding of the wake into A is applied in registers before the store
// of each entry, in both passes: column c receives sum_w coef * W[i, w] for
// the wake panels whose fold table names c, in wake order (the reference's
// order of the += in geometry.cpp:859-864).
//
// FP32 matrices: far field in FP32 arithmetic, near field the FP64 edge sums
// rounded once.
#include "device.h"
#include "timer.h"
#include "wavequeue.h"

namespace cpgpu {

namespace {

constexpr int kRows = 256;   // rows per workgroup (4 wavefronts)

__global__ __launch_bounds__(256) void wake_influence_kernel(const PanelRec* __restrict__ wake, int nw,
                                                             const double* __restrict__ cx,
                                                             const double* __restrict__ cy,
                                                             const double* __restrict__ cz, int n,
                                                             double* __restrict__ W, int64_t ldw) {
    int i = blockIdx.x * kRows + threadIdx.x;
    int w = blockIdx.y;
    double v = 0;
    if (i < n) v = dub_phi_inf_f64(wake[w], cx[i], cy[i], cz[i]);
    W[i + (int64_t)w * ldw] = v;
}

// Kutta folding of column j into the doublet influence at row i (FP64)
__device__ inline double fold_in(double dub, int i, int j, const int* __restrict__ fold_ptr,
                                 const int* __restrict__ fold_w, const double* __restrict__ fold_c,
                                 const double* __restrict__ W, int64_t ldw) {
    const int e0 = fold_ptr[j], e1 = fold_ptr[j + 1];
    for (int e = e0; e < e1; e++) {
#pragma clang fp contract(off)
        double t = W[i + (int64_t)fold_w[e] * ldw] * fold_c[e];
        dub = dub + t;
    }
    return dub;
}

// far-field passes: a workgroup owns 1024 rows (4 per lane, 256 apart) and
// 16 columns, so each column receives 8 KiB of contiguous rows per matrix
// from one workgroup (longer write runs than one 256-row tile) and each lane
// has four independent pairs in flight; rows from n to lda are zeroed, rows
// past lda do not exist
constexpr int kFarR = 4;
constexpr int kFarRows = kFarR * 256;
constexpr int kFarCols = 16;

// far-field pass, FP64
__global__ __launch_bounds__(256) void far_f64_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int n, int64_t lda, double* __restrict__ A, double* __restrict__ B,
    const int* __restrict__ fold_ptr, const int* __restrict__ fold_w, const double* __restrict__ fold_c,
    const double* __restrict__ W, int64_t ldw) {
    const int i0 = blockIdx.x * kFarRows + threadIdx.x;
    const int j0 = blockIdx.y * kFarCols;
    const int j1 = min(n, j0 + kFarCols);
    double px[kFarR], py[kFarR], pz[kFarR];
    for (int r = 0; r < kFarR; r++) {
        const int i = i0 + r * 256;
        const bool in = i < lda;
        px[r] = in ? cx[i] : 0.0;
        py[r] = in ? cy[i] : 0.0;
        pz[r] = in ? cz[i] : 0.0;
    }
    for (int j = j0; j < j1; j++) {
#pragma clang fp contract(off)
        const PanelRec& P = pan[j];
        for (int r = 0; r < kFarR; r++) {
            const int i = i0 + r * 256;
            if (i >= lda) break;
            double d0 = px[r] - P.c[0], d1 = py[r] - P.c[1], d2 = pz[r] - P.c[2];
            double r2 = sq_norm3(d0, d1, d2);
            // one reciprocal for both point formulas (area / (4 pi r) and
            // PN area / (4 pi r^3), bodyPanel.cpp:392-395, panel.cpp:737-740):
            // a few ulp from the reference's two divisions
            double inv = 1.0 / sqrt(r2);
            double PN = d0 * P.n[0] + d1 * P.n[1] + d2 * P.n[2];
            double src = P.area_4pi * inv;
            double dub = PN * P.area_4pi * (inv * inv * inv);
            dub = fold_in(dub, i, j, fold_ptr, fold_w, fold_c, W, ldw);
            const int64_t o = i + (int64_t)j * lda;
            const bool live = i < n;
            // the matrices are written once here and read by later kernels:
            // streaming stores keep them out of the caches
            __builtin_nontemporal_store(live ? dub : 0.0, &A[o]);
            __builtin_nontemporal_store(live ? src : 0.0, &B[o]);
        }
    }
}

// far-field pass, FP32: offset vector formed in FP64 then rounded, point
// formulas in FP32, folding in FP64, one rounding at the store
__global__ __launch_bounds__(256) void far_f32_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int n, int64_t lda, float* __restrict__ A, float* __restrict__ B,
    const int* __restrict__ fold_ptr, const int* __restrict__ fold_w, const double* __restrict__ fold_c,
    const double* __restrict__ W, int64_t ldw) {
    const int i0 = blockIdx.x * kFarRows + threadIdx.x;
    const int j0 = blockIdx.y * kFarCols;
    const int j1 = min(n, j0 + kFarCols);
    double px[kFarR], py[kFarR], pz[kFarR];
    for (int r = 0; r < kFarR; r++) {
        const int i = i0 + r * 256;
        const bool in = i < lda;
        px[r] = in ? cx[i] : 0.0;
        py[r] = in ? cy[i] : 0.0;
        pz[r] = in ? cz[i] : 0.0;
    }
    for (int j = j0; j < j1; j++) {
        const PanelRec& P = pan[j];
        const int e0 = fold_ptr[j], e1 = fold_ptr[j + 1];
        for (int r = 0; r < kFarR; r++) {
            const int i = i0 + r * 256;
            if (i >= lda) break;
            double d0 = px[r] - P.c[0], d1 = py[r] - P.c[1], d2 = pz[r] - P.c[2];
            float f0 = (float)d0, f1 = (float)d1, f2 = (float)d2;
            float rinv = rsqrtf((float)sq_norm3(d0, d1, d2));
            float a4 = (float)P.area_4pi;
            float PN = f0 * (float)P.n[0] + f1 * (float)P.n[1] + f2 * (float)P.n[2];
            float srcf = a4 * rinv;
            float dubf = PN * a4 * (rinv * rinv * rinv);
            if (e0 != e1) {
                double dub64 = dubf;
                for (int e = e0; e < e1; e++) dub64 += W[i + (int64_t)fold_w[e] * ldw] * fold_c[e];
                dubf = (float)dub64;
            }
            const int64_t o = i + (int64_t)j * lda;
            const bool live = i < n;
            __builtin_nontemporal_store(live ? dubf : 0.0f, &A[o]);
            __builtin_nontemporal_store(live ? srcf : 0.0f, &B[o]);
        }
    }
}

// Walks the 64-row tiles of column j that can hold a near-field row
// (bounding spheres: centre, radius per tile of collocation points) and
// hands each tile's rows with their squared distance r2 to visit(row, r2,
// valid); rows past n are passed with valid = false.
template <class V>
__device__ inline void near_tiles(const PanelRec& P, const double* __restrict__ cx, const double* __restrict__ cy,
                                  const double* __restrict__ cz, const double* __restrict__ tsph, int ntiles, int n,
                                  int lane, V&& visit) {
    // 5 longSide plus a margin above the is_far band and the rounding of
    // the sphere test: a tile it drops has no near row
    const double reach = 5.0 * P.long_side * (1.0 + 1e-9);
    for (int t0 = 0; t0 < ntiles; t0 += 64) {
        const int t = t0 + lane;
        bool cand = false;
        if (t < ntiles) {
            const double dx = tsph[4 * t] - P.c[0], dy = tsph[4 * t + 1] - P.c[1], dz = tsph[4 * t + 2] - P.c[2];
            const double lim = reach + tsph[4 * t + 3];
            cand = dx * dx + dy * dy + dz * dz <= lim * lim;
        }
        uint64_t cm = __ballot(cand);
        while (cm) {
            const int b = __ffsll((unsigned long long)cm) - 1;
            cm &= cm - 1;
            const int i = (t0 + b) * 64 + lane;
            const bool valid = i < n;
            double r2 = 0;
            if (valid) {
#pragma clang fp contract(off)
                const double d0 = cx[i] - P.c[0], d1 = cy[i] - P.c[1], d2 = cz[i] - P.c[2];
                r2 = sq_norm3(d0, d1, d2);
            }
            visit(i, r2, valid);
        }
    }
}

// near-field pass, FP64: one wavefront per column
__global__ __launch_bounds__(256) void near_f64_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, const double* __restrict__ tsph, int ntiles, int n, int64_t lda,
    double* __restrict__ A, double* __restrict__ B, const int* __restrict__ fold_ptr,
    const int* __restrict__ fold_w, const double* __restrict__ fold_c, const double* __restrict__ W, int64_t ldw) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;  // whole wavefronts
    const PanelRec& P = pan[j];
    auto eval = [&](int i, bool ok) {
        if (!ok) return;
        double src, dub;
        near_phi_tri_f64(P, cx[i], cy[i], cz[i], src, dub);
        dub = fold_in(dub, i, j, fold_ptr, fold_w, fold_c, W, ldw);
        const int64_t o = i + (int64_t)j * lda;
        A[o] = dub;
        B[o] = src;
    };
    RowQueue q;
    near_tiles(P, cx, cy, cz, tsph, ntiles, n, lane, [&](int i, double r2, bool valid) {
        queue_push(q, valid && !is_far(P, r2), i, lane, eval);
    });
    eval(q.p, lane < q.qn);
}

// near-field pass, FP32: the FP64 edge sums rounded once (FP32 edge sums
// lose up to 3e-4 of an entry where the three edge terms cancel, which would
// change the system the FP32 refinement converges to)
__global__ __launch_bounds__(256) void near_f32_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, const double* __restrict__ tsph, int ntiles, int n, int64_t lda,
    float* __restrict__ A, float* __restrict__ B, const int* __restrict__ fold_ptr,
    const int* __restrict__ fold_w, const double* __restrict__ fold_c, const double* __restrict__ W, int64_t ldw) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    const PanelRec& P = pan[j];
    auto eval = [&](int i, bool ok) {
        if (!ok) return;
        double src, dub;
        near_phi_tri_f64(P, cx[i], cy[i], cz[i], src, dub);
        dub = fold_in(dub, i, j, fold_ptr, fold_w, fold_c, W, ldw);
        const int64_t o = i + (int64_t)j * lda;
        A[o] = (float)dub;
        B[o] = (float)src;
    };
    RowQueue q;
    near_tiles(P, cx, cy, cz, tsph, ntiles, n, lane, [&](int i, double r2, bool valid) {
        queue_push(q, valid && !is_far(P, r2), i, lane, eval);
    });
    eval(q.p, lane < q.qn);
}

// FP64 value of the listed entries (body influence + Kutta folding, the
// same arithmetic as assemble_f64_kernel) minus the stored FP32 value.
__global__ __launch_bounds__(256) void near_correction_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int64_t lda, const float* __restrict__ A32, const float* __restrict__ B32,
    const int* __restrict__ fold_ptr, const int* __restrict__ fold_w, const double* __restrict__ fold_c,
    const double* __restrict__ W, int64_t ldw, int nnz, const int* __restrict__ rows, const int* __restrict__ cols,
    double* __restrict__ dA, double* __restrict__ dB) {
    int e = blockIdx.x * kRows + threadIdx.x;
    if (e >= nnz) return;
    const int i = rows[e], j = cols[e];
    const PanelRec& P = pan[j];
    double src, dub;
    panel_phi_inf_f64(P, cx[i], cy[i], cz[i], src, dub);
    for (int q = fold_ptr[j]; q < fold_ptr[j + 1]; q++) {
#pragma clang fp contract(off)
        double t = W[i + (int64_t)fold_w[q] * ldw] * fold_c[q];
        dub = dub + t;
    }
    const int64_t o = i + (int64_t)j * lda;
    dA[e] = dub - (double)A32[o];
    dB[e] = src - (double)B32[o];
}

}  // namespace

void launch_near_correction(const DevGeom& g, const double* W, int64_t ldw, const float* A32, const float* B32,
                            int nnz, const int* rows, const int* cols, double* dA, double* dB, hipStream_t s) {
    if (nnz == 0) return;
    hipLaunchKernelGGL(near_correction_kernel, dim3((unsigned)((nnz + kRows - 1) / kRows)), dim3(kRows), 0, s, g.body,
                       g.cx, g.cy, g.cz, g.lda, A32, B32, g.fold_ptr, g.fold_w, g.fold_c, W, ldw, nnz, rows, cols, dA,
                       dB);
    CP_LAUNCH_CHECK();
}

void launch_wake_influence(const DevGeom& g, double* W, int64_t ldw, hipStream_t s) {
    if (g.nw == 0) return;
    dim3 grid((unsigned)(g.lda / kRows), (unsigned)g.nw);
    hipLaunchKernelGGL(wake_influence_kernel, grid, dim3(kRows), 0, s, g.wake, g.nw, g.cx, g.cy, g.cz, g.n, W,
                       ldw);
    CP_LAUNCH_CHECK();
}

void launch_assemble_f64(const DevGeom& g, const double* W, int64_t ldw, double* A, double* B, hipStream_t s) {
    dim3 grid((unsigned)((g.lda + kFarRows - 1) / kFarRows), (unsigned)((g.n + kFarCols - 1) / kFarCols));
    {
        TimedScope ts("assemble_far_f64", s);
        hipLaunchKernelGGL(far_f64_kernel, grid, dim3(kRows), 0, s, g.body, g.cx, g.cy, g.cz, g.n, g.lda, A, B,
                           g.fold_ptr, g.fold_w, g.fold_c, W, ldw);
        CP_LAUNCH_CHECK();
    }
    TimedScope ts("assemble_near_f64", s);
    hipLaunchKernelGGL(near_f64_kernel, dim3((unsigned)((g.n + 3) / 4)), dim3(256), 0, s, g.body, g.cx, g.cy, g.cz,
                       g.tsph, g.ntiles, g.n, g.lda, A, B, g.fold_ptr, g.fold_w, g.fold_c, W, ldw);
    CP_LAUNCH_CHECK();
}

void launch_assemble_f32(const DevGeom& g, const double* W, int64_t ldw, float* A, float* B, hipStream_t s) {
    dim3 grid((unsigned)((g.lda + kFarRows - 1) / kFarRows), (unsigned)((g.n + kFarCols - 1) / kFarCols));
    {
        TimedScope ts("assemble_far_f32", s);
        hipLaunchKernelGGL(far_f32_kernel, grid, dim3(kRows), 0, s, g.body, g.cx, g.cy, g.cz, g.n, g.lda, A, B,
                           g.fold_ptr, g.fold_w, g.fold_c, W, ldw);
        CP_LAUNCH_CHECK();
    }
    TimedScope ts("assemble_near_f32", s);
    hipLaunchKernelGGL(near_f32_kernel, dim3((unsigned)((g.n + 3) / 4)), dim3(256), 0, s, g.body, g.cx, g.cy, g.cz,
                       g.tsph, g.ntiles, g.n, g.lda, A, B, g.fold_ptr, g.fold_w, g.fold_c, W, ldw);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
