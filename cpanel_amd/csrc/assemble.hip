// K1: dense source/doublet influence assembly for gfx950.
//
// Replaces geometry::setInfCoeff (geometry.cpp:809-891).  A and B are
// column-major (A[i + j*lda], column j = panel j, row i = collocation point
// i), the layout the reference writes.  Work mapping:
//   * one lane per row i: the collocation point lives in registers, loaded
//     once, coalesced, from SoA arrays;
//   * a wavefront walks a chunk of columns j; the panel record of column j is
//     the same for every lane, so it arrives through scalar loads;
//   * every column store is 64 consecutive rows (512 B FP64 / 256 B FP32) per
//     wavefront, fully coalesced.
// Kutta folding of the wake into A is applied in registers before the single
// store of each entry: column c receives sum_w coef * W[i, w] for the wake
// panels whose fold table names c, in wake order (the reference's order of
// the += in geometry.cpp:859-864).
#include "device.h"

namespace cpgpu {

namespace {

constexpr int kRows = 256;   // rows per workgroup (4 wavefronts)
constexpr int kCols = 32;    // columns per workgroup

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

__global__ __launch_bounds__(256) void assemble_f64_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int n, int64_t lda, double* __restrict__ A, double* __restrict__ B,
    const int* __restrict__ fold_ptr, const int* __restrict__ fold_w, const double* __restrict__ fold_c,
    const double* __restrict__ W, int64_t ldw) {
    const int i = blockIdx.x * kRows + threadIdx.x;
    const int j0 = blockIdx.y * kCols;
    const int j1 = min(n, j0 + kCols);
    const bool live = i < n;
    const double px = cx[i], py = cy[i], pz = cz[i];
    for (int j = j0; j < j1; j++) {
        const PanelRec& P = pan[j];
        double src, dub;
        double d0 = px - P.c[0], d1 = py - P.c[1], d2 = pz - P.c[2];
        double r2 = sq_norm3(d0, d1, d2);
        if (is_far(P, r2)) {
#pragma clang fp contract(off)
            double r = sqrt(r2);
            double PN = d0 * P.n[0] + d1 * P.n[1] + d2 * P.n[2];
            src = P.area / (4 * kPI * r);
            dub = PN * P.area / (4 * kPI * (r * r * r));
        } else {
            near_phi_f64(P, px, py, pz, src, dub);
        }
        const int e0 = fold_ptr[j], e1 = fold_ptr[j + 1];
        for (int e = e0; e < e1; e++) {
#pragma clang fp contract(off)
            double t = W[i + (int64_t)fold_w[e] * ldw] * fold_c[e];
            dub = dub + t;
        }
        const int64_t o = i + (int64_t)j * lda;
        A[o] = live ? dub : 0.0;
        B[o] = live ? src : 0.0;
    }
}

// FP32 assembly: far-field pairs in FP32 arithmetic (offset vector formed in
// FP64, then rounded), near-field and self pairs (|pjk| <= 5 longSide, where
// the edge formulas lose digits in FP32) in FP64, wake folding in FP64, one
// rounding to FP32 at the store.  The far/near split is the FP64 decision.
__global__ __launch_bounds__(256) void assemble_f32_kernel(
    const PanelRec* __restrict__ pan, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int n, int64_t lda, float* __restrict__ A, float* __restrict__ B,
    const int* __restrict__ fold_ptr, const int* __restrict__ fold_w, const double* __restrict__ fold_c,
    const double* __restrict__ W, int64_t ldw) {
    const int i = blockIdx.x * kRows + threadIdx.x;
    const int j0 = blockIdx.y * kCols;
    const int j1 = min(n, j0 + kCols);
    const bool live = i < n;
    const double px = cx[i], py = cy[i], pz = cz[i];
    for (int j = j0; j < j1; j++) {
        const PanelRec& P = pan[j];
        double d0 = px - P.c[0], d1 = py - P.c[1], d2 = pz - P.c[2];
        double r2 = sq_norm3(d0, d1, d2);
        float srcf, dubf;
        double dub64 = 0;
        bool far = is_far(P, r2);
        if (far) {
            float f0 = (float)d0, f1 = (float)d1, f2 = (float)d2;
            float rinv = rsqrtf((float)r2);
            float a4 = (float)P.area_4pi;
            float PN = f0 * (float)P.n[0] + f1 * (float)P.n[1] + f2 * (float)P.n[2];
            srcf = a4 * rinv;
            dubf = PN * a4 * (rinv * rinv * rinv);
            dub64 = dubf;
        } else {
            double s, d;
            near_phi_f64(P, px, py, pz, s, d);
            srcf = (float)s;
            dub64 = d;
        }
        const int e0 = fold_ptr[j], e1 = fold_ptr[j + 1];
        for (int e = e0; e < e1; e++) dub64 += W[i + (int64_t)fold_w[e] * ldw] * fold_c[e];
        dubf = (float)dub64;
        const int64_t o = i + (int64_t)j * lda;
        A[o] = live ? dubf : 0.0f;
        B[o] = live ? srcf : 0.0f;
    }
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
    dim3 grid((unsigned)(g.lda / kRows), (unsigned)((g.n + kCols - 1) / kCols));
    hipLaunchKernelGGL(assemble_f64_kernel, grid, dim3(kRows), 0, s, g.body, g.cx, g.cy, g.cz, g.n, g.lda, A, B,
                       g.fold_ptr, g.fold_w, g.fold_c, W, ldw);
    CP_LAUNCH_CHECK();
}

void launch_assemble_f32(const DevGeom& g, const double* W, int64_t ldw, float* A, float* B, hipStream_t s) {
    dim3 grid((unsigned)(g.lda / kRows), (unsigned)((g.n + kCols - 1) / kCols));
    hipLaunchKernelGGL(assemble_f32_kernel, grid, dim3(kRows), 0, s, g.body, g.cx, g.cy, g.cz, g.n, g.lda, A, B,
                       g.fold_ptr, g.fold_w, g.fold_c, W, ldw);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
