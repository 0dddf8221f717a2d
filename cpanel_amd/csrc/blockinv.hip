// Batched inversion of the strip block-Jacobi diagonal blocks on the device.
//
// Each block M (b x b, b <= 3072, column-major, FP64) is inverted in place by
// Gauss-Jordan elimination with partial pivoting, in a blocked right-looking
// form over panels of kPw pivot columns and without physical row swaps: the
// pivot of step k is the row of largest |M[i][k]| among the rows not yet
// used as pivots (the set of candidates is the same as after the swaps of
// the textbook form, so the pivot choice is partial pivoting).
//
//   panel kernel  one 1024-thread workgroup per block, one row per thread
//                 (two or three above 1024 rows), the rows' kPw panel entries in
//                 registers: runs the kPw pivot steps on the panel columns
//                 only.  Afterwards column
//                 k0+t holds column p_t of the accumulated transform T, which
//                 differs from I only in the pivot-row columns p_t.
//   update kernel every other column j:  M[:, j] += (T - I)[:, p] * M[p, j]
//                 with p the panel's pivot rows.  A workgroup owns whole
//                 columns, so it reads the pivot rows' entries of its columns
//                 into LDS before it overwrites them.
//   finish kernel with X the stored matrix after the last panel,
//                 inv(M)[i][p_s] = X[p_i][s]; rounds to FP32 into the packed
//                 block array the preconditioner applies.
//
// Per panel a block costs one read and one write of its b^2 entries and
// 2 b^2 kPw flops in the update, so a geometry's blocks (about 800 rows
// each) invert in a few milliseconds; the host form this replaces (FP64
// Gauss-Jordan on host threads) took hundreds.
#include "batched.h"
#include "device.h"

namespace cpgpu {

namespace {

constexpr int kPw = 16;        // pivot columns per panel
constexpr int kPanelT = 1024;  // panel-kernel workgroup: one or two rows per thread
constexpr int kMaxB = 3072;    // rows per block
constexpr int kUc = 16;        // columns per update workgroup
constexpr int kUt = 256;       // update / elementwise workgroup size

__global__ __launch_bounds__(kUt) void bjinv_gather_kernel(const float* __restrict__ A, int64_t lda,
                                                           const int* __restrict__ perm, const int* __restrict__ ptr,
                                                           const int64_t* __restrict__ off, double* __restrict__ X,
                                                           int* __restrict__ used, int* __restrict__ piv,
                                                           int* __restrict__ status) {
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const int64_t e = (int64_t)blockIdx.x * kUt + threadIdx.x;
    if (e == 0) status[q] = 0;
    if (e < b) {
        used[b0 + e] = 0;
        piv[b0 + e] = 0;
    }
    if (e >= (int64_t)b * b) return;
    const int r = (int)(e % b), c = (int)(e / b);
    X[off[q] + e] = (double)A[perm[b0 + r] + (int64_t)perm[b0 + c] * lda];
}

// argmax of (v, i) over a wavefront: larger v first, then smaller i
__device__ inline void wave_argmax(double& v, int& i) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        double ov = __shfl_xor(v, m, 64);
        int oi = __shfl_xor(i, m, 64);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
}

template <int RPT>
__global__ __launch_bounds__(kPanelT) void bjinv_panel_kernel(double* __restrict__ X, const int* __restrict__ ptr,
                                                              const int64_t* __restrict__ off, int k0,
                                                              int* __restrict__ used, int* __restrict__ piv,
                                                              int* __restrict__ status) {
    const int q = blockIdx.x;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    if (k0 >= b || status[q]) return;
    const int w = min(kPw, b - k0);
    double* M = X + off[q];
    const int tid = threadIdx.x;
    // rows tid + kPanelT * k, k < RPT
    bool usd[RPT];
    double r[RPT][kPw];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int i = tid + kPanelT * k;
        const bool row = i < b;
        usd[k] = row ? used[b0 + i] != 0 : true;
#pragma unroll
        for (int t = 0; t < kPw; t++) r[k][t] = (row && t < w) ? M[i + (int64_t)(k0 + t) * b] : 0.0;
    }

    __shared__ double sv[kPanelT / 64];
    __shared__ int si[kPanelT / 64];
    __shared__ double prow[kPw];
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int t = 0; t < kPw; t++) {
        if (t >= w) break;
        // this thread's best candidate (larger |value|, then smaller row)
        double v = -1.0;
        int vi = tid;
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const double a = usd[k] ? -1.0 : fabs(r[k][t]);
            if (a > v) {
                v = a;
                vi = tid + kPanelT * k;
            }
        }
        wave_argmax(v, vi);
        if (lane == 0) {
            sv[wv] = v;
            si[wv] = vi;
        }
        __syncthreads();
        double best = sv[0];
        int p = si[0];
        for (int u = 1; u < kPanelT / 64; u++) {
            const double ov = sv[u];
            const int oi = si[u];
            if (ov > best || (ov == best && oi < p)) {
                best = ov;
                p = oi;
            }
        }
        if (!(best > 0)) {  // singular (or NaN) block: the caller falls back
            if (tid == 0) status[q] = 1;
            return;
        }
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (tid + kPanelT * k == p) {
                usd[k] = true;
                const double inv = 1.0 / r[k][t];
                r[k][t] = 1.0;
#pragma unroll
                for (int j = 0; j < kPw; j++) {
                    r[k][j] *= inv;
                    prow[j] = r[k][j];
                }
            }
        if (tid == 0) piv[b0 + k0 + t] = p;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const int i = tid + kPanelT * k;
            if (i < b && i != p) {
                const double f = r[k][t];
                r[k][t] = 0.0;
#pragma unroll
                for (int j = 0; j < kPw; j++) r[k][j] = fma(-f, prow[j], r[k][j]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int i = tid + kPanelT * k;
        if (i < b) {
#pragma unroll
            for (int t = 0; t < kPw; t++)
                if (t < w) M[i + (int64_t)(k0 + t) * b] = r[k][t];
            used[b0 + i] = usd[k] ? 1 : 0;
        }
    }
}

__global__ __launch_bounds__(kUt) void bjinv_update_kernel(double* __restrict__ X, const int* __restrict__ ptr,
                                                           const int64_t* __restrict__ off, int k0,
                                                           const int* __restrict__ piv,
                                                           const int* __restrict__ status) {
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    if (k0 >= b || status[q]) return;
    const int w = min(kPw, b - k0);
    const int c0 = blockIdx.x * kUc;
    if (c0 >= b - w) return;
    double* M = X + off[q];
    __shared__ int pv[kPw];
    __shared__ double Y[kPw][kUc];
    __shared__ int64_t colo[kUc];
    const int tid = threadIdx.x;
    if (tid < kPw) pv[tid] = tid < w ? piv[b0 + k0 + tid] : 0;
    if (tid < kUc) {
        const int c = c0 + tid;
        const int j = c < k0 ? c : c + w;
        colo[tid] = c < b - w ? (int64_t)j * b : -1;
    }
    __syncthreads();
    for (int e = tid; e < kPw * kUc; e += kUt) {
        const int t = e / kUc, cc = e % kUc;
        Y[t][cc] = (t < w && colo[cc] >= 0) ? M[pv[t] + colo[cc]] : 0.0;
    }
    __syncthreads();
    for (int ii = tid; ii < b; ii += kUt) {
        double z[kPw];
#pragma unroll
        for (int t = 0; t < kPw; t++)
            z[t] = t < w ? M[ii + (int64_t)(k0 + t) * b] - (ii == pv[t] ? 1.0 : 0.0) : 0.0;
#pragma unroll
        for (int cc = 0; cc < kUc; cc++) {
            if (colo[cc] < 0) break;
            double acc = M[ii + colo[cc]];
#pragma unroll
            for (int t = 0; t < kPw; t++) acc = fma(z[t], Y[t][cc], acc);
            M[ii + colo[cc]] = acc;
        }
    }
}

__global__ __launch_bounds__(kUt) void bjinv_finish_kernel(const double* __restrict__ X, const int* __restrict__ ptr,
                                                           const int64_t* __restrict__ off,
                                                           const int* __restrict__ piv,
                                                           const int* __restrict__ status, float* __restrict__ out) {
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const int64_t e = (int64_t)blockIdx.x * kUt + threadIdx.x;
    if (e >= (int64_t)b * b || status[q]) return;
    const int i = (int)(e % b), s = (int)(e / b);
    out[off[q] + i + (int64_t)piv[b0 + s] * b] = (float)X[off[q] + piv[b0 + i] + (int64_t)s * b];
}

}  // namespace

void launch_bj_invert(const float* A32, int64_t lda, const int* perm, const int* ptr, const int64_t* off, int nblocks,
                      int max_b, double* X, int* used, int* piv, int* status, float* out, hipStream_t s) {
    if (max_b > kMaxB) throw HipError{hipErrorInvalidValue, "block-Jacobi block larger than 3072 rows"};
    if (nblocks <= 0) return;
    const dim3 gsq((unsigned)(((int64_t)max_b * max_b + kUt - 1) / kUt), (unsigned)nblocks);
    hipLaunchKernelGGL(bjinv_gather_kernel, gsq, dim3(kUt), 0, s, A32, lda, perm, ptr, off, X, used, piv, status);
    CP_LAUNCH_CHECK();
    const dim3 gup((unsigned)((max_b + kUc - 1) / kUc), (unsigned)nblocks);
    for (int k0 = 0; k0 < max_b; k0 += kPw) {
        if (max_b <= kPanelT)
            hipLaunchKernelGGL(bjinv_panel_kernel<1>, dim3((unsigned)nblocks), dim3(kPanelT), 0, s, X, ptr, off, k0,
                               used, piv, status);
        else if (max_b <= 2 * kPanelT)
            hipLaunchKernelGGL(bjinv_panel_kernel<2>, dim3((unsigned)nblocks), dim3(kPanelT), 0, s, X, ptr, off, k0,
                               used, piv, status);
        else
            hipLaunchKernelGGL(bjinv_panel_kernel<3>, dim3((unsigned)nblocks), dim3(kPanelT), 0, s, X, ptr, off, k0,
                               used, piv, status);
        CP_LAUNCH_CHECK();
        hipLaunchKernelGGL(bjinv_update_kernel, gup, dim3(kUt), 0, s, X, ptr, off, k0, piv, status);
        CP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(bjinv_finish_kernel, gsq, dim3(kUt), 0, s, X, ptr, off, piv, status, out);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
