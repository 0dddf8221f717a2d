// Batched inversion of the strip block-Jacobi diagonal blocks on the device.
//
// Each block M (b x b, b <= 3072, column-major, FP64) is inverted in place by
// Gauss-Jordan elimination with partial pivoting, in a blocked right-looking
// form without physical row swaps: the pivot of step k is the row of largest
// |M[i][k]| among the rows not yet used as pivots (the set of candidates is
// the same as after the swaps of the textbook form, so the pivot choice is
// partial pivoting).  The columns go in super-panels of kSw, each made of
// panels of kPw pivot columns:
//   panel kernel  one 512-thread workgroup per block, up to 6 rows per thread,
//                 the rows' kPw panel entries in registers: runs the kPw pivot
//                 steps on the panel columns only.  Afterwards column
//                 k0+t holds column p_t of the accumulated transform T, which
//                 differs from I only in the pivot-row columns p_t.  It then
//                 copies the pivot rows of the super-panel's other columns
//                 out, so that the update can split the rows over workgroups.
//   update kernel the other columns j of the super-panel:
//                 M[:, j] += (T - I)[:, p] * M[p, j] with p the panel's pivot
//                 rows, M[p, j] taken from the copy.
//   trail kernel  after the super-panel's last panel, its columns hold
//                 (T_S - I)[:, P] + I[:, P] for the super-panel's transform
//                 T_S = T_last ... T_first (a product of such factors differs
//                 from I only in the union P of their pivot columns), and every
//                 column outside it takes M[:, j] += (T_S - I)[:, P] M[P, j]
//                 at once: a rank-kSw product on the FP64 matrix cores, with
//                 M[P, j] copied out first (gather kernel) so that no
//                 workgroup reads a pivot row another has already updated.
//   finish kernel with X the stored matrix after the last panel,
//                 inv(M)[i][p_s] = X[p_i][s]; rounds to FP32 into the packed
//                 block array the preconditioner applies.
//
// Per super-panel a block costs one read and one write of its b^2 entries
// and 2 b^2 kSw flops; the rank-kPw form this replaces read and wrote every
// block once per panel (8x the traffic at kSw = 128), which held a 25,916-
// panel geometry's eleven 2,560-row blocks at ~120 ms.
#include "batched.h"
#include "device.h"

namespace cpgpu {

namespace {

constexpr int kPw = 16;        // pivot columns per panel
constexpr int kMaxB = 3072;    // rows per block
constexpr int kUc = 16;        // columns per update workgroup
constexpr int kUt = 256;       // update / elementwise workgroup size
constexpr int kSw = 128;       // columns per super-panel (a multiple of kPw)
static_assert(kSw % kPw == 0, "a super-panel is whole panels");
static_assert(3 * 1024 >= kMaxB, "the panel kernel holds every row of a block");
typedef double f64x4 __attribute__((ext_vector_type(4)));  // a v_mfma_f64_16x16x4f64 accumulator
typedef double f64x16 __attribute__((ext_vector_type(16)));  // a row's panel entries

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

// Y of block q, rows of kSw: Y[t * kSw + c] = M[p_t, j(c)] for the panel's
// pivots p_t and the window [lo, min(hi, b))'s columns other than the panel
// (j(c) = lo + c below k0, lo + c + w from there on).
template <int RPT, int kPanelT>
__global__ __launch_bounds__(kPanelT) void bjinv_panel_kernel(double* __restrict__ X, const int* __restrict__ ptr,
                                                              const int64_t* __restrict__ off, int k0, int lo, int hi,
                                                              int* __restrict__ used, int* __restrict__ piv,
                                                              int* __restrict__ status, double* __restrict__ Y,
                                                              int64_t ldy) {
    const int q = blockIdx.x;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    if (k0 >= b || status[q]) return;
    const int w = min(kPw, b - k0);
    double* M = X + off[q];
    const int tid = threadIdx.x;
    // rows tid + kPanelT * k, k < RPT; each row's kPw panel entries are one
    // vector value, indexed by the (wave-uniform) step with register moves
    bool usd[RPT];
    f64x16 r[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int i = tid + kPanelT * k;
        const bool row = i < b;
        usd[k] = row ? used[b0 + i] != 0 : true;
#pragma unroll
        for (int t = 0; t < kPw; t++) r[k][t] = (row && t < w) ? M[i + (int64_t)(k0 + t) * b] : 0.0;
    }

    constexpr int NW = kPanelT / 64;
    __shared__ double sv[NW];
    __shared__ int si[NW];
    __shared__ double prow[kPw];
    __shared__ int pvs[kPw];
    const int lane = tid & 63, wv = tid >> 6;
    // a compact loop: unrolled, the steps are straight-line code fetched once
    // per launch
#pragma unroll 1
    for (int t = 0; t < w; t++) {
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
        for (int u = 1; u < NW; u++) {
            const double ov = sv[u];
            const int oi = si[u];
            if (ov > best || (ov == best && oi < p)) {
                best = ov;
                p = oi;
            }
        }
        if (!(best > 0)) {  // singular (or NaN) block (uniform): the caller falls back
            if (tid == 0) status[q] = 1;
            return;
        }
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (tid + kPanelT * k == p) {
                usd[k] = true;
                const double inv = 1.0 / r[k][t];
                r[k][t] = 1.0;
                r[k] *= inv;
#pragma unroll
                for (int j = 0; j < kPw; j++) prow[j] = r[k][j];
            }
        if (tid == 0) {
            piv[b0 + k0 + t] = p;
            pvs[t] = p;
        }
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
    __syncthreads();  // pvs of the last step
    // the pivot rows of the window's other columns (this kernel does not
    // write them)
    const int ncol = min(hi, b) - lo - w;
    for (int e = tid; e < w * ncol; e += kPanelT) {
        const int t = e / ncol, c = e % ncol;
        const int j = lo + c < k0 ? lo + c : lo + c + w;
        Y[ldy * q + (int64_t)t * kSw + c] = M[pvs[t] + (int64_t)j * b];
    }
}

// the columns [lo, min(hi, b)) of each block except the panel [k0, k0 + w),
// rows in chunks of kUt: M[:, j] += (T - I)[:, p] M[p, j] with M[p, j] from
// the panel kernel's copy
__global__ __launch_bounds__(kUt) void bjinv_update_kernel(double* __restrict__ X, const int* __restrict__ ptr,
                                                           const int64_t* __restrict__ off, int k0, int lo, int hi,
                                                           const int* __restrict__ piv,
                                                           const int* __restrict__ status,
                                                           const double* __restrict__ Y, int64_t ldy) {
    const int q = blockIdx.z;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    if (k0 >= b || status[q]) return;
    const int w = min(kPw, b - k0);
    const int ncol = min(hi, b) - lo - w;  // the window's columns other than the panel
    const int c0 = blockIdx.x * kUc;
    if (c0 >= ncol || (int)blockIdx.y * kUt >= b) return;
    double* M = X + off[q];
    __shared__ int pv[kPw];
    __shared__ double Ys[kPw][kUc];
    __shared__ int64_t colo[kUc];
    const int tid = threadIdx.x;
    if (tid < kPw) pv[tid] = tid < w ? piv[b0 + k0 + tid] : 0;
    if (tid < kUc) {
        const int c = c0 + tid;
        const int j = lo + c < k0 ? lo + c : lo + c + w;
        colo[tid] = c < ncol ? (int64_t)j * b : -1;
    }
    for (int e = tid; e < kPw * kUc; e += kUt) {
        const int t = e / kUc, cc = e % kUc;
        Ys[t][cc] = (t < w && c0 + cc < ncol) ? Y[ldy * q + (int64_t)t * kSw + c0 + cc] : 0.0;
    }
    __syncthreads();
    const int ii = blockIdx.y * kUt + tid;
    if (ii >= b) return;
    double z[kPw];
#pragma unroll
    for (int t = 0; t < kPw; t++) z[t] = t < w ? M[ii + (int64_t)(k0 + t) * b] - (ii == pv[t] ? 1.0 : 0.0) : 0.0;
#pragma unroll
    for (int cc = 0; cc < kUc; cc++) {
        if (colo[cc] < 0) break;
        double acc = M[ii + colo[cc]];
#pragma unroll
        for (int t = 0; t < kPw; t++) acc = fma(z[t], Ys[t][cc], acc);
        M[ii + colo[cc]] = acc;
    }
}

// Y[q][t][c] = M[p_t, j(c)] for the super-panel at s0 (width wS = min(kSw,
// b - s0)): the pivot rows of every column outside it (j(c) = c below s0,
// c + wS from there on), rows t >= wS zero.  Y rows are b long.
__global__ __launch_bounds__(kUt) void bjinv_gather_rows_kernel(const double* __restrict__ X,
                                                                const int* __restrict__ ptr,
                                                                const int64_t* __restrict__ off, int s0,
                                                                const int* __restrict__ piv,
                                                                const int* __restrict__ status,
                                                                double* __restrict__ Y, int64_t ldy) {
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    if (s0 >= b || status[q]) return;
    const int wS = min(kSw, b - s0);
    const int nc = b - wS;
    const int64_t e = (int64_t)blockIdx.x * kUt + threadIdx.x;
    if (e >= (int64_t)kSw * nc) return;
    const int t = (int)(e / nc), c = (int)(e % nc);
    const int j = c < s0 ? c : c + wS;
    Y[ldy * q + (int64_t)t * b + c] = t < wS ? X[off[q] + piv[b0 + s0 + t] + (int64_t)j * b] : 0.0;
}

// The super-panel's rank-kSw update of the columns outside it, as tiles of
// out = Y^T Z^T (out rows: the columns c, out columns: the rows r of M) on
// v_mfma_f64_16x16x4f64, Z[r][t] = M[r, s0 + t] - (r == p_t).  The operand
// layout: lane l holds A[l%16][l/16] and B[l/16][l%16]; result register u of
// lane l holds out[4u + l/16][l%16], so each load and store of M covers 16
// consecutive rows of 4 columns.  A wavefront owns 64 columns x 64 rows, a
// workgroup four such row tiles.
__global__ __launch_bounds__(256, 2) void bjinv_trail_kernel(double* __restrict__ X, const int* __restrict__ ptr,
                                                             const int64_t* __restrict__ off, int s0,
                                                             const int* __restrict__ piv,
                                                             const int* __restrict__ status,
                                                             const double* __restrict__ Y, int64_t ldy) {
    const int q = blockIdx.z;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    if (s0 >= b || status[q]) return;
    const int wS = min(kSw, b - s0);
    const int nc = b - wS;
    const int cb = blockIdx.x * 64;
    if (cb >= nc || (int)blockIdx.y * 256 >= b) return;
    __shared__ int pv[kSw];
    if (threadIdx.x < kSw) pv[threadIdx.x] = (int)threadIdx.x < wS ? piv[b0 + s0 + threadIdx.x] : -1;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb = blockIdx.y * 256 + wave * 64;
    if (rb >= b) return;
    const int m = lane & 15, lk = lane >> 4;
    double* M = X + off[q];
    const double* Yq = Y + ldy * q;
    // this lane's columns (out rows 4u + lk of each 16-column tile ct) and
    // rows (out column m of each 16-row tile rt)
    int64_t co[4][4];
    bool cok[4][4];
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int c = cb + 16 * ct + 4 * u + lk;
            cok[ct][u] = c < nc;
            co[ct][u] = (int64_t)(c < s0 ? c : c + wS) * b;
        }
    int row[4];
    bool rok[4];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        row[rt] = rb + 16 * rt + m;
        rok[rt] = row[rt] < b;
    }
    f64x4 acc[4][4];
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int u = 0; u < 4; u++)
                acc[ct][rt][u] = (cok[ct][u] && rok[rt]) ? M[row[rt] + co[ct][u]] : 0.0;
    // chunks of kTu k steps of 4; the next chunk's operands are loaded
    // before the current chunk's products
    constexpr int kTu = 2;
    double a[kTu][4], z[kTu][4], an[kTu][4], zn[kTu][4];
    auto load = [&](int t0, double (&aa)[kTu][4], double (&zz)[kTu][4]) {
#pragma unroll
        for (int uu = 0; uu < kTu; uu++) {
            const int t = t0 + 4 * uu + lk;
            const bool tok = t < wS;
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                const int c = cb + 16 * ct + m;
                aa[uu][ct] = (tok && c < nc) ? Yq[(int64_t)t * b + c] : 0.0;
            }
#pragma unroll
            for (int rt = 0; rt < 4; rt++)
                zz[uu][rt] = (tok && rok[rt])
                                 ? M[row[rt] + (int64_t)(s0 + t) * b] - (row[rt] == pv[t] ? 1.0 : 0.0)
                                 : 0.0;
        }
    };
    load(0, a, z);
    for (int t0 = 0; t0 < wS; t0 += 4 * kTu) {
        const bool more = t0 + 4 * kTu < wS;
        if (more) load(t0 + 4 * kTu, an, zn);
#pragma unroll
        for (int uu = 0; uu < kTu; uu++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int rt = 0; rt < 4; rt++)
                    acc[ct][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[uu][ct], z[uu][rt], acc[ct][rt], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int uu = 0; uu < kTu; uu++)
#pragma unroll
                for (int x = 0; x < 4; x++) {
                    a[uu][x] = an[uu][x];
                    z[uu][x] = zn[uu][x];
                }
        }
    }
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (cok[ct][u] && rok[rt]) M[row[rt] + co[ct][u]] = acc[ct][rt][u];
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

int64_t bj_invert_rows_doubles(int nblocks, int max_b) { return (int64_t)std::max(nblocks, 1) * kSw * max_b; }

void launch_bj_invert(const float* A32, int64_t lda, const int* perm, const int* ptr, const int64_t* off, int nblocks,
                      int max_b, double* X, int* used, int* piv, int* status, double* Y, float* out, hipStream_t s) {
    if (max_b > kMaxB) throw HipError{hipErrorInvalidValue, "block-Jacobi block larger than 3072 rows"};
    if (nblocks <= 0) return;
    const int64_t ldy = (int64_t)kSw * max_b;
    const dim3 gsq((unsigned)(((int64_t)max_b * max_b + kUt - 1) / kUt), (unsigned)nblocks);
    hipLaunchKernelGGL(bjinv_gather_kernel, gsq, dim3(kUt), 0, s, A32, lda, perm, ptr, off, X, used, piv, status);
    CP_LAUNCH_CHECK();
    const dim3 gup((unsigned)((kSw + kUc - 1) / kUc), (unsigned)((max_b + kUt - 1) / kUt), (unsigned)nblocks);
    for (int s0 = 0; s0 < max_b; s0 += kSw) {
        for (int k0 = s0; k0 < std::min(s0 + kSw, max_b); k0 += kPw) {
            const int hi = s0 + kSw;
            if (max_b <= 512)
                hipLaunchKernelGGL((bjinv_panel_kernel<1, 512>), dim3((unsigned)nblocks), dim3(512), 0, s, X, ptr, off,
                                   k0, s0, hi, used, piv, status, Y, ldy);
            else if (max_b <= 1024)
                hipLaunchKernelGGL((bjinv_panel_kernel<1, 1024>), dim3((unsigned)nblocks), dim3(1024), 0, s, X, ptr,
                                   off, k0, s0, hi, used, piv, status, Y, ldy);
            else if (max_b <= 2048)
                hipLaunchKernelGGL((bjinv_panel_kernel<2, 1024>), dim3((unsigned)nblocks), dim3(1024), 0, s, X, ptr,
                                   off, k0, s0, hi, used, piv, status, Y, ldy);
            else
                hipLaunchKernelGGL((bjinv_panel_kernel<3, 1024>), dim3((unsigned)nblocks), dim3(1024), 0, s, X, ptr,
                                   off, k0, s0, hi, used, piv, status, Y, ldy);
            CP_LAUNCH_CHECK();
            hipLaunchKernelGGL(bjinv_update_kernel, gup, dim3(kUt), 0, s, X, ptr, off, k0, s0, hi, piv, status, Y,
                               ldy);
            CP_LAUNCH_CHECK();
        }
        // columns outside the super-panel: b - min(kSw, b - s0) in a block of
        // b > s0 rows, largest at b = max_b
        const int nc = max_b - std::min(kSw, max_b - s0);
        if (nc <= 0) continue;
        const dim3 gy((unsigned)(((int64_t)kSw * max_b + kUt - 1) / kUt), (unsigned)nblocks);
        hipLaunchKernelGGL(bjinv_gather_rows_kernel, gy, dim3(kUt), 0, s, X, ptr, off, s0, piv, status, Y, ldy);
        CP_LAUNCH_CHECK();
        const dim3 gt((unsigned)((nc + 63) / 64), (unsigned)((max_b + 255) / 256), (unsigned)nblocks);
        hipLaunchKernelGGL(bjinv_trail_kernel, gt, dim3(256), 0, s, X, ptr, off, s0, piv, status, Y, ldy);
        CP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(bjinv_finish_kernel, gsq, dim3(kUt), 0, s, X, ptr, off, piv, status, out);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
