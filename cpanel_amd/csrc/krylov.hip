// K2: FP64 BiCGSTAB for one right-hand side (cpCase::solveMatrixEq,
// cpCase.cpp:185-214), restating Eigen's BiCGSTAB with its defaults: Jacobi
// preconditioner, x0 = 0, restart when r0.r becomes too small.
//
// The matrix-vector product y = A x (A column-major N x N, FP64) is HBM bound:
// every iteration streams A twice.  GEMV mapping: a lane owns two consecutive
// rows (one 16-byte load per column), a workgroup owns 512 rows and a slice of
// columns (split-K); x[j] is uniform per wavefront and arrives through scalar
// loads.  Partial row sums of the S column slices are added in a fixed order by
// the consumer kernel, so results are bitwise reproducible run to run.
//
// Vector updates and dot products are fused into the reduction kernels;
// scalars alpha and omega stay on the device, rho/rho_old/beta come from the
// host, which reads r.r and r0.r once per iteration (convergence test and the
// restart rule).
#include "device.h"
#include "krylov.h"
#include "timer.h"

namespace cpgpu {

namespace {

constexpr int kThreads = 256;
constexpr int kRowsPerBlock = 512;  // two rows per lane

__global__ __launch_bounds__(256) void gemv_f64_partial_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ part, int chunk) {
    const int64_t i0 = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * 2;
    const int s = blockIdx.y;
    const int j0 = s * chunk;
    const int j1 = min(n, j0 + chunk);
    double a0 = 0, a1 = 0;
    const double* col = A + i0 + (int64_t)j0 * lda;
    int j = j0;
    for (; j + 8 <= j1; j += 8) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = *reinterpret_cast<const double2*>(col + (int64_t)u * lda);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            double xj = x[j + u];
            a0 = fma(v[u].x, xj, a0);
            a1 = fma(v[u].y, xj, a1);
        }
        col += 8 * lda;
    }
    for (; j < j1; j++) {
        double2 v = *reinterpret_cast<const double2*>(col);
        double xj = x[j];
        a0 = fma(v.x, xj, a0);
        a1 = fma(v.y, xj, a1);
        col += lda;
    }
    *reinterpret_cast<double2*>(part + (int64_t)s * lda + i0) = make_double2(a0, a1);
}

// Block-wide sum with a fixed reduction tree; result broadcast to all threads.
__device__ double block_sum(double v, double* sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] += sh[t + o];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Sum of np partials in a fixed order (every block computes the same value).
__device__ double sum_partials(const double* __restrict__ p, int np, double* sh) {
    double s = 0;
    for (int k = threadIdx.x; k < np; k += kThreads) s += p[k];
    return block_sum(s, sh);
}

__device__ inline double row_sum(const double* __restrict__ part, int S, int64_t lda, int64_t i) {
    double y = 0;
    for (int s = 0; s < S; s++) y += part[(int64_t)s * lda + i];
    return y;
}

// y = sign * sum_s part[s]  (plain GEMV result)
__global__ __launch_bounds__(256) void gemv_reduce_kernel(const double* __restrict__ part, int S, int64_t lda,
                                                          int n, double sign, double* __restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < lda) y[i] = (i < n) ? sign * row_sum(part, S, lda, i) : 0.0;
}

// v = A y (from partials); partial dot r0.v per block
__global__ __launch_bounds__(256) void reduce_v_kernel(const double* __restrict__ part, int S, int64_t lda, int n,
                                                       double* __restrict__ v, const double* __restrict__ r0,
                                                       double* __restrict__ dpart) {
    __shared__ double sh[kThreads];
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double d = 0;
    if (i < n) {
        double vi = row_sum(part, S, lda, i);
        v[i] = vi;
        d = r0[i] * vi;
    }
    d = block_sum(d, sh);
    if (threadIdx.x == 0) dpart[blockIdx.x] = d;
}

// t = A z (from partials); partial dots t.t and t.s per block
__global__ __launch_bounds__(256) void reduce_t_kernel(const double* __restrict__ part, int S, int64_t lda, int n,
                                                       double* __restrict__ t, const double* __restrict__ s,
                                                       double* __restrict__ dtt, double* __restrict__ dts) {
    __shared__ double sh[kThreads];
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double a = 0, b = 0;
    if (i < n) {
        double ti = row_sum(part, S, lda, i);
        t[i] = ti;
        a = ti * ti;
        b = ti * s[i];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        dtt[blockIdx.x] = a;
        dts[blockIdx.x] = b;
    }
}

// p = r + beta (p - omega v);  y = invd .* p
__global__ __launch_bounds__(256) void update_p_kernel(int n, double beta, const double* __restrict__ sc,
                                                       const double* __restrict__ r, double* __restrict__ p,
                                                       const double* __restrict__ v,
                                                       const double* __restrict__ invd, double* __restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    double omega = sc[SC_OMEGA];
    double pi = r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    y[i] = invd[i] * pi;
}

// alpha = rho / (r0.v);  s = r - alpha v;  z = invd .* s
__global__ __launch_bounds__(256) void update_s_kernel(int n, double rho, const double* __restrict__ dpart, int np,
                                                       double* __restrict__ sc, const double* __restrict__ r,
                                                       const double* __restrict__ v,
                                                       const double* __restrict__ invd, double* __restrict__ s,
                                                       double* __restrict__ z) {
    __shared__ double sh[kThreads];
    double r0v = sum_partials(dpart, np, sh);
    double alpha = rho / r0v;
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_ALPHA] = alpha;
    if (i >= n) return;
    double si = r[i] - alpha * v[i];
    s[i] = si;
    z[i] = invd[i] * si;
}

// omega = t.s / t.t;  x += alpha y + omega z;  r = s - omega t;
// partial dots r.r and r0.r per block
__global__ __launch_bounds__(256) void update_x_kernel(int n, const double* __restrict__ dtt,
                                                       const double* __restrict__ dts, int np,
                                                       double* __restrict__ sc, double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       double* __restrict__ r, const double* __restrict__ s,
                                                       const double* __restrict__ t, const double* __restrict__ r0,
                                                       double* __restrict__ drr, double* __restrict__ dr0r) {
    __shared__ double sh[kThreads];
    double tt = sum_partials(dtt, np, sh);
    double ts = sum_partials(dts, np, sh);
    double omega = tt > 0 ? ts / tt : 0.0;
    double alpha = sc[SC_ALPHA];
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double a = 0, b = 0;
    if (i < n) {
        x[i] += alpha * y[i] + omega * z[i];
        double ri = s[i] - omega * t[i];
        r[i] = ri;
        a = ri * ri;
        b = r0[i] * ri;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        drr[blockIdx.x] = a;
        dr0r[blockIdx.x] = b;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_OMEGA] = omega;
}

// final reduction of r.r and r0.r into the scalar block
__global__ __launch_bounds__(256) void finish_dots_kernel(const double* __restrict__ drr,
                                                          const double* __restrict__ dr0r, int np,
                                                          double* __restrict__ sc) {
    __shared__ double sh[kThreads];
    double a = sum_partials(drr, np, sh);
    double b = sum_partials(dr0r, np, sh);
    if (threadIdx.x == 0) {
        sc[SC_RR] = a;
        sc[SC_R0R] = b;
    }
}

// r = b - part (after GEMV of x), r0 = r, invd from the diagonal, partial r.r
__global__ __launch_bounds__(256) void residual_kernel(int n, const double* __restrict__ b,
                                                       const double* __restrict__ part, int S, int64_t lda,
                                                       double* __restrict__ r, double* __restrict__ r0,
                                                       double* __restrict__ drr, double* __restrict__ dr0r) {
    __shared__ double sh[kThreads];
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double a = 0;
    if (i < n) {
        double ri = b[i] - (part ? row_sum(part, S, lda, i) : 0.0);
        r[i] = ri;
        r0[i] = ri;
        a = ri * ri;
    }
    a = block_sum(a, sh);
    if (threadIdx.x == 0) {
        drr[blockIdx.x] = a;
        dr0r[blockIdx.x] = a;
    }
}

// Jacobi preconditioner of Eigen's DiagonalPreconditioner: 1/A_jj, 1 if zero
__global__ __launch_bounds__(256) void invdiag_kernel(int n, const double* __restrict__ A, int64_t lda,
                                                      double* __restrict__ invd) {
    int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= n) return;
    double d = A[j + j * lda];
    invd[j] = d != 0 ? 1 / d : 1.0;
}

__global__ void fill_kernel(double* __restrict__ p, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

Gemv64::Gemv64(int n, int64_t lda) : n_(n), lda_(lda) {
    int rowBlocks = (int)(lda / kRowsPerBlock);
    int target = 4096;
    int S = std::max(1, (target + rowBlocks - 1) / rowBlocks);
    int maxS = std::max(1, n / 64);
    S_ = std::min(S, maxS);
    chunk_ = (n + S_ - 1) / S_;
    S_ = (n + chunk_ - 1) / chunk_;
}

void Gemv64::partial(const double* A, const double* x, double* part, hipStream_t s) const {
    TimedScope ts("gemv_f64", s);
    dim3 grid((unsigned)(lda_ / kRowsPerBlock), (unsigned)S_);
    hipLaunchKernelGGL(gemv_f64_partial_kernel, grid, dim3(kThreads), 0, s, A, lda_, n_, x, part, chunk_);
    CP_LAUNCH_CHECK();
}

void Gemv64::apply(const double* A, const double* x, double* y, double sign, double* part, hipStream_t s) const {
    partial(A, x, part, s);
    hipLaunchKernelGGL(gemv_reduce_kernel, dim3((unsigned)(lda_ / kThreads)), dim3(kThreads), 0, s, part, S_, lda_,
                       n_, sign, y);
    CP_LAUNCH_CHECK();
}

Bicgstab64::Bicgstab64(int n, int64_t lda) : n_(n), lda_(lda), gemv_(n, lda) {
    nb_ = (int)(lda / kThreads);
    size_t vec = (size_t)lda * sizeof(double);
    CP_HIP(hipMalloc(&ws_, vec * kNumVec + (size_t)gemv_.slices() * vec + 8 * (size_t)nb_ * sizeof(double) +
                               64 * sizeof(double)));
    double* p = ws_;
    for (int k = 0; k < kNumVec; k++) {
        vec_[k] = p;
        p += lda;
    }
    part_ = p;
    p += (size_t)gemv_.slices() * lda;
    for (int k = 0; k < 8; k++) {
        dp_[k] = p;
        p += nb_;
    }
    sc_ = p;
    CP_HIP(hipHostMalloc(&host_sc_, 64 * sizeof(double), hipHostMallocDefault));
}

Bicgstab64::~Bicgstab64() {
    if (ws_) (void)hipFree(ws_);
    if (host_sc_) (void)hipHostFree(host_sc_);
}

void Bicgstab64::fetch(hipStream_t s) {
    CP_HIP(hipMemcpyAsync(host_sc_, sc_, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
    CP_HIP(hipStreamSynchronize(s));
}

SolveStats Bicgstab64::solve(const double* A, const double* b, double* x, double tol, int maxit, hipStream_t s,
                             const double* invd_in, const double* x0) {
    SolveStats st;
    const int n = n_;
    if (maxit <= 0) maxit = 2 * n;
    if (tol <= 0) tol = 2.220446049250313e-16;
    double *r = vec_[0], *r0 = vec_[1], *p = vec_[2], *v = vec_[3], *y = vec_[4], *z = vec_[5], *sv = vec_[6],
           *t = vec_[7], *invd = vec_[8];
    double *drr = dp_[0], *dr0r = dp_[1], *dv = dp_[2], *dtt = dp_[3], *dts = dp_[4];
    dim3 g((unsigned)nb_), blk(kThreads);
    if (invd_in) {
        CP_HIP(hipMemcpyAsync(invd, invd_in, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        hipLaunchKernelGGL(invdiag_kernel, g, blk, 0, s, n, A, lda_, invd);
        CP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(fill_kernel, g, blk, 0, s, x, 0.0, (int64_t)lda_);
    hipLaunchKernelGGL(fill_kernel, g, blk, 0, s, p, 0.0, (int64_t)lda_);
    hipLaunchKernelGGL(fill_kernel, g, blk, 0, s, v, 0.0, (int64_t)lda_);
    // r = b - A*0, r0 = r
    hipLaunchKernelGGL(residual_kernel, g, blk, 0, s, n, b, (const double*)nullptr, 0, lda_, r, r0, drr, dr0r);
    hipLaunchKernelGGL(finish_dots_kernel, dim3(1), blk, 0, s, drr, dr0r, nb_, sc_);
    double init[2] = {1.0, 1.0};  // alpha, omega
    CP_HIP(hipMemcpyAsync(sc_ + SC_ALPHA, &init[0], sizeof(double), hipMemcpyHostToDevice, s));
    CP_HIP(hipMemcpyAsync(sc_ + SC_OMEGA, &init[1], sizeof(double), hipMemcpyHostToDevice, s));
    fetch(s);
    double rr = host_sc_[SC_RR];
    double rhs_sq = rr;
    if (x0) {
        // initial guess: r = b - A x0 (Eigen's solveWithGuess); |b| stays the scale
        CP_HIP(hipMemcpyAsync(x, x0, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
        gemv_.partial(A, x, part_, s);
        hipLaunchKernelGGL(residual_kernel, g, blk, 0, s, n, b, (const double*)part_, gemv_.slices(), lda_, r, r0,
                           drr, dr0r);
        hipLaunchKernelGGL(finish_dots_kernel, dim3(1), blk, 0, s, drr, dr0r, nb_, sc_);
        CP_LAUNCH_CHECK();
        fetch(s);
        rr = host_sc_[SC_RR];
    }
    double r0_sq = rr;
    if (rhs_sq == 0) {
        st.rel_residual = 0;
        return st;
    }
    double rho = 1, alpha = 1, omega = 1;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = 2.220446049250313e-16, eps2 = eps * eps;
    double r0r = host_sc_[SC_R0R];
    int i = 0, restarts = 0;
    while (rr > tol2 && i < maxit) {
        double rho_old = rho;
        rho = r0r;
        if (std::fabs(rho) < eps2 * r0_sq) {
            // restart with a fresh r0 = b - A x (Eigen BiCGSTAB restart rule)
            gemv_.partial(A, x, part_, s);
            hipLaunchKernelGGL(residual_kernel, g, blk, 0, s, n, b, (const double*)part_, gemv_.slices(), lda_, r,
                               r0, drr, dr0r);
            hipLaunchKernelGGL(finish_dots_kernel, dim3(1), blk, 0, s, drr, dr0r, nb_, sc_);
            fetch(s);
            rho = r0_sq = host_sc_[SC_RR];
            if (restarts++ == 0) i = 0;
            st.restarts = restarts;
        }
        double beta = (rho / rho_old) * (alpha / omega);
        hipLaunchKernelGGL(update_p_kernel, g, blk, 0, s, n, beta, (const double*)sc_, r, p, v, invd, y);
        gemv_.partial(A, y, part_, s);
        hipLaunchKernelGGL(reduce_v_kernel, g, blk, 0, s, part_, gemv_.slices(), lda_, n, v, r0, dv);
        hipLaunchKernelGGL(update_s_kernel, g, blk, 0, s, n, rho, dv, nb_, sc_, r, v, invd, sv, z);
        gemv_.partial(A, z, part_, s);
        hipLaunchKernelGGL(reduce_t_kernel, g, blk, 0, s, part_, gemv_.slices(), lda_, n, t, sv, dtt, dts);
        hipLaunchKernelGGL(update_x_kernel, g, blk, 0, s, n, dtt, dts, nb_, sc_, x, y, z, r, sv, t, r0, drr, dr0r);
        hipLaunchKernelGGL(finish_dots_kernel, dim3(1), blk, 0, s, drr, dr0r, nb_, sc_);
        CP_LAUNCH_CHECK();
        fetch(s);
        rr = host_sc_[SC_RR];
        r0r = host_sc_[SC_R0R];
        alpha = host_sc_[SC_ALPHA];
        omega = host_sc_[SC_OMEGA];
        ++i;
    }
    st.iters = i;
    st.rel_residual = std::sqrt(rr / rhs_sq);
    return st;
}

}  // namespace cpgpu
