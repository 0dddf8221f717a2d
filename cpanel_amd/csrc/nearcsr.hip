// The near-field pairs of the mixed-precision path, built on the device.
//
// A pair (i, j) is near when |c_i - c_j| <= factor * L_j (collocation point
// i, panel j of centre c_j and longest side L_j); the mixed path keeps the
// FP64 value of these entries beside its FP32 matrices, and the packed bf16
// product restores the 32-row x 16-column tiles of A that hold one.  The
// pairs form a CSR by row with ascending columns, each entry also carrying
// its row, and the tiles a list per 32-row tile with ascending 16-column
// blocks.
//
// Every kernel here gives each lane a row (or a row tile) and walks the
// columns in ascending order, so the output order needs no sort, and the
// count and fill passes visit the pairs identically.  The distance test
// repeats the host's arithmetic (differences, then (x*x + y*y) + z*z, no
// contraction) and takes the square root only inside a 1e-13 relative band
// around the threshold, where the squared comparison could differ from it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"

namespace cpgpu {

namespace {

constexpr int kNearT = 64;   // rows per workgroup (one wavefront) and columns per LDS tile
constexpr int kScanT = 1024;  // the single-workgroup scan

// |c - p| <= L from d2 = |c - p|^2 formed as the host forms it
__device__ __forceinline__ bool within(double d2, double L, double lo, double hi) {
    if (d2 < lo) return true;
    if (d2 > hi) return false;
    return sqrt(d2) <= L;
}

// count (C = true: cnt[i] = row i's pairs) or fill (cols, rows from rowptr)
template <bool C>
__global__ __launch_bounds__(kNearT) void near_pairs_kernel(const PanelRec* __restrict__ body, int n, double factor,
                                                            int* __restrict__ cnt, const int* __restrict__ rowptr,
                                                            int* __restrict__ cols, int* __restrict__ rows) {
#pragma clang fp contract(off)
    __shared__ double tc[4][kNearT];  // column centres and factor * L
    const int i = blockIdx.x * kNearT + threadIdx.x;
    double px = 0, py = 0, pz = 0;
    if (i < n) {
        px = body[i].c[0];
        py = body[i].c[1];
        pz = body[i].c[2];
    }
    int k = 0, e = 0;
    if (!C && i < n) e = rowptr[i];
    for (int j0 = 0; j0 < n; j0 += kNearT) {
        __syncthreads();
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            tc[0][threadIdx.x] = body[jl].c[0];
            tc[1][threadIdx.x] = body[jl].c[1];
            tc[2][threadIdx.x] = body[jl].c[2];
            tc[3][threadIdx.x] = factor * body[jl].long_side;
        }
        __syncthreads();
        if (i >= n) continue;
        const int m = min(kNearT, n - j0);
        for (int u = 0; u < m; u++) {
            const double dx = px - tc[0][u], dy = py - tc[1][u], dz = pz - tc[2][u];
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double L = tc[3][u], L2 = L * L;
            if (within(d2, L, L2 * (1 - 1e-13), L2 * (1 + 1e-13))) {
                if (C) {
                    k++;
                } else {
                    cols[e] = j0 + u;
                    rows[e] = i;
                    e++;
                }
            }
        }
    }
    if (C && i < n) cnt[i] = k;
}

// out[0] = 0, out[t + 1] = in[0] + ... + in[t] for t < n (one workgroup)
__global__ __launch_bounds__(kScanT) void exclusive_scan_kernel(const int* __restrict__ in, int n,
                                                                int* __restrict__ out) {
    __shared__ int part[kScanT];
    const int per = (n + kScanT - 1) / kScanT;
    const int a = min(n, (int)threadIdx.x * per), b = min(n, a + per);
    int s = 0;
    for (int t = a; t < b; t++) s += in[t];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < kScanT; t++) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = part[threadIdx.x];
    if (threadIdx.x == 0) out[0] = 0;
    for (int t = a; t < b; t++) {
        run += in[t];
        out[t + 1] = run;
    }
}

// mark[kb * nrt + rt] = 1 for each pair's (row tile, column block)
__global__ void near_mark_kernel(int nnz, const int* __restrict__ rows, const int* __restrict__ cols, int nrt,
                                 unsigned char* __restrict__ mark) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    mark[(int64_t)(cols[e] / 16) * nrt + rows[e] / 32] = 1;
}

// per row tile rt (one lane each, columns blocks ascending): count (C) or
// write its marked blocks
template <bool C>
__global__ void near_tiles_kernel(const unsigned char* __restrict__ mark, int nrt, int nkb, int* __restrict__ cnt,
                                  const int* __restrict__ tptr, int* __restrict__ tkb) {
    const int rt = blockIdx.x * blockDim.x + threadIdx.x;
    if (rt >= nrt) return;
    int k = 0, e = C ? 0 : tptr[rt];
    for (int kb = 0; kb < nkb; kb++)
        if (mark[(int64_t)kb * nrt + rt]) {
            if (C)
                k++;
            else
                tkb[e++] = kb;
        }
    if (C) cnt[rt] = k;
}

}  // namespace

void launch_near_count(const DevGeom& g, double factor, int* cnt, hipStream_t s) {
    if (g.n <= 0) return;
    hipLaunchKernelGGL(near_pairs_kernel<true>, dim3((unsigned)((g.n + kNearT - 1) / kNearT)), dim3(kNearT), 0, s,
                       g.body, g.n, factor, cnt, nullptr, nullptr, nullptr);
    CP_LAUNCH_CHECK();
}

void launch_near_fill(const DevGeom& g, double factor, const int* rowptr, int* cols, int* rows, hipStream_t s) {
    if (g.n <= 0) return;
    hipLaunchKernelGGL(near_pairs_kernel<false>, dim3((unsigned)((g.n + kNearT - 1) / kNearT)), dim3(kNearT), 0, s,
                       g.body, g.n, factor, nullptr, rowptr, cols, rows);
    CP_LAUNCH_CHECK();
}

void launch_exclusive_scan(const int* in, int n, int* out, hipStream_t s) {
    hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(kScanT), 0, s, in, n, out);
    CP_LAUNCH_CHECK();
}

void launch_near_tile_count(int nnz, const int* rows, const int* cols, int nrt, int nkb, unsigned char* mark,
                            int* cnt, hipStream_t s) {
    if (nnz > 0) {
        hipLaunchKernelGGL(near_mark_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, nnz, rows, cols, nrt,
                           mark);
        CP_LAUNCH_CHECK();
    }
    if (nrt > 0) {
        hipLaunchKernelGGL(near_tiles_kernel<true>, dim3((unsigned)((nrt + 255) / 256)), dim3(256), 0, s, mark, nrt,
                           nkb, cnt, nullptr, nullptr);
        CP_LAUNCH_CHECK();
    }
}

void launch_near_tile_fill(const unsigned char* mark, int nrt, int nkb, const int* tptr, int* tkb, hipStream_t s) {
    if (nrt <= 0) return;
    hipLaunchKernelGGL(near_tiles_kernel<false>, dim3((unsigned)((nrt + 255) / 256)), dim3(256), 0, s, mark, nrt, nkb,
                       nullptr, tptr, tkb);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
