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
// count and fill passes visit the pairs identically.  A workgroup takes 64
// rows and a chunk of kNearChunk 64-column tiles (so that rows whose tile of
// panels is spread over the body, near many column tiles, share their work
// among workgroups), and skips each column tile that no lane's point comes
// within its bounding sphere's radius plus the tile's largest factor * L of
// (with a relative margin of 1e-9 over the rounding): no pair in such a tile
// can be near.  Counts are kept per (chunk, row), and a per-row prefix over
// the chunks gives each chunk's first slot, so the columns stay ascending.  The distance test
// repeats the host's arithmetic (differences, then (x*x + y*y) + z*z, no
// contraction) and takes the square root only inside a 1e-13 relative band
// around the threshold, where the squared comparison could differ from it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"

namespace cpgpu {

namespace {

constexpr int kNearT = 64;   // rows per workgroup (one wavefront) and columns per LDS tile
constexpr int kNearChunk = 8;  // column tiles per workgroup
constexpr int kScanT = 1024;  // the single-workgroup scan

// |c - p| <= L from d2 = |c - p|^2 formed as the host forms it
__device__ __forceinline__ bool within(double d2, double L, double lo, double hi) {
    if (d2 < lo) return true;
    if (d2 > hi) return false;
    return sqrt(d2) <= L;
}

// reach[t] = factor * the largest longSide of 64-panel tile t
__global__ void near_reach_kernel(const PanelRec* __restrict__ body, int n, double factor, int ntiles,
                                  double* __restrict__ reach) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    double L = 0;
    for (int j = t * kNearT; j < min(n, (t + 1) * kNearT); j++) L = fmax(L, body[j].long_side);
    reach[t] = factor * L;
}

// count (C = true: cnt[q * n + i] = row i's pairs in column chunk q) or
// fill (from slot rowptr[i] + cnt[q * n + i]: cols, rows); tsph: the
// 64-point tiles' bounding spheres (centre, radius)
template <bool C>
__global__ __launch_bounds__(kNearT) void near_pairs_kernel(const PanelRec* __restrict__ body, int n, double factor,
                                                            const double* __restrict__ tsph,
                                                            const double* __restrict__ reach, int ntiles,
                                                            int* __restrict__ cnt, const int* __restrict__ rowptr,
                                                            int* __restrict__ cols, int* __restrict__ rows) {
#pragma clang fp contract(off)
    __shared__ double tc[4][kNearT];  // column centres and factor * L
    const int q = blockIdx.y;
    const int i = blockIdx.x * kNearT + threadIdx.x;
    double px = 0, py = 0, pz = 0;
    if (i < n) {
        px = body[i].c[0];
        py = body[i].c[1];
        pz = body[i].c[2];
    }
    int k = 0, e = 0;
    if (!C && i < n) e = rowptr[i] + cnt[(int64_t)q * n + i];
    const int ct1 = min(ntiles, (q + 1) * kNearChunk);
    for (int ct = q * kNearChunk; ct < ct1; ct++) {
        const double ex = px - tsph[4 * ct], ey = py - tsph[4 * ct + 1], ez = pz - tsph[4 * ct + 2];
        const double lim = (tsph[4 * ct + 3] + reach[ct]) * (1 + 1e-9);
        const bool mine = i < n && ex * ex + ey * ey + ez * ez <= lim * lim;
        if (!__ballot(mine)) continue;  // uniform
        const int j0 = ct * kNearT;
        __syncthreads();
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            tc[0][threadIdx.x] = body[jl].c[0];
            tc[1][threadIdx.x] = body[jl].c[1];
            tc[2][threadIdx.x] = body[jl].c[2];
            tc[3][threadIdx.x] = factor * body[jl].long_side;
        }
        __syncthreads();
        if (!mine) continue;
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
    if (C && i < n) cnt[(int64_t)q * n + i] = k;
}

// per row i: cnt[q * n + i] becomes the row's count in the chunks before q,
// total[i] the row's count
__global__ void near_chunk_prefix_kernel(int n, int nq, int* __restrict__ cnt, int* __restrict__ total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int run = 0;
    for (int q = 0; q < nq; q++) {
        const int v = cnt[(int64_t)q * n + i];
        cnt[(int64_t)q * n + i] = run;
        run += v;
    }
    total[i] = run;
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

// mark[rt * nkb + kb] = 1 for each pair's (row tile, column block)
__global__ void near_mark_kernel(int nnz, const int* __restrict__ rows, const int* __restrict__ cols, int nkb,
                                 unsigned char* __restrict__ mark) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    mark[(int64_t)(rows[e] / 32) * nkb + cols[e] / 16] = 1;
}

// per row tile rt (one wavefront, its marks read 64 at a time in column
// block order): count (C) or write its marked blocks ascending
template <bool C>
__global__ __launch_bounds__(64) void near_tiles_kernel(const unsigned char* __restrict__ mark, int nrt, int nkb,
                                                        int* __restrict__ cnt, const int* __restrict__ tptr,
                                                        int* __restrict__ tkb) {
    const int rt = blockIdx.x;
    const int lane = threadIdx.x;
    if (rt >= nrt) return;
    const unsigned char* row = mark + (int64_t)rt * nkb;
    const unsigned long long below = (1ull << lane) - 1;
    int k = 0, e = C ? 0 : tptr[rt];
    for (int kb0 = 0; kb0 < nkb; kb0 += 64) {
        const int kb = kb0 + lane;
        const bool on = kb < nkb && row[kb] != 0;
        const unsigned long long bm = __ballot(on);
        if (!C && on) tkb[e + __popcll(bm & below)] = kb;
        k += __popcll(bm);
        e += __popcll(bm);
    }
    if (C && lane == 0) cnt[rt] = k;
}

}  // namespace

void launch_near_reach(const DevGeom& g, double factor, double* reach, hipStream_t s) {
    if (g.n <= 0) return;
    hipLaunchKernelGGL(near_reach_kernel, dim3((unsigned)((g.ntiles + 255) / 256)), dim3(256), 0, s, g.body, g.n,
                       factor, g.ntiles, reach);
    CP_LAUNCH_CHECK();
}

int near_chunks(const DevGeom& g) { return (g.ntiles + kNearChunk - 1) / kNearChunk; }

void launch_near_count(const DevGeom& g, double factor, const double* reach, int* cnt, int* total, hipStream_t s) {
    if (g.n <= 0) return;
    if (g.ntiles != (g.n + kNearT - 1) / kNearT) throw HipError{hipErrorInvalidValue, "tile spheres do not match"};
    const int nq = near_chunks(g);
    hipLaunchKernelGGL(near_pairs_kernel<true>, dim3((unsigned)g.ntiles, (unsigned)nq), dim3(kNearT), 0, s, g.body,
                       g.n, factor, g.tsph, reach, g.ntiles, cnt, nullptr, nullptr, nullptr);
    CP_LAUNCH_CHECK();
    hipLaunchKernelGGL(near_chunk_prefix_kernel, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, g.n, nq, cnt,
                       total);
    CP_LAUNCH_CHECK();
}

void launch_near_fill(const DevGeom& g, double factor, const double* reach, const int* cnt, const int* rowptr,
                      int* cols, int* rows, hipStream_t s) {
    if (g.n <= 0) return;
    if (g.ntiles != (g.n + kNearT - 1) / kNearT) throw HipError{hipErrorInvalidValue, "tile spheres do not match"};
    hipLaunchKernelGGL(near_pairs_kernel<false>, dim3((unsigned)g.ntiles, (unsigned)near_chunks(g)), dim3(kNearT), 0,
                       s, g.body, g.n, factor, g.tsph, reach, g.ntiles, const_cast<int*>(cnt), rowptr, cols, rows);
    CP_LAUNCH_CHECK();
}

void launch_exclusive_scan(const int* in, int n, int* out, hipStream_t s) {
    hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(kScanT), 0, s, in, n, out);
    CP_LAUNCH_CHECK();
}

void launch_near_tile_count(int nnz, const int* rows, const int* cols, int nrt, int nkb, unsigned char* mark,
                            int* cnt, hipStream_t s) {
    if (nnz > 0) {
        hipLaunchKernelGGL(near_mark_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, nnz, rows, cols, nkb,
                           mark);
        CP_LAUNCH_CHECK();
    }
    if (nrt > 0) {
        hipLaunchKernelGGL(near_tiles_kernel<true>, dim3((unsigned)nrt), dim3(64), 0, s, mark, nrt, nkb, cnt, nullptr,
                           nullptr);
        CP_LAUNCH_CHECK();
    }
}

void launch_near_tile_fill(const unsigned char* mark, int nrt, int nkb, const int* tptr, int* tkb, hipStream_t s) {
    if (nrt <= 0) return;
    hipLaunchKernelGGL(near_tiles_kernel<false>, dim3((unsigned)nrt), dim3(64), 0, s, mark, nrt, nkb, nullptr, tptr,
                       tkb);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
