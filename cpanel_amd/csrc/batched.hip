// K3: multi-right-hand-side solve, FP32 Krylov + FP64 iterative refinement.
//
// Up to 64 freestream conditions of one geometry share every pass over the
// matrix: the BiCGSTAB matrix-vector product becomes Y = A32 * P with P an
// N x 64 multivector, a dense contraction that runs on the matrix cores
// (v_mfma_f32_32x32x2f32, FP32 in, FP32 accumulate).  Each column keeps its
// own BiCGSTAB scalars (rho, alpha, omega), so every condition follows
// exactly the recurrence it would follow alone; the columns only share the
// reads of A.
//
// Refinement: x (FP64) is corrected by FP32 inner solves of A32 d = r/|r|;
// the residual r = b - A32 x is formed with FP64 accumulation on the FP64
// matrix cores (v_mfma_f64_16x16x4f64, A32 entries widened exactly).  The
// right-hand side b = -B32 sigma uses the same FP64 kernel.
//
// Layout: multivectors are row-major lda x 64 (the 64 values of one row are
// contiguous), A32/B32 column-major.  GEMM work is split over K into
// kslices; workgroup id -> (kslice = id % kslices, row block = id / kslices)
// so with kslices a multiple of 8 every XCD works on one K slice and keeps
// that slice of the multivector in its own L2.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>

#include "batched.h"
#include "device.h"
#include "timer.h"

namespace cpgpu {

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int KP = kRhsTile;   // 64 columns
constexpr int kT = 256;
constexpr int kRedRows = 64;   // rows per reduction block
constexpr int kRed64Rows = 16; // rows per block of the FP64 residual's reduction (its partials: lda / 16)

enum {
    Q_RHO = 0, Q_ALPHA, Q_OMEGA, Q_RR, Q_R0R, Q_R0V, Q_TT, Q_TS, Q_RR0, Q_BB, Q_RR64, Q_SCALE, Q_BETA, Q_IT, Q_CHK,
    Q_CHKIT, Q_MIN, Q_ITOL, Q_GMNJ, Q_NQ = 20  // Q_ITOL: inner residual reduction of the current sweep;
                                              // Q_GMNJ: GMRES iterations of the last sweep
};

// stagnation window of the inner FP32 solve: a column stops when the best
// residual seen so far fell by less than 2x over the last kStagWin
// iterations (BiCGSTAB plateaus are shorter; a longer stall means the FP32
// solve has reached its attainable accuracy and refinement takes over)
constexpr int kStagWin = 60;

// true when none of the 64 columns is active: the launch belongs to an inner
// iteration queued ahead of the host's convergence check and past the end of
// its sweep (every lane of the wavefront reads one flag)
__device__ inline bool gate_idle(const int* gate) { return __ballot(gate[threadIdx.x & 63] != 0) == 0; }

// ---------------------------------------------------------------- GEMMs
// Y_part[ks] = A32[:, K slice ks] * P[K slice ks, :] for 64 columns.
// RT row tiles of 32 rows per wavefront (4 wavefronts per workgroup), the
// whole 64-column width per wavefront (2 column tiles), K consumed 16 at a
// time.  The next K chunk's operands are loaded into registers before the
// current chunk's MFMAs are issued, so HBM latency overlaps the matrix cores.
// Operand layout of v_mfma_f32_32x32x2f32: lane l supplies A[row l%32][k l/32]
// and B[k l/32][col l%32]; result register j of lane l is row
// 8*(j/4) + 4*(l/32) + j%4, column l%32.
template <int RT>
__global__ __launch_bounds__(256) void gemm_f32_kernel(const float* __restrict__ A, int64_t lda, int n,
                                                       const float* __restrict__ P, float* __restrict__ part,
                                                       int kchunk, int KS, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    constexpr int U = 8;  // K steps of 2 per chunk (16 K values)
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)rb * 4 + wave) * (32 * RT);
    const int k0 = ks * kchunk, k1 = min(n, k0 + kchunk);
    const int lr = lane & 31, lk = lane >> 5;
    floatx16 acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; t++) acc[t][0] = acc[t][1] = floatx16{};
    const float* Ab = A + row0 + lr;
    const float* Pb = P + lr;
    float a[RT][U], b[2][U], an[RT][U], bn[2][U];
    auto load = [&](int k, float (&aa)[RT][U], float (&bb)[2][U]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t kk = k + 2 * u + lk;
            const bool ok = kk < k1;
            const int64_t kc = ok ? kk : 0;
#pragma unroll
            for (int t = 0; t < RT; t++) aa[t][u] = ok ? Ab[kc * lda + 32 * t] : 0.f;
            bb[0][u] = ok ? Pb[kc * KP] : 0.f;
            bb[1][u] = ok ? Pb[kc * KP + 32] : 0.f;
        }
    };
    if (k0 < k1) load(k0, a, b);
    for (int k = k0; k < k1; k += 2 * U) {
        const bool more = k + 2 * U < k1;
        if (more) load(k + 2 * U, an, bn);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int t = 0; t < RT; t++) {
                acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t][u], b[0][u], acc[t][0], 0, 0, 0);
                acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t][u], b[1][u], acc[t][1], 0, 0, 0);
            }
        if (more) {
#pragma unroll
            for (int u = 0; u < U; u++) {
#pragma unroll
                for (int t = 0; t < RT; t++) a[t][u] = an[t][u];
                b[0][u] = bn[0][u];
                b[1][u] = bn[1][u];
            }
        }
    }
    float* out = part + (int64_t)ks * lda * KP;
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int64_t r = row0 + 32 * t + 8 * (j >> 2) + 4 * lk + (j & 3);
            out[r * KP + lr] = acc[t][0][j];
            out[r * KP + 32 + lr] = acc[t][1][j];
        }
}

// Variant with 16-byte A loads: a wavefront owns 128 rows as 4 interleaved
// 32-row tiles (tile t = rows 4m+t, m = 0..31).  Lane l loads the float4 of
// rows 4(l%32)..4(l%32)+3 of column k + l/32; component t is that lane's A
// operand for tile t.  One load instruction feeds 8 MFMAs (4 row tiles x 2
// column tiles); the output rows are un-interleaved at the store.
template <int U>
__global__ __launch_bounds__(256) void gemm_f32_x4_kernel(const float* __restrict__ A, int64_t lda, int n,
                                                          const float* __restrict__ P, float* __restrict__ part,
                                                          int kchunk, int KS, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)rb * 4 + wave) * 128;
    const int k0 = ks * kchunk, k1 = min(n, k0 + kchunk);
    const int m = lane & 31, kh = lane >> 5;
    floatx16 acc[4][2];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t][0] = acc[t][1] = floatx16{};
    const float* Ab = A + row0 + 4 * m;
    const float* Pb = P + m;
    float4 a[U], an[U];
    float b[2][U], bn[2][U];
    auto load = [&](int k, float4 (&aa)[U], float (&bb)[2][U]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t kk = k + 2 * u + kh;
            const bool ok = kk < k1;
            const int64_t kc = ok ? kk : 0;
            float4 v = *reinterpret_cast<const float4*>(Ab + kc * lda);
            aa[u] = ok ? v : make_float4(0.f, 0.f, 0.f, 0.f);
            bb[0][u] = ok ? Pb[kc * KP] : 0.f;
            bb[1][u] = ok ? Pb[kc * KP + 32] : 0.f;
        }
    };
    if (k0 < k1) load(k0, a, b);
    for (int k = k0; k < k1; k += 2 * U) {
        const bool more = k + 2 * U < k1;
        if (more) load(k + 2 * U, an, bn);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const float av[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], b[0][u], acc[t][0], 0, 0, 0);
                acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], b[1][u], acc[t][1], 0, 0, 0);
            }
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                a[u] = an[u];
                b[0][u] = bn[0][u];
                b[1][u] = bn[1][u];
            }
        }
    }
    float* out = part + (int64_t)ks * lda * KP;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int rho = 8 * (j >> 2) + 4 * kh + (j & 3);
            const int64_t r = row0 + 4 * rho + t;
            out[r * KP + m] = acc[t][0][j];
            out[r * KP + 32 + m] = acc[t][1][j];
        }
}

// Same tiling as gemm_f32_x4_kernel with three register stages: the loads of
// chunk i+2 are in flight while chunk i is multiplied, so HBM latency up to
// two chunks' MFMA time is hidden by one wavefront.  Chunks past the K range
// load zeros (their MFMAs add nothing).
template <int U>
__global__ __launch_bounds__(256) void gemm_f32_x4s3_kernel(const float* __restrict__ A, int64_t lda, int n,
                                                            const float* __restrict__ P, float* __restrict__ part,
                                                            int kchunk, int KS, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)rb * 4 + wave) * 128;
    const int k0 = ks * kchunk, k1 = min(n, k0 + kchunk);
    const int m = lane & 31, kh = lane >> 5;
    floatx16 acc[4][2];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t][0] = acc[t][1] = floatx16{};
    const float* Ab = A + row0 + 4 * m;
    const float* Pb = P + m;
    constexpr int D = 2 * U;  // K values per chunk
    float4 a0[U], a1[U], a2[U];
    float b0[2][U], b1[2][U], b2[2][U];
    auto load = [&](int k, float4 (&aa)[U], float (&bb)[2][U]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t kk = k + 2 * u + kh;
            const bool ok = kk < k1;
            const int64_t kc = ok ? kk : 0;
            float4 v = *reinterpret_cast<const float4*>(Ab + kc * lda);
            aa[u] = ok ? v : make_float4(0.f, 0.f, 0.f, 0.f);
            bb[0][u] = ok ? Pb[kc * KP] : 0.f;
            bb[1][u] = ok ? Pb[kc * KP + 32] : 0.f;
        }
    };
    auto mma = [&](const float4 (&aa)[U], const float (&bb)[2][U]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const float av[4] = {aa[u].x, aa[u].y, aa[u].z, aa[u].w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], bb[0][u], acc[t][0], 0, 0, 0);
                acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], bb[1][u], acc[t][1], 0, 0, 0);
            }
        }
    };
    // chunks rounded up to a multiple of 3 (at most two all-zero chunks)
    const int nch = (k1 - k0 + D - 1) / D;
    const int kend = k0 + ((nch + 2) / 3) * 3 * D;
    load(k0, a0, b0);
    load(k0 + D, a1, b1);
    for (int k = k0; k < kend; k += 3 * D) {
        load(k + 2 * D, a2, b2);
        mma(a0, b0);
        load(k + 3 * D, a0, b0);
        mma(a1, b1);
        load(k + 4 * D, a1, b1);
        mma(a2, b2);
    }
    float* out = part + (int64_t)ks * lda * KP;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int rho = 8 * (j >> 2) + 4 * kh + (j & 3);
            const int64_t r = row0 + 4 * rho + t;
            out[r * KP + m] = acc[t][0][j];
            out[r * KP + 32 + m] = acc[t][1][j];
        }
}

// ---------------------------------------------------------------- bf16x3
// FP32 product from bf16 matrix cores: every FP32 value x is split as
// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi) (|x - hi - lo| <=
// 2^-17 |x|), and a*p ~ hi_a hi_p + hi_a lo_p + lo_a hi_p (the dropped lo*lo
// term is <= 2^-16 |a p|), accumulated in FP32 by v_mfma_f32_32x32x16_bf16.
// The inner Krylov solve only needs a 1e-3 reduction per refinement sweep,
// and the FP64 refinement residual uses the exact FP32 matrix, so the
// refined solution is unchanged; the product costs a third of the FP32
// matrix-core time and becomes bound by reading the matrix from HBM.
//
// Packed tiles (one per 32 rows x 16 K values) in two planes, hi and lo: in
// each plane a tile is 64 lanes x 16 bytes, lane l holding row l%32, K values
// 8(l/32)..8(l/32)+7 as 8 bf16: the operand layout of the MFMA, so a
// wavefront reads one contiguous 1 KiB tile per plane with one 16-byte load
// per lane.  The bf16x3 product reads both planes; the bf16x2 product (below)
// reads only the hi plane.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint32_t bf16_bits_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ inline void split8(const float (&v)[8], uint4& hi, uint4& lo) {
    uint32_t h[8], l[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        h[i] = bf16_bits_rne(v[i]);
        const float r = v[i] - __uint_as_float(h[i] << 16);
        l[i] = bf16_bits_rne(r);
    }
    hi = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
    lo = make_uint4(l[0] | (l[1] << 16), l[2] | (l[3] << 16), l[4] | (l[5] << 16), l[6] | (l[7] << 16));
}

// A (column-major, lda) -> tiles [row tile][K block] of the hi plane (out)
// and the lo plane (out + ntiles * 64); K values >= n are 0
__global__ __launch_bounds__(256) void pack_a_bf16x2_kernel(const float* __restrict__ A, int64_t lda, int n,
                                                            int nkb, int64_t ntiles, uint4* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t tile = g >> 6;
    if (tile >= ntiles) return;
    const int lane = (int)(g & 63);
    const int64_t rt = tile / nkb;
    const int kb = (int)(tile % nkb);
    const int64_t row = rt * 32 + (lane & 31);
    const int kbase = kb * 16 + 8 * (lane >> 5);
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kbase + i;
        v[i] = k < n ? A[row + (int64_t)k * lda] : 0.f;
    }
    uint4 hi, lo;
    split8(v, hi, lo);
    out[g] = hi;
    out[ntiles * 64 + g] = lo;
}

// P (row-major lda x 64) -> tiles [K block][column tile 0/1]; rows >= n are 0
__device__ __forceinline__ void pack_p_body(int bx, const float* __restrict__ P, int n, int nkb, uint4* __restrict__ out,
                                   const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    const int64_t g = (int64_t)bx * 256 + threadIdx.x;
    const int64_t tile = g >> 6;
    if (tile >= 2 * (int64_t)nkb) return;
    const int lane = (int)(g & 63);
    const int kb = (int)(tile >> 1), ct = (int)(tile & 1);
    const int col = 32 * ct + (lane & 31);
    const int kbase = kb * 16 + 8 * (lane >> 5);
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kbase + i;
        v[i] = k < n ? P[(int64_t)k * KP + col] : 0.f;
    }
    uint4 hi, lo;
    split8(v, hi, lo);
    out[2 * g] = hi;
    out[2 * g + 1] = lo;
}
__global__ __launch_bounds__(256) void pack_p_bf16x2_kernel(const float* __restrict__ P, int n, int nkb,
                                                            uint4* __restrict__ out, const int* __restrict__ gate) {
    pack_p_body(blockIdx.x, P, n, nkb, out, gate);
}

// Y_part[ks] = A~ P~ over the K blocks of slice ks.  Wavefront: 4 row tiles
// (128 rows) x 2 column tiles; workgroup: 512 rows.  Two register stages.
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(const uint4* __restrict__ Apk, int64_t plane, int nkb,
                                                          const uint4* __restrict__ Ppk, float* __restrict__ part,
                                                          int64_t lda, int kbchunk, int KS, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t rt0 = ((int64_t)rb * 4 + wave) * 4;
    const int kb0 = ks * kbchunk, kb1 = min(nkb, kb0 + kbchunk);
    floatx16 acc[4][2];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t][0] = acc[t][1] = floatx16{};
    uint4 ah[4], al[4], ph[2], pl[2], nah[4], nal[4], nph[2], npl[2];
    auto load = [&](int kb, uint4 (&h)[4], uint4 (&l)[4], uint4 (&qh)[2], uint4 (&ql)[2]) {
        const bool ok = kb < kb1;
        const int kc = ok ? kb : kb0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int64_t o = ((rt0 + t) * nkb + kc) * 64 + lane;
            h[t] = Apk[o];
            l[t] = Apk[plane + o];
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint4* src = Ppk + (((int64_t)kc * 2 + c) * 64 + lane) * 2;
            qh[c] = src[0];
            ql[c] = src[1];
        }
    };
    if (kb0 < kb1) load(kb0, ah, al, ph, pl);
    for (int kb = kb0; kb < kb1; kb++) {
        const bool more = kb + 1 < kb1;
        if (more) load(kb + 1, nah, nal, nph, npl);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const bf16x8 xh = __builtin_bit_cast(bf16x8, ah[t]);
            const bf16x8 xl = __builtin_bit_cast(bf16x8, al[t]);
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const bf16x8 yh = __builtin_bit_cast(bf16x8, ph[c]);
                const bf16x8 yl = __builtin_bit_cast(bf16x8, pl[c]);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, acc[t][c], 0, 0, 0);
            }
        }
        if (more) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                ah[t] = nah[t];
                al[t] = nal[t];
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                ph[c] = nph[c];
                pl[c] = npl[c];
            }
        }
    }
    float* out = part + (int64_t)ks * lda * KP;
    const int m = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int64_t r = (rt0 + t) * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
            out[r * KP + m] = acc[t][0][j];
            out[r * KP + 32 + m] = acc[t][1][j];
        }
}

// Generalised bf16x3 product: RT row tiles per wavefront (workgroup = 4 x RT
// x 32 rows), S register stages (the loads for K block kb + S - 1 are issued
// before block kb is multiplied).  K blocks past the slice load zeros.
template <int RT, int S>
__global__ __launch_bounds__(256) void gemm_bf16x3_s_kernel(const uint4* __restrict__ Apk, int64_t plane,
                                                            int nkb, const uint4* __restrict__ Ppk,
                                                            float* __restrict__ part, int64_t lda, int kbchunk,
                                                            int KS, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t rt0 = ((int64_t)rb * 4 + wave) * RT;
    const int kb0 = ks * kbchunk, kb1 = min(nkb, kb0 + kbchunk);
    floatx16 acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; t++) acc[t][0] = acc[t][1] = floatx16{};
    uint4 ah[S][RT], al[S][RT], ph[S][2], pl[S][2];
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    auto load = [&](int kb, uint4 (&h)[RT], uint4 (&l)[RT], uint4 (&qh)[2], uint4 (&ql)[2]) {
        if (kb < kb1) {
#pragma unroll
            for (int t = 0; t < RT; t++) {
                const int64_t o = ((rt0 + t) * nkb + kb) * 64 + lane;
                h[t] = Apk[o];
                l[t] = Apk[plane + o];
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const uint4* src = Ppk + (((int64_t)kb * 2 + c) * 64 + lane) * 2;
                qh[c] = src[0];
                ql[c] = src[1];
            }
        } else {
#pragma unroll
            for (int t = 0; t < RT; t++) h[t] = l[t] = z;
#pragma unroll
            for (int c = 0; c < 2; c++) qh[c] = ql[c] = z;
        }
    };
    auto mma = [&](const uint4 (&h)[RT], const uint4 (&l)[RT], const uint4 (&qh)[2], const uint4 (&ql)[2]) {
#pragma unroll
        for (int t = 0; t < RT; t++) {
            const bf16x8 xh = __builtin_bit_cast(bf16x8, h[t]);
            const bf16x8 xl = __builtin_bit_cast(bf16x8, l[t]);
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const bf16x8 yh = __builtin_bit_cast(bf16x8, qh[c]);
                const bf16x8 yl = __builtin_bit_cast(bf16x8, ql[c]);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, acc[t][c], 0, 0, 0);
            }
        }
    };
    const int nsteps = kb1 - kb0;
    const int kbend = kb0 + (nsteps > 0 ? (nsteps + S - 1) / S * S : 0);
#pragma unroll
    for (int st = 0; st < S - 1; st++) load(kb0 + st, ah[st], al[st], ph[st], pl[st]);
    for (int kb = kb0; kb < kbend; kb += S) {
#pragma unroll
        for (int st = 0; st < S; st++) {
            const int nx = (st + S - 1) % S;
            load(kb + st + S - 1, ah[nx], al[nx], ph[nx], pl[nx]);
            mma(ah[st], al[st], ph[st], pl[st]);
        }
    }
    float* out = part + (int64_t)ks * lda * KP;
    const int m = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int64_t r = (rt0 + t) * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
            out[r * KP + m] = acc[t][0][j];
            out[r * KP + 32 + m] = acc[t][1][j];
        }
}

// Near-field restoration of the bf16x2 product, per 32 x 16 tile: for row
// tile rt and each K block kb listed for it (the tiles holding a near-field
// entry |c_i - c_j| <= longSide_j, about 1% of them), the lo plane tile
// bf16(A32 - bf16(A32)) times the hi plane of P.  Those tiles then carry A32
// to 2^-16 like the bf16x3 product, the rest A32 to bf16.  One wavefront per
// row tile writes its 32 rows of the extra partial slice (zeros when nothing
// is listed), so the slice is complete for the reduction.
__device__ inline void lo_tile_row(const uint4* __restrict__ Alo, int nkb, const int* __restrict__ tptr,
                                   const int* __restrict__ tkb, const uint4* __restrict__ Ppk,
                                   float* __restrict__ out, int64_t rt, int64_t nrt) {
    if (rt >= nrt) return;
    const int lane = threadIdx.x & 63;
    floatx16 acc0 = floatx16{}, acc1 = floatx16{};
    const int e1 = tptr[rt + 1];
    for (int e = tptr[rt]; e < e1; e++) {
        const int kb = tkb[e];
        const bf16x8 xl = __builtin_bit_cast(bf16x8, Alo[(rt * nkb + kb) * 64 + lane]);
        const bf16x8 y0 = __builtin_bit_cast(bf16x8, Ppk[(((int64_t)kb * 2 + 0) * 64 + lane) * 2]);
        const bf16x8 y1 = __builtin_bit_cast(bf16x8, Ppk[(((int64_t)kb * 2 + 1) * 64 + lane) * 2]);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, y0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, y1, acc1, 0, 0, 0);
    }
    const int m = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int64_t r = rt * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
        out[r * KP + m] = acc0[j];
        out[r * KP + 32 + m] = acc1[j];
    }
}

// bf16x2 product: A rounded once to bf16 (the hi plane alone, 2 bytes per
// entry), the multivector split hi + lo as above, so a*p ~ hi_a (hi_p + lo_p)
// with FP32 accumulation.  Outside the near field the inner Krylov solve
// sees A to bf16 precision (|a - hi_a| <= 2^-9 |a|); the near-field entries,
// which carry the solution, get A32 back to 2^-16 through lo_tile_row (the
// lo plane on the tiles that hold them, one more partial slice of the same
// fixed-order reduction).  Half the matrix bytes of bf16x3; the FP64
// refinement residual is unchanged, so the refined solution is too.
// RT row tiles per wavefront (workgroup = 4 x RT x 32 rows), S register stages.
// Workgroups past the product's (lda / (128 RT)) * KS run the near-field
// restoration (lo_tile_row, 4 row tiles each) on the CUs and wave slots the
// product leaves free, writing the extra partial slice in the same launch.
template <int RT, int S>
__device__ __forceinline__ void gemm_bf16x2_s_body(int bid, const uint4* __restrict__ Apk, int nkb,
                                          const uint4* __restrict__ Ppk, float* __restrict__ part, int64_t lda,
                                          int kbchunk, int KS, int ngemm, int64_t plane,
                                          const int* __restrict__ tptr, const int* __restrict__ tkb,
                                          const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    if (bid >= ngemm) {
        lo_tile_row(Apk + plane, nkb, tptr, tkb, Ppk, part + (int64_t)KS * lda * KP,
                    (int64_t)(bid - ngemm) * 4 + (threadIdx.x >> 6), lda / 32);
        return;
    }
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t rt0 = ((int64_t)rb * 4 + wave) * RT;
    const int kb0 = ks * kbchunk, kb1 = min(nkb, kb0 + kbchunk);
    floatx16 acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; t++) acc[t][0] = acc[t][1] = floatx16{};
    uint4 ah[S][RT], ph[S][2], pl[S][2];
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    auto load = [&](int kb, uint4 (&h)[RT], uint4 (&qh)[2], uint4 (&ql)[2]) {
        if (kb < kb1) {
#pragma unroll
            for (int t = 0; t < RT; t++) {
                // each A tile is read once per product: stream it past the caches
                const u32x4 v = __builtin_nontemporal_load(
                    reinterpret_cast<const u32x4*>(Apk + ((rt0 + t) * nkb + kb) * 64 + lane));
                h[t] = make_uint4(v.x, v.y, v.z, v.w);
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const uint4* src = Ppk + (((int64_t)kb * 2 + c) * 64 + lane) * 2;
                qh[c] = src[0];
                ql[c] = src[1];
            }
        } else {
#pragma unroll
            for (int t = 0; t < RT; t++) h[t] = z;
#pragma unroll
            for (int c = 0; c < 2; c++) qh[c] = ql[c] = z;
        }
    };
    auto mma = [&](const uint4 (&h)[RT], const uint4 (&qh)[2], const uint4 (&ql)[2]) {
#pragma unroll
        for (int t = 0; t < RT; t++) {
            const bf16x8 xh = __builtin_bit_cast(bf16x8, h[t]);
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const bf16x8 yh = __builtin_bit_cast(bf16x8, qh[c]);
                const bf16x8 yl = __builtin_bit_cast(bf16x8, ql[c]);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, acc[t][c], 0, 0, 0);
            }
        }
    };
    const int nsteps = kb1 - kb0;
    const int kbend = kb0 + (nsteps > 0 ? (nsteps + S - 1) / S * S : 0);
#pragma unroll
    for (int st = 0; st < S - 1; st++) load(kb0 + st, ah[st], ph[st], pl[st]);
    for (int kb = kb0; kb < kbend; kb += S) {
#pragma unroll
        for (int st = 0; st < S; st++) {
            const int nx = (st + S - 1) % S;
            load(kb + st + S - 1, ah[nx], ph[nx], pl[nx]);
            mma(ah[st], ph[st], pl[st]);
        }
    }
    float* out = part + (int64_t)ks * lda * KP;
    const int m = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int64_t r = (rt0 + t) * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
            out[r * KP + m] = acc[t][0][j];
            out[r * KP + 32 + m] = acc[t][1][j];
        }
}
template <int RT, int S>
__global__ __launch_bounds__(256) void gemm_bf16x2_s_kernel(const uint4* __restrict__ Apk, int nkb,
                                                            const uint4* __restrict__ Ppk, float* __restrict__ part,
                                                            int64_t lda, int kbchunk, int KS, int ngemm,
                                                            int64_t plane, const int* __restrict__ tptr,
                                                            const int* __restrict__ tkb,
                                                            const int* __restrict__ gate) {
    gemm_bf16x2_s_body<RT, S>(blockIdx.x, Apk, nkb, Ppk, part, lda, kbchunk, KS, ngemm, plane, tptr, tkb, gate);
}

__global__ __launch_bounds__(256) void gemm_f64acc_kernel(const float* __restrict__ A, int64_t lda, int n,
                                                          const double* __restrict__ X, double* __restrict__ part,
                                                          int kchunk, int KS) {
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = (int64_t)rb * 256 + wave * 64;
    const int k0 = ks * kchunk, k1 = min(n, k0 + kchunk);
    const int lr = lane & 15, lk = lane >> 4;
    doublex4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = doublex4{0, 0, 0, 0};
    const float* Ab = A + row0 + lr;
    const double* Xb = X + lr;
    for (int k = k0; k < k1; k += 4) {
        const int64_t kk = k + lk;
        const bool ok = kk < k1;
        double a[4], b[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            a[t] = ok ? (double)Ab[kk * lda + 16 * t] : 0.0;
            b[t] = ok ? Xb[kk * KP + 16 * t] : 0.0;
        }
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
                acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
    }
    double* out = part + (int64_t)ks * lda * KP;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // v_mfma_f64_16x16x4f64 result: register j of lane l holds
                // row 4*j + l/16, column l%16 of the 16x16 tile
                const int64_t r = row0 + 16 * rt + 4 * j + lk;
                out[r * KP + 16 * ct + lr] = acc[rt][ct][j];
            }
}

// Pipelined form: a wavefront owns 64 rows as 4 interleaved 16-row tiles
// (tile t = rows 4m + t, m = 0..15); lane l loads the float4 of rows
// 4(l%16)..+3 at column k + l/16 and its components feed the 4 row tiles.
// U K steps of 4 per chunk; the next chunk's loads are issued before the
// current chunk's MFMAs.
template <class T>
struct Vec4Of;
template <>
struct Vec4Of<float> {
    typedef float4 type;
};
template <>
struct Vec4Of<double> {
    typedef double4 type;
};

template <int U, class T = float>
__global__ __launch_bounds__(256, 2) void gemm_f64acc_p_kernel(const T* __restrict__ A, int64_t lda, int n,
                                                            const double* __restrict__ X, double* __restrict__ part,
                                                            int kchunk, int KS) {
    typedef typename Vec4Of<T>::type V4;
    const int bid = blockIdx.x;
    const int ks = bid % KS, rb = bid / KS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = (int64_t)rb * 256 + wave * 64;
    const int k0 = ks * kchunk, k1 = min(n, k0 + kchunk);
    const int m = lane & 15, lk = lane >> 4;
    doublex4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = doublex4{0, 0, 0, 0};
    const T* Ab = A + row0 + 4 * m;
    const double* Xb = X + m;
    V4 a[U], an[U];
    double b[U][4], bn[U][4];
    auto load = [&](int k, V4 (&aa)[U], double (&bb)[U][4]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t kk = k + 4 * u + lk;
            const bool ok = kk < k1;
            const int64_t kc = ok ? kk : 0;
            V4 v = *reinterpret_cast<const V4*>(Ab + kc * lda);
            if (!ok) v.x = v.y = v.z = v.w = 0;
            aa[u] = v;
#pragma unroll
            for (int c = 0; c < 4; c++) bb[u][c] = ok ? Xb[kc * KP + 16 * c] : 0.0;
        }
    };
    if (k0 < k1) load(k0, a, b);
    for (int k = k0; k < k1; k += 4 * U) {
        const bool more = k + 4 * U < k1;
        if (more) load(k + 4 * U, an, bn);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const double av[4] = {(double)a[u].x, (double)a[u].y, (double)a[u].z, (double)a[u].w};
#pragma unroll
            for (int rt = 0; rt < 4; rt++)
#pragma unroll
                for (int ct = 0; ct < 4; ct++)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rt], b[u][ct], acc[rt][ct], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                a[u] = an[u];
#pragma unroll
                for (int c = 0; c < 4; c++) b[u][c] = bn[u][c];
            }
        }
    }
    double* out = part + (int64_t)ks * lda * KP;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t r = row0 + 4 * (4 * j + lk) + rt;
                out[r * KP + 16 * ct + m] = acc[rt][ct][j];
            }
}

// ---------------------------------------------------------------- helpers
// Per-column block partial sums: thread t -> column t%64, row group t/64.
__device__ void col_partial(double v, double* sh, double* dst) {
    sh[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < KP) {
        double s = sh[threadIdx.x] + sh[threadIdx.x + 64];
        s += sh[threadIdx.x + 128];
        s += sh[threadIdx.x + 192];
        dst[threadIdx.x] = s;
    }
    __syncthreads();
}

__global__ void sigma_rows_kernel(int n, int64_t lda, int kc, const double* __restrict__ nrm,
                                  const double* __restrict__ vinf4, double* __restrict__ S) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= lda * KP) return;
    int64_t i = idx / KP;
    int c = (int)(idx % KP);
    double v = 0;
    if (i < n && c < kc)
        v = -(vinf4[4 * c] * nrm[i] + vinf4[4 * c + 1] * nrm[lda + i] + vinf4[4 * c + 2] * nrm[2 * lda + i]);
    S[idx] = v;
}

// S(i, d) = n_d(i) for d < 3, 0 elsewhere (the normals as three columns)
__global__ void normal_cols_kernel(int n, int64_t lda, const double* __restrict__ nrm, double* __restrict__ S) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= lda * KP) return;
    int64_t i = idx / KP;
    int d = (int)(idx % KP);
    S[idx] = (i < n && d < 3) ? nrm[d * lda + i] : 0.0;
}

// bn(i, d) = Y(i, d) for d < 3 (Y = -B N, lda x 64)
__global__ void keep3_kernel(int64_t lda, const double* __restrict__ Y, double* __restrict__ bn) {
    int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= lda) return;
    for (int d = 0; d < 3; d++) bn[4 * i + d] = Y[i * KP + d];
    bn[4 * i + 3] = 0.0;
}

// R(i, c) = -(bn(i, :) . Vinf_c) = -B sigma_c for the steady source
// strengths sigma_c = -Vinf_c . n (cpCase::setSourceStrengths), c < kc
__global__ void rhs_bn_kernel(int n, int64_t lda, int kc, const double* __restrict__ bn,
                              const double* __restrict__ vinf4, double* __restrict__ R) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= lda * KP) return;
    int64_t i = idx / KP;
    int c = (int)(idx % KP);
    double v = 0;
    if (i < n && c < kc)
        v = -(bn[4 * i] * vinf4[4 * c] + bn[4 * i + 1] * vinf4[4 * c + 1] + bn[4 * i + 2] * vinf4[4 * c + 2]);
    R[idx] = v;
}

// Y = Badd + sign * (sum_ks part + D X) ; partial ||Y_c||^2.  D is the
// optional sparse FP64 near-field correction (CSR by row).
// Y = Badd + sign * (sum_ks part + near-field correction), rows >= n zeroed,
// and the per-block partials of ||Y_c||^2.  A workgroup owns kRed64Rows rows
// (four per 64-lane group, lane = column); a row's near-field entries (~120)
// run as four independent chains so the gathers of X overlap.
__global__ __launch_bounds__(256) void reduce64_kernel(const double* __restrict__ part, int KS, int64_t lda, int n,
                                                       const double* __restrict__ Badd, double sign,
                                                       double* __restrict__ Y, double* __restrict__ dpart,
                                                       const int* __restrict__ rowptr, const int* __restrict__ cols,
                                                       const double* __restrict__ vals,
                                                       const double* __restrict__ X) {
    __shared__ double sh[kT];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    double acc = 0;
    for (int rr = g; rr < kRed64Rows; rr += 4) {
        int64_t i = (int64_t)blockIdx.x * kRed64Rows + rr;
        int64_t o = i * KP + c;
        double s = 0;
#pragma unroll 8
        for (int q = 0; q < KS; q++) s += part[(int64_t)q * lda * KP + o];
        if (vals && i < n) {
            const int e1 = rowptr[i + 1];
            int e = rowptr[i];
            double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
            for (; e + 4 <= e1; e += 4) {
                t0 += vals[e] * X[(int64_t)cols[e] * KP + c];
                t1 += vals[e + 1] * X[(int64_t)cols[e + 1] * KP + c];
                t2 += vals[e + 2] * X[(int64_t)cols[e + 2] * KP + c];
                t3 += vals[e + 3] * X[(int64_t)cols[e + 3] * KP + c];
            }
            for (; e < e1; e++) t0 += vals[e] * X[(int64_t)cols[e] * KP + c];
            s += (t0 + t1) + (t2 + t3);
        }
        double y = (Badd ? Badd[o] : 0.0) + sign * s;
        if (i >= n) y = 0;
        Y[o] = y;
        acc += y * y;
    }
    col_partial(acc, sh, dpart + (int64_t)blockIdx.x * KP);
}

// Y = sum_ks part (FP32, fixed order); rows >= n zeroed
__global__ __launch_bounds__(256) void reduce32_kernel(const float* __restrict__ part, int KS, int64_t lda, int n,
                                                       float* __restrict__ Y) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= lda * KP) return;
    float s = 0;
#pragma unroll 8
    for (int q = 0; q < KS; q++) s += part[(int64_t)q * lda * KP + idx];
    Y[idx] = (idx / KP) < n ? s : 0.f;
}

// rho/beta step at the top of an inner iteration; deactivates converged or
// broken-down columns (breakdown is left to the next refinement sweep)
__device__ inline void scal_beta_col(double* __restrict__ s, int* __restrict__ act_c) {
    if (!*act_c) return;
    if (s[Q_RR] <= s[Q_ITOL] * s[Q_ITOL] * s[Q_RR0]) {
        *act_c = 0;
        return;
    }
    const double it = s[Q_IT];
    s[Q_MIN] = fmin(s[Q_MIN], s[Q_RR]);
    if (it - s[Q_CHKIT] >= kStagWin) {
        if (s[Q_MIN] > 0.25 * s[Q_CHK]) {  // best |r| fell by less than 2x over the window
            *act_c = 0;
            return;
        }
        s[Q_CHK] = s[Q_MIN];
        s[Q_CHKIT] = it;
    }
    s[Q_IT] = it + 1;
    double rho_old = s[Q_RHO];
    double rho = s[Q_R0R];
    const double eps = 1.1920928955078125e-07;
    if (!(fabs(rho) > eps * eps * s[Q_RR0]) || !(s[Q_OMEGA] != 0)) {
        *act_c = 0;
        return;
    }
    s[Q_BETA] = (rho / rho_old) * (s[Q_ALPHA] / s[Q_OMEGA]);
    s[Q_RHO] = rho;
}

__global__ void scal_beta_kernel(double* __restrict__ sc, int* __restrict__ act) {
    int c = threadIdx.x;
    if (c >= KP) return;
    scal_beta_col(sc + c * Q_NQ, act + c);
}

// sum of per-block column partials (fixed order) into sc[c][q], then the
// scalar step that consumes them on the active columns (op: COLSUM_ALPHA
// alpha = rho / (r0 . v); COLSUM_OMEGA omega = (t . s) / (t . t);
// COLSUM_BETA the next iteration's rho/beta step)
// 1024 threads: column c = t % 64, group g = t / 64 sums blocks b = g mod 16;
// the 16 group sums are then added in a fixed order
enum { COLSUM_NONE = 0, COLSUM_ALPHA = 1, COLSUM_OMEGA = 2, COLSUM_BETA = 3 };
__global__ __launch_bounds__(1024) void colsum_kernel(const double* __restrict__ dpart, int nblk, int nq, int q0,
                                                      int q1, double* __restrict__ sc, int op,
                                                      int* __restrict__ act) {
    __shared__ double sh[1024];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (int q = 0; q < nq; q++) {
        double s = 0;
        const double* d = dpart + (int64_t)q * nblk * KP;
        // blocks g, g + 16, ... added in that order; eight loads in flight
        // (a masked slot adds +0.0, which leaves a sum started at +0 as it is)
        for (int b = g; b < nblk; b += 128) {
            double v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = b + 16 * k < nblk ? d[(int64_t)(b + 16 * k) * KP + c] : 0.0;
#pragma unroll
            for (int k = 0; k < 8; k++) s += v[k];
        }
        sh[threadIdx.x] = s;
        __syncthreads();
        if (g == 0) {
            double t = 0;
            for (int k = 0; k < 16; k++) t += sh[k * 64 + c];
            sc[c * Q_NQ + (q == 0 ? q0 : q1)] = t;
        }
        __syncthreads();
    }
    if (g != 0 || op == COLSUM_NONE) return;
    double* s = sc + c * Q_NQ;
    if (op == COLSUM_BETA)
        scal_beta_col(s, act + c);
    else if (op == COLSUM_ALPHA && act[c])
        s[Q_ALPHA] = s[Q_RHO] / s[Q_R0V];
    else if (op == COLSUM_OMEGA && act[c])
        s[Q_OMEGA] = s[Q_TT] > 0 ? s[Q_TS] / s[Q_TT] : 0.0;
}

// inner init: r = R/|R| (FP32), r0 = r, p = v = d = 0, partial r.r
__global__ __launch_bounds__(256) void inner_init_kernel(int n, const double* __restrict__ R,
                                                         const double* __restrict__ sc, const int* __restrict__ act,
                                                         float* __restrict__ r, float* __restrict__ r0,
                                                         float* __restrict__ p, float* __restrict__ v,
                                                         float* __restrict__ d, double* __restrict__ dpart) {
    __shared__ double sh[kT];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    double scale = sc[c * Q_NQ + Q_SCALE];
    double inv = scale > 0 ? 1.0 / scale : 0.0;
    double acc = 0;
    for (int rr = g; rr < kRedRows; rr += 4) {
        int64_t i = (int64_t)blockIdx.x * kRedRows + rr;
        int64_t o = i * KP + c;
        float ri = (act[c] && i < n) ? (float)(R[o] * inv) : 0.f;
        r[o] = ri;
        r0[o] = ri;
        p[o] = 0.f;
        v[o] = 0.f;
        d[o] = 0.f;
        acc += (double)ri * ri;
    }
    col_partial(acc, sh, dpart + (int64_t)blockIdx.x * KP);
}

// per-column scalar steps (one block of 64 threads)
__global__ void scal_init_kernel(double* __restrict__ sc) {
    int c = threadIdx.x;
    if (c >= KP) return;
    double* s = sc + c * Q_NQ;
    s[Q_RR0] = s[Q_RR];
    s[Q_R0R] = s[Q_RR];
    s[Q_RHO] = 1;
    s[Q_ALPHA] = 1;
    s[Q_OMEGA] = 1;
    s[Q_IT] = 0;
    s[Q_CHK] = s[Q_RR];
    s[Q_CHKIT] = 0;
    s[Q_MIN] = s[Q_RR];
}

// p = r + beta (p - omega v); y = invd p (Jacobi), or p also written in
// block order to pp (row i -> iperm[i]) for the block-Jacobi product
__global__ __launch_bounds__(256) void inner_p_kernel(int64_t rows, int n, const double* __restrict__ sc,
                                                      const int* __restrict__ act, const float* __restrict__ r,
                                                      float* __restrict__ p, const float* __restrict__ v,
                                                      const float* __restrict__ invd, float* __restrict__ y,
                                                      const int* __restrict__ iperm, float* __restrict__ pp) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= rows * KP) return;
    int c = (int)(idx & 63);
    if (!act[c]) return;
    float beta = (float)sc[c * Q_NQ + Q_BETA], omega = (float)sc[c * Q_NQ + Q_OMEGA];
    float pi = r[idx] + beta * (p[idx] - omega * v[idx]);
    p[idx] = pi;
    if (invd) y[idx] = invd[idx >> 6] * pi;
    const int64_t i = idx >> 6;
    if (iperm && i < n) pp[(int64_t)iperm[i] * KP + c] = pi;
}

// v = sum parts; partial r0.v
__global__ __launch_bounds__(256) void inner_v_kernel(const float* __restrict__ part, int KS, int64_t lda, int n,
                                                      const int* __restrict__ act, float* __restrict__ v,
                                                      const float* __restrict__ r0, double* __restrict__ dpart) {
    __shared__ double sh[kT];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    double acc = 0;
    if (act[c]) {
        for (int rr = g; rr < kRedRows; rr += 4) {
            int64_t i = (int64_t)blockIdx.x * kRedRows + rr;
            int64_t o = i * KP + c;
            float s = 0;
#pragma unroll 8
            for (int q = 0; q < KS; q++) s += part[(int64_t)q * lda * KP + o];
            if (i >= n) s = 0;
            v[o] = s;
            acc += (double)r0[o] * s;
        }
    }
    col_partial(acc, sh, dpart + (int64_t)blockIdx.x * KP);
}

// s = r - alpha v; z = invd s (Jacobi), or s also written in block order
__global__ __launch_bounds__(256) void inner_s_kernel(int64_t rows, int n, const double* __restrict__ sc,
                                                      const int* __restrict__ act, const float* __restrict__ r,
                                                      const float* __restrict__ v, const float* __restrict__ invd,
                                                      float* __restrict__ s, float* __restrict__ z,
                                                      const int* __restrict__ iperm, float* __restrict__ pp) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= rows * KP) return;
    int c = (int)(idx & 63);
    if (!act[c]) return;
    float alpha = (float)sc[c * Q_NQ + Q_ALPHA];
    float si = r[idx] - alpha * v[idx];
    s[idx] = si;
    if (invd) z[idx] = invd[idx >> 6] * si;
    const int64_t i = idx >> 6;
    if (iperm && i < n) pp[(int64_t)iperm[i] * KP + c] = si;
}

// t = sum parts; partial t.t and t.s
__global__ __launch_bounds__(256) void inner_t_kernel(const float* __restrict__ part, int KS, int64_t lda, int n,
                                                      const int* __restrict__ act, float* __restrict__ t,
                                                      const float* __restrict__ s, double* __restrict__ dtt,
                                                      double* __restrict__ dts) {
    __shared__ double sh[kT];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    double a = 0, b = 0;
    if (act[c]) {
        for (int rr = g; rr < kRedRows; rr += 4) {
            int64_t i = (int64_t)blockIdx.x * kRedRows + rr;
            int64_t o = i * KP + c;
            float x = 0;
#pragma unroll 8
            for (int q = 0; q < KS; q++) x += part[(int64_t)q * lda * KP + o];
            if (i >= n) x = 0;
            t[o] = x;
            a += (double)x * x;
            b += (double)x * s[o];
        }
    }
    col_partial(a, sh, dtt + (int64_t)blockIdx.x * KP);
    col_partial(b, sh, dts + (int64_t)blockIdx.x * KP);
}

// d += alpha y + omega z; r = s - omega t; partial r.r and r0.r
__global__ __launch_bounds__(256) void inner_x_kernel(int n, const double* __restrict__ sc,
                                                      const int* __restrict__ act, float* __restrict__ d,
                                                      const float* __restrict__ y, const float* __restrict__ z,
                                                      float* __restrict__ r, const float* __restrict__ s,
                                                      const float* __restrict__ t, const float* __restrict__ r0,
                                                      double* __restrict__ drr, double* __restrict__ dr0r) {
    __shared__ double sh[kT];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    double a = 0, b = 0;
    if (act[c]) {
        float alpha = (float)sc[c * Q_NQ + Q_ALPHA], omega = (float)sc[c * Q_NQ + Q_OMEGA];
        for (int rr = g; rr < kRedRows; rr += 4) {
            int64_t i = (int64_t)blockIdx.x * kRedRows + rr;
            int64_t o = i * KP + c;
            d[o] += alpha * y[o] + omega * z[o];
            float ri = s[o] - omega * t[o];
            r[o] = ri;
            a += (double)ri * ri;
            b += (double)r0[o] * ri;
        }
    }
    col_partial(a, sh, drr + (int64_t)blockIdx.x * KP);
    col_partial(b, sh, dr0r + (int64_t)blockIdx.x * KP);
}

// X += scale_c * d  (FP64 accumulation of the FP32 correction)
// ---------------------------------------------------------------- GMRES
// Right-preconditioned GMRES per column, the inner solve used with the strip
// block-Jacobi preconditioner: one product per iteration against BiCGSTAB's
// two, and the minimal residual over the Krylov space (a numpy restatement
// of the 4,956-panel case needs 86 products per column for the four
// refinement sweeps where the batched BiCGSTAB takes 182).  The basis
// vectors u_q are stored unnormalised in block order (the preconditioner's
// row order) with their norms s_q, v_q = u_q / s_q.  Per column a state
// block of kGmStride doubles holds the Hessenberg columns (already rotated),
// the Givens rotations, the rotated right-hand side g, the norms, the
// solution coefficients y and the iteration count.  A sweep that reaches
// the Krylov dimension simply ends: the FP64 refinement is the restart.
constexpr int kGmMax = 48;                      // Krylov dimension bound per sweep
constexpr int GM_H = 0;                         // column j of H at GM_H + j * (kGmMax + 1)
constexpr int GM_CS = GM_H + (kGmMax + 1) * kGmMax;
constexpr int GM_SN = GM_CS + kGmMax;
constexpr int GM_G = GM_SN + kGmMax;            // kGmMax + 1
constexpr int GM_S = GM_G + kGmMax + 1;         // kGmMax + 1 norms s_q
constexpr int GM_Y = GM_S + kGmMax + 1;         // kGmMax coefficients
constexpr int GM_NJ = GM_Y + kGmMax;            // iterations taken
constexpr int kGmStride = GM_NJ + 1;

// u_0 = R / scale in block order on the sweeping columns; ||u_0||^2 partials
__global__ __launch_bounds__(256) void gm_init_kernel(int n, const double* __restrict__ R,
                                                      const double* __restrict__ sc, const int* __restrict__ act,
                                                      const int* __restrict__ iperm, float* __restrict__ U0,
                                                      double* __restrict__ dpart) {
    __shared__ double sh[kT];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const double scale = sc[c * Q_NQ + Q_SCALE];
    const double inv = scale > 0 ? 1.0 / scale : 0.0;
    double acc = 0;
    const bool on = act[c] != 0;
    // rows g, g + 4, ... in order, eight at a time
    for (int rr0 = g; rr0 < kRedRows; rr0 += 32) {
        double r8[8];
        int p8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int64_t i = (int64_t)blockIdx.x * kRedRows + rr0 + 4 * k;
            const bool ok = rr0 + 4 * k < kRedRows && i < n;
            r8[k] = ok && on ? R[i * KP + c] : 0.0;
            p8[k] = ok ? iperm[i] : -1;
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (p8[k] >= 0) {
                const float u = on ? (float)(r8[k] * inv) : 0.f;
                U0[(int64_t)p8[k] * KP + c] = u;
                acc += (double)u * u;
            }
    }
    col_partial(acc, sh, dpart + (int64_t)blockIdx.x * KP);
}

// sum of the per-block partials of quantity q (fixed order) for column c,
// with 1024 threads: g = t / 64 sums blocks g, g + 16, ... (sixteen running
// sums over blocks g + 256 i + 16 r, r = 0..15, so sixteen loads are in
// flight); thread group 0 adds the 16 group sums and returns the total
// (others: 0)
__device__ inline double colsum1024(const double* __restrict__ d, int nblk, double* sh) {
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    double r[16];
#pragma unroll
    for (int k = 0; k < 16; k++) r[k] = 0;
    int b = g;
    for (; b + 240 < nblk; b += 256)
#pragma unroll
        for (int k = 0; k < 16; k++) r[k] += d[(int64_t)(b + 16 * k) * KP + c];
    // the rest into r[0] in block order, eight loads in flight (a masked slot
    // adds +0.0)
    for (; b < nblk; b += 128) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = b + 16 * k < nblk ? d[(int64_t)(b + 16 * k) * KP + c] : 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) r[0] += v[k];
    }
#pragma unroll
    for (int w = 8; w >= 1; w >>= 1)
#pragma unroll
        for (int k = 0; k < w; k++) r[k] += r[k + w];
    const double s = r[0];
    sh[threadIdx.x] = s;
    __syncthreads();
    double t = 0;
    if (g == 0)
        for (int k = 0; k < 16; k++) t += sh[k * 64 + c];
    __syncthreads();
    return t;
}

// s_0 = ||u_0||, g_0 = s_0, no iterations yet
__global__ __launch_bounds__(1024) void gm_start_kernel(const double* __restrict__ dpart, int nblk,
                                                        double* __restrict__ st) {
    __shared__ double sh[1024];
    const double t = colsum1024(dpart, nblk, sh);
    if (threadIdx.x >= KP) return;
    double* S = st + threadIdx.x * kGmStride;
    S[GM_S] = sqrt(t);
    S[GM_G] = sqrt(t);
    S[GM_NJ] = 0;
}

// w = sum_ks part (the product A~ z_j, row order) stored as u_{j+1} in block
// order; a thread sums four columns (skipped when none of them is active;
// an inactive column's u_{j+1} is never read)
__device__ __forceinline__ void gm_w_body(int bx, const float* __restrict__ part, int KS, int64_t lda, int n,
                                 const int* __restrict__ act, const int* __restrict__ iperm, float* __restrict__ W) {
    const int64_t idx = (int64_t)bx * kT + threadIdx.x;  // float4 index
    if (idx >= (int64_t)n * (KP / 4)) return;
    const int c = 4 * (int)(idx & 15);
    if (!(act[c] | act[c + 1] | act[c + 2] | act[c + 3])) return;
    const floatx4* P4 = reinterpret_cast<const floatx4*>(part);
    floatx4 v = floatx4{};
#pragma unroll 8
    for (int q = 0; q < KS; q++) v += P4[(int64_t)q * lda * (KP / 4) + idx];
    *reinterpret_cast<floatx4*>(W + (int64_t)iperm[idx >> 4] * KP + c) = v;
}
__global__ __launch_bounds__(256) void gm_w_kernel(const float* __restrict__ part, int KS, int64_t lda, int n,
                                                   const int* __restrict__ act, const int* __restrict__ iperm,
                                                   float* __restrict__ W) {
    gm_w_body(blockIdx.x, part, KS, lda, n, act, iperm, W);
}

// dots u_t . u_{j+1}: one wavefront per (chunk of kGmDotRows block-order
// positions, t <= j), lane = column; 32 position pairs in flight, added in
// position order; each wavefront writes its own partial (no block reduction)
constexpr int kGmDotRows = 256;
__device__ __forceinline__ void gm_dots_body(int bx, int by, int nchunk, int n, const int* __restrict__ act,
                                    const float* __restrict__ U, int64_t us, int j, double* __restrict__ dpart) {
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int t = by * 4 + g;
    if (t > j) return;  // whole wavefront
    const int64_t chunk = bx;
    double acc = 0;
    if (act[c]) {
        const float* ut = U + (int64_t)t * us + c;
        const float* w = U + (int64_t)(j + 1) * us + c;
        const int64_t p1 = std::min((int64_t)n, (chunk + 1) * kGmDotRows);
        int64_t p = chunk * kGmDotRows;
        for (; p + 32 <= p1; p += 32) {
            float a[32], b[32];
#pragma unroll
            for (int k = 0; k < 32; k++) {
                a[k] = ut[(p + k) * KP];
                b[k] = w[(p + k) * KP];
            }
#pragma unroll
            for (int k = 0; k < 32; k++) acc += (double)a[k] * b[k];
        }
        for (; p + 8 <= p1; p += 8) {
            float a[8], b[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                a[k] = ut[(p + k) * KP];
                b[k] = w[(p + k) * KP];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) acc += (double)a[k] * b[k];
        }
        for (; p < p1; p++) acc += (double)ut[p * KP] * w[p * KP];
    }
    dpart[((int64_t)t * nchunk + chunk) * KP + c] = acc;
}
__global__ __launch_bounds__(256) void gm_dots_kernel(int n, const int* __restrict__ act, const float* __restrict__ U,
                                                      int64_t us, int j, double* __restrict__ dpart) {
    gm_dots_body(blockIdx.x, blockIdx.y, gridDim.x, n, act, U, us, j, dpart);
}

// Hessenberg entries h_{q,j} = (u_q . w) / s_q, one quantity q per block
__device__ __forceinline__ void gm_hcol_body(int q, double* sh, const double* __restrict__ dpart, int nblk,
                                    const int* __restrict__ act, int j, double* __restrict__ st,
                                    float* __restrict__ coef) {
    const double t = colsum1024(dpart + (int64_t)q * nblk * KP, nblk, sh);
    if (threadIdx.x >= KP || !act[threadIdx.x]) return;
    double* S = st + threadIdx.x * kGmStride;
    const double h = t / S[GM_S + q];
    S[GM_H + j * (kGmMax + 1) + q] = h;
    coef[q * KP + threadIdx.x] = (float)(h / S[GM_S + q]);
}
__global__ __launch_bounds__(1024) void gm_hcol_kernel(const double* __restrict__ dpart, int nblk,
                                                       const int* __restrict__ act, int j, double* __restrict__ st,
                                                       float* __restrict__ coef) {
    __shared__ double sh[1024];
    gm_hcol_body(blockIdx.x, sh, dpart, nblk, act, j, st, coef);
}

// u_{j+1} = w - sum_{q <= j} h_{q,j} v_q on the active columns (coef[t][c] =
// h_{t,j} / s_t), and the partials of ||u_{j+1}||^2: kGmUpRounds groups of
// kGmUpPos block-order positions per thread (all rounds' positions loaded
// together), four basis vectors at a time (64 loads in flight), subtracted in
// t order, 4 * kGmUpRounds * kGmUpPos positions and one partial per
// workgroup
constexpr int kGmUpPos = 8;
constexpr int kGmUpRounds = 2;
__device__ __forceinline__ void gm_update_body(int bx, double* sh, int n, const int* __restrict__ act, float* __restrict__ U,
                                      int64_t us, int j, const float* __restrict__ coef, double* __restrict__ dpart) {
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    constexpr int NP = kGmUpRounds * kGmUpPos;
    double acc = 0;
    if (act[c]) {
        // round r's positions p0(r) .. p0(r) + kGmUpPos - 1, all rounds at once
        int64_t pp[NP];
#pragma unroll
        for (int r = 0; r < kGmUpRounds; r++)
#pragma unroll
            for (int k = 0; k < kGmUpPos; k++)
                pp[r * kGmUpPos + k] = (((int64_t)bx * kGmUpRounds + r) * 4 + g) * kGmUpPos + k;
        float* w = U + (int64_t)(j + 1) * us + c;
        float v[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) v[k] = pp[k] < n ? w[pp[k] * KP] : 0.f;
        // four basis vectors per batch (the last batch masked), subtracted in t order
        for (int t = 0; t <= j; t += 4) {
            float f[4], x[4][NP];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const bool tq = t + q <= j;
                f[q] = tq ? coef[(t + q) * KP + c] : 0.f;
                const float* ut = U + (int64_t)(t + q) * us + c;
#pragma unroll
                for (int k = 0; k < NP; k++) x[q][k] = tq && pp[k] < n ? ut[pp[k] * KP] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (t + q <= j)
#pragma unroll
                    for (int k = 0; k < NP; k++) v[k] -= f[q] * x[q][k];
        }
#pragma unroll
        for (int k = 0; k < NP; k++)
            if (pp[k] < n) {
                w[pp[k] * KP] = v[k];
                acc += (double)v[k] * v[k];
            }
    }
    col_partial(acc, sh, dpart + (int64_t)bx * KP);
}
__global__ __launch_bounds__(256) void gm_update_kernel(int n, const int* __restrict__ act, float* __restrict__ U,
                                                        int64_t us, int j, const float* __restrict__ coef,
                                                        double* __restrict__ dpart) {
    __shared__ double sh[kT];
    gm_update_body(blockIdx.x, sh, n, act, U, us, j, coef, dpart);
}

// h_{j+1,j} = ||u_{j+1}||; the earlier rotations and a new one applied to
// column j; g rotated; a column stops when |g_{j+1}| <= tol s_0, at the
// Krylov bound m, or on a breakdown (h_{j+1,j} = 0: the solution is exact);
// tol is the column's sc[Q_ITOL]
__device__ __forceinline__ void gm_givens_col(double t, int* __restrict__ act, int j, int m, const double* __restrict__ sc,
                                     double* __restrict__ st) {
    double* S = st + threadIdx.x * kGmStride;
    double* h = S + GM_H + j * (kGmMax + 1);
    const double hn = sqrt(t);
    // rotation i maps (h_i, h_{i+1}) to (cs h_i + sn h_{i+1}, -sn h_i + cs h_{i+1});
    // hc carries the rotated h_i, and the rotations' operands are loaded eight
    // at a time (h_{i+1} is still the unrotated value when rotation i reads it)
    double hc = h[0];
    for (int i0 = 0; i0 < j; i0 += 8) {
        double c8[8], s8[8], n8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool ok = i0 + k < j;
            c8[k] = ok ? S[GM_CS + i0 + k] : 0.0;
            s8[k] = ok ? S[GM_SN + i0 + k] : 0.0;
            n8[k] = ok ? h[i0 + k + 1] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (i0 + k < j) {
                const double cs = c8[k], sn = s8[k];
                const double a = cs * hc + sn * n8[k];
                hc = -sn * hc + cs * n8[k];
                h[i0 + k] = a;
            }
    }
    const double den = sqrt(hc * hc + hn * hn);
    const double cs = den > 0 ? hc / den : 1.0, sn = den > 0 ? hn / den : 0.0;
    S[GM_CS + j] = cs;
    S[GM_SN + j] = sn;
    h[j] = den;
    h[j + 1] = 0;
    const double gj = S[GM_G + j];
    S[GM_G + j + 1] = -sn * gj;
    S[GM_G + j] = cs * gj;
    S[GM_S + j + 1] = hn;
    S[GM_NJ] = j + 1;
    const double tol = sc[threadIdx.x * Q_NQ + Q_ITOL];
    if (!(fabs(S[GM_G + j + 1]) > tol * S[GM_S]) || j + 1 >= m || !(hn > 0)) act[threadIdx.x] = 0;
}
// gflag (or null): set to whether any column is still active (wavefront 0
// holds the 64 columns)
__device__ __forceinline__ void gm_givens_body(double* sh, const double* __restrict__ dpart, int nblk, int* __restrict__ act,
                                      int j, int m, const double* __restrict__ sc, double* __restrict__ st,
                                      int* __restrict__ gflag) {
    const double t = colsum1024(dpart, nblk, sh);
    if (threadIdx.x < KP && act[threadIdx.x]) gm_givens_col(t, act, j, m, sc, st);
    if (gflag && threadIdx.x < KP) {
        const bool any = __ballot(act[threadIdx.x] != 0) != 0;
        if (threadIdx.x == 0) *gflag = any ? 1 : 0;
    }
}
__global__ __launch_bounds__(1024) void gm_givens_kernel(const double* __restrict__ dpart, int nblk,
                                                         int* __restrict__ act, int j, int m,
                                                         const double* __restrict__ sc, double* __restrict__ st) {
    __shared__ double sh[1024];
    gm_givens_body(sh, dpart, nblk, act, j, m, sc, st, nullptr);
}

// y = H^-1 g over the iterations each column took (upper triangular); the
// count of those iterations goes to sc[Q_GMNJ] for the host
// (y kept in LDS, one column per thread; the H entries of a row are loaded
// eight at a time).  yc[q][c] = (float)(y_q / s_q) and yc[kGmMax][c] = the
// column's iteration count: the combination's coefficients, one row per q
__global__ __launch_bounds__(KP) void gm_backsub_kernel(double* __restrict__ st, double* __restrict__ sc,
                                                        float* __restrict__ yc) {
    __shared__ double ys[kGmMax][KP];
    const int c = threadIdx.x;
    if (c >= KP) return;
    double* S = st + c * kGmStride;
    const int nj = (int)S[GM_NJ];
    sc[c * Q_NQ + Q_GMNJ] = nj;
    for (int i = nj - 1; i >= 0; i--) {
        double v = S[GM_G + i];
        for (int k0 = i + 1; k0 < nj; k0 += 8) {
            double h8[8];
#pragma unroll
            for (int k = 0; k < 8; k++) h8[k] = k0 + k < nj ? S[GM_H + (k0 + k) * (kGmMax + 1) + i] : 0.0;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k0 + k < nj) v -= h8[k] * ys[k0 + k][c];
        }
        const double d = S[GM_H + i * (kGmMax + 1) + i];
        ys[i][c] = d != 0 ? v / d : 0.0;
    }
    // the combination reads y_q / s_q
    for (int i = 0; i < nj; i++) {
        const double y = ys[i][c] / S[GM_S + i];
        S[GM_Y + i] = y;
        yc[i * KP + c] = (float)y;
    }
    yc[kGmMax * KP + c] = (float)nj;
}

// T = sum_q y_q v_q in block order (zero on columns without iterations;
// yc as gm_backsub_kernel leaves it)
__global__ __launch_bounds__(256) void gm_combine_kernel(int n, const float* __restrict__ U, int64_t us,
                                                         const float* __restrict__ yc, float* __restrict__ T) {
    const int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= (int64_t)n * KP) return;
    const int c = (int)(idx & 63);
    const int nj = (int)yc[kGmMax * KP + c];
    float v = 0.f;
    // q = 0, 1, ... in order, eight loads in flight
    for (int q0 = 0; q0 < nj; q0 += 8) {
        float y8[8], u8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool ok = q0 + k < nj;
            y8[k] = ok ? yc[(q0 + k) * KP + c] : 0.f;
            u8[k] = ok ? U[(int64_t)(q0 + k) * us + idx] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (q0 + k < nj) v += y8[k] * u8[k];
    }
    T[idx] = v;
}

__global__ __launch_bounds__(256) void correct_kernel(int64_t rows, const double* __restrict__ sc,
                                                      const int* __restrict__ sweep, double* __restrict__ X,
                                                      const float* __restrict__ d) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= rows * KP) return;
    int c = (int)(idx & 63);
    if (!sweep[c]) return;
    X[idx] += sc[c * Q_NQ + Q_SCALE] * (double)d[idx];
}

__global__ void invd32_kernel(int n, const float* __restrict__ A, int64_t lda, float* __restrict__ invd) {
    int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (j >= lda) return;
    float d = j < n ? A[j + j * lda] : 0.f;
    invd[j] = d != 0 ? 1.f / d : 1.f;
}

// row-major lda x 64 multivector from kc column-major columns (rest zero)
__global__ void gather_cols_kernel(int n, int64_t lda, int kc, const double* __restrict__ src, int64_t ld,
                                   double* __restrict__ dst) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= lda * KP) return;
    int64_t i = idx / KP;
    int c = (int)(idx % KP);
    dst[idx] = (i < n && c < kc) ? src[(int64_t)c * ld + i] : 0.0;
}

__global__ void scatter_x_kernel(int n, int kc, const double* __restrict__ X, double* __restrict__ x, int64_t ldx) {
    int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    int64_t i = idx / KP;
    int c = (int)(idx % KP);
    if (i >= n || c >= kc) return;
    x[(int64_t)c * ldx + i] = X[idx];
}

__global__ void sigma_kernel(int n, const double* __restrict__ nrm, int64_t lda, const double* __restrict__ v4,
                             double* __restrict__ sig) {
    int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= lda) return;
    sig[i] = i < n ? -(v4[0] * nrm[i] + v4[1] * nrm[lda + i] + v4[2] * nrm[2 * lda + i]) : 0.0;
}

// ---------------------------------------------------------------- block Jacobi
constexpr int kBjCols = 8;  // multivector columns per workgroup

// Y[rows of block q] = inv_q * P[rows of block q]; the block's gathered rows of
// P (kBjCols columns) are staged in LDS and read as broadcasts, the inverse is
// column-major so consecutive lanes (rows) read consecutive addresses.
__global__ __launch_bounds__(256) void bj_apply_kernel(const int* __restrict__ perm, const int* __restrict__ ptr,
                                                       const int64_t* __restrict__ off, const float* __restrict__ inv,
                                                       const float* __restrict__ P, float* __restrict__ Y, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    extern __shared__ float Pl[];  // b x kBjCols
    const int q = blockIdx.x, c0 = blockIdx.y * kBjCols;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const float* M = inv + off[q];
    for (int s = threadIdx.x; s < b; s += kT) {
        const float* src = P + (int64_t)perm[b0 + s] * KP + c0;
#pragma unroll
        for (int c = 0; c < kBjCols; c++) Pl[s * kBjCols + c] = src[c];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < b; r += kT) {
        float acc[kBjCols];
#pragma unroll
        for (int c = 0; c < kBjCols; c++) acc[c] = 0.f;
        for (int s2 = 0; s2 < b; s2++) {
            const float m = M[(int64_t)s2 * b + r];
#pragma unroll
            for (int c = 0; c < kBjCols; c++) acc[c] += m * Pl[s2 * kBjCols + c];
        }
        float* dst = Y + (int64_t)perm[b0 + r] * KP + c0;
#pragma unroll
        for (int c = 0; c < kBjCols; c++) dst[c] = acc[c];
    }
}

// Matrix-core form: Y_q = inv_q * P_q as a 32x32x2 FP32 MFMA product per
// wavefront (32 block rows x 32 multivector columns), the K range = the block.
// Workgroup (x, q): rows [64x, 64x + 64) of block q, all 64 columns; wave w
// takes row tile w/2, column tile w%2.  Rows and K values past the block end
// load zeros and are not stored.
__global__ __launch_bounds__(256) void bj_apply_mfma_kernel(const int* __restrict__ perm, const int* __restrict__ ptr,
                                                            const int64_t* __restrict__ off,
                                                            const float* __restrict__ inv, const float* __restrict__ P,
                                                            float* __restrict__ Y, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    constexpr int U = 8;  // K steps of 2 per chunk
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = blockIdx.x * 64 + 32 * (wave >> 1);
    if (r0 >= b) return;  // no LDS, no barrier: idle waves leave
    const int c = 32 * (wave & 1) + (lane & 31);
    const int kh = lane >> 5;
    const int r = r0 + (lane & 31);
    const bool rok = r < b;
    const float* M = inv + off[q];
    const int* pq = perm + b0;
    floatx16 acc = floatx16{};
    for (int k = 0; k < b; k += 2 * U) {
        float a[U], x[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kk = k + 2 * u + kh;
            const bool ok = kk < b;
            a[u] = (ok && rok) ? M[(int64_t)kk * b + r] : 0.f;
            x[u] = ok ? P[(int64_t)pq[kk] * KP + c] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], x[u], acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int rr = r0 + 8 * (j >> 2) + 4 * kh + (j & 3);
        if (rr < b) Y[(int64_t)pq[rr] * KP + c] = acc[j];
    }
}

// Block-Jacobi inverses in matrix-core operand order: block q's b x b
// inverse (column-major at inv + off[q]) as ceil(b/32) row tiles by
// ceil(b/8) K chunks of 256 floats at pk + poff[q]; in tile (t, c) lane l
// holds, as one float4, rows 32 t + l % 32 and columns 8 c + 2 u + l / 32
// (u = 0..3), zero outside the block.
__global__ __launch_bounds__(256) void bj_pack_kernel(const float* __restrict__ inv, const int* __restrict__ ptr,
                                                      const int64_t* __restrict__ off,
                                                      const int64_t* __restrict__ poff, float* __restrict__ pk) {
    const int q = blockIdx.y;
    const int b = ptr[q + 1] - ptr[q];
    const int nkc = (b + 7) / 8;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)((b + 31) / 32) * nkc * 256) return;
    const int t = (int)(idx >> 8), e = (int)(idx & 255);
    const int lane = e >> 2, u = e & 3;
    const int row = (t / nkc) * 32 + (lane & 31);
    const int k = (t % nkc) * 8 + 2 * u + (lane >> 5);
    pk[poff[q] + idx] = (row < b && k < b) ? inv[off[q] + (int64_t)k * b + row] : 0.f;
}

// Block-Jacobi product from the block-ordered multivector PP (inner_p /
// inner_s / the GMRES basis hold it): Y(perm row) = inv_q * PP rows of block
// q, on FP32 matrix cores (v_mfma_f32_32x32x2f32).  A workgroup owns 32 rows
// of one block and all 64 columns; its four wavefronts take every fourth K
// chunk of 8 (one 1 KB float4 load of the packed inverse each), U chunks
// ahead of the matrix-core ops, and add their two 32 x 32 accumulators
// through LDS.  Every inverse entry is read once per product.
template <int U>
__global__ __launch_bounds__(256) void bj_apply_pk_kernel(const int* __restrict__ perm, const int* __restrict__ ptr,
                                                          const int64_t* __restrict__ poff,
                                                          const float* __restrict__ pk,
                                                          const float* __restrict__ PP, float* __restrict__ Y,
                                                          const double* __restrict__ cscale, int cstride,
                                                          const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    __shared__ float red[4][32][64];
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const int rt = blockIdx.x;
    if (rt * 32 >= b) return;  // whole workgroup, before any barrier
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cl = lane & 31, kh = lane >> 5;
    const int nkc = (b + 7) / 8;
    const floatx4* T = (const floatx4*)(pk + poff[q]) + (int64_t)rt * nkc * 64 + lane;
    const float* Pq = PP + (int64_t)b0 * KP;
    floatx16 acc0 = floatx16{}, acc1 = floatx16{};
    floatx4 a[U], an[U];
    float x0[U][4], x1[U][4], xn0[U][4], xn1[U][4];
    auto load = [&](int c0, floatx4 (&aa)[U], float (&y0)[U][4], float (&y1)[U][4]) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + 4 * u;
            const bool cok = c < nkc;
            aa[u] = cok ? __builtin_nontemporal_load(T + (int64_t)c * 64) : floatx4{};
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int k = 8 * c + 2 * v + kh;
                const bool ok = cok && k < b;
                y0[u][v] = ok ? Pq[(int64_t)k * KP + cl] : 0.f;
                y1[u][v] = ok ? Pq[(int64_t)k * KP + 32 + cl] : 0.f;
            }
        }
    };
    load(wave, a, x0, x1);
    for (int c = wave; c < nkc; c += 4 * U) {
        const bool more = c + 4 * U < nkc;
        if (more) load(c + 4 * U, an, xn0, xn1);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u][v], x0[u][v], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u][v], x1[u][v], acc1, 0, 0, 0);
            }
        if (more) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                a[u] = an[u];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    x0[u][v] = xn0[u][v];
                    x1[u][v] = xn1[u][v];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
        red[wave][j][lane] = acc0[j];
        red[wave][16 + j][lane] = acc1[j];
    }
    __syncthreads();
    // wavefront w finishes accumulator entries 8 w .. 8 w + 7
    float f0 = 1.f, f1 = 1.f;  // optional per-column factor 1 / cscale[c * cstride] (0 where that is 0)
    if (cscale) {
        const double s0 = cscale[(int64_t)cl * cstride], s1 = cscale[(int64_t)(32 + cl) * cstride];
        f0 = s0 != 0 ? (float)(1.0 / s0) : 0.f;
        f1 = s1 != 0 ? (float)(1.0 / s1) : 0.f;
    }
    const int* pq = perm + b0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int jj = 8 * wave + i;  // 0..15: acc0 entry jj, 16..31: acc1 entry jj - 16
        const float v = red[0][jj][lane] + red[1][jj][lane] + red[2][jj][lane] + red[3][jj][lane];
        const int j = jj & 15;
        const int rr = rt * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
        if (rr < b) {
            if (jj < 16)
                Y[(int64_t)pq[rr] * KP + cl] = v * f0;
            else
                Y[(int64_t)pq[rr] * KP + 32 + cl] = v * f1;
        }
    }
}

// The same inverses as bf16 hi/lo pairs in the bf16 matrix-core operand
// order: block q as ceil(b/32) row tiles by ceil(b/16) K chunks of 2 KiB at
// pk + poff[q] (floats), each chunk the hi plane then the lo plane (64 lanes
// x 16 bytes each); lane l holds row 32 t + l % 32, columns 16 c + 8 (l / 32)
// + 0..7, zero outside the block.
__global__ __launch_bounds__(256) void bj_pack_bf_kernel(const float* __restrict__ inv, const int* __restrict__ ptr,
                                                         const int64_t* __restrict__ off,
                                                         const int64_t* __restrict__ poff, float* __restrict__ pk) {
    const int q = blockIdx.y;
    const int b = ptr[q + 1] - ptr[q];
    const int nkc = (b + 15) / 16;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;  // (chunk, lane)
    if (g >= (int64_t)((b + 31) / 32) * nkc * 64) return;
    const int t = (int)(g >> 6), lane = (int)(g & 63);
    const int row = (t / nkc) * 32 + (lane & 31);
    const int kbase = (t % nkc) * 16 + 8 * (lane >> 5);
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kbase + i;
        v[i] = (row < b && k < b) ? inv[off[q] + (int64_t)k * b + row] : 0.f;
    }
    uint4 hi, lo;
    split8(v, hi, lo);
    uint4* T = reinterpret_cast<uint4*>(pk + poff[q]) + (int64_t)t * 128 + lane;
    T[0] = hi;
    T[64] = lo;
}

// Block-Jacobi product on bf16 matrix cores (v_mfma_f32_32x32x16_bf16): the
// PP rows are split hi + lo as they are loaded, so inv * PP ~ hi_inv (hi_p +
// lo_p) (+ lo_inv hi_p when LO): LO gives FP32-level accuracy from the same
// bytes as the FP32 inverses, !LO reads half of them (the inverse to bf16).
// A workgroup owns RT 32-row tiles of one block and all 64 columns; its four
// wavefronts take every fourth K chunk of 16 (each multivector load feeds the
// RT tiles) and add their accumulators through LDS.
template <bool LO, int RT>
__device__ __forceinline__ void bj_apply_bf_body(int bx, int q, const int* __restrict__ perm,
                                                 const int* __restrict__ ptr, const int64_t* __restrict__ poff,
                                                 const float* __restrict__ pk, const float* __restrict__ PP,
                                                 float* __restrict__ Y, const double* __restrict__ cscale,
                                                 int cstride, const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    __shared__ float red[4][RT * 32][64];
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const int rt0 = bx * RT;
    if (rt0 * 32 >= b) return;  // whole workgroup, before any barrier
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cl = lane & 31, kh = lane >> 5;
    const int nkc = (b + 15) / 16, ntile = (b + 31) / 32;
    const u32x4* T = reinterpret_cast<const u32x4*>(pk + poff[q]) + (int64_t)rt0 * nkc * 128 + lane;
    const float* Pq = PP + (int64_t)b0 * KP;
    floatx16 acc[RT][2];
#pragma unroll
    for (int r = 0; r < RT; r++) acc[r][0] = acc[r][1] = floatx16{};
    u32x4 ah[RT], al[RT], nah[RT], nal[RT];
    float x0[8], x1[8], nx0[8], nx1[8];
    auto load = [&](int c, u32x4 (&h)[RT], u32x4 (&l)[RT], float (&y0)[8], float (&y1)[8]) {
        const bool cok = c < nkc;
#pragma unroll
        for (int r = 0; r < RT; r++) {
            const bool ok = cok && rt0 + r < ntile;  // tiles past the block's last are not stored
            h[r] = ok ? __builtin_nontemporal_load(T + ((int64_t)r * nkc + c) * 128) : u32x4{};
            if (LO) l[r] = ok ? __builtin_nontemporal_load(T + ((int64_t)r * nkc + c) * 128 + 64) : u32x4{};
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = 16 * c + 8 * kh + i;
            const bool ok = cok && k < b;
            y0[i] = ok ? Pq[(int64_t)k * KP + cl] : 0.f;
            y1[i] = ok ? Pq[(int64_t)k * KP + 32 + cl] : 0.f;
        }
    };
    load(wave, ah, al, x0, x1);
    for (int c = wave; c < nkc; c += 4) {
        const bool more = c + 4 < nkc;
        if (more) load(c + 4, nah, nal, nx0, nx1);
        uint4 h0, l0, h1, l1;
        split8(x0, h0, l0);
        split8(x1, h1, l1);
#pragma unroll
        for (int r = 0; r < RT; r++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, ah[r]);
            acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, l0), acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, l1), acc[r][1], 0, 0, 0);
            if (LO) {
                const bf16x8 alo = __builtin_bit_cast(bf16x8, al[r]);
                acc[r][0] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, __builtin_bit_cast(bf16x8, h0), acc[r][0], 0, 0, 0);
                acc[r][1] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, __builtin_bit_cast(bf16x8, h1), acc[r][1], 0, 0, 0);
            }
            acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, h0), acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, h1), acc[r][1], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int r = 0; r < RT; r++) {
                ah[r] = nah[r];
                if (LO) al[r] = nal[r];
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                x0[i] = nx0[i];
                x1[i] = nx1[i];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            red[wave][r * 32 + j][lane] = acc[r][0][j];
            red[wave][r * 32 + 16 + j][lane] = acc[r][1][j];
        }
    __syncthreads();
    float f0 = 1.f, f1 = 1.f;  // optional per-column factor 1 / cscale[c * cstride] (0 where that is 0)
    if (cscale) {
        const double s0 = cscale[(int64_t)cl * cstride], s1 = cscale[(int64_t)(32 + cl) * cstride];
        f0 = s0 != 0 ? (float)(1.0 / s0) : 0.f;
        f1 = s1 != 0 ? (float)(1.0 / s1) : 0.f;
    }
    const int* pq = perm + b0;
    // wavefront w finishes entries 8 RT w .. 8 RT w + 8 RT - 1 of the 32 RT
    // (tile r: entries 32 r + 0..15 of acc[r][0], 32 r + 16..31 of acc[r][1])
#pragma unroll
    for (int i = 0; i < 8 * RT; i++) {
        const int jj = 8 * RT * wave + i;
        const float v = red[0][jj][lane] + red[1][jj][lane] + red[2][jj][lane] + red[3][jj][lane];
        const int r = jj >> 5, e = jj & 31, j = e & 15;
        const int rr = (rt0 + r) * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
        if (rr < b) {
            if (e < 16)
                Y[(int64_t)pq[rr] * KP + cl] = v * f0;
            else
                Y[(int64_t)pq[rr] * KP + 32 + cl] = v * f1;
        }
    }
}
// The single-geometry launch keeps its own copy of the body: calling the
// shared body (bj_apply_bf_body, above) from it measured 56 -> 66 us per
// product at 25,916 panels, with the same memory instructions but different
// scheduling.
template <bool LO, int RT>
__global__ __launch_bounds__(256) void bj_apply_bf_kernel(const int* __restrict__ perm, const int* __restrict__ ptr,
                                                          const int64_t* __restrict__ poff,
                                                          const float* __restrict__ pk,
                                                          const float* __restrict__ PP, float* __restrict__ Y,
                                                          const double* __restrict__ cscale, int cstride,
                                                          const int* __restrict__ gate) {
    if (gate && gate_idle(gate)) return;
    __shared__ float red[4][RT * 32][64];
    const int q = blockIdx.y;
    const int b0 = ptr[q], b = ptr[q + 1] - b0;
    const int rt0 = blockIdx.x * RT;
    if (rt0 * 32 >= b) return;  // whole workgroup, before any barrier
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cl = lane & 31, kh = lane >> 5;
    const int nkc = (b + 15) / 16, ntile = (b + 31) / 32;
    const u32x4* T = reinterpret_cast<const u32x4*>(pk + poff[q]) + (int64_t)rt0 * nkc * 128 + lane;
    const float* Pq = PP + (int64_t)b0 * KP;
    floatx16 acc[RT][2];
#pragma unroll
    for (int r = 0; r < RT; r++) acc[r][0] = acc[r][1] = floatx16{};
    u32x4 ah[RT], al[RT], nah[RT], nal[RT];
    float x0[8], x1[8], nx0[8], nx1[8];
    auto load = [&](int c, u32x4 (&h)[RT], u32x4 (&l)[RT], float (&y0)[8], float (&y1)[8]) {
        const bool cok = c < nkc;
#pragma unroll
        for (int r = 0; r < RT; r++) {
            const bool ok = cok && rt0 + r < ntile;  // tiles past the block's last are not stored
            h[r] = ok ? __builtin_nontemporal_load(T + ((int64_t)r * nkc + c) * 128) : u32x4{};
            if (LO) l[r] = ok ? __builtin_nontemporal_load(T + ((int64_t)r * nkc + c) * 128 + 64) : u32x4{};
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = 16 * c + 8 * kh + i;
            const bool ok = cok && k < b;
            y0[i] = ok ? Pq[(int64_t)k * KP + cl] : 0.f;
            y1[i] = ok ? Pq[(int64_t)k * KP + 32 + cl] : 0.f;
        }
    };
    load(wave, ah, al, x0, x1);
    for (int c = wave; c < nkc; c += 4) {
        const bool more = c + 4 < nkc;
        if (more) load(c + 4, nah, nal, nx0, nx1);
        uint4 h0, l0, h1, l1;
        split8(x0, h0, l0);
        split8(x1, h1, l1);
#pragma unroll
        for (int r = 0; r < RT; r++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, ah[r]);
            acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, l0), acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, l1), acc[r][1], 0, 0, 0);
            if (LO) {
                const bf16x8 alo = __builtin_bit_cast(bf16x8, al[r]);
                acc[r][0] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, __builtin_bit_cast(bf16x8, h0), acc[r][0], 0, 0, 0);
                acc[r][1] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, __builtin_bit_cast(bf16x8, h1), acc[r][1], 0, 0, 0);
            }
            acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, h0), acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, h1), acc[r][1], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int r = 0; r < RT; r++) {
                ah[r] = nah[r];
                if (LO) al[r] = nal[r];
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                x0[i] = nx0[i];
                x1[i] = nx1[i];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            red[wave][r * 32 + j][lane] = acc[r][0][j];
            red[wave][r * 32 + 16 + j][lane] = acc[r][1][j];
        }
    __syncthreads();
    float f0 = 1.f, f1 = 1.f;  // optional per-column factor 1 / cscale[c * cstride] (0 where that is 0)
    if (cscale) {
        const double s0 = cscale[(int64_t)cl * cstride], s1 = cscale[(int64_t)(32 + cl) * cstride];
        f0 = s0 != 0 ? (float)(1.0 / s0) : 0.f;
        f1 = s1 != 0 ? (float)(1.0 / s1) : 0.f;
    }
    const int* pq = perm + b0;
    // wavefront w finishes entries 8 RT w .. 8 RT w + 8 RT - 1 of the 32 RT
    // (tile r: entries 32 r + 0..15 of acc[r][0], 32 r + 16..31 of acc[r][1])
#pragma unroll
    for (int i = 0; i < 8 * RT; i++) {
        const int jj = 8 * RT * wave + i;
        const float v = red[0][jj][lane] + red[1][jj][lane] + red[2][jj][lane] + red[3][jj][lane];
        const int r = jj >> 5, e = jj & 31, j = e & 15;
        const int rr = (rt0 + r) * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
        if (rr < b) {
            if (e < 16)
                Y[(int64_t)pq[rr] * KP + cl] = v * f0;
            else
                Y[(int64_t)pq[rr] * KP + 32 + cl] = v * f1;
        }
    }
}

// ---------------------------------------------------------------- groups
// One GMRES iteration stage for several geometries in one launch: workgroup
// id -> (geometry p, its own workgroup id) through the prefix sums of the
// stage's per-geometry grids, then the same body as the single-geometry
// kernel on that geometry's arrays.  Each geometry's workgroups do exactly
// what its own launch would, so its results do not depend on the group.
}  // namespace

// Device pass: the descriptors' pointers are declared global-memory
// pointers (every one of them is a device allocation), so the grouped
// kernels access through them with global instructions; loaded from memory
// as plain pointers they would be flat ones.  Same layout in both passes.
#if defined(__HIP_DEVICE_COMPILE__)
#define CP_GLOBAL __attribute__((address_space(1)))
#else
#define CP_GLOBAL
#endif
struct GmProb {
    int n, nkb, kbchunk, kslices, nslices, ngemm, variant, m, ndc, nuc, bj_gx, bj_nblocks, bj_lo;
    int64_t lda, plane, us;
    CP_GLOBAL const uint4* apk;
    CP_GLOBAL uint4* ppk;
    CP_GLOBAL float* part32;
    CP_GLOBAL const int* tptr;
    CP_GLOBAL const int* tkb;
    CP_GLOBAL float* gm_u;
    CP_GLOBAL double* gm_st;
    CP_GLOBAL double* gm_dpart;
    CP_GLOBAL float* gm_coef;
    CP_GLOBAL double* dpart;
    CP_GLOBAL int* act;
    CP_GLOBAL const double* sc;
    CP_GLOBAL const int* iperm;
    CP_GLOBAL float* y;
    CP_GLOBAL const int* bj_perm;
    CP_GLOBAL const int* bj_ptr;
    CP_GLOBAL const int64_t* bj_poff;
    CP_GLOBAL const float* bj_pk;
    CP_GLOBAL int* gflag;  // any column of this geometry still active after the iteration
};

namespace {

struct GroupLaunch {
    const GmProb* probs;
    int G;  // geometries in this launch
    int j;  // Krylov iteration
    int idx[kMaxGroup];      // launch slot -> index into probs
    int pre[kMaxGroup + 1];  // workgroup prefix of the slots
};

// Slot of this workgroup: the number of slot starts (after the first) at or
// below its id.  Unrolled and branch-free, so the descriptor's pointers count
// as loaded from unclobbered kernel-argument memory and the bodies access
// global memory with global (not flat) instructions.
__device__ __forceinline__ const GmProb& group_of(const GroupLaunch& L, int& local) {
    const int bid = blockIdx.x;
    int p = 0;
#pragma unroll
    for (int i = 1; i < kMaxGroup; i++) p += (i < L.G && bid >= L.pre[i]) ? 1 : 0;
    local = bid - L.pre[p];
    return L.probs[L.idx[p]];
}

// every slot of one launch has the preconditioner form LO
template <bool LO>
__global__ __launch_bounds__(256) void grp_bj_kernel(GroupLaunch L) {
    int l;
    const GmProb& P = group_of(L, l);
    bj_apply_bf_body<LO, 2>(l % P.bj_gx, l / P.bj_gx, P.bj_perm, P.bj_ptr, P.bj_poff, P.bj_pk,
                            P.gm_u + (int64_t)L.j * P.us, P.y, P.gm_st + GM_S + L.j, kGmStride, P.act);
}

__global__ __launch_bounds__(256) void grp_pack_kernel(GroupLaunch L) {
    int l;
    const GmProb& P = group_of(L, l);
    pack_p_body(l, P.y, P.n, P.nkb, P.ppk, P.act);
}

// every slot of one launch has the product variant of (RT, S)
template <int RT, int S>
__global__ __launch_bounds__(256) void grp_gemm_kernel(GroupLaunch L) {
    int l;
    const GmProb& P = group_of(L, l);
    gemm_bf16x2_s_body<RT, S>(l, P.apk, P.nkb, P.ppk, P.part32, P.lda, P.kbchunk, P.kslices, P.ngemm,
                              P.plane, P.tptr, P.tkb, P.act);
}

__global__ __launch_bounds__(256) void grp_w_kernel(GroupLaunch L) {
    int l;
    const GmProb& P = group_of(L, l);
    gm_w_body(l, P.part32, P.nslices, P.lda, P.n, P.act, P.iperm, P.gm_u + (int64_t)(L.j + 1) * P.us);
}

__global__ __launch_bounds__(256) void grp_dots_kernel(GroupLaunch L) {
    int l;
    const GmProb& P = group_of(L, l);
    gm_dots_body(l % P.ndc, l / P.ndc, P.ndc, P.n, P.act, P.gm_u, P.us, L.j, P.gm_dpart);
}

__global__ __launch_bounds__(1024) void grp_hcol_kernel(GroupLaunch L) {
    __shared__ double sh[1024];
    int l;
    const GmProb& P = group_of(L, l);
    gm_hcol_body(l, sh, P.gm_dpart, P.ndc, P.act, L.j, P.gm_st, P.gm_coef);
}

__global__ __launch_bounds__(256) void grp_update_kernel(GroupLaunch L) {
    __shared__ double sh[kT];
    int l;
    const GmProb& P = group_of(L, l);
    gm_update_body(l, sh, P.n, P.act, P.gm_u, P.us, L.j, P.gm_coef, P.dpart);
}

__global__ __launch_bounds__(1024) void grp_givens_kernel(GroupLaunch L) {
    __shared__ double sh[1024];
    int l;
    const GmProb& P = group_of(L, l);
    gm_givens_body(sh, P.dpart, P.nuc, P.act, L.j, P.m, P.sc, P.gm_st, P.gflag);
}

inline unsigned nblocks(int64_t work) { return (unsigned)((work + kT - 1) / kT); }

}  // namespace

int packed_kblocks(int n) { return (n + 15) / 16; }

size_t packed_a_bytes(int64_t lda, int n) { return (size_t)(lda / 32) * packed_kblocks(n) * 2048; }

int64_t bj_packed_floats(int b) { return (int64_t)((b + 31) / 32) * ((b + 15) / 16) * 512; }

void launch_bj_pack(const float* inv, const int* ptr, const int64_t* off, const int64_t* poff, int nblocks, int max_b,
                    float* pk, int kind, hipStream_t s) {
    if (nblocks <= 0) return;
    const int64_t tiles32 = (int64_t)((max_b + 31) / 32);
    if (kind == kBjPackFp32) {
        const dim3 grid((unsigned)(tiles32 * ((max_b + 7) / 8)), (unsigned)nblocks);  // 256 floats per chunk
        hipLaunchKernelGGL(bj_pack_kernel, grid, dim3(256), 0, s, inv, ptr, off, poff, pk);
    } else {
        const dim3 grid((unsigned)((tiles32 * ((max_b + 15) / 16) * 64 + 255) / 256), (unsigned)nblocks);
        hipLaunchKernelGGL(bj_pack_bf_kernel, grid, dim3(256), 0, s, inv, ptr, off, poff, pk);
    }
    CP_LAUNCH_CHECK();
}

void launch_pack_a(const float* A32, int64_t lda, int n, void* out, hipStream_t s) {
    const int nkb = packed_kblocks(n);
    const int64_t ntiles = (lda / 32) * (int64_t)nkb;
    hipLaunchKernelGGL(pack_a_bf16x2_kernel, dim3((unsigned)((ntiles * 64 + 255) / 256)), dim3(256), 0, s, A32, lda,
                       n, nkb, ntiles, (uint4*)out);
    CP_LAUNCH_CHECK();
}

void launch_sigma(int n, int64_t lda, const double* normals, const double* vinf4, double* sigma, hipStream_t s) {
    hipLaunchKernelGGL(sigma_kernel, dim3(nblocks(lda)), dim3(kT), 0, s, n, normals, lda, vinf4, sigma);
    CP_LAUNCH_CHECK();
}

size_t BatchedSolver::workspace_bytes(int64_t lda) {
    const size_t mv64 = (size_t)lda * KP * sizeof(double);
    const size_t mv32 = (size_t)lda * KP * sizeof(float);
    const int KSmax = 32;
    const int nrb = (int)(lda / kRedRows);
    // 24 sub-buffers, each rounded up to 256 bytes; the FP32 partials hold
    // one slice more than the K split (the bf16x2 near-field slice)
    return 4 * mv64 + 11 * mv32 + (size_t)lda * sizeof(float) + (KSmax + 1) * mv32 + KSmax * mv64 +
           (size_t)lda * 4 * sizeof(double) + (kGmMax + 1) * mv32 + (size_t)KP * kGmStride * sizeof(double) +
           (size_t)(kGmMax + 1) * nrb * KP * sizeof(double) + (size_t)(kGmMax + 1) * KP * sizeof(float) +
           (size_t)4 * nrb * KP * sizeof(double) + KP * Q_NQ * sizeof(double) + KP * sizeof(int) + 26 * 256;
}

// per-column scalars, then 64-int blocks: the sweep flags and two read-back
// buffers of the active flags (the inner loop checks one iteration behind)
size_t BatchedSolver::host_bytes() { return KP * Q_NQ * sizeof(double) + 3 * KP * sizeof(int); }

BatchedSolver::BatchedSolver(int n, int64_t lda, int capacity, void* ws, void* host, hipStream_t s)
    : n_(n), cap_(capacity), lda_(lda) {
    nrb_ = (int)(lda / kRedRows);
    const size_t mv64 = (size_t)lda * KP * sizeof(double);
    const size_t mv32 = (size_t)lda * KP * sizeof(float);
    const int KSmax = 32;
    ws_ = ws;
    char* p = (char*)ws_;
    auto take = [&](size_t b) {
        char* q = p;
        p += (b + 255) / 256 * 256;
        return q;
    };
    X_ = (double*)take(mv64);
    Bv_ = (double*)take(mv64);
    R_ = (double*)take(mv64);
    S_ = (double*)take(mv64);
    float** f32[9] = {&r_, &r0_, &p_, &v_, &sv_, &t_, &y_, &z_, &d_};
    for (auto q : f32) *q = (float*)take(mv32);
    invd_ = (float*)take((size_t)lda * sizeof(float));
    ppk_ = (void*)take(mv32);  // packed_kblocks(n) * 4 KiB <= lda / 16 * 4 KiB
    pp_ = (float*)take(mv32);
    part32_ = (float*)take((KSmax + 1) * mv32);
    part64_ = (double*)take(KSmax * mv64);
    dpart_ = (double*)take((size_t)4 * nrb_ * KP * sizeof(double));
    bn_ = (double*)take((size_t)lda * 4 * sizeof(double));
    gm_u_ = (float*)take((kGmMax + 1) * mv32);
    gm_st_ = (double*)take((size_t)KP * kGmStride * sizeof(double));
    gm_dpart_ = (double*)take((size_t)(kGmMax + 1) * nrb_ * KP * sizeof(double));
    gm_coef_ = (float*)take((size_t)(kGmMax + 1) * KP * sizeof(float));
    sc_ = (double*)take(KP * Q_NQ * sizeof(double));
    act_ = (int*)take(KP * sizeof(int));
    CP_HIP(hipMemsetAsync(y_, 0, mv32, s));
    CP_HIP(hipMemsetAsync(z_, 0, mv32, s));
    CP_HIP(hipMemsetAsync(pp_, 0, mv32, s));
    host_sc_ = static_cast<double*>(host);
    host_act_ = reinterpret_cast<int*>(host_sc_ + KP * Q_NQ);
    for (hipEvent_t& e : act_ev_) CP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // default shape before tuning: 8 K slices (one per XCD), more for small
    // matrices
    int rb = (int)(lda / 256);
    shape_.rt = 2;
    shape_.kslices = rb >= 64 ? 8 : (rb >= 16 ? 16 : 32);
    // CPGPU_F64_DEPTH=2|3: K steps of 4 per load stage of the FP64-matrix residual
    if (const char* e = std::getenv("CPGPU_F64_DEPTH")) {
        int d = std::atoi(e);
        if (d >= 2 && d <= 3) f64_depth_ = d;
    }
    // CPGPU_F64_KS=1..32 pins the K split of the FP64-accumulated products
    if (const char* e = std::getenv("CPGPU_F64_KS")) {
        const int k = std::atoi(e);
        if (k >= 1 && k <= 32) f64_ks_ = k;
    }
    // CPGPU_F64_SIMPLE=1 selects the unpipelined FP64-accumulated product (tests)
    if (const char* e = std::getenv("CPGPU_F64_SIMPLE")) f64_pipelined_ = e[0] != '1';
    // CPGPU_NO_BF16X3=1 keeps the inner product in FP32 matrix-core arithmetic
    if (const char* e = std::getenv("CPGPU_NO_BF16X3")) bf16x3_ = e[0] != '1';
    // CPGPU_INNER=bicgstab keeps BiCGSTAB as the inner solver with the
    // block-Jacobi preconditioner; CPGPU_GMRES_M sets the Krylov dimension
    if (const char* e = std::getenv("CPGPU_INNER")) gmres_ = std::string(e) != "bicgstab";
    if (const char* e = std::getenv("CPGPU_GMRES_M")) {
        const int v = std::atoi(e);
        if (v >= 4 && v <= kGmMax) gm_m_ = v;
    }
    // CPGPU_GRAPH=1 replays the GMRES iterations from captured graphs
    // (default: launched kernel by kernel, which measured faster)
    if (const char* e = std::getenv("CPGPU_GRAPH")) graphs_ = e[0] == '1';
    // CPGPU_NO_BF16X2=1 keeps the bf16x2 product out of the tuner's candidates
    if (const char* e = std::getenv("CPGPU_NO_BF16X2")) bf16x2_ = e[0] != '1';
    // CPGPU_BJ_VALU=1 selects the vector-ALU block-Jacobi product (tests)
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            cus_ = cus;
    }
    if (const char* e = std::getenv("CPGPU_BJ_VALU")) bj_mfma_ = e[0] != '1';
    // CPGPU_BJ_PREC=bf16x3: the bf16 block-Jacobi product reads the lo plane
    // of the inverses too (FP32-level accuracy; default: the hi plane alone,
    // the inverses to bf16 -- a preconditioner changes only the inner
    // iterations, and on the 25,916-panel wing not even those)
    if (const char* e = std::getenv("CPGPU_BJ_PREC")) bj_lo_ = std::string(e) == "bf16x3";
    // CPGPU_BJ_RT=1|2: 32-row tiles per workgroup of the bf16 block-Jacobi product
    if (const char* e = std::getenv("CPGPU_BJ_RT")) bj_rt_ = std::atoi(e) == 1 ? 1 : 2;
    // CPGPU_BJ_U: K chunks per wavefront per load batch of the FP32
    // block-Jacobi product (1..3)
    if (const char* e = std::getenv("CPGPU_BJ_U")) {
        const int u = std::atoi(e);
        if (u >= 1 && u <= 3) bj_u_ = u;
    }
    // CPGPU_GEMM_SHAPE="variant,rt,kslices" pins the shape (tests, tuning runs)
    if (const char* e = std::getenv("CPGPU_GEMM_SHAPE")) {
        int v = 0, r = 2, k = 8;
        if (std::sscanf(e, "%d,%d,%d", &v, &r, &k) == 3 && v >= 0 && v <= 14 && (r == 1 || r == 2 || r == 4) &&
            (k == 4 || k == 5 || k == 6 || k == 8 || k == 10 || k == 12 || k == 16 || k == 32)) {
            shape_.variant = v;
            shape_.rt = r;
            shape_.kslices = k;
            tuned_ = true;
        }
    }
}

BatchedSolver::IterPlan::~IterPlan() {
    for (auto& it : items)
        if (it.first) (void)hipGraphExecDestroy(it.first);
}

// the owner synchronises the stream before a solver is destroyed, and the
// cache is cleared only between sweeps (after the residual's synchronisation),
// so no replay of these graphs is in flight
void BatchedSolver::gm_plan_clear() {
    gm_plans_.clear();
    gm_plan_key_.clear();
}

BatchedSolver::~BatchedSolver() {
    gm_plan_clear();
    for (hipEvent_t& e : act_ev_)
        if (e) (void)hipEventDestroy(e);
}

void BatchedSolver::launch_gemm32(const GemmShape& sh, const float* A32, const float* P, hipStream_t s, int parts) {
    const int KS = sh.kslices;
    int kchunk = (int)round_up((n_ + KS - 1) / KS, 16);
    if (sh.variant >= 5 && (parts & 1)) {  // the bf16 variants read P as hi/lo tiles
        if (!apk_) throw HipError{hipErrorInvalidValue, "bf16 GEMM without packed A"};
        TimedScope tp("pack_p", s);
        hipLaunchKernelGGL(pack_p_bf16x2_kernel, dim3((unsigned)((2 * (int64_t)nkb_ * 64 + 255) / 256)), dim3(256), 0,
                           s, P, n_, nkb_, (uint4*)ppk_, gate_);
        CP_LAUNCH_CHECK();
    }
    if (!(parts & 2)) return;
    if (sh.variant == 5) {  // bf16x3 on packed tiles
        if (!apk_) throw HipError{hipErrorInvalidValue, "bf16x3 GEMM without packed A"};
        const int nkb = nkb_;
        const int kbchunk = (nkb + KS - 1) / KS;
        dim3 grid((unsigned)((lda_ / 512) * KS));
        TimedScope ts("gemm_f32", s);
        hipLaunchKernelGGL(gemm_bf16x3_kernel, grid, dim3(256), 0, s, (const uint4*)apk_, plane_, nkb,
                           (const uint4*)ppk_, part32_, lda_, kbchunk, KS, gate_);
        CP_LAUNCH_CHECK();
        return;
    }
    if (sh.variant >= 6 && sh.variant <= 9) {  // bf16x3, RT row tiles, S stages
        if (!apk_) throw HipError{hipErrorInvalidValue, "bf16x3 GEMM without packed A"};
        const int nkb = nkb_;
        const int kbchunk = (nkb + KS - 1) / KS;
        const int rows_wg = (sh.variant == 6 || sh.variant == 7) ? 256 : 512;
        dim3 grid((unsigned)((lda_ / rows_wg) * KS));
        TimedScope ts("gemm_f32", s);
        const uint4* Ap = (const uint4*)apk_;
        const uint4* Pp = (const uint4*)ppk_;
        switch (sh.variant) {
            case 6:
                hipLaunchKernelGGL((gemm_bf16x3_s_kernel<2, 3>), grid, dim3(256), 0, s, Ap, plane_, nkb, Pp, part32_,
                                   lda_, kbchunk, KS, gate_);
                break;
            case 7:
                hipLaunchKernelGGL((gemm_bf16x3_s_kernel<2, 4>), grid, dim3(256), 0, s, Ap, plane_, nkb, Pp, part32_,
                                   lda_, kbchunk, KS, gate_);
                break;
            case 8:
                hipLaunchKernelGGL((gemm_bf16x3_s_kernel<4, 3>), grid, dim3(256), 0, s, Ap, plane_, nkb, Pp, part32_,
                                   lda_, kbchunk, KS, gate_);
                break;
            default:
                hipLaunchKernelGGL((gemm_bf16x3_s_kernel<4, 2>), grid, dim3(256), 0, s, Ap, plane_, nkb, Pp, part32_,
                                   lda_, kbchunk, KS, gate_);
                break;
        }
        CP_LAUNCH_CHECK();
        return;
    }
    if (sh.variant >= 10 && sh.variant <= 14) {  // bf16x2 + near-field restoration slice
        if (!apk_ || !tptr_)
            throw HipError{hipErrorInvalidValue, "bf16x2 GEMM without packed A or near-field tile lists"};
        const int nkb = nkb_;
        const int kbchunk = (nkb + KS - 1) / KS;
        const int rows_wg = (sh.variant == 13 || sh.variant == 14) ? 256 : 512;
        const int ngemm = (int)((lda_ / rows_wg) * KS);
        dim3 grid((unsigned)(ngemm + lda_ / 128));  // + 4 row tiles per workgroup
        const uint4* Ap = (const uint4*)apk_;
        const uint4* Pp = (const uint4*)ppk_;
        TimedScope ts("gemm_f32", s);
#define CP_BF16X2(RT, S)                                                                                         \
    hipLaunchKernelGGL((gemm_bf16x2_s_kernel<RT, S>), grid, dim3(256), 0, s, Ap, nkb, Pp, part32_, lda_, kbchunk, \
                       KS, ngemm, plane_, tptr_, tkb_, gate_)
        switch (sh.variant) {
            case 10: CP_BF16X2(4, 2); break;
            case 11: CP_BF16X2(4, 3); break;
            case 12: CP_BF16X2(4, 4); break;
            case 13: CP_BF16X2(2, 2); break;
            default: CP_BF16X2(2, 3); break;
        }
#undef CP_BF16X2
        CP_LAUNCH_CHECK();
        return;
    }
    TimedScope ts("gemm_f32", s);
    if (sh.variant >= 1) {  // 16-byte A loads, 512 rows per workgroup
        dim3 grid((unsigned)((lda_ / 512) * KS));
        if (sh.variant == 1)
            hipLaunchKernelGGL(gemm_f32_x4_kernel<8>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk, KS, gate_);
        else if (sh.variant == 2)
            hipLaunchKernelGGL(gemm_f32_x4_kernel<4>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk, KS, gate_);
        else if (sh.variant == 3)
            hipLaunchKernelGGL(gemm_f32_x4s3_kernel<4>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk,
                               KS, gate_);
        else
            hipLaunchKernelGGL(gemm_f32_x4s3_kernel<2>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk,
                               KS, gate_);
        CP_LAUNCH_CHECK();
        return;
    }
    const int rows_per_block = 4 * 32 * sh.rt;
    dim3 grid((unsigned)((lda_ / rows_per_block) * KS));
    switch (sh.rt) {
        case 1: hipLaunchKernelGGL(gemm_f32_kernel<1>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk, KS, gate_); break;
        case 2: hipLaunchKernelGGL(gemm_f32_kernel<2>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk, KS, gate_); break;
        default: hipLaunchKernelGGL(gemm_f32_kernel<4>, grid, dim3(256), 0, s, A32, lda_, n_, P, part32_, kchunk, KS, gate_); break;
    }
    CP_LAUNCH_CHECK();
}

void BatchedSolver::escalate_block_jacobi() {
    if (bj_level_ == 0 && bj_kind_ == kBjPackBf16 && !bj_lo_)
        bj_level_ = 1;  // the lo plane of the inverses: FP32-accurate operands
    else
        bj_level_ = 2;  // point Jacobi (the BiCGSTAB inner solve)
    if (bj_level_ == 2) bj_ = false;
    ++bj_escalations_;
}

void BatchedSolver::tune(const float* A32, hipStream_t s) {
    if (tuned_) return;
    tuned_ = true;
    static std::mutex mu;
    static std::map<int64_t, GemmShape> cache;  // per leading dimension (and bf16x3 use), process wide
    const bool use3 = apk_ && bf16x3_, use2 = use3 && bf16x2_ && tptr_;
    const int64_t ckey = 4 * lda_ + (use3 ? 1 : 0) + (use2 ? 2 : 0);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(ckey);
        if (it != cache.end()) {
            shape_ = it->second;
            return;
        }
    }
    // candidates: row tiles per wavefront x K slices; the row-block count must
    // divide lda (a multiple of 512).  With the bf16x2 product available only
    // its variants are timed: it reads half the bytes of bf16x3 and a quarter
    // of FP32 and has won at every size measured, so timing the others only
    // lengthens the first solve of every new size (a serving stream meets
    // many); CPGPU_NO_BF16X2=1 brings them back.
    std::vector<GemmShape> cand;
    for (int rt : {1, 2, 4})
        for (int ks : {4, 8, 16, 32}) {
            if (use2 || lda_ % (128 * rt) != 0) continue;
            GemmShape g;
            g.rt = rt;
            g.kslices = ks;
            cand.push_back(g);
        }
    // (variant 5 is the same tiling as 9; 9 is the one timed)
    // K splits: powers of two, plus 5, 6, 10, 12, which let the row-block
    // count times the split land near one or two workgroups per CU
    for (int v : {1, 2, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14})
        for (int ks : {4, 5, 6, 8, 10, 12, 16, 32}) {
            if (v < 10 && (ks == 5 || ks == 6 || ks == 10 || ks == 12)) continue;
            if (v >= 5 && !use3) continue;
            if ((v >= 10) != use2) continue;
            GemmShape g;
            g.variant = v;
            g.rt = (v == 13 || v == 14) ? 2 : 4;
            g.kslices = ks;
            cand.push_back(g);
        }
    hipEvent_t e0, e1;
    CP_HIP(hipEventCreate(&e0));
    CP_HIP(hipEventCreate(&e1));
    KernelTimer* saved = current_timer();
    set_current_timer(nullptr);
    float best = 1e30f;
    GemmShape bestg = shape_;
    // each candidate is timed with the split-K reduction its K slices imply
    auto run = [&](const GemmShape& g) {
        launch_gemm32(g, A32, p_, s);
        hipLaunchKernelGGL(reduce32_kernel, dim3(nblocks(lda_ * KP)), dim3(kT), 0, s, (const float*)part32_,
                           slices(g), lda_, n_, v_);
        CP_LAUNCH_CHECK();
    };
    for (const GemmShape& g : cand) {
        run(g);  // warm
        CP_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; r++) run(g);
        CP_HIP(hipEventRecord(e1, s));
        CP_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CP_HIP(hipEventElapsedTime(&ms, e0, e1));
        tune_log_.push_back({g, ms / 3});
        if (ms < best) {
            best = ms;
            bestg = g;
        }
    }
    set_current_timer(saved);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the first solver to finish tuning a key fixes its shape for the process:
    // one that tuned the same key concurrently takes that shape too, so
    // every solve of a leading dimension uses one split-K order (results do
    // not depend on which solver ran it)
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(ckey);
    if (it != cache.end())
        shape_ = it->second;
    else
        shape_ = cache[ckey] = bestg;
}

void BatchedSolver::matvec32(const float* A32, const float* P, float* Y, hipStream_t s) {
    const int KS = slices(shape_);
    launch_gemm32(shape_, A32, P, s);
    if (Y) {
        hipLaunchKernelGGL(reduce32_kernel, dim3(nblocks(lda_ * KP)), dim3(kT), 0, s, (const float*)part32_, KS, lda_,
                           n_, Y);
        CP_LAUNCH_CHECK();
    }
}

void BatchedSolver::copy64(const double* B, double* Y, hipStream_t s) {
    // the reduction with no K slices: Y = B (rows >= n zeroed) and the
    // per-block column partials of ||Y_c||^2
    hipLaunchKernelGGL(reduce64_kernel, dim3((unsigned)(lda_ / kRed64Rows)), dim3(kT), 0, s, (const double*)nullptr,
                       0, lda_, n_, B,
                       1.0, Y, dpart_, (const int*)nullptr, (const int*)nullptr, (const double*)nullptr,
                       (const double*)nullptr);
    CP_LAUNCH_CHECK();
}

void BatchedSolver::matvec64acc(const float* A32, const double* X, double* Y, const double* Badd, double sign,
                                hipStream_t s, const double* corr, const double* A64) {
    // K slices: the fewest whose workgroups (two resident per CU) fill the
    // slots of every round to within 3% of the best count up to 32 (each
    // slice adds an lda x 64 FP64 partial to write and reduce)
    const int rb = (int)(lda_ / 256);
    const int slots = 2 * cus_;
    auto fill = [&](int ks) {
        const int64_t wg = (int64_t)rb * ks;
        return (double)wg / (double)(((wg + slots - 1) / slots) * slots);
    };
    double best = 0;
    for (int ks = 1; ks <= 32; ks++) best = std::max(best, fill(ks));
    int KS = 1;
    while (KS < 32 && fill(KS) < best - 0.03) KS++;
    if (f64_ks_ > 0) KS = f64_ks_;  // CPGPU_F64_KS
    int kchunk = (int)round_up((n_ + KS - 1) / KS, 16);
    dim3 grid((unsigned)(rb * KS));
    {
        TimedScope ts("resid_f64", s);
        if (A64 && f64_depth_ == 3)
            hipLaunchKernelGGL((gemm_f64acc_p_kernel<3, double>), grid, dim3(256), 0, s, A64, lda_, n_, X, part64_,
                               kchunk, KS);
        else if (A64)
            hipLaunchKernelGGL((gemm_f64acc_p_kernel<2, double>), grid, dim3(256), 0, s, A64, lda_, n_, X, part64_,
                               kchunk, KS);
        else if (f64_pipelined_)
            hipLaunchKernelGGL(gemm_f64acc_p_kernel<2>, grid, dim3(256), 0, s, A32, lda_, n_, X, part64_, kchunk, KS);
        else
            hipLaunchKernelGGL(gemm_f64acc_kernel, grid, dim3(256), 0, s, A32, lda_, n_, X, part64_, kchunk, KS);
    }
    CP_LAUNCH_CHECK();
    const double* vals = (corr && near_rowptr_ && !A64) ? corr : nullptr;
    hipLaunchKernelGGL(reduce64_kernel, dim3((unsigned)(lda_ / kRed64Rows)), dim3(kT), 0, s, (const double*)part64_,
                       KS, lda_, n_,
                       Badd, sign, Y, dpart_, near_rowptr_, near_cols_, vals, X);
    CP_LAUNCH_CHECK();
}

// Y = M^-1 PP (PP block-ordered, Y row-ordered), column c scaled by
// 1 / cscale[c * cstride] when cscale is given
void BatchedSolver::bj_apply_pk(const float* PP, float* Y, const double* cscale, int cstride, hipStream_t s) {
    const dim3 grid((unsigned)((bj_maxb_ + 31) / 32), (unsigned)bj_nblocks_);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(kT), 0, s, bj_perm_, bj_ptr_, bj_poff_, bj_pk_, PP, Y, cscale, cstride,
                           gate_);
    };
    if (bj_kind_ == kBjPackBf16 && bj_rt_ == 2) {
        const dim3 grid2((unsigned)((bj_maxb_ + 63) / 64), (unsigned)bj_nblocks_);
        auto kern = (bj_lo_ || bj_level_ >= 1) ? bj_apply_bf_kernel<true, 2> : bj_apply_bf_kernel<false, 2>;
        hipLaunchKernelGGL(kern, grid2, dim3(kT), 0, s, bj_perm_, bj_ptr_, bj_poff_, bj_pk_, PP, Y, cscale, cstride,
                           gate_);
    } else if (bj_kind_ == kBjPackBf16)
        go((bj_lo_ || bj_level_ >= 1) ? bj_apply_bf_kernel<true, 1> : bj_apply_bf_kernel<false, 1>);
    else if (bj_u_ == 1)
        go(bj_apply_pk_kernel<1>);
    else if (bj_u_ == 3)
        go(bj_apply_pk_kernel<3>);
    else
        go(bj_apply_pk_kernel<2>);
    CP_LAUNCH_CHECK();
}

void BatchedSolver::apply_block_jacobi(const float* P, float* Y, hipStream_t s) {
    TimedScope ts("bj_apply", s);
    if (bj_mfma_ && bj_iperm_) {  // P was also written in block order to pp_
        bj_apply_pk((const float*)pp_, Y, nullptr, 0, s);
    } else if (bj_mfma_) {
        dim3 grid((unsigned)((bj_maxb_ + 63) / 64), (unsigned)bj_nblocks_);
        hipLaunchKernelGGL(bj_apply_mfma_kernel, grid, dim3(kT), 0, s, bj_perm_, bj_ptr_, bj_off_, bj_inv_, P, Y, gate_);
    } else {
        dim3 grid((unsigned)bj_nblocks_, (unsigned)(KP / kBjCols));
        hipLaunchKernelGGL(bj_apply_kernel, grid, dim3(kT), (size_t)bj_maxb_ * kBjCols * sizeof(float), s, bj_perm_,
                           bj_ptr_, bj_off_, bj_inv_, P, Y, gate_);
    }
    CP_LAUNCH_CHECK();
}

int BatchedSolver::gm_bound(const BatchedOpts& o) const {
    int m = gm_m_;
    if (o.max_iter > 0) m = std::min(m, o.max_iter);
    return m;
}

bool BatchedSolver::gmres_ready() const { return gmres_ && bj_ && bj_mfma_ && bj_iperm_; }

bool BatchedSolver::group_capable() const {
    return gmres_ready() && bj_kind_ == kBjPackBf16 && bj_rt_ == 2 && shape_.variant >= 10 && shape_.variant <= 14 &&
           apk_ && tptr_;
}

// One refinement sweep's inner solve by right-preconditioned GMRES (see the
// gm_* kernels): gm_begin forms u_0 = R / |R| on the sweeping columns (act_),
// gm_iterations runs at most gm_bound() iterations, gm_end leaves
// d_ = M^-1 V y and each column's iteration count in sc_[Q_GMNJ].
void BatchedSolver::gm_begin(hipStream_t s) {
    hipLaunchKernelGGL(gm_init_kernel, dim3((unsigned)nrb_), dim3(kT), 0, s, n_, (const double*)R_,
                       (const double*)sc_, (const int*)act_, bj_iperm_, gm_u_, dpart_);
    hipLaunchKernelGGL(gm_start_kernel, dim3(1), dim3(1024), 0, s, (const double*)dpart_, nrb_, gm_st_);
    CP_LAUNCH_CHECK();
}

void BatchedSolver::gm_end(hipStream_t s) {
    // d = M^-1 sum_q y_q v_q
    hipLaunchKernelGGL(gm_backsub_kernel, dim3(1), dim3(KP), 0, s, gm_st_, sc_, gm_coef_);
    hipLaunchKernelGGL(gm_combine_kernel, dim3(nblocks((int64_t)n_ * KP)), dim3(kT), 0, s, n_, (const float*)gm_u_,
                       lda_ * KP, (const float*)gm_coef_, pp_);
    CP_LAUNCH_CHECK();
    apply_block_jacobi(nullptr, d_, s);
    gm_counted_ = true;
}

// The host reads iteration it's active flags while iteration it is queued
// behind them, as in the BiCGSTAB loop.
void BatchedSolver::gm_iterations(const float* A32, const BatchedOpts& o, hipStream_t s) {
    const int64_t us = lda_ * KP;
    const int m = gm_bound(o);
    int* hact[2] = {host_act_, host_act_ + 2 * KP};  // host_act_ + KP holds the sweep flags
    auto top = [&](int k) {
        CP_HIP(hipMemcpyAsync(hact[k & 1], act_, KP * sizeof(int), hipMemcpyDeviceToHost, s));
        CP_HIP(hipEventRecord(act_ev_[k & 1], s));
    };
    // iteration j in four segments: 0 z = M^-1 v_j (row order), 1 packing of
    // z, 2 the product w = A~ z, 3 Gram-Schmidt, Hessenberg column, Givens
    auto segment = [&](int j, int k) {
        if (k == 0) {
            TimedScope tb("bj_apply", s);
            bj_apply_pk((const float*)(gm_u_ + j * us), y_, (const double*)(gm_st_ + GM_S + j), kGmStride, s);
            return;
        }
        if (k == 1 || k == 2) {
            launch_gemm32(shape_, A32, y_, s, k);
            return;
        }
        const int KS = slices(shape_);
        const int ndc = (int)((n_ + kGmDotRows - 1) / kGmDotRows);     // position chunks of the dots
        const int nuc = (int)((n_ + 4 * kGmUpRounds * kGmUpPos - 1) / (4 * kGmUpRounds * kGmUpPos));  // update WGs
        hipLaunchKernelGGL(gm_w_kernel, dim3(nblocks((int64_t)n_ * (KP / 4))), dim3(kT), 0, s, (const float*)part32_, KS,
                           lda_, n_, (const int*)act_, bj_iperm_, gm_u_ + (j + 1) * us);
        hipLaunchKernelGGL(gm_dots_kernel, dim3((unsigned)ndc, (unsigned)((j + 4) / 4)), dim3(kT), 0, s, n_,
                           (const int*)act_, (const float*)gm_u_, us, j, gm_dpart_);
        hipLaunchKernelGGL(gm_hcol_kernel, dim3((unsigned)(j + 1)), dim3(1024), 0, s, (const double*)gm_dpart_, ndc,
                           (const int*)act_, j, gm_st_, gm_coef_);
        hipLaunchKernelGGL(gm_update_kernel, dim3((unsigned)nuc), dim3(kT), 0, s, n_, (const int*)act_, gm_u_, us, j,
                           (const float*)gm_coef_, dpart_);
        hipLaunchKernelGGL(gm_givens_kernel, dim3(1), dim3(1024), 0, s, (const double*)dpart_, nuc, act_, j, m,
                           (const double*)sc_, gm_st_);
        CP_LAUNCH_CHECK();
    };
    auto body = [&](int j) {
        for (int k = 0; k < 4; k++) segment(j, k);
    };
    // iteration j's launches depend only on j and on state fixed for the
    // solver's life (or named in the key), so they are captured once as
    // graphs and replayed: at most three launch calls per iteration instead
    // of eight (one when nothing is timed)
    const bool use_graph = graphs_;
    unsigned timed = 0;  // segments launched directly so that their timer records them
    if (KernelTimer* tm = current_timer())
        timed = (unsigned)tm->wants("bj_apply") | (unsigned)tm->wants("pack_p") << 1 |
                (unsigned)tm->wants("gemm_f32") << 2;
    if (use_graph) {
        const std::vector<uintptr_t> key = {(uintptr_t)A32, (uintptr_t)shape_.variant, (uintptr_t)shape_.rt,
                                            (uintptr_t)shape_.kslices, (uintptr_t)bj_level_, (uintptr_t)bj_lo_,
                                            (uintptr_t)bj_rt_, (uintptr_t)bj_kind_, (uintptr_t)bj_u_,
                                            (uintptr_t)bj_maxb_, (uintptr_t)bj_nblocks_, (uintptr_t)bj_perm_,
                                            (uintptr_t)bj_ptr_, (uintptr_t)bj_poff_, (uintptr_t)bj_pk_,
                                            (uintptr_t)bj_iperm_, (uintptr_t)apk_, (uintptr_t)tptr_,
                                            (uintptr_t)tkb_, (uintptr_t)m, (uintptr_t)timed};
        if (key != gm_plan_key_) {
            gm_plan_clear();
            gm_plan_key_ = key;
        }
    }
    // captures segments k0..k1-1 of iteration j; null when the runtime
    // refuses (the solver then launches on the stream from here on)
    auto capture = [&](int j, int k0, int k1) -> hipGraphExec_t {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        CP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        bool captured = false;
        std::string what;
        try {
            for (int k = k0; k < k1; k++) segment(j, k);
            captured = true;
        } catch (const HipError& err) {
            what = err.where;
        }
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (captured && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (captured && e == hipSuccess) return exec;
        (void)hipGetLastError();
        std::fprintf(stderr, "cpgpu: GMRES iteration graph not captured (%s%s%s); launching on the stream\n",
                     hipGetErrorString(e), what.empty() ? "" : ", ", what.c_str());
        graphs_ = false;
        return nullptr;
    };
    auto run = [&](int j) {
        if (!use_graph || !graphs_) {
            body(j);
            return;
        }
        if ((int)gm_plans_.size() <= j) gm_plans_.resize(j + 1);
        std::unique_ptr<IterPlan>& P = gm_plans_[j];
        if (!P) {
            std::unique_ptr<IterPlan> np(new IterPlan);
            int k0 = 0;
            for (int k = 0; k <= 4; k++) {
                if (k < 4 && !(k < 3 && (timed >> k & 1u))) continue;  // segment k joins the open graph
                if (k > k0) {
                    hipGraphExec_t ex = capture(j, k0, k);
                    if (!ex) {  // refused: nothing of iteration j has run yet
                        body(j);
                        return;
                    }
                    np->items.emplace_back(ex, -1);
                }
                if (k < 4) np->items.emplace_back(nullptr, k);
                k0 = k + 1;
            }
            P = std::move(np);
        }
        for (const auto& it : P->items) {
            if (it.first)
                CP_HIP(hipGraphLaunch(it.first, s));
            else
                segment(j, it.second);
        }
    };
    gate_ = act_;
    top(0);
    int it = 0;
    for (;;) {
        const bool more = it < m;
        if (more) {
            run(it);
            top(it + 1);
        }
        CP_HIP(hipEventSynchronize(act_ev_[it & 1]));
        const int* fl = hact[it & 1];
        int nact = 0;
        for (int c = 0; c < KP; c++) nact += fl[c] != 0;
        if (!more || nact == 0) {
            // the iteration queued behind this check found no active column:
            // its launches did no work, so they are not timed
            if (more)
                if (KernelTimer* t = current_timer())
                    for (const char* nm : {"gemm_f32", "pack_p", "bj_apply"})
                        if (t->wants(nm)) t->discard(nm, 1);
            break;
        }
        ++it;
    }
    gate_ = nullptr;
}

void BatchedSolver::solve(const float* A32, const float* B32, const double* normals, const double* vinf4, int k,
                          const BatchedOpts& o, double* x, int64_t ldx, SolveStats* stats, hipStream_t s) {
    for (int c0 = 0; c0 < k; c0 += KP) {
        int kc = std::min(KP, k - c0);
        tile_solve(A32, B32, normals, vinf4 + 4 * c0, c0, kc, o, x + (int64_t)c0 * ldx, ldx, stats + c0, s);
    }
}

void BatchedSolver::solve_rhs(const float* A32, const float* B32, const double* sigma, const double* rhs_add,
                              const double* x0, int k, const BatchedOpts& o, double* x, int64_t ldx,
                              SolveStats* stats, hipStream_t s) {
    for (int c0 = 0; c0 < k; c0 += KP) {
        int kc = std::min(KP, k - c0);
        const int64_t off = (int64_t)c0 * ldx;
        tile_solve(A32, B32, nullptr, nullptr, c0, kc, o, x + off, ldx, stats + c0, s, sigma + off,
                   rhs_add ? rhs_add + off : nullptr, x0 ? x0 + off : nullptr);
    }
}

// ------------------------------------------------ refinement loop phases
// One 64-column tile's solve is: tile_begin (right-hand side, initial guess,
// |b|^2 read-back), tile_begin_host after the stream synchronised, then per
// sweep tile_residual + synchronisation + tile_check (which columns sweep),
// tile_sweep_open, the inner solve, tile_sweep_close; finally tile_end.
// tile_solve runs them for one geometry, solve_group for several.

void BatchedSolver::colsum_q(const double* dp, int nq, int q0, int q1, int op, hipStream_t s, int nblk) {
    hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(1024), 0, s, dp, nblk > 0 ? nblk : nrb_, nq, q0, q1, sc_, op,
                       act_);
    CP_LAUNCH_CHECK();
}

void BatchedSolver::tile_begin(Tile& t, hipStream_t s) {
    const unsigned nel = nblocks(lda_ * KP);
    // RHS b = add - B sigma with FP64 accumulation (B: the FP64 matrix, or
    // B32 + the near-field correction)
    if (!t.sig_cm && !t.add_cm) {
        // steady sources sigma_c = -Vinf_c . n: b_c is a combination of the
        // three columns -B n_x, -B n_y, -B n_z, computed once per matrix
        const int mode = b64_ ? 1 : 0;
        if (bn_mode_ != mode) {
            hipLaunchKernelGGL(normal_cols_kernel, dim3(nel), dim3(kT), 0, s, n_, lda_, t.normals, S_);
            matvec64acc(t.B32, S_, R_, nullptr, -1.0, s, near_dB_, b64_);
            hipLaunchKernelGGL(keep3_kernel, dim3(nblocks(lda_)), dim3(kT), 0, s, lda_, (const double*)R_, bn_);
            CP_LAUNCH_CHECK();
            bn_mode_ = mode;
        }
        hipLaunchKernelGGL(rhs_bn_kernel, dim3(nel), dim3(kT), 0, s, n_, lda_, t.kc, (const double*)bn_, t.vinf4, R_);
        copy64(R_, Bv_, s);
    } else {
        if (t.sig_cm)
            hipLaunchKernelGGL(gather_cols_kernel, dim3(nel), dim3(kT), 0, s, n_, lda_, t.kc, t.sig_cm, t.ldx, S_);
        else
            hipLaunchKernelGGL(sigma_rows_kernel, dim3(nel), dim3(kT), 0, s, n_, lda_, t.kc, t.normals, t.vinf4, S_);
        if (t.add_cm)
            hipLaunchKernelGGL(gather_cols_kernel, dim3(nel), dim3(kT), 0, s, n_, lda_, t.kc, t.add_cm, t.ldx, R_);
        matvec64acc(t.B32, S_, Bv_, t.add_cm ? R_ : nullptr, -1.0, s, near_dB_, b64_);
    }
    colsum_q(dpart_, 1, Q_BB, 0, COLSUM_NONE, s, (int)(lda_ / kRed64Rows));
    hipLaunchKernelGGL(invd32_kernel, dim3(nblocks(lda_)), dim3(kT), 0, s, n_, t.A32, lda_, invd_);
    if (autotune_) tune(t.A32, s);
    if (t.x0_cm)
        hipLaunchKernelGGL(gather_cols_kernel, dim3(nel), dim3(kT), 0, s, n_, lda_, t.kc, t.x0_cm, t.ldx, X_);
    else
        CP_HIP(hipMemsetAsync(X_, 0, (size_t)lda_ * KP * sizeof(double), s));
    CP_HIP(hipMemcpyAsync(host_sc_, sc_, KP * Q_NQ * sizeof(double), hipMemcpyDeviceToHost, s));
}

void BatchedSolver::tile_begin_host(Tile& t) {
    t.bb.assign(KP, 0.0);
    t.done.assign(KP, 1);
    t.rel.assign(KP, 0.0);
    t.prev.assign(KP, 0.0);
    t.sweep = 0;
    for (int c = 0; c < t.kc; c++) {
        t.bb[c] = host_sc_[c * Q_NQ + Q_BB];
        t.done[c] = t.bb[c] == 0;
        t.stats[c] = SolveStats();
    }
    gm_counted_ = false;
}

// residual R = b - A X in FP64 (R = b while X is still zero) and its column
// norms to the host
void BatchedSolver::tile_residual(Tile& t, hipStream_t s) {
    if (t.sweep == 0 && !t.x0_cm)
        copy64(Bv_, R_, s);
    else
        matvec64acc(t.A32, X_, R_, Bv_, -1.0, s, near_dA_, a64_);
    colsum_q(dpart_, 1, Q_RR64, 0, COLSUM_NONE, s, (int)(lda_ / kRed64Rows));
    CP_HIP(hipMemcpyAsync(host_sc_, sc_, KP * Q_NQ * sizeof(double), hipMemcpyDeviceToHost, s));
}

bool BatchedSolver::tile_check(Tile& t) {
    int* host_sweep = host_act_ + KP;
    if (gm_counted_) {  // the GMRES sweep just finished left its iteration counts
        for (int c = 0; c < t.kc; c++) t.stats[c].iters += (int)host_sc_[c * Q_NQ + Q_GMNJ];
        gm_counted_ = false;
    }
    const double tol2 = t.o.tol * t.o.tol;
    bool any = false;
    for (int c = 0; c < KP; c++) {
        host_sweep[c] = 0;
        if (c >= t.kc || t.bb[c] == 0) continue;
        double rr = host_sc_[c * Q_NQ + Q_RR64];
        t.rel[c] = std::sqrt(rr / t.bb[c]);
        t.stats[c].rel_residual = t.rel[c];
        if (rr <= tol2 * t.bb[c]) t.done[c] = 1;
        if (!t.done[c]) {
            host_sweep[c] = 1;
            any = true;
            host_sc_[c * Q_NQ + Q_SCALE] = std::sqrt(rr);
            // the caller's inner reduction, or kInnerTolTarget relaxed to
            // what this column still needs (half of tol / rel)
            host_sc_[c * Q_NQ + Q_ITOL] = t.o.inner_tol > 0
                                              ? t.o.inner_tol
                                              : std::min(0.5, std::max(kInnerTolTarget, 0.5 * t.o.tol / t.rel[c]));
        } else {
            host_sc_[c * Q_NQ + Q_SCALE] = 0;
        }
    }
    // a sweep whose inner solve did not halve the residual of every column
    // it ran on: the preconditioner is too weak for this geometry (bf16
    // rounding of an ill-conditioned block's inverse can leave the
    // preconditioned operator with eigenvalues near zero, and GMRES then
    // stagnates); strengthen it for the remaining sweeps and later solves
    if (t.sweep > 0 && bj_ && any) {
        bool stalled = false;
        for (int c = 0; c < t.kc; c++)
            if (host_sweep[c] && !(t.rel[c] <= 0.5 * t.prev[c])) stalled = true;
        if (stalled) escalate_block_jacobi();
    }
    for (int c = 0; c < t.kc; c++) t.prev[c] = t.rel[c];
    if (!any || t.sweep >= t.o.max_refine) return false;
    ++t.sweep;
    return true;
}

// the sweep's per-column scale and inner tolerance, and its columns
void BatchedSolver::tile_sweep_open(Tile& t, hipStream_t s) {
    (void)t;
    CP_HIP(hipMemcpyAsync(sc_, host_sc_, KP * Q_NQ * sizeof(double), hipMemcpyHostToDevice, s));
    CP_HIP(hipMemcpyAsync(act_, host_act_ + KP, KP * sizeof(int), hipMemcpyHostToDevice, s));
}

// X += |R| d on the columns of this sweep
void BatchedSolver::tile_sweep_close(Tile& t, hipStream_t s) {
    const int* host_sweep = host_act_ + KP;
    CP_HIP(hipMemcpyAsync(act_, host_sweep, KP * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(correct_kernel, dim3(nblocks(lda_ * KP)), dim3(kT), 0, s, lda_, (const double*)sc_,
                       (const int*)act_, X_, (const float*)d_);
    CP_LAUNCH_CHECK();
    for (int c = 0; c < t.kc; c++)
        if (host_sweep[c]) t.stats[c].refinements++;
}

void BatchedSolver::tile_end(Tile& t, hipStream_t s) {
    hipLaunchKernelGGL(scatter_x_kernel, dim3(nblocks(lda_ * KP)), dim3(kT), 0, s, n_, t.kc, (const double*)X_, t.x,
                       t.ldx);
    CP_LAUNCH_CHECK();
}

// inner FP32 solve of A32 d = R / |R| on the sweeping columns by BiCGSTAB
// (point Jacobi, or the block-Jacobi product without the GMRES form)
void BatchedSolver::bicgstab_sweep(Tile& t, hipStream_t s) {
    const int64_t rows = lda_;
    const unsigned nel = nblocks(rows * KP);
    const int KS = slices(shape_);  // partial slices the reductions sum
        hipLaunchKernelGGL(inner_init_kernel, dim3((unsigned)nrb_), dim3(kT), 0, s, n_, (const double*)R_,
                           (const double*)sc_, (const int*)act_, r_, r0_, p_, v_, d_, dpart_);
        colsum_q(dpart_, 1, Q_RR, 0, COLSUM_NONE, s);
        hipLaunchKernelGGL(scal_init_kernel, dim3(1), dim3(KP), 0, s, sc_);
        int it = 0;
        const int maxit = t.o.max_iter > 0 ? t.o.max_iter : 2 * n_;
        // The host reads iteration it's active flags while iteration it is
        // already queued behind them, so the device never waits for the
        // host's check.  An iteration queued past the end of its sweep finds
        // no active column: its heavy kernels return at once (gate_) and the
        // per-column updates are masked as for any finished column.
        int* hact[2] = {host_act_, host_act_ + 2 * KP};  // host_act_ + KP holds the sweep flags
        auto top = [&](int k) {  // flag read-back of iteration k (its rho/beta step ends iteration k - 1)
            if (k == 0) hipLaunchKernelGGL(scal_beta_kernel, dim3(1), dim3(KP), 0, s, sc_, act_);
            CP_HIP(hipMemcpyAsync(hact[k & 1], act_, KP * sizeof(int), hipMemcpyDeviceToHost, s));
            CP_HIP(hipEventRecord(act_ev_[k & 1], s));
        };
        auto body = [&]() {  // one BiCGSTAB iteration on the active columns
            const int* ip = (bj_ && bj_mfma_) ? bj_iperm_ : nullptr;
            hipLaunchKernelGGL(inner_p_kernel, dim3(nel), dim3(kT), 0, s, rows, n_, (const double*)sc_,
                               (const int*)act_, (const float*)r_, p_, (const float*)v_,
                               bj_ ? (const float*)nullptr : (const float*)invd_, y_, ip, pp_);
            if (bj_) apply_block_jacobi(p_, y_, s);
            matvec32(t.A32, y_, nullptr, s);
            {
                TimedScope tr("reduce_f32", s);
                hipLaunchKernelGGL(inner_v_kernel, dim3((unsigned)nrb_), dim3(kT), 0, s, (const float*)part32_, KS,
                                   lda_, n_, (const int*)act_, v_, (const float*)r0_, dpart_);
            }
            colsum_q(dpart_, 1, Q_R0V, 0, COLSUM_ALPHA, s);
            hipLaunchKernelGGL(inner_s_kernel, dim3(nel), dim3(kT), 0, s, rows, n_, (const double*)sc_,
                               (const int*)act_, (const float*)r_, (const float*)v_,
                               bj_ ? (const float*)nullptr : (const float*)invd_, sv_, z_, ip, pp_);
            if (bj_) apply_block_jacobi(sv_, z_, s);
            matvec32(t.A32, z_, nullptr, s);
            {
                TimedScope tr("reduce_f32", s);
                hipLaunchKernelGGL(inner_t_kernel, dim3((unsigned)nrb_), dim3(kT), 0, s, (const float*)part32_, KS,
                                   lda_, n_, (const int*)act_, t_, (const float*)sv_, dpart_,
                                   dpart_ + (size_t)nrb_ * KP);
            }
            colsum_q(dpart_, 2, Q_TT, Q_TS, COLSUM_OMEGA, s);
            hipLaunchKernelGGL(inner_x_kernel, dim3((unsigned)nrb_), dim3(kT), 0, s, n_, (const double*)sc_,
                               (const int*)act_, d_, (const float*)y_, (const float*)z_, r_, (const float*)sv_,
                               (const float*)t_, (const float*)r0_, dpart_, dpart_ + (size_t)nrb_ * KP);
            colsum_q(dpart_, 2, Q_RR, Q_R0R, COLSUM_BETA, s);  // and the next iteration's rho/beta step
            CP_LAUNCH_CHECK();
        };
        gate_ = act_;
        top(0);
        for (;;) {
            const bool more = it < maxit;
            if (more) {
                body();
                top(it + 1);
            }
            CP_HIP(hipEventSynchronize(act_ev_[it & 1]));
            const int* fl = hact[it & 1];
            int nact = 0;
            for (int c = 0; c < KP; c++) nact += fl[c] != 0;
            if (!more || nact == 0) {
                // the iteration queued behind this check found no active
                // column: its launches did no work, so they are not timed
                if (more)
                    if (KernelTimer* t = current_timer())
                        for (const char* nm : {"gemm_f32", "pack_p", "reduce_f32", "bj_apply"}) t->discard(nm, 2);
                break;
            }
            for (int c = 0; c < t.kc; c++)
                if (fl[c]) t.stats[c].iters++;
            ++it;
        }
        gate_ = nullptr;
}

void BatchedSolver::tile_solve(const float* A32, const float* B32, const double* normals, const double* vinf4,
                               int c0, int kc, const BatchedOpts& o, double* x, int64_t ldx, SolveStats* stats,
                               hipStream_t s, const double* sig_cm, const double* add_cm, const double* x0_cm) {
    (void)c0;
    Tile t;
    t.A32 = A32;
    t.B32 = B32;
    t.normals = normals;
    t.vinf4 = vinf4;
    t.kc = kc;
    t.o = o;
    t.x = x;
    t.ldx = ldx;
    t.stats = stats;
    t.sig_cm = sig_cm;
    t.add_cm = add_cm;
    t.x0_cm = x0_cm;
    tile_begin(t, s);
    CP_HIP(hipStreamSynchronize(s));
    tile_begin_host(t);
    for (;;) {
        tile_residual(t, s);
        CP_HIP(hipStreamSynchronize(s));
        if (!tile_check(t)) break;
        tile_sweep_open(t, s);
        if (gmres_ready()) {
            gm_begin(s);
            gm_iterations(A32, o, s);
            gm_end(s);
        } else {
            bicgstab_sweep(t, s);
        }
        tile_sweep_close(t, s);
    }
    tile_end(t, s);
}

size_t BatchedSolver::group_bytes(int G) { return (size_t)round_up((int64_t)G * sizeof(GmProb), 256) + 256; }

void BatchedSolver::fill_prob(GmProb& P, const BatchedOpts& o, int* flag) const {
    P = GmProb{};
    const int KS = shape_.kslices;
    P.n = n_;
    P.nkb = nkb_;
    P.kslices = KS;
    P.nslices = slices(shape_);
    P.kbchunk = (nkb_ + KS - 1) / KS;
    P.ngemm = (int)((lda_ / ((shape_.variant == 13 || shape_.variant == 14) ? 256 : 512)) * KS);
    P.variant = shape_.variant;
    P.m = gm_bound(o);
    P.ndc = (int)((n_ + kGmDotRows - 1) / kGmDotRows);
    P.nuc = (int)((n_ + 4 * kGmUpRounds * kGmUpPos - 1) / (4 * kGmUpRounds * kGmUpPos));
    P.bj_gx = (bj_maxb_ + 63) / 64;
    P.bj_nblocks = bj_nblocks_;
    P.bj_lo = (bj_lo_ || bj_level_ >= 1) ? 1 : 0;
    P.lda = lda_;
    P.plane = plane_;
    P.us = lda_ * KP;
    P.apk = (decltype(P.apk))apk_;
    P.ppk = (decltype(P.ppk))ppk_;
    P.part32 = (decltype(P.part32))part32_;
    P.tptr = (decltype(P.tptr))tptr_;
    P.tkb = (decltype(P.tkb))tkb_;
    P.gm_u = (decltype(P.gm_u))gm_u_;
    P.gm_st = (decltype(P.gm_st))gm_st_;
    P.gm_dpart = (decltype(P.gm_dpart))gm_dpart_;
    P.gm_coef = (decltype(P.gm_coef))gm_coef_;
    P.dpart = (decltype(P.dpart))dpart_;
    P.act = (decltype(P.act))act_;
    P.sc = (decltype(P.sc))sc_;
    P.iperm = (decltype(P.iperm))bj_iperm_;
    P.y = (decltype(P.y))y_;
    P.bj_perm = (decltype(P.bj_perm))bj_perm_;
    P.bj_ptr = (decltype(P.bj_ptr))bj_ptr_;
    P.bj_poff = (decltype(P.bj_poff))bj_poff_;
    P.bj_pk = (decltype(P.bj_pk))bj_pk_;
    P.gflag = (decltype(P.gflag))flag;
}

// GMRES iterations of several geometries' sweeps in lockstep: iteration j of
// every geometry still active is one launch per stage.  probs: the group's
// descriptors on the device (already uploaded), hp: the same on the host;
// flags: one int per slot.  As in gm_iterations the host reads the flags of
// iteration it - 1 while iteration it is queued, and drops a geometry from
// the launches once its flag reads 0 (a launch it is still in returns at
// once for it).
void BatchedSolver::group_iterations(const std::vector<BatchedSolver*>& sv, const GmProb* probs,
                                     const std::vector<GmProb>& hp, int* flags, hipStream_t s) {
    const int G = (int)sv.size();
    int mmax = 0;
    for (const GmProb& P : hp) mmax = std::max(mmax, P.m);
    BatchedSolver* s0 = sv[0];
    int* hfl[2] = {s0->host_act_, s0->host_act_ + 2 * KP};  // s0->host_act_ + KP holds its sweep flags
    std::vector<int> cur(G);
    for (int q = 0; q < G; q++) cur[q] = q;
    GroupLaunch L{};
    L.probs = probs;
    auto launch = [&](const std::vector<int>& slots, int j, auto count, auto go) {
        int acc = 0, g = 0;
        for (int q : slots) {
            const int c = count(hp[q]);
            if (c <= 0) continue;
            L.idx[g] = q;
            L.pre[g] = acc;
            acc += c;
            g++;
        }
        if (g == 0) return;
        L.G = g;
        L.j = j;
        L.pre[g] = acc;
        go((unsigned)acc);
        CP_LAUNCH_CHECK();
    };
    auto run = [&](int j) {
        for (int lo = 0; lo < 2; lo++) {  // one block-Jacobi launch per preconditioner form present
            std::vector<int> sl;
            for (int q : cur)
                if (hp[q].bj_lo == lo) sl.push_back(q);
            auto cnt = [](const GmProb& P) { return P.bj_gx * P.bj_nblocks; };
            if (lo)
                launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL(grp_bj_kernel<true>, dim3(g), dim3(kT), 0, s, L); });
            else
                launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL(grp_bj_kernel<false>, dim3(g), dim3(kT), 0, s, L); });
        }
        launch(cur, j, [](const GmProb& P) { return (int)((2 * (int64_t)P.nkb * 64 + 255) / 256); },
               [&](unsigned g) { hipLaunchKernelGGL(grp_pack_kernel, dim3(g), dim3(kT), 0, s, L); });
        for (int v = 10; v <= 14; v++) {  // one product launch per variant present
            std::vector<int> sl;
            for (int q : cur)
                if (hp[q].variant == v) sl.push_back(q);
            if (sl.empty()) continue;
            auto cnt = [](const GmProb& P) { return P.ngemm + (int)(P.lda / 128); };
            switch (v) {
                case 10: launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL((grp_gemm_kernel<4, 2>), dim3(g), dim3(kT), 0, s, L); }); break;
                case 11: launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL((grp_gemm_kernel<4, 3>), dim3(g), dim3(kT), 0, s, L); }); break;
                case 12: launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL((grp_gemm_kernel<4, 4>), dim3(g), dim3(kT), 0, s, L); }); break;
                case 13: launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL((grp_gemm_kernel<2, 2>), dim3(g), dim3(kT), 0, s, L); }); break;
                default: launch(sl, j, cnt, [&](unsigned g) { hipLaunchKernelGGL((grp_gemm_kernel<2, 3>), dim3(g), dim3(kT), 0, s, L); }); break;
            }
        }
        launch(cur, j, [](const GmProb& P) { return (int)(((int64_t)P.n * (KP / 4) + kT - 1) / kT); },
               [&](unsigned g) { hipLaunchKernelGGL(grp_w_kernel, dim3(g), dim3(kT), 0, s, L); });
        launch(cur, j, [j](const GmProb& P) { return P.ndc * ((j + 4) / 4); },
               [&](unsigned g) { hipLaunchKernelGGL(grp_dots_kernel, dim3(g), dim3(kT), 0, s, L); });
        launch(cur, j, [j](const GmProb&) { return j + 1; },
               [&](unsigned g) { hipLaunchKernelGGL(grp_hcol_kernel, dim3(g), dim3(1024), 0, s, L); });
        launch(cur, j, [](const GmProb& P) { return P.nuc; },
               [&](unsigned g) { hipLaunchKernelGGL(grp_update_kernel, dim3(g), dim3(kT), 0, s, L); });
        launch(cur, j, [](const GmProb&) { return 1; },
               [&](unsigned g) { hipLaunchKernelGGL(grp_givens_kernel, dim3(g), dim3(1024), 0, s, L); });
    };
    auto top = [&](int k) {
        CP_HIP(hipMemcpyAsync(hfl[k & 1], flags, (size_t)G * sizeof(int), hipMemcpyDeviceToHost, s));
        CP_HIP(hipEventRecord(s0->act_ev_[k & 1], s));
    };
    int it = 0;
    for (;;) {
        const bool more = it < mmax;
        if (more) {
            run(it);
            top(it + 1);
        }
        if (it > 0) {  // every slot sweeps some column entering iteration 0
            CP_HIP(hipEventSynchronize(s0->act_ev_[it & 1]));
            const int* fl = hfl[it & 1];
            std::vector<int> nx;
            for (int q : cur)
                if (fl[q]) nx.push_back(q);
            cur.swap(nx);
            if (cur.empty()) break;  // iteration it (if queued) found nothing active
        }
        if (!more) break;
        ++it;
    }
}

void BatchedSolver::solve_group(std::vector<GroupItem>& items, void* dev, hipStream_t s) {
    const int G = (int)items.size();
    if (G <= 0) return;
    if (G > kMaxGroup) throw HipError{hipErrorInvalidValue, "solve_group: more than kMaxGroup geometries"};
    std::vector<Tile> T(G);
    for (int p = 0; p < G; p++) {
        const GroupItem& it = items[p];
        if (it.k <= 0 || it.k > KP) throw HipError{hipErrorInvalidValue, "solve_group: 1..64 conditions per geometry"};
        Tile& t = T[p];
        t.A32 = it.A32;
        t.B32 = it.B32;
        t.normals = it.normals;
        t.vinf4 = it.vinf4;
        t.kc = it.k;
        t.o = *it.o;
        t.x = it.x;
        t.ldx = it.ldx;
        t.stats = it.stats;
    }
    for (int p = 0; p < G; p++) items[p].sv->tile_begin(T[p], s);
    CP_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < G; p++) items[p].sv->tile_begin_host(T[p]);
    GmProb* dprobs = static_cast<GmProb*>(dev);
    int* dflags = reinterpret_cast<int*>(static_cast<char*>(dev) + round_up((int64_t)G * sizeof(GmProb), 256));
    std::vector<int> live(G, 1);
    std::vector<GmProb> hp;  // uploaded descriptors: kept until the next synchronisation
    for (;;) {
        for (int p = 0; p < G; p++)
            if (live[p]) items[p].sv->tile_residual(T[p], s);
        CP_HIP(hipStreamSynchronize(s));
        std::vector<int> grp, solo;
        for (int p = 0; p < G; p++) {
            if (!live[p]) continue;
            if (!items[p].sv->tile_check(T[p])) {
                live[p] = 0;
                continue;
            }
            (items[p].sv->group_capable() ? grp : solo).push_back(p);
        }
        if (grp.size() == 1) {  // a group of one runs its own launches
            solo.push_back(grp[0]);
            grp.clear();
        }
        if (grp.empty() && solo.empty()) break;
        for (int p : solo) {
            BatchedSolver* sv = items[p].sv;
            sv->tile_sweep_open(T[p], s);
            if (sv->gmres_ready()) {
                sv->gm_begin(s);
                sv->gm_iterations(T[p].A32, T[p].o, s);
                sv->gm_end(s);
            } else {
                sv->bicgstab_sweep(T[p], s);
            }
            sv->tile_sweep_close(T[p], s);
        }
        if (!grp.empty()) {
            std::vector<BatchedSolver*> sv;
            hp.assign(grp.size(), GmProb{});
            for (size_t q = 0; q < grp.size(); q++) {
                BatchedSolver* b = items[grp[q]].sv;
                b->tile_sweep_open(T[grp[q]], s);
                b->gm_begin(s);
                b->fill_prob(hp[q], T[grp[q]].o, dflags + q);
                sv.push_back(b);
            }
            CP_HIP(hipMemcpyAsync(dprobs, hp.data(), hp.size() * sizeof(GmProb), hipMemcpyHostToDevice, s));
            group_iterations(sv, dprobs, hp, dflags, s);
            for (size_t q = 0; q < grp.size(); q++) {
                sv[q]->gm_end(s);
                sv[q]->tile_sweep_close(T[grp[q]], s);
            }
        }
    }
    for (int p = 0; p < G; p++) items[p].sv->tile_end(T[p], s);
}

}  // namespace cpgpu
