// CPU replay of the HIP runtime, for running the library's host code and its
// kernels on the CPU under AddressSanitizer, UndefinedBehaviorSanitizer and
// ThreadSanitizer (test infrastructure; never part of the product build).
// Device memory is host memory, streams and events are no-ops (every call
// completes before it returns), and kernels run on the host, one fiber per
// GPU thread, so that __syncthreads and __shared__ keep their meaning
// (workgroups run one at a time).  Wavefront intrinsics (ballot, lane
// permutes, shuffles) are lockstep exchanges between the 64 fibers of a
// wavefront; they assume what the kernels guarantee: every lane of a
// wavefront calls them the same number of times (wave-uniform control), and
// a kernel's __syncthreads come before its wavefront exchanges diverge
// between wavefronts.  Matrix-core builtins are not provided: the
// translation units that use them are not replayed (stubs.cpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__ __restrict
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3s { unsigned x, y, z; };
extern uint3s threadIdx, blockIdx;
extern dim3 blockDim, gridDim;

struct double2 { double x, y; };
inline double2 make_double2(double a, double b) { return double2{a, b}; }
template <class T>
inline void __builtin_nontemporal_store(T v, T* p) { *p = v; }

using std::min;
using std::max;
inline int64_t min(int64_t a, int b) { return a < b ? a : b; }
inline int64_t min(int a, int64_t b) { return a < b ? a : b; }
inline int64_t max(int64_t a, int b) { return a > b ? a : b; }
inline int64_t max(int a, int64_t b) { return a > b ? a : b; }

inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
void rp_syncthreads();
#define __syncthreads() rp_syncthreads()

// every lane of the calling wavefront deposits v at slot `dst`; returns the
// wavefront's 64 slots once all its lanes have deposited
const uint64_t* rp_wave_exchange(uint64_t v, int dst);
int rp_lane();
inline uint64_t __ballot(int pred) {
    const uint64_t* s = rp_wave_exchange(pred ? 1 : 0, rp_lane());
    uint64_t m = 0;
    for (int l = 0; l < 64; l++)
        if (s[l]) m |= 1ull << l;
    return m;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
// the lanes run one after another between switches, so a plain update is atomic
inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v) {
    const unsigned long long o = *p;
    *p = o | v;
    return o;
}
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __builtin_amdgcn_ds_permute(int addr, int v) {
    return (int)(uint32_t)rp_wave_exchange((uint32_t)v, (addr >> 2) & 63)[rp_lane()];
}
inline int __builtin_amdgcn_ds_bpermute(int addr, int v) {
    return (int)(uint32_t)rp_wave_exchange((uint32_t)v, rp_lane())[(addr >> 2) & 63];
}
inline double __shfl_xor(double v, int mask, int width = 64) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    const uint64_t* s = rp_wave_exchange(b, rp_lane());
    const int src = (rp_lane() & ~(width - 1)) | ((rp_lane() ^ mask) & (width - 1));
    double r;
    std::memcpy(&r, &s[src], 8);
    return r;
}

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
typedef struct rp_stream* hipStream_t;
typedef struct rp_event* hipEvent_t;
typedef struct rp_graph_exec* hipGraphExec_t;  // the solver that captures graphs is not replayed
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum { hipStreamNonBlocking = 1, hipHostMallocDefault = 0, hipEventDisableTiming = 2 };

hipError_t hipMalloc(void** p, size_t n);
template <class T> hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
template <class T> hipError_t hipHostMalloc(T** p, size_t n, unsigned flags = 0) { return hipHostMalloc((void**)p, n, flags); }
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, hipMemcpyKind k,
                            hipStream_t st = nullptr);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipMemset(void* d, int v, size_t n);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int* d);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b);
hipError_t hipGetSymbolAddress(void** p, const void* sym);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f);

void rp_launch(dim3 g, dim3 b, const std::function<void()>& body);
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
    rp_launch(dim3(grid), dim3(block), [&]() { kernel(__VA_ARGS__); })
