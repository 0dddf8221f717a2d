// CPU replay runtime (see hip/hip_runtime.h): HIP calls on host memory and
// kernels on fibers.  A fiber is a stack and a saved stack pointer; rp_swap
// saves the callee-saved registers and the FP control words of the running
// context and resumes the other.  A workgroup's threads start as fibers and
// run in turn until each reaches __syncthreads (which swaps back to the
// scheduler) or returns; rounds repeat until every fiber has returned.
#include "hip/hip_runtime.h"

#include <cstdio>
#include <vector>

uint3s threadIdx, blockIdx;
dim3 blockDim, gridDim;

extern "C" void rp_swap(void** save_sp, void* load_sp);
asm(R"(
.text
.globl rp_swap
.type rp_swap,@function
rp_swap:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.globl rp_trampoline
.type rp_trampoline,@function
rp_trampoline:
    movq %r12, %rdi
    call *%r13
    ud2
)");
extern "C" void rp_trampoline();

namespace {
constexpr size_t kStack = 256 * 1024;
struct Fiber {
    void* sp = nullptr;
    char* stack = nullptr;
    bool done = false;
    unsigned tx, ty, tz;
    unsigned phase = 0;  // wavefront exchanges made
    int wave = 0, lane = 0;
};
// double-buffered exchange slots per wavefront (16 wavefronts of 64 lanes):
// exchange k writes buffer k & 1 in one round and reads it in the next,
// while a lane that already moved on writes exchange k + 1 into the other
constexpr int kMaxWaves = 16;
uint64_t slots[kMaxWaves][2][64];
std::vector<Fiber> fibers;
void* sched_sp = nullptr;
int cur = -1;
const std::function<void()>* body_ptr = nullptr;
bool in_kernel = false;

extern "C" void rp_fiber_main(void* arg) {
    Fiber* f = (Fiber*)arg;
    (*body_ptr)();
    f->done = true;
    rp_swap(&f->sp, sched_sp);
    __builtin_trap();
}

void init_fiber(Fiber& f) {
    if (!f.stack) f.stack = (char*)aligned_alloc(64, kStack);
    uintptr_t top = ((uintptr_t)(f.stack + kStack) & ~(uintptr_t)15) - 16;
    uint64_t* s = (uint64_t*)(top - 64);
    uint32_t* ctl = (uint32_t*)s;
    ctl[0] = 0x1F80;             // mxcsr
    ((uint16_t*)ctl)[2] = 0x037F;  // x87 control word
    s[1] = 0;                      // r15
    s[2] = 0;                      // r14
    s[3] = (uint64_t)(void*)&rp_fiber_main;  // r13
    s[4] = (uint64_t)(void*)&f;    // r12
    s[5] = 0;                      // rbx
    s[6] = 0;                      // rbp
    s[7] = (uint64_t)(void*)&rp_trampoline;  // return address
    f.sp = s;
    f.done = false;
}
}  // namespace

const uint64_t* rp_wave_exchange(uint64_t v, int dst) {
    Fiber& f = fibers[cur];
    uint64_t* buf = slots[f.wave][f.phase & 1];
    buf[dst] = v;
    f.phase++;
    rp_swap(&f.sp, sched_sp);
    return buf;
}

int rp_lane() { return fibers[cur].lane; }

void rp_syncthreads() {
    if (!in_kernel || cur < 0) return;
    Fiber& f = fibers[cur];
    rp_swap(&f.sp, sched_sp);
}

void rp_launch(dim3 g, dim3 b, const std::function<void()>& body) {
    const unsigned nt = b.x * b.y * b.z;
    if (nt == 0 || g.x * g.y * g.z == 0) return;
    if (nt > 1024) { fprintf(stderr, "replay: block of %u threads\n", nt); abort(); }
    if (fibers.size() < nt) fibers.resize(nt);
    gridDim = g;
    blockDim = b;
    body_ptr = &body;
    in_kernel = true;
    for (unsigned bz = 0; bz < g.z; bz++)
        for (unsigned by = 0; by < g.y; by++)
            for (unsigned bx = 0; bx < g.x; bx++) {
                blockIdx = uint3s{bx, by, bz};
                unsigned t = 0;
                for (unsigned z = 0; z < b.z; z++)
                    for (unsigned y = 0; y < b.y; y++)
                        for (unsigned x = 0; x < b.x; x++, t++) {
                            init_fiber(fibers[t]);
                            fibers[t].phase = 0;
                            fibers[t].wave = (int)(t / 64);
                            fibers[t].lane = (int)(t % 64);
                            fibers[t].tx = x;
                            fibers[t].ty = y;
                            fibers[t].tz = z;
                        }
                for (;;) {
                    bool any = false;
                    for (unsigned i = 0; i < nt; i++) {
                        Fiber& f = fibers[i];
                        if (f.done) continue;
                        threadIdx = uint3s{f.tx, f.ty, f.tz};
                        cur = (int)i;
                        rp_swap(&sched_sp, f.sp);
                        cur = -1;
                        any = true;
                    }
                    if (!any) break;
                }
            }
    in_kernel = false;
}

// ---- memory and streams
static int last_err = 0;
hipError_t hipMalloc(void** p, size_t n) {
    *p = aligned_alloc(256, (n + 255) / 256 * 256);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) { free(p); return hipSuccess; }
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, hipMemcpyKind, hipStream_t) {
    if (w > dp || w > sp) return hipErrorInvalidValue;
    for (size_t r = 0; r < h; r++) memmove((char*)d + r * dp, (const char*)s + r * sp, w);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { memset(d, v, n); return hipSuccess; }
hipError_t hipMemset(void* d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
struct rp_stream { int x; };
struct rp_event { int x; };
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new rp_stream{0}; return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
// a 16 GiB device
hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
    *free_b = *total_b = (size_t)16 << 30;
    return hipSuccess;
}
hipError_t hipGetLastError() { int e = last_err; last_err = 0; return e; }
const char* hipGetErrorString(hipError_t) { return "replay error"; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new rp_event{0}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
