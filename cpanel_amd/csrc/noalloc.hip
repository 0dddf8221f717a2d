// An allocation-free form of the FP64 product y = A x (cp_matvec kernel 0)
// for callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The caller
// passes device buffers for x and y; the split-K partial sums of the GEMV
// live in a file-scope device array instead of a workspace allocation.
//
// The array is one per GPU (each device holds its own copy at its own
// address, found once per device with hipGetSymbolAddress) and keeps the
// last call's contents; every partial slice a call reads it has written
// first, so nothing needs clearing.  Calls that overlap on one GPU race on
// it.
#include <mutex>

#include "capi_internal.h"

using cpgpu::capi::fail;
using cpgpu::capi::guarded_g;

// Gemv64 splits K into S slices of lda partial rows each, S * lda <= (4096 +
// lda / 512) * 512 doubles: 2.5 Mi doubles (20 MiB) hold every leading
// dimension up to 524,288
constexpr size_t kScratchDoubles = (size_t)5 << 19;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device (device d)
double* scratch_on(int d) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[d]) {
        void* p = nullptr;
        CP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[d] = static_cast<double*>(p);
    }
    return g_scratch[d];
}

}  // namespace

extern "C" cp_status cp_matvec_f64_noalloc(cp_geom* g, int32_t k, const double* X, int64_t ldx, double* Y,
                                           int64_t ldy) {
    if (!g || k <= 0 || !X || !Y) return CP_BAD_ARG;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    if (c->device >= kMaxDevices) return fail(c, CP_BAD_ARG, "device index beyond the scratch table");
    const int n = g->dev.n;
    const int64_t lda = g->dev.lda;
    if (ldx < n || ldy < lda)
        return fail(c, CP_BAD_ARG, "ldx must be >= n_body and ldy >= the geometry's lda (" + std::to_string(lda) + ")");
    return guarded_g(g, [&]() -> cp_status {
        if (!g->have64) return fail(c, CP_NOT_ASSEMBLED, "FP64 matrices not assembled");
        const Gemv64 gv(n, lda);  // the same split as cp_matvec: bit-identical results
        if ((size_t)gv.slices() * (size_t)lda > kScratchDoubles)
            return fail(c, CP_BAD_ARG, "split-K partials need " + std::to_string((size_t)gv.slices() * lda) +
                                           " doubles, beyond the static scratch of " +
                                           std::to_string(kScratchDoubles));
        double* part = scratch_on(c->device);
        for (int q = 0; q < k; q++)
            gv.apply(g->A64.as<double>(), X + (int64_t)q * ldx, Y + (int64_t)q * ldy, 1.0, part, g->stream);
        CP_HIP(hipStreamSynchronize(g->stream));
        return CP_OK;
    });
}
