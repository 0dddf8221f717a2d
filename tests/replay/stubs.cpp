// The batched mixed-precision path (batched.hip, blockinv.hip: matrix-core
// builtins) is not replayed: its solver entry points throw, which the C ABI
// reports as CP_HIP_ERROR.  launch_sigma (used by the FP64 solve) is restated
// on the host; the bf16 packing is skipped.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "capi_internal.h"
namespace cpgpu {
[[noreturn]] static void nope(const char* w) { throw HipError{1, std::string("replay: not emulated: ") + w}; }
BatchedSolver::BatchedSolver(int, int64_t, int, void*, void*, hipStream_t) { nope("BatchedSolver"); }
size_t BatchedSolver::workspace_bytes(int64_t) { return 256; }
size_t BatchedSolver::host_bytes() { return 256; }
BatchedSolver::~BatchedSolver() {}
BatchedSolver::IterPlan::~IterPlan() {}
size_t BatchedSolver::group_bytes(int) { return 256; }
void BatchedSolver::solve_group(std::vector<GroupItem>&, void*, hipStream_t) { nope("solve_group"); }
void BatchedSolver::solve(const float*, const float*, const double*, const double*, int, const BatchedOpts&, double*,
                          int64_t, SolveStats*, hipStream_t) { nope("solve"); }
void BatchedSolver::solve_rhs(const float*, const float*, const double*, const double*, const double*, int,
                              const BatchedOpts&, double*, int64_t, SolveStats*, hipStream_t) { nope("solve_rhs"); }
void BatchedSolver::matvec32(const float*, const float*, float*, hipStream_t) { nope("matvec32"); }
void BatchedSolver::matvec64acc(const float*, const double*, double*, const double*, double, hipStream_t,
                                const double*, const double*) { nope("matvec64acc"); }
void launch_sigma(int n, int64_t lda, const double* nrm, const double* v4, double* sig, hipStream_t) {
    for (int64_t i = 0; i < lda; i++)
        sig[i] = i < n ? -(v4[0] * nrm[i] + v4[1] * nrm[lda + i] + v4[2] * nrm[2 * lda + i]) : 0.0;
}
int packed_kblocks(int n) { return (n + 15) / 16; }
size_t packed_a_bytes(int64_t, int) { return 256; }
int64_t bj_packed_floats(int b) { return (int64_t)b * b; }
void launch_bj_pack(const float*, const int*, const int64_t*, const int64_t*, int, int, float*, int, hipStream_t) { nope("bj_pack"); }
// the bf16 tile planes feed only the batched products, which are not
// replayed
void launch_pack_a(const float*, int64_t, int, void*, hipStream_t) {}
int64_t bj_invert_rows_doubles(int, int) { return 1; }
void launch_bj_invert(const float*, int64_t, const int*, const int*, const int64_t*, int, int, double*, int*, int*, int*,
                      double*, float*, hipStream_t) { nope("bj_invert"); }
}  // namespace cpgpu

// the device octree build needs hipCUB's radix sort and scan, which the
// replay does not provide: report "not built", so the host build runs
namespace cpgpu {
namespace capi {
bool build_tree_device(cp_ctx*, int, const double*, const double*, const double*, int, hipStream_t, cp_tree*) {
    return false;
}
}  // namespace capi
}  // namespace cpgpu

// the spatial order of the panel sums' near pass (hipCUB on the device):
// the same 30-bit Morton keys over the bounding cube, stably sorted on the host
namespace cpgpu {
size_t morton_sort_bytes(int) { return 0; }
void sort_points_morton(int m, const double* pts, void*, int* perm, hipStream_t) {
    if (m <= 0) return;
    double mn[3], mx[3];
    for (int d = 0; d < 3; d++) mn[d] = mx[d] = pts[d];
    for (int i = 0; i < m; i++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], pts[3 * i + d]);
            mx[d] = std::max(mx[d], pts[3 * i + d]);
        }
    double c[3], maxD = 0;
    for (int d = 0; d < 3; d++) {
        c[d] = 0.5 * (mn[d] + mx[d]);
        const double e = mx[d] - mn[d];
        maxD = d == 0 ? e : (e > maxD ? e : maxD);
    }
    const double h = 0.51 * maxD;
    auto spread = [](uint32_t v) {
        v &= 0x3ff;
        v = (v | (v << 16)) & 0x030000ffu;
        v = (v | (v << 8)) & 0x0300f00fu;
        v = (v | (v << 4)) & 0x030c30c3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    std::vector<uint32_t> key(m);
    for (int i = 0; i < m; i++) {
        uint32_t q[3] = {0, 0, 0};
        if (h > 0)
            for (int d = 0; d < 3; d++) {
                const double u = (pts[3 * i + d] - (c[d] - h)) / (2 * h) * 1024.0;
                q[d] = (uint32_t)std::fmin(std::fmax(u, 0.0), 1023.0);
            }
        key[i] = (spread(q[0]) << 2) | (spread(q[1]) << 1) | spread(q[2]);
        perm[i] = i;
    }
    std::stable_sort(perm, perm + m, [&](int a, int b) { return key[a] < key[b]; });
}
}  // namespace cpgpu
