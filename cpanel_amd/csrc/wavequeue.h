// A wavefront-wide queue of work items (one int per lane): items marked by
// a ballot join it, compacted with lane permutes, and every time 64 are
// queued the whole wavefront evaluates them, so work found sparsely (a few
// lanes at a time) runs with all 64 lanes busy.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cpgpu {

// the queued items, one per lane in lanes [0, qn)
struct RowQueue {
    int p = 0;
    int qn = 0;  // wave-uniform
};

// lanes with `take` append `row`; when 64 are queued, eval(item, true) runs
// on every lane (one item each) and the rest stays queued; flush with
// eval(q.p, lane < q.qn) at the end
template <class F>
__device__ inline void queue_push(RowQueue& q, bool take, int row, int lane, F&& eval) {
    const uint64_t m = __ballot(take);
    if (m == 0) return;
    const int c = __popcll(m);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int rn = __popcll(m & below);
    // a permutation: the marked items to lanes [0, c), the others above
    const int dst = take ? rn : c + (lane - rn);
    const int b = __builtin_amdgcn_ds_permute(dst << 2, row);
    // the new items after the queued ones (lanes [qn, qn + c), wrapping)
    const int sh = __builtin_amdgcn_ds_bpermute(((lane - q.qn) & 63) << 2, b);
    const int merged = lane < q.qn ? q.p : sh;
    if (q.qn + c >= 64) {
        eval(merged, true);
        q.p = sh;  // lanes [0, qn + c - 64): the items that did not fit
        q.qn = q.qn + c - 64;
    } else {
        q.p = merged;
        q.qn += c;
    }
}

}  // namespace cpgpu
