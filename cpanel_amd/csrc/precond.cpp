// Strip block-Jacobi plan and block inversion (see precond.h).
#include "precond.h"

#include <algorithm>
#include <atomic>
#include <cmath>

namespace cpgpu {

void build_strip_blocks(const Geometry& G, int target, int hard_max, BlockPlan& plan) {
    const int n = (int)G.body.size();
    // strips: spanwise intervals of the trailing-edge edges of every wake line
    std::vector<std::pair<double, double>> strips;
    for (const WakeInfo& W : G.wakes)
        for (const WakeLineInfo& wl : W.lines) strips.push_back({std::min(wl.p1.y, wl.p2.y), std::max(wl.p1.y, wl.p2.y)});
    std::sort(strips.begin(), strips.end());
    std::vector<std::vector<int>> groups;
    if (strips.size() >= 2) {
        groups.resize(strips.size());
        std::vector<double> lo(strips.size());
        for (size_t k = 0; k < strips.size(); k++) lo[k] = strips[k].first;
        for (int i = 0; i < n; i++) {
            double y = G.body[i].center.y;
            // last strip whose lower bound is <= y (tips fall into the end strips)
            size_t k = (size_t)(std::upper_bound(lo.begin(), lo.end(), y) - lo.begin());
            k = k == 0 ? 0 : k - 1;
            groups[k].push_back(i);
        }
    } else {
        // no wake lines: spanwise slabs of the y-sorted panels
        std::vector<int> idx(n);
        for (int i = 0; i < n; i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return G.body[a].center.y < G.body[b].center.y; });
        for (int a = 0; a < n; a += target) groups.emplace_back(idx.begin() + a, idx.begin() + std::min(n, a + target));
    }
    // merge consecutive strips up to the target size; split oversize strips
    plan.perm.clear();
    plan.ptr.assign(1, 0);
    std::vector<int> cur;
    auto flush = [&]() {
        if (cur.empty()) return;
        std::sort(cur.begin(), cur.end());
        for (size_t a = 0; a < cur.size(); a += hard_max) {
            size_t e = std::min(cur.size(), a + (size_t)hard_max);
            plan.perm.insert(plan.perm.end(), cur.begin() + a, cur.begin() + e);
            plan.ptr.push_back((int)plan.perm.size());
        }
        cur.clear();
    };
    for (auto& g : groups) {
        if (g.empty()) continue;
        if (!cur.empty() && (int)(cur.size() + g.size()) > target) flush();
        cur.insert(cur.end(), g.begin(), g.end());
    }
    flush();
    plan.off.assign(plan.nblocks() + 1, 0);
    plan.max_b = 0;
    for (int q = 0; q < plan.nblocks(); q++) {
        int b = plan.ptr[q + 1] - plan.ptr[q];
        plan.max_b = std::max(plan.max_b, b);
        plan.off[q + 1] = plan.off[q] + (int64_t)b * b;
    }
}

namespace {

// Gauss-Jordan inversion with partial pivoting of a row-major b x b matrix.
bool gauss_jordan(int b, double* M) {
    std::vector<std::pair<int, int>> swaps;
    for (int k = 0; k < b; k++) {
        int p = k;
        double best = std::fabs(M[(size_t)k * b + k]);
        for (int i = k + 1; i < b; i++) {
            double v = std::fabs(M[(size_t)i * b + k]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (!(best > 0)) return false;
        if (p != k) {
            std::swap_ranges(M + (size_t)k * b, M + (size_t)k * b + b, M + (size_t)p * b);
            swaps.push_back({k, p});
        }
        double* rk = M + (size_t)k * b;
        const double piv = rk[k];
        rk[k] = 1.0;
        const double inv = 1.0 / piv;
        for (int j = 0; j < b; j++) rk[j] *= inv;
        for (int i = 0; i < b; i++) {
            if (i == k) continue;
            double* ri = M + (size_t)i * b;
            const double f = ri[k];
            if (f == 0) continue;
            ri[k] = 0.0;
            for (int j = 0; j < b; j++) ri[j] -= f * rk[j];
        }
    }
    // inverse of the row-permuted matrix -> undo with column swaps, reversed
    for (auto it = swaps.rbegin(); it != swaps.rend(); ++it)
        for (int i = 0; i < b; i++) std::swap(M[(size_t)i * b + it->first], M[(size_t)i * b + it->second]);
    return true;
}

}  // namespace

bool invert_blocks(const BlockPlan& plan, std::vector<double>& blocks) {
    std::atomic<bool> ok{true};
    const int nb = plan.nblocks();
    std::atomic<int> next{0};
    parallel_for(std::max(nb, 256), 0, [&](int, int) {
        // dynamic assignment: blocks differ in size
        for (;;) {
            int q = next.fetch_add(1);
            if (q >= nb) break;
            int b = plan.ptr[q + 1] - plan.ptr[q];
            double* M = blocks.data() + plan.off[q];
            // column-major in, row-major work, column-major out
            std::vector<double> R((size_t)b * b);
            for (int r = 0; r < b; r++)
                for (int s = 0; s < b; s++) R[(size_t)r * b + s] = M[(size_t)s * b + r];
            if (!gauss_jordan(b, R.data())) ok = false;
            for (int r = 0; r < b; r++)
                for (int s = 0; s < b; s++) M[(size_t)s * b + r] = R[(size_t)r * b + s];
        }
    });
    return ok;
}

}  // namespace cpgpu
