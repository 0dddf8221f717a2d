// Strip block-Jacobi plan and block inversion (see precond.h).
#include "precond.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>

namespace cpgpu {

int bj_block_target(int n) {
    int target = std::min(kBjHardMax, std::max(768, n / 10));
    if (const char* e = std::getenv("CPGPU_BJ_TARGET")) {
        const int t = std::atoi(e);
        if (t >= 512 && t <= kBjHardMax) target = t;
    }
    return target;
}

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

}  // namespace cpgpu
