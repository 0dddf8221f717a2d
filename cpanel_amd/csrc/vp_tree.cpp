// Particle octree build and flattening (see vp_tree.h).
#include "vp_tree.h"

#include <cmath>

namespace cpgpu {

namespace {

#pragma clang fp contract(off)

struct BNode {
    double o[3];
    double h;
    int level;
    int child[8];
    std::vector<int> members;
    double ex, ey, ez, sx, sy, sz, radius;
};

struct Builder {
    const double* xyz;
    const double* alpha;
    int max_members;
    int depth_cap = 60;  // coincident particles would split forever
    std::vector<BNode> B;
    int max_level = 0;

    int make(const double o[3], double h, int level) {
        BNode b;
        b.o[0] = o[0];
        b.o[1] = o[1];
        b.o[2] = o[2];
        b.h = h;
        b.level = level;
        for (int& c : b.child) c = -1;
        b.ex = b.ey = b.ez = b.sx = b.sy = b.sz = b.radius = 0;
        B.push_back(std::move(b));
        if (level > max_level) max_level = level;
        return (int)B.size() - 1;
    }

    // top-down split of the particles reaching node id (stable per octant)
    void split(int id, std::vector<int>& idx) {
        if ((int)idx.size() <= max_members || B[id].level >= depth_cap) {
            B[id].members.swap(idx);
            return;
        }
        std::vector<int> oct[8];
        const double ox = B[id].o[0], oy = B[id].o[1], oz = B[id].o[2];
        for (int i : idx) {
            int c = 0;
            if (xyz[3 * i] > ox) c |= 4;
            if (xyz[3 * i + 1] > oy) c |= 2;
            if (xyz[3 * i + 2] > oz) c |= 1;
            oct[c].push_back(i);
        }
        std::vector<int>().swap(idx);
        for (int c = 0; c < 8; c++) {
            if (oct[c].empty()) continue;
            const double h = B[id].h;
            double co[3] = {B[id].o[0], B[id].o[1], B[id].o[2]};
            co[0] += h * (c & 4 ? 0.5 : -0.5);
            co[1] += h * (c & 2 ? 0.5 : -0.5);
            co[2] += h * (c & 1 ? 0.5 : -0.5);
            int ch = make(co, 0.5 * h, B[id].level + 1);
            B[id].child[c] = ch;
            split(ch, oct[c]);
        }
    }

    bool leaf(int id) const {
        for (int c : B[id].child)
            if (c >= 0) return false;
        return true;
    }

    // pseudo-particle of a node: strength-magnitude-weighted position, summed
    // strength, radius |half| * 2 * 1.5 (particleFMM.cpp:33-55, 120-170)
    void expansion(int id, bool root) {
        for (int c : B[id].child)
            if (c >= 0) expansion(c, false);
        if (root) return;
        BNode& b = B[id];
        double wx = 0, wy = 0, wz = 0, tot = 0, sx = 0, sy = 0, sz = 0;
        auto add = [&](double px, double py, double pz, double ax, double ay, double az) {
            double nrm = std::sqrt(ax * ax + ay * ay + az * az);
            wx += px * nrm;
            wy += py * nrm;
            wz += pz * nrm;
            tot += nrm;
        };
        if (leaf(id)) {
            for (int i : b.members)
                add(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], alpha[3 * i], alpha[3 * i + 1], alpha[3 * i + 2]);
            for (int i : b.members) {
                sx += alpha[3 * i];
                sy += alpha[3 * i + 1];
                sz += alpha[3 * i + 2];
            }
        } else {
            for (int c : b.child)
                if (c >= 0) add(B[c].ex, B[c].ey, B[c].ez, B[c].sx, B[c].sy, B[c].sz);
            for (int c : b.child)
                if (c >= 0) {
                    sx += B[c].sx;
                    sy += B[c].sy;
                    sz += B[c].sz;
                }
        }
        b.ex = wx / tot;
        b.ey = wy / tot;
        b.ez = wz / tot;
        b.sx = sx;
        b.sy = sy;
        b.sz = sz;
        b.radius = std::sqrt(b.h * b.h + b.h * b.h + b.h * b.h) * 2 * 1.5;
    }

    void emit(int id, BhTreeHost& T) {
        const int me = (int)T.nodes.size();
        T.nodes.emplace_back();
        const BNode& b = B[id];
        BhNode nd{};
        nd.ex = b.ex;
        nd.ey = b.ey;
        nd.ez = b.ez;
        nd.width = 2 * b.h;
        nd.sx = b.sx;
        nd.sy = b.sy;
        nd.sz = b.sz;
        nd.radius = b.radius;
        if (leaf(id)) {
            nd.first = (int32_t)T.order.size();
            nd.count = (int32_t)b.members.size();
            for (int i : b.members) T.order.push_back(i);
        } else {
            nd.first = 0;
            nd.count = -1;
            for (int c : b.child)
                if (c >= 0) emit(c, T);
        }
        nd.skip = (int32_t)T.nodes.size();
        T.nodes[me] = nd;
    }
};

}  // namespace

void build_bh_tree(int n, const double* xyz, const double* alpha, const double* radius, int max_members,
                   BhTreeHost& T) {
    T = BhTreeHost();
    if (n <= 0) return;
    // octree::setDimensions (octree.h:70-80)
    double mn[3] = {xyz[0], xyz[1], xyz[2]}, mx[3] = {xyz[0], xyz[1], xyz[2]};
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double v = xyz[3 * i + d];
            if (v < mn[d]) mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
    double c[3], maxD = 0;
    for (int d = 0; d < 3; d++) {
        c[d] = 0.5 * (mn[d] + mx[d]);
        double e = mx[d] - mn[d];
        maxD = d == 0 ? e : (e > maxD ? e : maxD);
    }
    Builder b;
    b.xyz = xyz;
    b.alpha = alpha;
    b.max_members = max_members;
    b.B.reserve((size_t)n / 2 + 16);
    int root = b.make(c, 0.51 * maxD, 1);  // root level 1 (octree.h:181)
    std::vector<int> all(n);
    for (int i = 0; i < n; i++) all[i] = i;
    b.split(root, all);
    b.expansion(root, true);
    b.emit(root, T);
    T.levels = b.max_level;
    T.origin[0] = c[0];
    T.origin[1] = c[1];
    T.origin[2] = c[2];
    T.half = 0.51 * maxD;
    T.slot.assign(n, 0);
    T.mpos.resize(3 * (size_t)n);
    T.mstr.resize(3 * (size_t)n);
    T.mrad.resize(n);
    for (int s = 0; s < n; s++) {
        int i = T.order[s];
        T.slot[i] = s;
        for (int d = 0; d < 3; d++) {
            T.mpos[3 * (size_t)s + d] = xyz[3 * (size_t)i + d];
            T.mstr[3 * (size_t)s + d] = alpha[3 * (size_t)i + d];
        }
        T.mrad[s] = radius[i];
    }
}

}  // namespace cpgpu
