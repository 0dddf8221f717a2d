// Product topology builder: .tri connectivity -> resolved panel tables.
//
// Restates the order-dependent rules of the reference so that panel order,
// folding table and clusters are identical:
//   * wake node welding to body nodes within the shortest edge (L-inf norm),
//     first body node wins (geometry.cpp:631-667);
//   * edge::checkTE runs after every panel-to-edge attachment (edge.cpp:48-109);
//   * a wake line is created on every TE attachment (wakePanel.cpp:120-152);
//   * body panels are collected with surfaces prepended (geometry.cpp:452-458);
//   * neighbours are linked in edge creation order (geometry.cpp:418-421);
//   * tip flags then clusters in body order (geometry.cpp:473-484).
// Implementation is array based: edges through a hash map, wake welding
// through a uniform grid instead of the all-pairs scan.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <fstream>
#include <functional>
#include <limits>
#include <thread>
#include <unordered_map>

#include "geometry.h"

// Host precompute follows the reference operation order without FMA
// contraction: CHTLS fits on near-flat clusters are ill conditioned and
// amplify last-bit differences.
#pragma clang fp contract(off)

namespace cpgpu {

namespace {

const double kPi = 3.14159265358979323846;

inline Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 scale(double s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline Vec3 unit(Vec3 a) {
    double s = dot(a, a);
    if (s > 0) {
        double r = std::sqrt(s);
        return {a.x / r, a.y / r, a.z / r};
    }
    return a;
}

// panel::setGeom (panel.cpp:33-104)
void setup_panel(PanelGeom& p) {
    p.long_side = 0;
    for (int i = 0; i < p.nv; i++) {
        double l = len(sub(p.v[(i + 1) % p.nv], p.v[i]));
        if (l > p.long_side) p.long_side = l;
    }
    if (p.nv == 3) {
        Vec3 a = sub(p.v[1], p.v[0]), b = sub(p.v[2], p.v[0]);
        Vec3 s = add(add(p.v[0], p.v[1]), p.v[2]);
        p.center = {s.x / 3.0, s.y / 3.0, s.z / 3.0};
        double th = std::acos(dot(a, b) / (len(a) * len(b)));
        p.area = 0.5 * len(a) * len(b) * std::sin(th);
        p.normal = unit(cross(a, b));
    } else {
        Vec3 a = sub(p.v[1], p.v[0]), b = sub(p.v[2], p.v[1]), c = sub(p.v[3], p.v[2]);
        Vec3 pp = add(b, c), q = add(a, b);
        Vec3 m1 = add(p.v[1], scale(0.5, pp)), m2 = add(p.v[0], scale(0.5, q));
        p.center = scale(0.5, add(m1, m2));
        p.area = 0.5 * len(cross(pp, q));
        p.normal = unit(cross(a, b));
    }
    p.bez = p.normal;
}

struct EdgeRec {
    int n1, n2;
    int nbody = 0, body[2] = {-1, -1};
    int nwake = 0, wake0 = -1;
    bool te = false;
};

struct Builder {
    const int32_t* ids;
    std::vector<Vec3> nodes;
    std::vector<int> tri;
    std::vector<PanelGeom> pan;     // file order
    std::vector<char> is_wake;
    std::vector<int> group;         // surface / wake group index
    std::vector<std::array<int, 3>> pedge;
    std::vector<EdgeRec> edges;
    // edge index by node pair: open addressing over a power-of-two table
    // (key = lo << 32 | hi + 1, 0 = empty), edges numbered by first use
    std::vector<uint64_t> ekey;
    std::vector<int> eval;
    int eshift = 60;  // 64 - log2(table size): Fibonacci hashing takes the top bits
    void edges_reserve(size_t n) {
        size_t cap = 16;
        eshift = 60;
        while (cap < 2 * n) {
            cap *= 2;
            eshift--;
        }
        ekey.assign(cap, 0);
        eval.assign(cap, -1);
    }
    size_t eslot(uint64_t key) const { return (size_t)((key * 0x9E3779B97F4A7C15ull) >> eshift); }
    // per panel state
    std::vector<char> upper, lower, te_panel;
    std::vector<int> up_pan, lo_pan;
    std::vector<char> wte;
    std::vector<std::vector<int>> nbr;
    // groups
    std::vector<char> surf_lifting;
    std::vector<int> surf_of_group;
    struct WakeBuild {
        std::vector<int> panels;
        std::vector<WakeLineInfo> lines;  // upper/lower in FILE panel index here
        double yMin, yMax, x0, xf, z0, zf;
    };
    std::vector<WakeBuild> wk;

    int edge_of(int a, int b) {
        uint64_t lo = (uint64_t)std::min(a, b), hi = (uint64_t)std::max(a, b);
        const uint64_t key = ((lo << 32) | hi) + 1;
        if (2 * (edges.size() + 1) > ekey.size()) {  // keep the load under one half
            std::vector<uint64_t> ok = std::move(ekey);
            std::vector<int> ov = std::move(eval);
            edges_reserve(ok.size());
            for (size_t q = 0; q < ok.size(); q++)
                if (ok[q]) {
                    size_t h = eslot(ok[q]);
                    while (ekey[h]) h = (h + 1) & (ekey.size() - 1);
                    ekey[h] = ok[q];
                    eval[h] = ov[q];
                }
        }
        size_t h = eslot(key);
        while (ekey[h]) {
            if (ekey[h] == key) return eval[h];
            h = (h + 1) & (ekey.size() - 1);
        }
        EdgeRec e;
        e.n1 = a;
        e.n2 = b;
        edges.push_back(e);
        ekey[h] = key;
        eval[h] = (int)edges.size() - 1;
        return (int)edges.size() - 1;
    }

    void add_line(WakeBuild& w, const WakeLineInfo& wl) {
        w.lines.push_back(wl);
        std::sort(w.lines.begin(), w.lines.end(),
                  [](const WakeLineInfo& a, const WakeLineInfo& b) { return a.y < b.y; });
    }

    WakeLineInfo line_for(int up, int lo) {
        WakeLineInfo wl;
        wl.upper = up;
        wl.lower = lo;
        int te = -1;
        for (int k = 0; k < 3; k++)
            if (edges[pedge[up][k]].te) te = pedge[up][k];
        wl.p1 = nodes[edges[te].n1];
        wl.p2 = nodes[edges[te].n2];
        wl.y = (wl.p1.y + wl.p2.y) / 2.0;
        return wl;
    }

    void te_check(int ei) {
        EdgeRec& e = edges[ei];
        if (e.nbody != 2) return;
        if (e.nwake == 1) {
            e.te = true;
            const PanelGeom& w = pan[e.wake0];
            Vec3 v1 = sub(pan[e.body[0]].center, w.center), v2 = sub(pan[e.body[1]].center, w.center);
            double t1 = std::acos(dot(v1, w.normal) / (len(v1) * len(w.normal)));
            double t2 = std::acos(dot(v2, w.normal) / (len(v2) * len(w.normal)));
            int up = t1 < t2 ? e.body[0] : e.body[1];
            int lo = t1 < t2 ? e.body[1] : e.body[0];
            int wi = e.wake0;
            upper[up] = 1;
            lower[lo] = 1;
            wte[wi] = 1;
            if (up_pan[wi] < 0 || pan[up].center.x < w.center.x) {
                up_pan[wi] = up;
                lo_pan[wi] = lo;
            }
            add_line(wk[group[wi]], line_for(up_pan[wi], lo_pan[wi]));
        } else {
            const PanelGeom& a = pan[e.body[0]];
            const PanelGeom& b = pan[e.body[1]];
            double ang = std::acos(dot(a.normal, b.normal));
            if (ang > 4.7 * kPi / 6 && ids[e.body[0]] == ids[e.body[1]]) {
                e.te = true;
                for (int k = 0; k < 2; k++) {
                    te_panel[e.body[k]] = 1;
                    surf_lifting[group[e.body[k]]] = 1;
                }
            }
        }
    }
};

int fact(int i) {
    int r = 1;
    for (int k = 2; k <= i; k++) r *= k;
    return r;
}

}  // namespace

void parallel_for(int n, int nthreads, const std::function<void(int, int)>& body) {
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 256 || nthreads == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        int a = (int)((int64_t)n * t / nthreads), b = (int)((int64_t)n * (t + 1) / nthreads);
        th.emplace_back([&body, a, b] { body(a, b); });
    }
    for (auto& x : th) x.join();
}

bool read_tri_file(const std::string& path, std::vector<double>& xyz, std::vector<int32_t>& tri,
                   std::vector<int32_t>& ids, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    long nn, nt;
    if (!(f >> nn >> nt) || nn <= 0 || nt <= 0) {
        err = "bad .tri header";
        return false;
    }
    xyz.resize(3 * nn);
    for (auto& v : xyz)
        if (!(f >> v)) {
            err = "truncated node list";
            return false;
        }
    tri.resize(3 * nt);
    for (auto& v : tri) {
        long x;
        if (!(f >> x)) {
            err = "truncated connectivity";
            return false;
        }
        v = (int32_t)(x - 1);  // 1-based in the file (geometry.cpp:213)
    }
    ids.resize(nt);
    for (auto& v : ids)
        if (!(f >> v)) {
            err = "truncated surface IDs";
            return false;
        }
    return true;
}

namespace {

// The two buffer rows of the vortex-particle mode (geometry.cpp:267-397,
// 1276-1342), appended after the body panels: first-row quads become wake
// panels of one new wake (their attachment to the trailing-edge edge makes
// the wake lines and upper/lower flags through te_check, as edge::addWakePan
// does), second-row quads go to W with the topology the step loop needs.
bool build_vp_rows(Builder& B, const VpParams& vp, VpBuffer& W, std::vector<int>& vp_wake, std::string& err) {
    std::vector<int> te;
    for (int e = 0; e < (int)B.edges.size(); e++)
        if (B.edges[e].te && B.edges[e].nbody == 2) te.push_back(e);
    if (te.size() < 2) {
        err = "vortex-particle mode: fewer than two trailing-edge edges";
        return false;
    }
    double dt = vp.dt;
    if (!(dt > 0)) {  // geometry::calcTimeStep
        double sum = 0;
        for (int e : te) sum += len(sub(B.nodes[B.edges[e].n2], B.nodes[B.edges[e].n1]));
        double avg = sum / (double)te.size();
        dt = avg / vp.V;
    }
    W = VpBuffer();
    W.dt = dt;
    const int nn0 = (int)B.nodes.size();
    // a node's trailing edges in edge creation order (cpNode::getTrailingEdges)
    std::vector<std::vector<int>> node_te(nn0);
    for (int e : te) {
        node_te[B.edges[e].n1].push_back(e);
        node_te[B.edges[e].n2].push_back(e);
    }
    // cpNode::nodeWakeProjAngle: normals of the trailing edges' body panels,
    // the running sum divided by 3 after each edge, normalised
    auto bisector = [&](int n) {
        Vec3 avg{0, 0, 0};
        for (int e : node_te[n]) {
            const EdgeRec& E = B.edges[e];
            for (int k = 0; k < E.nbody; k++) avg = add(avg, B.pan[E.body[k]].normal);
            avg = {avg.x / 3.0, avg.y / 3.0, avg.z / 3.0};
        }
        return unit(avg);
    };
    const double cw = 1.3;  // cpNode::c_w
    std::vector<int> f1(nn0, -1), f2(nn0, -1);
    auto project = [&](int n) {  // firstProjNode / secProjNode, first use only
        if (f1[n] >= 0) return;
        const Vec3 b = bisector(n), p = B.nodes[n];
        double s1 = cw * dt;
        s1 = s1 * vp.V;
        double s2 = (cw + 1) * dt;
        s2 = s2 * vp.V;
        f1[n] = (int)B.nodes.size();
        B.nodes.push_back({p.x + s1 * b.x, p.y + s1 * b.y, p.z + s1 * b.z});
        f2[n] = (int)B.nodes.size();
        B.nodes.push_back({p.x + s2 * b.x, p.y + s2 * b.y, p.z + s2 * b.z});
    };
    Builder::WakeBuild wk;
    B.wk.push_back(wk);
    const int grp = (int)B.wk.size() - 1;
    // wake::addPanel bounds over both rows, in creation order
    auto bounds = [&](const PanelGeom& p) {
        Builder::WakeBuild& w = B.wk[grp];
        if (w.panels.empty() && W.panels.empty() && vp_wake.empty()) {
            Vec3 q = p.v[0];
            w.yMin = w.yMax = q.y;
            w.x0 = w.xf = q.x;
            w.z0 = w.zf = q.z;
        }
        for (int k = 0; k < p.nv; k++) {
            Vec3 q = p.v[k];
            if (q.y > w.yMax) w.yMax = q.y;
            else if (q.y < w.yMin) w.yMin = q.y;
            if (q.x > w.xf) w.xf = q.x;
            else if (q.x < w.x0) w.x0 = q.x;
            if (q.z > w.zf) w.zf = q.z;
            else if (q.z < w.z0) w.z0 = q.z;
        }
    };
    auto quad = [&](std::array<int, 4> q) {
        PanelGeom p;
        p.nv = 4;
        for (int k = 0; k < 4; k++) {
            p.node[k] = q[k];
            p.v[k] = B.nodes[q[k]];
        }
        setup_panel(p);
        return p;
    };
    std::vector<std::vector<int>> edge_w2;  // second-row panels on each edge
    std::vector<std::array<int, 4>> w2_edges;
    for (int e : te) {
        const int n1 = B.edges[e].n1, n2 = B.edges[e].n2;
        project(n1);
        project(n2);
        const int a1 = f1[n1], a2 = f1[n2], b1 = f2[n1], b2 = f2[n2];
        // first row (n1, n2, n2', n1')
        PanelGeom p1 = quad({n1, n2, a2, a1});
        bounds(p1);
        const int pi = (int)B.pan.size();
        B.pan.push_back(p1);
        B.is_wake.push_back(1);
        B.group.push_back(grp);
        B.pedge.push_back({-1, -1, -1});
        B.upper.push_back(0);
        B.lower.push_back(0);
        B.te_panel.push_back(0);
        B.up_pan.push_back(-1);
        B.lo_pan.push_back(-1);
        B.wte.push_back(0);
        B.nbr.emplace_back();
        vp_wake.push_back(pi);
        B.wk[grp].panels.push_back(pi);
        for (int k = 0; k < 4; k++) {
            const int ei = B.edge_of(p1.node[k], p1.node[(k + 1) % 4]);
            EdgeRec& E = B.edges[ei];
            if (E.nwake == 0) E.wake0 = pi;
            E.nwake++;
            B.te_check(ei);
        }
        // second row (n1', n2', n2'', n1'')
        PanelGeom p2 = quad({a1, a2, b2, b1});
        bounds(p2);
        const int wi = (int)W.panels.size();
        W.panels.push_back(p2);
        W.parent.push_back((int)vp_wake.size() - 1);
        std::array<int, 4> pe;
        for (int k = 0; k < 4; k++) {
            pe[k] = B.edge_of(p2.node[k], p2.node[(k + 1) % 4]);
            if ((int)edge_w2.size() < (int)B.edges.size()) edge_w2.resize(B.edges.size());
            edge_w2[pe[k]].push_back(wi);
            B.edges[pe[k]].nwake++;
        }
        w2_edges.push_back(pe);
    }
    // second-row topology: wakePanel::edgesInOrder / pointsInOrder,
    // getOtherWakePan across the side edges, edge first use
    auto mid = [&](int ei) {
        const Vec3 a = B.nodes[B.edges[ei].n1], b = B.nodes[B.edges[ei].n2];
        const Vec3 v = sub(b, a);
        return add(a, scale(0.5, v));
    };
    auto other = [&](int ei, int self) {
        for (int o : edge_w2[ei])
            if (o != self) return o;
        return -1;
    };
    std::vector<char> used(B.edges.size(), 0);
    const int m = (int)W.panels.size();
    W.pio.resize(m);
    W.eio_nbr.assign(m, {-1, -1, -1, -1});
    W.eio_first.assign(m, {0, 0, 0, 0});
    W.rad_nbr.resize(m);
    for (int i = 0; i < m; i++) {
        const PanelGeom& p = W.panels[i];
        const std::array<int, 4>& pe = w2_edges[i];
        std::array<int, 4> eio{pe[0], -1, pe[2], -1};
        const Vec3 a = sub(mid(pe[0]), p.center);
        const Vec3 c = cross(a, p.normal);
        const Vec3 d = add(p.center, c);
        const double d1 = len(sub(mid(pe[1]), d)), d3 = len(sub(mid(pe[3]), d));
        if (d3 < d1) {
            eio[3] = pe[3];
            eio[1] = pe[1];
        } else {
            eio[3] = pe[1];
            eio[1] = pe[3];
        }
        auto has = [&](int ei, int n) { return B.edges[ei].n1 == n || B.edges[ei].n2 == n; };
        int q[4];
        {
            const int x1 = B.edges[eio[0]].n1, x2 = B.edges[eio[0]].n2;
            if (has(eio[3], x1)) {
                q[0] = x1;
                q[1] = x2;
            } else {
                q[0] = x2;
                q[1] = x1;
            }
            const int y1 = B.edges[eio[2]].n1, y2 = B.edges[eio[2]].n2;
            if (has(eio[1], y1)) {
                q[2] = y1;
                q[3] = y2;
            } else {
                q[2] = y2;
                q[3] = y1;
            }
        }
        for (int k = 0; k < 4; k++) W.pio[i][k] = B.nodes[q[k]];
        for (int j = 1; j < 4; j++) {
            if (j != 2) W.eio_nbr[i][j] = other(eio[j], i);
            if (!used[eio[j]]) {
                used[eio[j]] = 1;
                W.eio_first[i][j] = 1;
            }
        }
        W.rad_nbr[i] = {other(pe[1], i), other(pe[3], i)};
    }
    return true;
}

}  // namespace

bool build_geometry(int nn, const double* xyz, int nt, const int32_t* tri_in, const int32_t* ids, Geometry& G,
                    std::string& err, const VpParams* vp) {
    Builder B;
    B.ids = ids;
    B.nodes.resize(nn);
    for (int i = 0; i < nn; i++) B.nodes[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    B.tri.assign(tri_in, tri_in + 3 * (size_t)nt);
    for (int i = 0; i < 3 * nt; i++)
        if (B.tri[i] < 0 || B.tri[i] >= nn) {
            err = "connectivity references a missing node";
            return false;
        }

    // --- weld wake nodes onto body nodes (geometry.cpp:215-244, 631-698)
    int wakeTriStart = nt, wakeNodeStart = nn;
    bool hasWake = false;
    for (int i = 0; i < nt; i++) {
        if (ids[i] > 1000) hasWake = true;
        if (i > 0 && ids[i] > 1000 && ids[i - 1] < 1000) {
            wakeTriStart = i;
            wakeNodeStart = std::min(B.tri[3 * i], std::min(B.tri[3 * i + 1], B.tri[3 * i + 2]));
        }
    }
    if (hasWake && wakeNodeStart < nn) {
        double tol = 100000;
        for (int i = 0; i < nt; i++)
            for (int j = 0; j < 3; j++) {
                double l = len(sub(B.nodes[B.tri[3 * i + (j + 1) % 3]], B.nodes[B.tri[3 * i + j]]));
                tol = std::min(tol, l);
            }
        // uniform grid of cell size tol over the body nodes; the smallest body
        // index within L-inf distance < tol replaces the wake node
        auto cell = [&](double v) { return (int64_t)std::floor(v / tol); };
        auto key = [](int64_t a, int64_t b, int64_t c) {
            return ((uint64_t)(a & 0x1FFFFF) << 42) | ((uint64_t)(b & 0x1FFFFF) << 21) | (uint64_t)(c & 0x1FFFFF);
        };
        std::unordered_map<uint64_t, std::vector<int>> grid;
        for (int i = 0; i < wakeNodeStart; i++)
            grid[key(cell(B.nodes[i].x), cell(B.nodes[i].y), cell(B.nodes[i].z))].push_back(i);
        std::vector<int> repl(nn, -1);
        for (int j = wakeNodeStart; j < nn; j++) {
            Vec3 q = B.nodes[j];
            int best = -1;
            int64_t cx = cell(q.x), cy = cell(q.y), cz = cell(q.z);
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        auto it = grid.find(key(cx + dx, cy + dy, cz + dz));
                        if (it == grid.end()) continue;
                        for (int i : it->second) {
                            Vec3 d = sub(B.nodes[i], q);
                            double linf = std::max(std::fabs(d.x), std::max(std::fabs(d.y), std::fabs(d.z)));
                            if (linf < tol && (best < 0 || i < best)) best = i;
                        }
                    }
            repl[j] = best;
        }
        for (int i = wakeTriStart; i < nt; i++)
            for (int j = 0; j < 3; j++) {
                int r = repl[B.tri[3 * i + j]];
                if (r >= 0) B.tri[3 * i + j] = r;
            }
    }

    // --- panels and incremental TE detection (geometry.cpp:712-747)
    B.pan.resize(nt);
    B.is_wake.assign(nt, 0);
    B.group.assign(nt, -1);
    B.pedge.resize(nt);
    B.upper.assign(nt, 0);
    B.lower.assign(nt, 0);
    B.te_panel.assign(nt, 0);
    B.up_pan.assign(nt, -1);
    B.lo_pan.assign(nt, -1);
    B.wte.assign(nt, 0);
    B.nbr.resize(nt);
    // each panel's own geometry (no order dependence) on host threads; the
    // edges, groups and trailing-edge checks below follow the file order
    parallel_for(nt, 0, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            PanelGeom& p = B.pan[i];
            p.nv = 3;
            for (int k = 0; k < 3; k++) {
                p.node[k] = B.tri[3 * i + k];
                p.v[k] = B.nodes[p.node[k]];
            }
            setup_panel(p);
        }
    });
    B.edges_reserve(3 * (size_t)nt);
    for (int i = 0; i < nt; i++) {
        PanelGeom& p = B.pan[i];
        for (int k = 0; k < 3; k++) B.pedge[i][k] = B.edge_of(p.node[k], p.node[(k + 1) % 3]);
        bool newGroup = (i == 0 || ids[i] != ids[i - 1]);
        if (ids[i] <= 1000) {
            if (newGroup) B.surf_lifting.push_back(0);
            B.group[i] = (int)B.surf_lifting.size() - 1;
            for (int k = 0; k < 3; k++) {
                EdgeRec& e = B.edges[B.pedge[i][k]];
                if (e.nbody < 2) e.body[e.nbody] = i;
                e.nbody++;
                B.te_check(B.pedge[i][k]);
            }
        } else {
            B.is_wake[i] = 1;
            if (newGroup) {
                Builder::WakeBuild w;
                B.wk.push_back(w);
            }
            B.group[i] = (int)B.wk.size() - 1;
            for (int k = 0; k < 3; k++) {
                EdgeRec& e = B.edges[B.pedge[i][k]];
                if (e.nwake == 0) e.wake0 = i;
                e.nwake++;
                B.te_check(B.pedge[i][k]);
            }
            // wake::addPanel bounds (wake.cpp:42-84)
            Builder::WakeBuild& w = B.wk.back();
            if (w.panels.empty()) {
                Vec3 q = p.v[0];
                w.yMin = w.yMax = q.y;
                w.x0 = w.xf = q.x;
                w.z0 = w.zf = q.z;
            }
            for (int k = 0; k < 3; k++) {
                Vec3 q = p.v[k];
                if (q.y > w.yMax) w.yMax = q.y;
                else if (q.y < w.yMin) w.yMin = q.y;
                if (q.x > w.xf) w.xf = q.x;
                else if (q.x < w.x0) w.x0 = q.x;
                if (q.z > w.zf) w.zf = q.z;
                else if (q.z < w.z0) w.z0 = q.z;
            }
            w.panels.push_back(i);
        }
    }
    for (auto& e : B.edges)
        if (e.nbody > 2) {
            err = "non-manifold body edge (more than two body panels)";
            return false;
        }
    for (auto& e : B.edges)
        if (e.nbody == 2) {
            B.nbr[e.body[0]].push_back(e.body[1]);
            B.nbr[e.body[1]].push_back(e.body[0]);
        }
    std::vector<int> vp_wake;  // first-row panels of the vortex-particle mode
    if (vp) {
        if (!B.wk.empty()) {
            err = "vortex-particle mode needs a mesh without prescribed wake panels";
            return false;
        }
        if (!build_vp_rows(B, *vp, G.w2, vp_wake, err)) return false;
        G.vp = true;
    }
    const int np = (int)B.pan.size();  // file triangles plus first-row quads

    // --- wake merge (geometry.cpp:424-447): same x0, z0, xf, zf within 1e-2
    std::vector<int> wake_final(B.wk.size(), -1);
    std::vector<Builder::WakeBuild> wakes;
    {
        std::vector<char> alive(B.wk.size(), 1);
        std::vector<Builder::WakeBuild> wk = B.wk;
        for (size_t i = 0; i < wk.size(); i++) {
            if (!alive[i]) continue;
            for (size_t j = i + 1; j < wk.size(); j++) {
                if (!alive[j]) continue;
                const double eps = 1e-2;
                if (std::fabs(wk[j].x0 - wk[i].x0) < eps && std::fabs(wk[j].z0 - wk[i].z0) < eps &&
                    std::fabs(wk[j].xf - wk[i].xf) < eps && std::fabs(wk[j].zf - wk[i].zf) < eps) {
                    for (int pi : wk[j].panels) wk[i].panels.push_back(pi);
                    for (auto& wl : wk[j].lines) B.add_line(wk[i], wl);
                    wk[i].yMin = std::min(wk[i].yMin, wk[j].yMin);
                    wk[i].yMax = std::max(wk[i].yMax, wk[j].yMax);
                    alive[j] = 0;
                    wake_final[j] = (int)i;
                }
            }
        }
        std::vector<int> newIdx(wk.size(), -1);
        for (size_t i = 0; i < wk.size(); i++)
            if (alive[i]) {
                newIdx[i] = (int)wakes.size();
                wakes.push_back(wk[i]);
            }
        for (size_t i = 0; i < wk.size(); i++) wake_final[i] = alive[i] ? newIdx[i] : newIdx[wake_final[i]];
    }

    // --- panel order: surfaces prepended, wakes in order
    std::vector<std::vector<int>> surfPanels(B.surf_lifting.size());
    for (int i = 0; i < np; i++)
        if (!B.is_wake[i]) surfPanels[B.group[i]].push_back(i);
    std::vector<int> border;
    for (int s = (int)surfPanels.size() - 1; s >= 0; s--)
        border.insert(border.end(), surfPanels[s].begin(), surfPanels[s].end());
    std::vector<int> worder;
    for (auto& w : wakes) worder.insert(worder.end(), w.panels.begin(), w.panels.end());
    int nb = (int)border.size(), nw = (int)worder.size();
    if (nb == 0) {
        err = "mesh has no body panels";
        return false;
    }
    std::vector<int> bidx(np, -1), widx(np, -1);
    for (int i = 0; i < nb; i++) bidx[border[i]] = i;
    for (int j = 0; j < nw; j++) widx[worder[j]] = j;
    for (auto& w : wakes)
        if (w.lines.size() < 2) {
            err = "wake with fewer than two wake lines (no trailing edge found)";
            return false;
        }

    // --- tip flags (bodyPanel::setTipFlag, bodyPanel.cpp:58-95)
    std::vector<char> tip(np, 0);
    for (int fi : border) {
        if (!B.surf_lifting[B.group[fi]]) continue;
        int count = 0;
        Vec3 n = B.pan[fi].normal;
        for (int o : B.nbr[fi]) {
            Vec3 m = B.pan[o].normal;
            double d = dot(n, m) / (len(n) * len(m));
            d = std::min(1.0, std::max(-1.0, d));
            if (std::acos(d) < 1e-8 && std::asin(m.x) > -kPi / 12) count++;
        }
        tip[fi] = count >= 2;
    }

    // --- clusters (bodyPanel::setCluster, bodyPanel.cpp:402-487)
    G.cluster.assign(16 * (size_t)nb, -1);
    G.ts_order.assign(nb, 3);
    // each panel's cluster depends only on the finished topology and fills
    // only its own slots: built on host threads
    std::atomic<bool> too_big{false};
    parallel_for(nb, 0, [&](int bi0, int bi1) {
        for (int bi = bi0; bi < bi1; bi++) {
            int fi = border[bi];
            int dim = tip[fi] ? 2 : 3;
            int order = 3;
            size_t buffer = 10;
            size_t nObs = (size_t)(fact(order + dim) / (fact(dim) * fact(order)) - 1) + buffer;
            size_t nPan = (nObs + 1) / 2;
            std::vector<int> cl{fi};
            size_t first = 0;
            bool upF = B.upper[fi], loF = B.lower[fi];
            Vec3 n = B.pan[fi].normal;
            auto accept = [&](int o) {
                if ((upF && B.lower[o]) || (loF && B.upper[o])) return false;
                if (tip[fi] != tip[o]) return false;
                double d = dot(B.pan[o].normal, n);
                d = std::min(1.0, std::max(-1.0, d));
                return std::acos(d) < 5 * kPi / 6 && std::find(cl.begin(), cl.end(), o) == cl.end() && o != fi;
            };
            while (cl.size() < nPan) {
                std::vector<int> add;
                for (size_t i = first; i < cl.size(); i++)
                    for (int o : B.nbr[cl[i]])
                        if (accept(o)) {
                            if (B.upper[o]) upF = true;
                            else if (B.lower[o]) loF = true;
                            if (std::find(add.begin(), add.end(), o) == add.end()) add.push_back(o);
                        }
                if (add.empty()) {
                    while (cl.size() < nPan + 1 && order >= 1) {
                        if (buffer > 0) buffer--;
                        else order--;
                        nObs = (size_t)(fact(order + dim) / (fact(dim) * fact(order))) + buffer;
                        nPan = (nObs + 1) / 2;
                    }
                    if (nPan + 1 < cl.size()) cl.resize(nPan + 1);
                    break;
                }
                first = cl.size();
                for (int o : add) {
                    cl.push_back(o);
                    if (cl.size() == nPan + 1) break;
                }
            }
            if (cl.size() > 16) {
                too_big = true;
                continue;
            }
            for (size_t k = 0; k < cl.size(); k++) G.cluster[16 * (size_t)bi + k] = bidx[cl[k]];
            G.ts_order[bi] = order;
        }
    });
    if (too_big) {
        err = "cluster larger than 16 panels";
        return false;
    }

    // --- outputs
    G.n_nodes = (int)B.nodes.size();
    G.nodes = B.nodes;
    G.body.resize(nb);
    G.body_tri = border;
    G.tip.resize(nb);
    G.upper.resize(nb);
    G.lower.resize(nb);
    G.body_te_edges.assign(6 * (size_t)nb, -1);
    for (int i = 0; i < nb; i++) {
        int fi = border[i];
        G.body[i] = B.pan[fi];
        G.tip[i] = tip[fi];
        G.upper[i] = B.upper[fi];
        G.lower[i] = B.lower[fi];
        for (int k = 0; k < 3; k++) {
            const EdgeRec& e = B.edges[B.pedge[fi][k]];
            if (e.te) {
                G.body_te_edges[6 * i + 2 * k] = e.n1;
                G.body_te_edges[6 * i + 2 * k + 1] = e.n2;
            }
        }
    }
    G.wake.resize(nw);
    G.wake_tri = worder;
    for (int& t : G.wake_tri)
        if (t >= nt) t = -1;  // first-row quads of the vortex-particle mode
    G.wake_te.resize(nw);
    G.wake_upper.resize(nw);
    for (int j = 0; j < nw; j++) {
        int fi = worder[j];
        G.wake[j] = B.pan[fi];
        G.wake_te[j] = B.wte[fi];
        G.wake_upper[j] = B.up_pan[fi] >= 0 ? bidx[B.up_pan[fi]] : -1;
    }
    G.wakes.resize(wakes.size());
    for (size_t w = 0; w < wakes.size(); w++) {
        WakeInfo& o = G.wakes[w];
        o.yMin = wakes[w].yMin;
        o.yMax = wakes[w].yMax;
        o.x0 = wakes[w].x0;
        o.xf = wakes[w].xf;
        o.z0 = wakes[w].z0;
        o.zf = wakes[w].zf;
        for (int pi : wakes[w].panels) o.panels.push_back(widx[pi]);
        o.lines = wakes[w].lines;
        for (auto& wl : o.lines) {
            wl.upper = bidx[wl.upper];
            wl.lower = bidx[wl.lower];
        }
    }
    // Kutta folding table (wakePanel::interpPanels, wakePanel.cpp:48-81)
    G.fold_idx.assign(4 * (size_t)nw, 0);
    G.fold_coef.assign(nw, 0);
    for (int j = 0; j < nw; j++) {
        int fi = worder[j];
        const std::vector<WakeLineInfo>& L = G.wakes[wake_final[B.group[fi]]].lines;
        double cy = B.pan[fi].center.y;
        size_t k1 = 0;
        if (cy < L[1].y) k1 = 0;
        else if (cy >= L.back().y) k1 = L.size() - 2;
        else
            for (size_t i = 1; i + 1 < L.size(); i++)
                if (L[i].y <= cy && L[i + 1].y > cy) k1 = i;
        const WakeLineInfo &a = L[k1], &b = L[k1 + 1];
        G.fold_coef[j] = (cy - a.y) / (b.y - a.y);
        G.fold_idx[4 * j + 0] = a.upper;
        G.fold_idx[4 * j + 1] = a.lower;
        G.fold_idx[4 * j + 2] = b.upper;
        G.fold_idx[4 * j + 3] = b.lower;
    }
    return true;
}

void panel_setup(PanelGeom& p) { setup_panel(p); }

}  // namespace cpgpu
