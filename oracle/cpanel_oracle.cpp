// FP64 CPU oracle for the CPanel steady solve path.
//
// TEST INFRASTRUCTURE ONLY: a restatement of the reference algorithms used to
// check the HIP product path.  The product never links or calls this code.
// Reference citations are file:line in jakeddavis/CPanel (CPanel/ directory).
//
// Arithmetic is FP64 throughout and follows the reference's operation order
// where that order is visible in the source (Eigen 3-vector dot/norm are
// ((x*x+y*y)+z*z)).  The only intentional difference is the BiCGSTAB stopping
// rule: the reference runs Eigen's default (tol = machine eps, 2N iterations);
// the oracle accepts a tolerance so that large cases finish (the converged
// solutions agree to cond(A)*tol).
#include "cpanel_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI = 3.14159265358979323846;

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Eigen normalize(): divide by sqrt(squaredNorm) when it is positive
inline V3 normalized(V3 a) {
    double s = dot(a, a);
    if (s > 0) {
        double r = std::sqrt(s);
        return {a.x / r, a.y / r, a.z / r};
    }
    return a;
}
inline bool isnan3(V3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }

// ---------------------------------------------------------------------------
// Panel geometry and pair kernels (panel.cpp, bodyPanel.cpp)
// ---------------------------------------------------------------------------
struct PanelGeo {
    int nv;
    V3 p[4];
    V3 center, normal;
    double area, longSide;
};

// panel::setGeom (panel.cpp:33-104).  Edge lengths come from the edges, which
// have the same length as the node differences used here.
void set_geom(PanelGeo& g) {
    g.longSide = 0;
    for (int i = 0; i < g.nv; i++) {
        double l = norm(g.p[(i + 1) % g.nv] - g.p[i]);
        if (l > g.longSide) g.longSide = l;
    }
    if (g.nv == 3) {
        V3 a = g.p[1] - g.p[0], b = g.p[2] - g.p[0];
        g.center = (g.p[0] + g.p[1] + g.p[2]) / 3.0;
        double theta = std::acos(dot(a, b) / (norm(a) * norm(b)));
        g.area = 0.5 * norm(a) * norm(b) * std::sin(theta);
        g.normal = normalized(cross(a, b));
    } else {
        V3 a = g.p[1] - g.p[0], b = g.p[2] - g.p[1], c = g.p[3] - g.p[2];
        V3 pp = b + c, q = a + b;
        V3 m1 = g.p[1] + 0.5 * pp, m2 = g.p[0] + 0.5 * q;
        g.center = 0.5 * (m1 + m2);
        g.area = 0.5 * norm(cross(pp, q));
        g.normal = normalized(cross(a, b));
    }
}

// panel::getLocalSys (panel.cpp:156-168)
inline void local_sys(const PanelGeo& g, V3& l, V3& m, V3& n) {
    l = normalized(g.p[0] - g.center);
    m = cross(g.normal, l);
    n = g.normal;
}

// panel::vortexPhi (panel.cpp:713-735)
inline double vortex_phi(double PN, double Al, V3 a, V3 b, V3 s, V3 l, V3 m) {
    const double eps = 1e-15;
    double PA = dot(a, cross(l, cross(a, s)));
    double sm = dot(s, m);
    double PB = PA - Al * sm;
    double na = norm(a), nb = norm(b);
    double num = sm * PN * (nb * PA - na * PB);
    double denom = PA * PB + PN * PN * na * nb * (sm * sm);
    if (denom == 0 && std::fabs(PN) < eps) return PN >= 0 ? 0.5 * PI : -0.5 * PI;
    return std::atan2(num, denom);
}

// bodyPanel::srcSidePhi (bodyPanel.cpp:361-373)
inline double src_side_phi(double PN, double Al, double phiV, V3 a, V3 b, V3 s) {
    double A = norm(a), B = norm(b), S = norm(s);
    double GL = 0;
    if (std::fabs(A + B - S) > 0 && S > 0) GL = 1 / S * std::log(std::fabs((A + B + S) / (A + B - S)));
    return Al * GL - PN * phiV;
}

// bodyPanel::panelPhiInf (bodyPanel.cpp:132-190).  phiSrc/phiDub accumulate
// onto their input values as in the reference (they start at 0 in setInfCoeff).
void panel_phi_inf(const PanelGeo& g, V3 poi, double& phiSrc, double& phiDub) {
    V3 pjk = poi - g.center;
    bool itself = norm(pjk) < 1e-10;
    V3 l, m, n;
    local_sys(g, l, m, n);
    double PN = dot(pjk, n);
    if (norm(pjk) / g.longSide > 5) {
        phiSrc = g.area / (4 * PI * norm(pjk));             // pntSrcPhi :392-395
        phiDub = PN * g.area / (4 * PI * std::pow(norm(pjk), 3));  // pntDubPhi panel.cpp:737-740
    } else {
        double phiV = 0;
        for (int i = 0; i < g.nv; i++) {
            V3 p1 = g.p[i], p2 = g.p[(i + 1) % g.nv];
            V3 a = poi - p1, b = poi - p2, s = p2 - p1;
            double Al = dot(n, cross(s, a));
            if (!itself) {
                phiV = vortex_phi(PN, Al, a, b, s, l, m);
                phiDub += phiV;
            }
            phiSrc += src_side_phi(PN, Al, phiV, a, b, s);
        }
        phiSrc /= (4 * PI);
        if (!itself)
            phiDub /= (4 * PI);
        else
            phiDub = -0.5;
    }
}

// panel::dubPhiInf (panel.cpp:208-252)
double dub_phi_inf(const PanelGeo& g, V3 poi) {
    V3 pjk = poi - g.center;
    V3 l, m, n;
    local_sys(g, l, m, n);
    double PN = dot(pjk, n);
    if (norm(pjk) < 0.0000001) return -0.5;
    if (norm(pjk) / g.longSide > 5) return PN * g.area / (4 * PI * std::pow(norm(pjk), 3));
    double phi = 0;
    for (int i = 0; i < g.nv; i++) {
        V3 p1 = g.p[i], p2 = g.p[(i + 1) % g.nv];
        V3 a = poi - p1, b = poi - p2, s = p2 - p1;
        double Al = dot(n, cross(s, a));
        phi += vortex_phi(PN, Al, a, b, s, l, m);
    }
    return phi / (4 * PI);
}

const double CORE = 0.05;  // panel::core (panel.h:67)

// panel::vortexV (panel.cpp:707-711)
inline V3 vortex_v(V3 a, V3 b, V3 s) {
    double na = norm(a), nb = norm(b);
    double cs = CORE * norm(s);
    return (na + nb) * cross(a, b) / (na * nb * (na * nb + dot(a, b)) + cs * cs);
}

// panel::pntDubV (panel.cpp:773-776)
inline V3 pnt_dub_v(const PanelGeo& g, V3 n, V3 pjk) {
    double r = norm(pjk);
    return g.area * (3 * dot(pjk, n) * pjk - (r * r) * n) / (4 * PI * std::pow(r, 5));
}

// panel::dubVInf (panel.cpp:666-705)
V3 dub_v_inf(const PanelGeo& g, V3 poi) {
    V3 pjk = poi - g.center;
    V3 l, m, n;
    local_sys(g, l, m, n);
    if (norm(pjk) / g.longSide > 5) return pnt_dub_v(g, n, pjk);
    V3 vel = {0, 0, 0};
    for (int i = 0; i < g.nv; i++) {
        V3 p1 = g.p[i], p2 = g.p[(i + 1) % g.nv];
        vel = vel + vortex_v(poi - p1, poi - p2, p2 - p1);
    }
    return vel / (4 * PI);
}

// bodyPanel::srcSideV (bodyPanel.cpp:375-390)
inline V3 src_side_v(double PN, double Al, V3 a, V3 b, V3 s, V3 l, V3 m, V3 n) {
    double A = norm(a), B = norm(b), S = norm(s);
    double GL = 0;
    if (std::fabs(A + B - S) > 0 && S > 0) GL = 1 / S * std::log(std::fabs((A + B + S) / (A + B - S)));
    double CJK = vortex_phi(PN, Al, a, b, s, l, m);
    return GL * (dot(s, m) * l - dot(s, l) * m) + CJK * n;
}

// bodyPanel::panelVInf (bodyPanel.cpp:305-348).  Far field: the point source
// velocity is negated (bodyPanel.cpp:314) while the near field is not.
void panel_v_inf(const PanelGeo& g, V3 poi, V3& vSrc, V3& vDub) {
    V3 pjk = poi - g.center;
    V3 l, m, n;
    local_sys(g, l, m, n);
    double PN = dot(pjk, n);
    if (norm(pjk) / g.longSide > 5) {
        double r = norm(pjk);
        vSrc = -1.0 * (g.area * pjk / (4 * PI * std::pow(r, 3)));
        vDub = pnt_dub_v(g, n, pjk);
        return;
    }
    vSrc = {0, 0, 0};
    vDub = {0, 0, 0};
    for (int i = 0; i < g.nv; i++) {
        V3 p1 = g.p[i], p2 = g.p[(i + 1) % g.nv];
        V3 a = poi - p1, b = poi - p2, s = a - b;
        double Al = dot(n, cross(s, a));
        vDub = vDub + vortex_v(a, b, s);
        vSrc = vSrc + src_side_v(PN, Al, a, b, s, l, m, n);
    }
    vDub = vDub / (4 * PI);
    vSrc = vSrc / (4 * PI);
}

PanelGeo make_panel(int nv, const double* v) {
    PanelGeo g;
    g.nv = nv;
    for (int i = 0; i < nv; i++) g.p[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    set_geom(g);
    return g;
}

// ---------------------------------------------------------------------------
// CHTLS least-squares derivative weights (chtlsnd.cpp:20-321), unconstrained
// branch only (every caller on this path passes V0 = 0).
// ---------------------------------------------------------------------------
int factorial(int i) {  // chtlsnd.cpp:306-321
    if (i == 0) return 1;
    int out = i, mult = i - 1;
    while (mult != 0) {
        out *= mult;
        mult--;
    }
    return out;
}

typedef std::vector<std::vector<int>> IMat;

// chtlsnd::derivSequence (chtlsnd.cpp:212-249)
IMat deriv_sequence(int q, int N) {
    IMat d;
    if (q == 0) {
        d.push_back(std::vector<int>(N, 0));
        return d;
    }
    for (int i = 0; i <= q; i++) {
        if (N > 1) {
            IMat nxt = deriv_sequence(q - i, N - 1);
            for (auto& r : nxt) {
                std::vector<int> t = r;
                t.resize(N);
                t[N - 1] = i;
                d.push_back(t);
            }
        } else {
            d.push_back(std::vector<int>(1, i));
        }
    }
    return d;
}

// chtlsnd::sortBySum (chtlsnd.cpp:251-275): stable insertion before the first
// row with a strictly larger sum.
IMat sort_by_sum(const IMat& m) {
    auto sum = [](const std::vector<int>& r) {
        int s = 0;
        for (int v : r) s += v;
        return s;
    };
    IMat out;
    out.push_back(m[0]);
    for (size_t i = 1; i < m.size(); i++) {
        bool flag = false;
        for (size_t j = 0; j < out.size(); j++) {
            if (sum(m[i]) < sum(out[j])) {
                out.insert(out.begin() + j, m[i]);
                flag = true;
                break;
            }
        }
        if (!flag) out.push_back(m[i]);
    }
    return out;
}

// Householder QR solve M X = R (M square t x t, R t x ncol, column-major),
// following Eigen's unblocked householder_qr_inplace and
// HouseholderQR::solve: Q^T applied reflector by reflector, then upper
// triangular back substitution.  No rank check, so an exactly singular M
// produces inf/NaN as Eigen does.
void householder_solve(int t, std::vector<double>& M, int ncol, std::vector<double>& R) {
    auto Mi = [&](int i, int j) -> double& { return M[i + (size_t)j * t]; };
    auto Ri = [&](int i, int j) -> double& { return R[i + (size_t)j * t]; };
    std::vector<double> hc(t), tmp(std::max(t, ncol));
    const double tol = std::numeric_limits<double>::min();
    for (int k = 0; k < t; k++) {
        int rem = t - k;
        // makeHouseholderInPlace on column k, rows k..t-1
        double tailSq = 0;
        for (int i = k + 1; i < t; i++) tailSq += Mi(i, k) * Mi(i, k);
        double c0 = Mi(k, k), tau, beta;
        if (rem == 1) tailSq = 0;
        if (tailSq <= tol) {
            tau = 0;
            beta = c0;
            for (int i = k + 1; i < t; i++) Mi(i, k) = 0;
        } else {
            beta = std::sqrt(c0 * c0 + tailSq);
            if (c0 >= 0) beta = -beta;
            for (int i = k + 1; i < t; i++) Mi(i, k) = Mi(i, k) / (c0 - beta);
            tau = (beta - c0) / beta;
        }
        hc[k] = tau;
        Mi(k, k) = beta;
        // applyHouseholderOnTheLeft to the trailing columns
        int cols = t - k - 1;
        if (cols > 0) {
            if (rem == 1) {
                for (int j = k + 1; j < t; j++) Mi(k, j) *= (1 - tau);
            } else if (tau != 0) {
                for (int j = k + 1; j < t; j++) {
                    double s = 0;
                    for (int i = k + 1; i < t; i++) s += Mi(i, k) * Mi(i, j);
                    s += Mi(k, j);
                    Mi(k, j) -= tau * s;
                    for (int i = k + 1; i < t; i++) Mi(i, j) -= tau * Mi(i, k) * s;
                }
            }
        }
    }
    // apply Q^T = H_{t-1} ... H_0 to R
    for (int k = 0; k < t; k++) {
        int rem = t - k;
        double tau = hc[k];
        if (rem == 1) {
            for (int j = 0; j < ncol; j++) Ri(k, j) *= (1 - tau);
        } else if (tau != 0) {
            for (int j = 0; j < ncol; j++) {
                double s = 0;
                for (int i = k + 1; i < t; i++) s += Mi(i, k) * Ri(i, j);
                s += Ri(k, j);
                Ri(k, j) -= tau * s;
                for (int i = k + 1; i < t; i++) Ri(i, j) -= tau * Mi(i, k) * s;
            }
        }
    }
    // back substitution with the upper triangle
    for (int j = 0; j < ncol; j++) {
        for (int i = t - 1; i >= 0; i--) {
            double xi = Ri(i, j) / Mi(i, i);
            Ri(i, j) = xi;
            for (int r = 0; r < i; r++) Ri(r, j) -= xi * Mi(r, i);
        }
    }
}

// chtlsnd::chtlsnd (chtlsnd.cpp:20-209), V0 == 0 branch.  Xf nf x N and
// Xb nb x N row-major; returns F (t x nf, column-major in a flat vector) and
// the derivative order sums per row.
void chtls(int N, const double* X0, int nf, const double* Xf, int order, int nb,
           const double* Xb, const double* Vb, std::vector<double>& F, int& tout) {
    std::vector<double> dXf(nf * N), dXb(nb * N);
    double rmax = 0;
    for (int i = 0; i < nf; i++) {
        double s = 0;
        for (int k = 0; k < N; k++) {
            dXf[i * N + k] = Xf[i * N + k] - X0[k];
            s += dXf[i * N + k] * dXf[i * N + k];
        }
        double r = std::sqrt(s);
        if (r > rmax) rmax = r;
    }
    for (auto& v : dXf) v /= rmax;
    for (int i = 0; i < nb; i++)
        for (int k = 0; k < N; k++) dXb[i * N + k] = Xb[i * N + k] - X0[k];
    for (auto& v : dXb) v /= rmax;
    std::vector<double> Wf(nf), Wb(nb);
    for (int i = 0; i < nf; i++) {
        double s = 0;
        for (int k = 0; k < N; k++) s += dXf[i * N + k] * dXf[i * N + k];
        Wf[i] = 1 / std::sqrt(s);
    }
    for (int i = 0; i < nb; i++) {
        double s = 0;
        for (int k = 0; k < N; k++) s += dXb[i * N + k] * dXb[i * N + k];
        Wb[i] = 1 / std::sqrt(s);
    }
    IMat ms = sort_by_sum(deriv_sequence(order, N));
    ms.erase(ms.begin());
    int t = (int)ms.size();
    std::vector<int> sums(t);
    for (int k = 0; k < t; k++) {
        int s = 0;
        for (int v : ms[k]) s += v;
        sums[k] = s;
    }
    std::vector<double> A((size_t)nf * t), B((size_t)nb * t);
    for (int k = 0; k < t; k++) {
        double fact = 1;
        for (int i = 0; i < N; i++) fact /= factorial(ms[k][i]);
        for (int i = 0; i < nf; i++) {
            double prod = 1;
            for (int ii = 0; ii < N; ii++) prod *= std::pow(dXf[i * N + ii], ms[k][ii]);
            A[i + (size_t)k * nf] = fact * prod;
        }
        for (int i = 0; i < nb; i++) {
            double asum = 0;
            for (int a = 0; a < N; a++) {
                if (ms[k][a] != 0) {
                    std::vector<int> ps = ms[k];
                    ps[a] -= 1;
                    double prod = 1;
                    for (int ii = 0; ii < N; ii++) prod *= std::pow(dXb[i * N + ii], ps[ii]);
                    asum += Vb[i * N + a] * ms[k][a] * prod;
                }
            }
            B[i + (size_t)k * nb] = fact * asum;
        }
    }
    // M = A^T Wf A + B^T Wb B ; RHS = A^T Wf
    std::vector<double> M((size_t)t * t, 0.0), R((size_t)t * nf, 0.0);
    for (int c = 0; c < t; c++)
        for (int r = 0; r < t; r++) {
            double s = 0;
            for (int i = 0; i < nf; i++) s += A[i + (size_t)r * nf] * Wf[i] * A[i + (size_t)c * nf];
            double s2 = 0;
            for (int i = 0; i < nb; i++) s2 += B[i + (size_t)r * nb] * Wb[i] * B[i + (size_t)c * nb];
            M[r + (size_t)c * t] = s + s2;
        }
    for (int j = 0; j < nf; j++)
        for (int r = 0; r < t; r++) R[r + (size_t)j * t] = A[j + (size_t)r * nf] * Wf[j];
    householder_solve(t, M, nf, R);
    for (int i = 0; i < t; i++) {
        double sc = std::pow(rmax, sums[i]);
        for (int j = 0; j < nf; j++) R[i + (size_t)j * t] /= sc;
    }
    F.swap(R);
    tout = t;
}

// ---------------------------------------------------------------------------
// Topology (geometry.cpp, edge.cpp, wakePanel.cpp, wake.cpp, bodyPanel.cpp)
// ---------------------------------------------------------------------------
struct Edge {
    int n1, n2;
    std::vector<int> body, wake;
    bool TE = false;
};

struct Panel {
    PanelGeo g;
    V3 bez;
    int nodes[4];
    int edges[4];
    int surfID;
    bool isWake;
    int parent;  // surface or wake index
    // body
    std::vector<int> neighbors, cluster;
    int TSorder = 3;
    bool upper = false, lower = false, TEpanel = false, tip = false;
    int TEedge = -1, index = -1;
    // wake
    int upperPan = -1, lowerPan = -1;
    bool wTE = false;
    // state
    double mu = 0, sigma = 0, phi = 0, Cp = 0;
    V3 vel = {0, 0, 0};
};

struct WakeLine {
    int upper, lower;
    V3 p1, p2, pMid, normal;
    double y() const { return pMid.y; }
};

struct Surface {
    int id;
    std::vector<int> panels;
    bool LS = false;
};

struct Wake {
    int id;
    std::vector<int> panels;
    std::vector<WakeLine> lines;
    std::vector<int> TEpanels;
    double yMin = 0, yMax = 0, x0 = 0, xf = 0, z0 = 0, zf = 0;
    V3 normal;
    double CL = 0, CD = 0;
    std::vector<double> yLoc, Cl, Cd;
};

}  // namespace

struct or_geom {
    std::vector<V3> nodes;
    std::vector<char> nodeTE;
    std::vector<Edge> edges;
    std::unordered_map<uint64_t, int> edgeMap;
    std::vector<Panel> P;  // all panels in file order
    std::vector<Surface> surfs;
    std::vector<Wake> wakes;
    std::vector<int> bPanels, wPanels;
};

namespace {

// geometry::findEdge (geometry.cpp:771-787): same result as the linear scan,
// found through a map keyed by the unordered node pair.
int find_edge(or_geom& G, int a, int b) {
    uint64_t lo = (uint64_t)std::min(a, b), hi = (uint64_t)std::max(a, b);
    uint64_t key = (lo << 32) | hi;
    auto it = G.edgeMap.find(key);
    if (it != G.edgeMap.end()) return it->second;
    Edge e;
    e.n1 = a;
    e.n2 = b;
    G.edges.push_back(e);
    int id = (int)G.edges.size() - 1;
    G.edgeMap[key] = id;
    return id;
}

// wake::addWakeLine (wake.cpp:134-138)
void add_wake_line(Wake& w, const WakeLine& wl) {
    w.lines.push_back(wl);
    std::sort(w.lines.begin(), w.lines.end(), [](const WakeLine& a, const WakeLine& b) { return a.y() < b.y(); });
}

// wakeLine::setDimensions (wakeLine.cpp:31-46)
WakeLine make_wake_line(const or_geom& G, int upper, int lower, V3 normal) {
    WakeLine wl;
    wl.upper = upper;
    wl.lower = lower;
    wl.normal = normal;
    const Panel& up = G.P[upper];
    int te = -1;
    for (int i = 0; i < up.g.nv; i++)
        if (G.edges[up.edges[i]].TE) te = up.edges[i];
    wl.p1 = G.nodes[G.edges[te].n1];
    wl.p2 = G.nodes[G.edges[te].n2];
    wl.pMid = (wl.p1 + wl.p2) / 2.0;
    return wl;
}

// wakePanel::setParentPanels (wakePanel.cpp:120-152)
void set_parent_panels(or_geom& G, int w, int upper, int lower) {
    Panel& W = G.P[w];
    G.P[upper].upper = true;
    G.P[lower].lower = true;
    if (!W.wTE) {
        W.wTE = true;
        G.wakes[W.parent].TEpanels.push_back(w);
    }
    if (W.upperPan == -1) {
        W.upperPan = upper;
        W.lowerPan = lower;
    } else if (G.P[upper].g.center.x < W.g.center.x) {
        W.upperPan = upper;
        W.lowerPan = lower;
    }
    add_wake_line(G.wakes[W.parent], make_wake_line(G, W.upperPan, W.lowerPan, W.g.normal));
}

// edge::checkTE (edge.cpp:48-109)
void check_te(or_geom& G, int ei) {
    Edge& e = G.edges[ei];
    if (e.body.size() != 2) return;
    if (e.wake.size() == 1) {
        e.TE = true;
        const Panel& b0 = G.P[e.body[0]];
        const Panel& b1 = G.P[e.body[1]];
        const Panel& w0 = G.P[e.wake[0]];
        V3 nrm = w0.g.normal;
        V3 v1 = b0.g.center - w0.g.center, v2 = b1.g.center - w0.g.center;
        double th1 = std::acos(dot(v1, nrm) / (norm(v1) * norm(nrm)));
        double th2 = std::acos(dot(v2, nrm) / (norm(v2) * norm(nrm)));
        int up, lo;
        if (th1 < th2) {
            up = e.body[0];
            lo = e.body[1];
        } else {
            up = e.body[1];
            lo = e.body[0];
        }
        int n1 = e.n1, n2 = e.n2, w = e.wake[0];
        set_parent_panels(G, w, up, lo);
        G.nodeTE[n1] = 1;
        G.nodeTE[n2] = 1;
    } else {
        const Panel& b0 = G.P[e.body[0]];
        const Panel& b1 = G.P[e.body[1]];
        double angle = std::acos(dot(b0.g.normal, b1.g.normal));
        if (angle > 4.7 * PI / 6 && b0.surfID == b1.surfID) {
            e.TE = true;
            G.nodeTE[e.n1] = 1;
            G.nodeTE[e.n2] = 1;
            for (int k = 0; k < 2; k++) {  // bodyPanel::setTEpanel (bodyPanel.cpp:47-53)
                Panel& b = G.P[e.body[k]];
                b.TEpanel = true;
                G.surfs[b.parent].LS = true;
                b.TEedge = ei;
            }
        }
    }
}

// wake::addPanel (wake.cpp:42-84)
void wake_add_panel(or_geom& G, Wake& w, int pi) {
    const Panel& p = G.P[pi];
    if (w.panels.empty()) {
        V3 pnt = G.nodes[p.nodes[0]];
        w.yMax = w.yMin = pnt.y;
        w.x0 = w.xf = pnt.x;
        w.z0 = w.zf = pnt.z;
        w.normal = p.g.normal;
    }
    for (int i = 0; i < p.g.nv; i++) {
        V3 pnt = G.nodes[p.nodes[i]];
        if (pnt.y > w.yMax)
            w.yMax = pnt.y;
        else if (pnt.y < w.yMin)
            w.yMin = pnt.y;
        if (pnt.x > w.xf)
            w.xf = pnt.x;
        else if (pnt.x < w.x0)
            w.x0 = pnt.x;
        if (pnt.z > w.zf)
            w.zf = pnt.z;
        else if (pnt.z < w.z0)
            w.z0 = pnt.z;
    }
    w.panels.push_back(pi);
}

// bodyPanel::setTipFlag (bodyPanel.cpp:58-95)
void set_tip_flag(or_geom& G, int bi) {
    Panel& b = G.P[bi];
    if (!G.surfs[b.parent].LS) {
        b.tip = false;
        return;
    }
    int count = 0;
    for (int nb : b.neighbors) {
        V3 nN = G.P[nb].g.normal;
        double d = dot(b.g.normal, nN) / (norm(b.g.normal) * norm(nN));
        if (d > 1)
            d = 1;
        else if (d < -1)
            d = -1;
        double angle = std::acos(d);
        if (angle < 1e-8 && std::asin(nN.x) > -PI / 12) count++;
    }
    if (count >= 2) b.tip = true;
}

// bodyPanel::clusterTest (bodyPanel.cpp:617-640)
bool cluster_test(const or_geom& G, const Panel& self, int selfIdx, int other, double angle, bool upFlag,
                  bool lowFlag) {
    const Panel& o = G.P[other];
    if ((upFlag && o.lower) || (lowFlag && o.upper)) return false;
    if (self.tip != o.tip) return false;
    double d = dot(o.g.normal, self.g.normal);
    if (d > 1)
        d = 1;
    else if (d < -1)
        d = -1;
    return std::acos(d) < angle &&
           std::find(self.cluster.begin(), self.cluster.end(), other) == self.cluster.end() && other != selfIdx;
}

// bodyPanel::setCluster (bodyPanel.cpp:402-487)
void set_cluster(or_geom& G, int bi) {
    Panel& b = G.P[bi];
    int dim = b.tip ? 2 : 3;
    size_t buffer = 10;
    size_t nObs = (size_t)(factorial(b.TSorder + dim) / (factorial(dim) * factorial(b.TSorder)) - 1) + buffer;
    size_t nPanels = (nObs + 1) / 2;
    size_t oldSize = b.cluster.size();
    b.cluster.push_back(bi);
    bool upFlag = b.upper, lowFlag = b.lower;
    while (b.cluster.size() < nPanels) {
        std::vector<int> toAdd;
        for (size_t i = oldSize; i < b.cluster.size(); i++) {
            const std::vector<int>& temp = G.P[b.cluster[i]].neighbors;
            for (size_t j = 0; j < temp.size(); j++) {
                if (cluster_test(G, b, bi, temp[j], 5 * PI / 6, upFlag, lowFlag)) {
                    if (G.P[temp[j]].upper)
                        upFlag = true;
                    else if (G.P[temp[j]].lower)
                        lowFlag = true;
                    if (std::find(toAdd.begin(), toAdd.end(), temp[j]) == toAdd.end()) toAdd.push_back(temp[j]);
                }
            }
        }
        if (toAdd.empty() && b.cluster.size() < nPanels) {
            while (b.cluster.size() < nPanels + 1 && b.TSorder >= 1) {
                if (buffer > 0)
                    buffer -= 1;
                else
                    b.TSorder -= 1;
                nObs = (size_t)(factorial(b.TSorder + dim) / (factorial(dim) * factorial(b.TSorder))) + buffer;
                nPanels = (nObs + 1) / 2;
            }
            if (!b.cluster.empty() && nPanels + 1 < b.cluster.size())
                b.cluster.erase(b.cluster.begin() + (long)(nPanels + 1), b.cluster.end());
            break;
        } else {
            oldSize = b.cluster.size();
            for (size_t i = 0; i < toAdd.size(); i++) {
                b.cluster.push_back(toAdd[i]);
                if (b.cluster.size() == nPanels + 1) break;
            }
        }
    }
}

// geometry::readTri / createSurfaces / correctWakeConnectivity, steady branch
// (geometry.cpp:132-484, 631-747)
or_geom* build_geom(int nn, const double* xyz, int nt, const int32_t* tri_in, const int32_t* ids) {
    or_geom* G = new or_geom();
    G->nodes.resize(nn);
    for (int i = 0; i < nn; i++) G->nodes[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    G->nodeTE.assign(nn, 0);
    std::vector<int> tri(tri_in, tri_in + 3 * (size_t)nt);

    // wake start detection (geometry.cpp:215-234)
    size_t wakeNodeStart = nn, wakeTriStart = nt;
    bool anyWake = false;
    for (int i = 0; i < nt; i++) {
        if (ids[i] > 1000) anyWake = true;
        if (i > 0 && ids[i] > 1000 && ids[i - 1] < 1000) {
            int mn = std::min(tri[3 * i], std::min(tri[3 * i + 1], tri[3 * i + 2]));
            wakeNodeStart = mn;
            wakeTriStart = i;
        }
    }
    if (anyWake) {
        // correctWakeConnectivity (geometry.cpp:631-667); tolerance is the
        // shortest triangle edge (geometry.cpp:669-698)
        double tol = 100000;
        for (int i = 0; i < nt; i++)
            for (int j = 0; j < 3; j++) {
                V3 p1 = G->nodes[tri[3 * i + j]];
                V3 p2 = G->nodes[tri[3 * i + (j < 2 ? j + 1 : 0)]];
                double l = norm(p2 - p1);
                if (l < tol) tol = l;
            }
        std::vector<std::pair<int, int>> reps;
        for (size_t i = 0; i < wakeNodeStart; i++)
            for (size_t j = wakeNodeStart; j < (size_t)nn; j++) {
                V3 v = G->nodes[i] - G->nodes[j];
                double linf = std::max(std::fabs(v.x), std::max(std::fabs(v.y), std::fabs(v.z)));
                if (linf < tol) reps.push_back({(int)j, (int)i});
            }
        for (size_t i = wakeTriStart; i < (size_t)nt; i++)
            for (int j = 0; j < 3; j++)
                for (auto& r : reps)
                    if (tri[3 * i + j] == r.first) tri[3 * i + j] = r.second;
    }

    // createSurfaces (geometry.cpp:712-747); panel ctors add themselves to
    // their edges, and each addition runs edge::checkTE
    G->P.reserve(nt);
    for (int i = 0; i < nt; i++) {
        Panel p;
        p.g.nv = 3;
        for (int k = 0; k < 3; k++) {
            p.nodes[k] = tri[3 * i + k];
            p.g.p[k] = G->nodes[p.nodes[k]];
        }
        for (int k = 0; k < 3; k++) p.edges[k] = find_edge(*G, p.nodes[k], p.nodes[(k + 1) % 3]);
        set_geom(p.g);
        p.bez = p.g.normal;
        p.surfID = ids[i];
        p.isWake = ids[i] > 1000;
        bool newGroup = (i == 0 || ids[i] != ids[i - 1]);
        int pi = (int)G->P.size();
        if (!p.isWake) {
            if (newGroup) {
                Surface s;
                s.id = ids[i];
                G->surfs.push_back(s);
            }
            p.parent = (int)G->surfs.size() - 1;
            G->P.push_back(p);
            for (int k = 0; k < 3; k++) {
                G->edges[p.edges[k]].body.push_back(pi);
                check_te(*G, p.edges[k]);
            }
            G->surfs.back().panels.push_back(pi);
        } else {
            if (newGroup) {
                Wake w;
                w.id = ids[i];
                G->wakes.push_back(w);
            }
            p.parent = (int)G->wakes.size() - 1;
            G->P.push_back(p);
            for (int k = 0; k < 3; k++) {
                G->edges[p.edges[k]].wake.push_back(pi);
                check_te(*G, p.edges[k]);
            }
            wake_add_panel(*G, G->wakes.back(), pi);
        }
    }

    // edge::setNeighbors in edge creation order (geometry.cpp:418-421)
    for (auto& e : G->edges) {
        if (e.body.size() == 2) {
            G->P[e.body[0]].neighbors.push_back(e.body[1]);
            G->P[e.body[1]].neighbors.push_back(e.body[0]);
        }
    }

    // wake merge (geometry.cpp:424-447, wake.cpp:86-132)
    if (G->wakes.size() > 1) {
        std::vector<Wake> newWakes;
        std::vector<char> alive(G->wakes.size(), 1);
        bool merged = false;
        std::vector<int> remap(G->wakes.size());
        for (size_t i = 0; i < G->wakes.size(); i++) remap[i] = (int)i;
        for (size_t i = 0; i < G->wakes.size(); i++) {
            if (!alive[i]) continue;
            for (size_t j = i; j < G->wakes.size(); j++) {
                if (j == i || !alive[j]) continue;
                Wake& a = G->wakes[i];
                Wake& o = G->wakes[j];
                double eps = 1e-2;
                if (std::fabs(o.x0 - a.x0) < eps && std::fabs(o.z0 - a.z0) < eps && std::fabs(o.xf - a.xf) < eps &&
                    std::fabs(o.zf - a.zf) < eps) {
                    for (int pi : o.panels) {
                        a.panels.push_back(pi);
                        G->P[pi].parent = (int)i;
                    }
                    for (auto& wl : o.lines) add_wake_line(a, wl);
                    if (o.yMin < a.yMin) a.yMin = o.yMin;
                    if (o.yMax > a.yMax) a.yMax = o.yMax;
                    alive[j] = 0;
                    merged = true;
                }
            }
        }
        if (merged) {
            std::vector<int> newIdx(G->wakes.size(), -1);
            for (size_t i = 0; i < G->wakes.size(); i++)
                if (alive[i]) {
                    newIdx[i] = (int)newWakes.size();
                    newWakes.push_back(G->wakes[i]);
                }
            for (auto& p : G->P)
                if (p.isWake) p.parent = newIdx[p.parent];
            G->wakes.swap(newWakes);
        }
    }

    // collect panels: surfaces are PREPENDED (geometry.cpp:452-458)
    for (size_t s = 0; s < G->surfs.size(); s++) {
        const std::vector<int>& sp = G->surfs[s].panels;
        G->bPanels.insert(G->bPanels.begin(), sp.begin(), sp.end());
    }
    for (auto& w : G->wakes)
        for (int pi : w.panels) G->wPanels.push_back(pi);

    for (int bi : G->bPanels) set_tip_flag(*G, bi);
    for (int bi : G->bPanels) set_cluster(*G, bi);
    for (size_t i = 0; i < G->bPanels.size(); i++) G->P[G->bPanels[i]].index = (int)i;
    return G;
}

// wakePanel::interpPanels (wakePanel.cpp:48-81)
void interp_panels(const or_geom& G, int w, int idx[4], double& coef) {
    const Panel& W = G.P[w];
    const std::vector<WakeLine>& L = G.wakes[W.parent].lines;
    const WakeLine *wl1 = nullptr, *wl2 = nullptr;
    double cy = W.g.center.y;
    if (cy < L[1].y()) {
        wl1 = &L[0];
        wl2 = &L[1];
    } else if (cy >= L.back().y()) {
        wl1 = &L[L.size() - 2];
        wl2 = &L[L.size() - 1];
    } else {
        for (size_t i = 1; i < L.size() - 1; i++)
            if (L[i].y() <= cy && L[i + 1].y() > cy) {
                wl1 = &L[i];
                wl2 = &L[i + 1];
            }
    }
    coef = (cy - wl1->y()) / (wl2->y() - wl1->y());
    idx[0] = G.P[wl1->upper].index;
    idx[1] = G.P[wl1->lower].index;
    idx[2] = G.P[wl2->upper].index;
    idx[3] = G.P[wl2->lower].index;
}

// wake::wakeStrength (wake.cpp:478-506)
double wake_strength(const or_geom& G, const Wake& w, double y) {
    const std::vector<WakeLine>& L = w.lines;
    const WakeLine *wl1 = nullptr, *wl2 = nullptr;
    if (y < L[1].y()) {
        wl1 = &L[0];
        wl2 = &L[1];
    } else if (y >= L.back().y()) {
        wl1 = &L[L.size() - 2];
        wl2 = &L[L.size() - 1];
    } else {
        for (size_t i = 1; i < L.size() - 1; i++)
            if (L[i].y() <= y && L[i + 1].y() > y) {
                wl1 = &L[i];
                wl2 = &L[i + 1];
            }
    }
    double interp = (y - wl1->y()) / (wl2->y() - wl1->y());
    double s1 = G.P[wl1->upper].mu - G.P[wl1->lower].mu;
    double s2 = G.P[wl2->upper].mu - G.P[wl2->lower].mu;
    return s1 + interp * (s2 - s1);
}

double wake_potential(const or_geom& G, V3 pnt) {  // geometry.cpp:1229-1237
    double pot = 0;
    for (int wi : G.wPanels) pot += -G.P[wi].mu * dub_phi_inf(G.P[wi].g, pnt);  // wakePanel::panelPhi :84-87
    return pot;
}

// wake::pntInWake (wake.cpp:601-640)
V3 pnt_in_wake(const or_geom& G, const Wake& w, double x, double y) {
    V3 yDir = {0, 1, 0};
    for (int wi : w.panels) {
        const Panel& W = G.P[wi];
        if (!W.wTE) continue;
        const Panel& up = G.P[W.upperPan];
        for (int k = 0; k < up.g.nv; k++) {
            const Edge& e = G.edges[up.edges[k]];
            if (!e.TE) continue;
            V3 p1 = G.nodes[e.n1], p2 = G.nodes[e.n2];
            if ((p1.y <= y && p2.y >= y) || (p1.y >= y && p2.y <= y)) {
                double t = (y - p1.y) / (p2.y - p1.y);
                V3 pnt = p1 + t * (p2 - p1);
                V3 out = -1.0 * cross(W.g.normal, yDir);
                if (out.x < 0) out = -1.0 * out;
                double scale = (x - pnt.x) / out.x;
                return pnt + scale * out;
            }
        }
    }
    return {0, 0, 0};
}

// wake::Vradial (wake.cpp:508-597)
double v_radial(const or_geom& G, const Wake& w, V3 pWake) {
    double r, theta = PI / 4, dZmax = 0.3, delZ;
    V3 POI;
    POI.x = pWake.x;
    if (pWake.y >= 0)
        r = w.yMax - pWake.y;
    else
        r = pWake.y - w.yMin;
    delZ = r * std::sin(theta);
    if (delZ > dZmax) {
        delZ = dZmax;
        theta = std::asin(dZmax / r);
    }
    if (pWake.y >= 0)
        POI.y = w.yMax - r * std::cos(theta);
    else
        POI.y = w.yMin + r * std::cos(theta);
    POI.z = pWake.z + r * std::sin(theta);
    const int nP = 6;
    double dz = 0.5 * delZ;
    double step = 2 * dz / (nP - 1);
    double phiPOI = wake_potential(G, POI);
    double dPy[nP], dPz[nP], dY[nP], dZ[nP];
    for (int i = 0; i < nP; i++) {
        double delta = -dz + i * step;
        V3 p1 = POI + delta * V3{0, 1, 0};
        V3 p2 = POI + delta * V3{0, 0, 1};
        dPy[i] = wake_potential(G, p1) - phiPOI;
        dPz[i] = wake_potential(G, p2) - phiPOI;
        dY[i] = p1.y - POI.y;
        dZ[i] = p2.z - POI.z;
    }
    double x0 = 0;
    std::vector<double> F;
    int t;
    chtls(1, &x0, nP, dY, 3, 0, nullptr, nullptr, F, t);
    double v = 0;
    for (int i = 0; i < nP; i++) v += F[0 + (size_t)i * t] * dPy[i];
    chtls(1, &x0, nP, dZ, 3, 0, nullptr, nullptr, F, t);
    double ww = 0;
    for (int i = 0; i < nP; i++) ww += F[0 + (size_t)i * t] * dPz[i];
    return std::sqrt(v * v + ww * ww);
}

// wake::trefftzPlane (wake.cpp:146-188)
void trefftz_plane(const or_geom& G, Wake& w, double Vinf, double Sref) {
    int nPnts = 164;
    w.yLoc.assign(nPnts + 1, 0);
    w.Cl.assign(nPnts + 1, 0);
    w.Cd.assign(nPnts + 1, 0);
    std::vector<double> wv(nPnts + 1, 0), dPhi(nPnts + 1, 0);
    w.yLoc[0] = w.yMin;
    w.yLoc[nPnts] = w.yMax;
    double step = (w.yMax - w.yMin) / nPnts;
    double xT = w.x0 + 2 * (w.xf - w.x0) / 3;
    for (int i = 1; i < nPnts; i++) {
        w.yLoc[i] = w.yMin + i * step;
        V3 pW = pnt_in_wake(G, w, xT, w.yLoc[i]);
        wv[i] = v_radial(G, w, pW);
        dPhi[i] = -wake_strength(G, w, w.yLoc[i]);
        w.Cl[i] = 2 * dPhi[i] / (Vinf * Sref);
        w.Cd[i] = dPhi[i] * wv[i] / (Vinf * Vinf * Sref);
    }
    w.CL = 0;
    w.CD = 0;
    int i = 0;
    while (i < (int)w.Cl.size() - 2) {
        w.CL += 1.0 / 3 * step * (w.Cl[i] + 4 * w.Cl[i + 1] + w.Cl[i + 2]);
        w.CD += 1.0 / 3 * step * (w.Cd[i] + 4 * w.Cd[i + 1] + w.Cd[i + 2]);
        i += 2;
    }
}

// local frame helpers for velocity2D (panel.cpp:179-201)
inline V3 global2local(const PanelGeo& g, V3 v, bool translate) {
    V3 l, m, n;
    local_sys(g, l, m, n);
    V3 t = translate ? v - g.center : v;
    return {dot(l, t), dot(m, t), dot(n, t)};
}
inline V3 local2global(const PanelGeo& g, V3 v, bool translate) {
    V3 l, m, n;
    local_sys(g, l, m, n);
    V3 r = {l.x * v.x + m.x * v.y + n.x * v.z, l.y * v.x + m.y * v.y + n.y * v.z, l.z * v.x + m.z * v.y + n.z * v.z};
    return translate ? r + g.center : r;
}

// bodyPanel::velocity2D (bodyPanel.cpp:526-560)
V3 velocity2d(const or_geom& G, const Panel& b, V3 pnt, double pntPot) {
    std::vector<int> cl = b.cluster;
    if (pnt.x == b.g.center.x && pnt.y == b.g.center.y && pnt.z == b.g.center.z) cl.erase(cl.begin());
    int nf = (int)cl.size();
    V3 pl = global2local(b.g, pnt, true);
    double X0[2] = {pl.x, pl.y};
    std::vector<double> Xf(2 * nf), df(nf);
    for (int i = 0; i < nf; i++) {
        V3 c = global2local(b.g, G.P[cl[i]].g.center, true);
        Xf[2 * i] = c.x;
        Xf[2 * i + 1] = c.y;
        df[i] = G.P[cl[i]].phi - pntPot;
    }
    const double nan = std::numeric_limits<double>::quiet_NaN();
    // no neighbours or a zero-order fit: the reference reads past the end of
    // an empty Eigen matrix here (undefined); the oracle returns NaN so the
    // caller's last-resort fallback applies
    if (nf == 0 || b.TSorder < 1) return V3{nan, nan, nan};
    std::vector<double> F;
    int t;
    chtls(2, X0, nf, Xf.data(), b.TSorder, 0, nullptr, nullptr, F, t);
    double v0 = 0, v1 = 0;
    for (int i = 0; i < nf; i++) {
        v0 += F[0 + (size_t)i * t] * df[i];
        v1 += F[1 + (size_t)i * t] * df[i];
    }
    return local2global(b.g, V3{v0, v1, 0}, false);
}

// bodyPanel::velocity3D (bodyPanel.cpp:562-589)
V3 velocity3d(const or_geom& G, const Panel& b, V3 pnt, double pntPot) {
    std::vector<int> cl = b.cluster;
    if (pnt.x == b.g.center.x && pnt.y == b.g.center.y && pnt.z == b.g.center.z) cl.erase(cl.begin());
    int nf = (int)cl.size();
    std::vector<double> Xf(3 * nf), Vb(3 * nf), df(nf);
    for (int i = 0; i < nf; i++) {
        const Panel& c = G.P[cl[i]];
        Xf[3 * i] = c.g.center.x;
        Xf[3 * i + 1] = c.g.center.y;
        Xf[3 * i + 2] = c.g.center.z;
        Vb[3 * i] = c.bez.x;
        Vb[3 * i + 1] = c.bez.y;
        Vb[3 * i + 2] = c.bez.z;
        df[i] = c.phi - pntPot;
    }
    double X0[3] = {pnt.x, pnt.y, pnt.z};
    if (nf == 0 || b.TSorder < 1) {  // see velocity2d
        const double nan = std::numeric_limits<double>::quiet_NaN();
        return V3{nan, nan, nan};
    }
    std::vector<double> F;
    int t;
    chtls(3, X0, nf, Xf.data(), b.TSorder, nf, Xf.data(), Vb.data(), F, t);
    V3 v = {0, 0, 0};
    for (int i = 0; i < nf; i++) {
        v.x += F[0 + (size_t)i * t] * df[i];
        v.y += F[1 + (size_t)i * t] * df[i];
        v.z += F[2 + (size_t)i * t] * df[i];
    }
    return v;
}

// bodyPanel::pntVelocity (bodyPanel.cpp:490-524)
V3 pnt_velocity(const or_geom& G, const Panel& b, V3 pnt, double pot, double PG, V3 Vinf) {
    V3 vel;
    if (b.tip) {
        vel = velocity2d(G, b, pnt, pot);
    } else {
        vel = velocity3d(G, b, pnt, pot);
        if (isnan3(vel)) vel = velocity2d(G, b, pnt, pot);
    }
    if (isnan3(vel)) vel = cross(cross(b.bez, Vinf), b.bez);
    vel.x /= PG;
    return vel;
}

// cpCase::windToBody (cpCase.cpp:142-166)
void wind_to_body(double V, double a, double b, double T[9], V3& Vinf) {
    a *= PI / 180;
    b *= PI / 180;
    T[0] = std::cos(a) * std::cos(b);
    T[1] = std::cos(a) * std::sin(b);
    T[2] = -std::sin(a);
    T[3] = -std::sin(b);
    T[4] = std::cos(b);
    T[5] = 0;
    T[6] = std::sin(a) * std::cos(b);
    T[7] = std::sin(a) * std::sin(b);
    T[8] = std::cos(a);
    Vinf = {T[0] * V, T[3] * V, T[6] * V};
}

int nthreads_or_default(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// column-major GEMV y = A x (rows split over threads)
void gemv(int n, const double* A, int64_t lda, const double* x, double* y, int nth) {
#pragma omp parallel num_threads(nth)
    {
        int tid = 0, nt = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        int r0 = (int)((int64_t)n * tid / nt), r1 = (int)((int64_t)n * (tid + 1) / nt);
        for (int i = r0; i < r1; i++) y[i] = 0;
        for (int j = 0; j < n; j++) {
            const double* col = A + (int64_t)j * lda;
            double xj = x[j];
            for (int i = r0; i < r1; i++) y[i] += col[i] * xj;
        }
    }
}

double dotv(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

void post_process(or_geom* G, const or_case_params* p, or_case_result* res, V3 Vinf, const double T[9]) {
    double PG = std::sqrt(1 - p->mach * p->mach);
    // wake mu and potentials (cpCase.cpp:205-213, wakePanel.cpp:94-100)
    for (int bi : G->bPanels) {
        Panel& b = G->P[bi];
        b.phi = dot(Vinf, b.g.center) - b.mu;  // panel::setPotential panel.cpp:106-111
    }
    for (int wi : G->wPanels) {
        Panel& w = G->P[wi];
        int idx[4];
        double c;
        interp_panels(*G, wi, idx, c);
        const std::vector<int>& bp = G->bPanels;
        w.mu = (1 - c) * G->P[bp[idx[0]]].mu + (c - 1) * G->P[bp[idx[1]]].mu + c * G->P[bp[idx[2]]].mu -
               c * G->P[bp[idx[3]]].mu;
        w.phi = dot(Vinf, w.g.center) - w.mu;
    }
    // compVelocity (cpCase.cpp:342-365)
    V3 Fb = {0, 0, 0}, CM = {0, 0, 0};
    V3 cg = {p->cg[0], p->cg[1], p->cg[2]};
    for (int bi : G->bPanels) {
        Panel& b = G->P[bi];
        b.vel = pnt_velocity(*G, b, b.g.center, b.phi, PG, Vinf);
        double vn = norm(b.vel) / p->V;
        b.Cp = 1 - vn * vn;  // bodyPanel::computeCp :591-594
        Fb = Fb + (-b.Cp * b.g.area * b.bez) / p->Sref;
        V3 r = b.g.center - cg;  // bodyPanel::computeMoments :610-615
        V3 F = -b.Cp * b.bez;
        F = b.g.area * F;
        V3 mom = cross(r, F);
        CM.x += mom.x / (p->Sref * p->bref);
        CM.y += mom.y / (p->Sref * p->cref);
        CM.z += mom.z / (p->Sref * p->bref);
    }
    // bodyToWind: T^T * F
    V3 Fw = {T[0] * Fb.x + T[3] * Fb.y + T[6] * Fb.z, T[1] * Fb.x + T[4] * Fb.y + T[7] * Fb.z,
             T[2] * Fb.x + T[5] * Fb.y + T[8] * Fb.z};
    res->Fbody[0] = Fb.x;
    res->Fbody[1] = Fb.y;
    res->Fbody[2] = Fb.z;
    res->Fwind[0] = Fw.x;
    res->Fwind[1] = Fw.y;
    res->Fwind[2] = Fw.z;
    res->CM[0] = CM.x;
    res->CM[1] = CM.y;
    res->CM[2] = CM.z;
    // trefftzPlaneAnalysis (cpCase.cpp:541-552)
    res->CL = 0;
    res->CDi = 0;
    for (auto& w : G->wakes) {
        trefftz_plane(*G, w, p->V, p->Sref);
        res->CL += w.CL / PG;
        res->CDi += w.CD / (PG * PG);
    }
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

void or_panel_phi_inf(int nv, const double* v, const double* poi, double* phi_src, double* phi_dub) {
    PanelGeo g = make_panel(nv, v);
    double s = 0, d = 0;
    panel_phi_inf(g, V3{poi[0], poi[1], poi[2]}, s, d);
    *phi_src = s;
    *phi_dub = d;
}

double or_dub_phi_inf(int nv, const double* v, const double* poi) {
    PanelGeo g = make_panel(nv, v);
    return dub_phi_inf(g, V3{poi[0], poi[1], poi[2]});
}

void or_panel_v_inf(int nv, const double* v, const double* poi, double* vsrc, double* vdub) {
    PanelGeo g = make_panel(nv, v);
    V3 s, d;
    panel_v_inf(g, V3{poi[0], poi[1], poi[2]}, s, d);
    vsrc[0] = s.x;
    vsrc[1] = s.y;
    vsrc[2] = s.z;
    vdub[0] = d.x;
    vdub[1] = d.y;
    vdub[2] = d.z;
}

void or_dub_v_inf(int nv, const double* v, const double* poi, double* vdub) {
    PanelGeo g = make_panel(nv, v);
    V3 d = dub_v_inf(g, V3{poi[0], poi[1], poi[2]});
    vdub[0] = d.x;
    vdub[1] = d.y;
    vdub[2] = d.z;
}

or_geom* or_geom_create(int n_nodes, const double* xyz, int n_tris, const int32_t* tri, const int32_t* ids) {
    return build_geom(n_nodes, xyz, n_tris, tri, ids);
}

or_geom* or_geom_read_tri(const char* path) {
    std::ifstream f(path);
    if (!f) return nullptr;
    int nn, nt;
    if (!(f >> nn >> nt)) return nullptr;
    std::vector<double> xyz(3 * (size_t)nn);
    for (auto& v : xyz)
        if (!(f >> v)) return nullptr;
    std::vector<int32_t> tri(3 * (size_t)nt), ids(nt);
    for (auto& v : tri) {
        long long x;
        if (!(f >> x)) return nullptr;
        v = (int32_t)(x - 1);
    }
    for (auto& v : ids)
        if (!(f >> v)) return nullptr;
    return build_geom(nn, xyz.data(), nt, tri.data(), ids.data());
}

void or_geom_destroy(or_geom* g) { delete g; }
int or_geom_n_body(const or_geom* g) { return (int)g->bPanels.size(); }
int or_geom_n_wake(const or_geom* g) { return (int)g->wPanels.size(); }
int or_geom_n_wakelines(const or_geom* g) {
    int n = 0;
    for (auto& w : g->wakes) n += (int)w.lines.size();
    return n;
}

void or_geom_body_info(const or_geom* g, int32_t* tri_index, double* center, double* normal, double* area,
                       double* long_side, int32_t* tip_flag) {
    for (size_t i = 0; i < g->bPanels.size(); i++) {
        const Panel& p = g->P[g->bPanels[i]];
        if (tri_index) tri_index[i] = g->bPanels[i];
        if (center) {
            center[3 * i] = p.g.center.x;
            center[3 * i + 1] = p.g.center.y;
            center[3 * i + 2] = p.g.center.z;
        }
        if (normal) {
            normal[3 * i] = p.g.normal.x;
            normal[3 * i + 1] = p.g.normal.y;
            normal[3 * i + 2] = p.g.normal.z;
        }
        if (area) area[i] = p.g.area;
        if (long_side) long_side[i] = p.g.longSide;
        if (tip_flag) tip_flag[i] = p.tip ? 1 : 0;
    }
}

void or_geom_wake_info(const or_geom* g, int32_t* tri_index, int32_t* idx4, double* coef) {
    for (size_t j = 0; j < g->wPanels.size(); j++) {
        if (tri_index) tri_index[j] = g->wPanels[j];
        int idx[4];
        double c;
        interp_panels(*g, g->wPanels[j], idx, c);
        if (idx4)
            for (int k = 0; k < 4; k++) idx4[4 * j + k] = idx[k];
        if (coef) coef[j] = c;
    }
}

void or_geom_clusters(const or_geom* g, int32_t* out16, int32_t* ts_order) {
    for (size_t i = 0; i < g->bPanels.size(); i++) {
        const Panel& p = g->P[g->bPanels[i]];
        for (int k = 0; k < 16; k++)
            out16[16 * i + k] = k < (int)p.cluster.size() ? g->P[p.cluster[k]].index : -1;
        if (ts_order) ts_order[i] = p.TSorder;
    }
}

void or_assemble(const or_geom* g, double* A, double* B, int64_t lda, int nthreads) {
    int nb = (int)g->bPanels.size();
    int nth = nthreads_or_default(nthreads);
    std::vector<V3> ctr(nb);
    for (int i = 0; i < nb; i++) ctr[i] = g->P[g->bPanels[i]].g.center;
    // body-body loop (geometry.cpp:823-841): column j = panel j, row i = centre i
#pragma omp parallel for schedule(dynamic, 16) num_threads(nth)
    for (int j = 0; j < nb; j++) {
        const PanelGeo& pg = g->P[g->bPanels[j]].g;
        double* Aj = A + (int64_t)j * lda;
        double* Bj = B + (int64_t)j * lda;
        for (int i = 0; i < nb; i++) {
            double s = 0, d = 0;
            panel_phi_inf(pg, ctr[i], s, d);
            Bj[i] = s;
            Aj[i] = d;
        }
    }
    // wake folding (geometry.cpp:854-872), wake panels in order per row
    int nw = (int)g->wPanels.size();
    std::vector<int> idx(4 * (size_t)nw);
    std::vector<double> cf(nw);
    for (int j = 0; j < nw; j++) interp_panels(*g, g->wPanels[j], &idx[4 * j], cf[j]);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int i = 0; i < nb; i++) {
        for (int j = 0; j < nw; j++) {
            double inf = dub_phi_inf(g->P[g->wPanels[j]].g, ctr[i]);
            double c = cf[j];
            A[i + (int64_t)idx[4 * j + 0] * lda] += inf * (1 - c);
            A[i + (int64_t)idx[4 * j + 1] * lda] += inf * (c - 1);
            A[i + (int64_t)idx[4 * j + 2] * lda] += inf * c;
            A[i + (int64_t)idx[4 * j + 3] * lda] -= inf * c;
        }
    }
}

// The body-body loop of or_assemble (geometry.cpp:823-841) over columns
// j0..j1-1 only: a bounded sample of the assembly for timing.
void or_assemble_cols(const or_geom* g, double* A, double* B, int64_t lda, int j0, int j1, int nthreads) {
    int nb = (int)g->bPanels.size();
    int nth = nthreads_or_default(nthreads);
    std::vector<V3> ctr(nb);
    for (int i = 0; i < nb; i++) ctr[i] = g->P[g->bPanels[i]].g.center;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nth)
    for (int j = j0; j < j1; j++) {
        const PanelGeo& pg = g->P[g->bPanels[j]].g;
        double* Aj = A + (int64_t)(j - j0) * lda;
        double* Bj = B + (int64_t)(j - j0) * lda;
        for (int i = 0; i < nb; i++) {
            double s = 0, d = 0;
            panel_phi_inf(pg, ctr[i], s, d);
            Bj[i] = s;
            Aj[i] = d;
        }
    }
}

// Eigen BiCGSTAB (Eigen/src/IterativeLinearSolvers/BiCGSTAB.h), restated:
// diagonal preconditioner, x0 = 0, restart when rho becomes too small.
int or_bicgstab(int n, const double* A, int64_t lda, const double* b, double* x, double tol, int maxit,
                int nthreads, double* err) {
    int nth = nthreads_or_default(nthreads);
    if (tol <= 0) tol = std::numeric_limits<double>::epsilon();
    if (maxit <= 0) maxit = 2 * n;
    std::vector<double> invd(n);
    for (int j = 0; j < n; j++) {
        double d = A[j + (int64_t)j * lda];
        invd[j] = d != 0 ? 1 / d : 1;
    }
    std::vector<double> r(n), r0(n), v(n, 0), p(n, 0), y(n), z(n), s(n), t(n), tmp(n);
    for (int i = 0; i < n; i++) x[i] = 0;
    for (int i = 0; i < n; i++) r[i] = b[i];  // r = b - A*0
    r0 = r;
    double r0_sq = dotv(n, r0.data(), r0.data());
    double rhs_sq = dotv(n, b, b);
    if (rhs_sq == 0) {
        *err = 0;
        return 0;
    }
    double rho = 1, alpha = 1, w = 1;
    double tol2 = tol * tol * rhs_sq;
    double eps = std::numeric_limits<double>::epsilon();
    double eps2 = eps * eps;
    int i = 0, restarts = 0;
    while (dotv(n, r.data(), r.data()) > tol2 && i < maxit) {
        double rho_old = rho;
        rho = dotv(n, r0.data(), r.data());
        if (std::fabs(rho) < eps2 * r0_sq) {
            gemv(n, A, lda, x, tmp.data(), nth);
            for (int k = 0; k < n; k++) r[k] = b[k] - tmp[k];
            r0 = r;
            rho = r0_sq = dotv(n, r.data(), r.data());
            if (restarts++ == 0) i = 0;
        }
        double beta = (rho / rho_old) * (alpha / w);
        for (int k = 0; k < n; k++) p[k] = r[k] + beta * (p[k] - w * v[k]);
        for (int k = 0; k < n; k++) y[k] = invd[k] * p[k];
        gemv(n, A, lda, y.data(), v.data(), nth);
        alpha = rho / dotv(n, r0.data(), v.data());
        for (int k = 0; k < n; k++) s[k] = r[k] - alpha * v[k];
        for (int k = 0; k < n; k++) z[k] = invd[k] * s[k];
        gemv(n, A, lda, z.data(), t.data(), nth);
        double tmpn = dotv(n, t.data(), t.data());
        if (tmpn > 0)
            w = dotv(n, t.data(), s.data()) / tmpn;
        else
            w = 0;
        for (int k = 0; k < n; k++) x[k] += alpha * y[k] + w * z[k];
        for (int k = 0; k < n; k++) r[k] = s[k] - w * t[k];
        ++i;
    }
    *err = std::sqrt(dotv(n, r.data(), r.data()) / rhs_sq);
    return i;
}

int or_case_run(or_geom* g, const double* A, const double* B, int64_t lda, const or_case_params* p, double tol,
                int maxit, int nthreads, or_case_result* res, double* mu, double* sigma, double* phi, double* vel,
                double* cp, double* wake_mu, double* yloc, double* span_cl, double* span_cd) {
    int nb = (int)g->bPanels.size();
    int nth = nthreads_or_default(nthreads);
    double T[9];
    V3 Vinf;
    wind_to_body(p->V, p->alpha_deg, p->beta_deg, T, Vinf);
    // setSourceStrengths (cpCase.cpp:174-182), bodyPanel::setSigma (:97-100)
    std::vector<double> sig(nb), rhs(nb), x(nb);
    for (int i = 0; i < nb; i++) {
        Panel& b = g->P[g->bPanels[i]];
        b.sigma = -dot(Vinf, b.g.normal) + 0;
        sig[i] = b.sigma;
    }
    // solveMatrixEq (cpCase.cpp:185-214): RHS = -B*sigma
    gemv(nb, B, lda, sig.data(), rhs.data(), nth);
    for (auto& v : rhs) v = -v;
    double err = 0;
    int it = or_bicgstab(nb, A, lda, rhs.data(), x.data(), tol, maxit, nth, &err);
    res->solve_err = err;
    res->solve_iters = it;
    for (int i = 0; i < nb; i++) g->P[g->bPanels[i]].mu = x[i];
    post_process(g, p, res, Vinf, T);
    for (int i = 0; i < nb; i++) {
        const Panel& b = g->P[g->bPanels[i]];
        if (mu) mu[i] = b.mu;
        if (sigma) sigma[i] = b.sigma;
        if (phi) phi[i] = b.phi;
        if (vel) {
            vel[3 * i] = b.vel.x;
            vel[3 * i + 1] = b.vel.y;
            vel[3 * i + 2] = b.vel.z;
        }
        if (cp) cp[i] = b.Cp;
    }
    if (wake_mu)
        for (size_t j = 0; j < g->wPanels.size(); j++) wake_mu[j] = g->P[g->wPanels[j]].mu;
    if (!g->wakes.empty() && yloc) {
        const Wake& w = g->wakes[0];
        for (size_t k = 0; k < w.yLoc.size(); k++) {
            yloc[k] = w.yLoc[k];
            if (span_cl) span_cl[k] = w.Cl[k];
            if (span_cd) span_cd[k] = w.Cd[k];
        }
    }
    return err <= 1e-10 ? 0 : 1;  // converged flag threshold (cpCase.cpp:199-202)
}

void or_case_post(or_geom* g, const or_case_params* p, const double* mu_in, or_case_result* res, double* phi,
                  double* vel, double* cp, double* wake_mu) {
    int nb = (int)g->bPanels.size();
    double T[9];
    V3 Vinf;
    wind_to_body(p->V, p->alpha_deg, p->beta_deg, T, Vinf);
    for (int i = 0; i < nb; i++) {
        Panel& b = g->P[g->bPanels[i]];
        b.sigma = -dot(Vinf, b.g.normal);
        b.mu = mu_in[i];
    }
    post_process(g, p, res, Vinf, T);
    res->solve_err = 0;
    res->solve_iters = 0;
    for (int i = 0; i < nb; i++) {
        const Panel& b = g->P[g->bPanels[i]];
        if (phi) phi[i] = b.phi;
        if (vel) {
            vel[3 * i] = b.vel.x;
            vel[3 * i + 1] = b.vel.y;
            vel[3 * i + 2] = b.vel.z;
        }
        if (cp) cp[i] = b.Cp;
    }
    if (wake_mu)
        for (size_t j = 0; j < g->wPanels.size(); j++) wake_mu[j] = g->P[g->wPanels[j]].mu;
}

double or_wake_potential(const or_geom* g, const double* pnt) {
    return wake_potential(*g, V3{pnt[0], pnt[1], pnt[2]});
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Panel -> particle stretching (bodyPanel::partStretching, bodyPanel.cpp:697-729;
// wakePanel::partStretching, wakePanel.cpp:266-289).  M(i, j) = d u_j / d x_i
// (the reference's "dudx dvdx dwdx / dudy ... " layout); the stretching is
// M * alpha.  The formulas are restated as the reference writes them,
// including its frames: the near-field source gradient is evaluated in the
// panel's local frame and applied to the global strength, and the far-field
// point source/doublet gradients use the panel centre's z where a relative z
// would be expected.
// ---------------------------------------------------------------------------
namespace {

struct M3 {
    double a[3][3];
};

inline M3 m3_zero() {
    M3 m;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m.a[i][j] = 0;
    return m;
}

inline V3 to_local(const PanelGeo& g, V3 p) {
    V3 l, m, n;
    local_sys(g, l, m, n);
    V3 d = p - g.center;
    return {dot(l, d), dot(m, d), dot(n, d)};
}

// bodyPanel::velocityGradientPointSource (bodyPanel.cpp:733-763)
M3 vg_point_source(const PanelGeo& g, double sigma, V3 poi) {
    M3 M;
    double x = g.center.x, y = g.center.y, z = g.center.z, x0 = poi.x, y0 = poi.y, z0 = poi.z;
    double sdd = std::pow(std::pow(x - x0, 2) + std::pow(y - y0, 2) + z * z, 2.5);
    double c = sigma * g.area / (4 * PI);
    M.a[0][0] = (-2 * (x - x0) * (x - x0) + (y - y0) * (y - y0) + z * z) / sdd;
    M.a[1][0] = -3 * (x - x0) * (y - y0) / sdd;
    M.a[2][0] = -3 * (x - x0) * z / sdd;
    M.a[0][1] = -3 * (x - x0) * (y - y0) / sdd;
    M.a[1][1] = ((x - x0) * (x - x0) - (y - y0) * (y - y0) + z * z) / sdd;
    M.a[2][1] = -3 * z * (y - y0) / sdd;
    M.a[0][2] = -3 * (x - x0) * (z - z0) / sdd;
    M.a[1][2] = -3 * (y - y0) * (z - z0) / sdd;
    M.a[2][2] = (-3 * z * (z - z0) + ((x - x0) * (x - x0) + (y - y0) * (y - y0) + z * z)) / sdd;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M.a[i][j] *= c;
    return M;
}

// bodyPanel::velocityGradientTriSource / QuadSource (bodyPanel.cpp:766-1063):
// the Katz source-panel velocity differentiated in local coordinates, one
// term per edge i -> i+1
M3 vg_panel_source(const PanelGeo& g, double sigma, V3 poi) {
    M3 M = m3_zero();
    const int nv = g.nv;
    double X[4], Y[4];
    for (int k = 0; k < nv; k++) {
        V3 q = to_local(g, g.p[k]);
        X[k] = q.x;
        Y[k] = q.y;
    }
    V3 P = to_local(g, poi);
    double x = P.x, y = P.y, z = P.z;
    double r[4], drx[4], dry[4], drz[4], e[4], h[4];
    for (int k = 0; k < nv; k++) {
        double q = (x - X[k]) * (x - X[k]) + (y - Y[k]) * (y - Y[k]) + z * z;
        r[k] = std::pow(q, 0.5);
        drx[k] = (x - X[k]) / std::pow(q, 0.5);
        dry[k] = (y - Y[k]) / std::pow(q, 0.5);
        drz[k] = z / std::pow(q, 0.5);
        e[k] = (x - X[k]) * (x - X[k]) + z * z;
        h[k] = (x - X[k]) * (y - Y[k]);
    }
    for (int i = 0; i < nv; i++) {
        int j = (i + 1) % nv;
        double d = std::pow((X[j] - X[i]) * (X[j] - X[i]) + (Y[j] - Y[i]) * (Y[j] - Y[i]), 0.5);
        double S = r[i] + r[j];
        double G[3] = {drx[i] + drx[j], dry[i] + dry[j], drz[i] + drz[j]};
        for (int k = 0; k < 3; k++) {
            M.a[k][0] += (Y[j] - Y[i]) * (S + d) * G[k] * (2 * d) / d / (S - d) / std::pow(S + d, 2);
            M.a[k][1] += (X[i] - X[j]) * (S + d) * G[k] * (2 * d) / d / (S - d) / std::pow(S + d, 2);
        }
        double mm = (Y[j] - Y[i]) / (X[j] - X[i]);
        for (int t = 0; t < 2; t++) {
            int q = t == 0 ? i : j;
            double sg = t == 0 ? 1.0 : -1.0;
            double me = mm * e[q] - h[q];
            double zr = z * r[q];
            double f = 1 / (1 + std::pow(me / zr, 2));
            M.a[0][2] += sg * f * ((zr * (mm * 2 * (x - X[q]) - (y - Y[q])) - me * (z * drx[q])) / std::pow(zr, 2));
            M.a[1][2] += sg * f * ((-z * r[q] * (x - X[q]) - me * (z * dry[q])) / std::pow(zr, 2));
            M.a[2][2] += sg * f * ((zr * mm * (2 * z) - me * (r[q] + z * drz[q])) / std::pow(zr, 2));
        }
    }
    double c = sigma / (4 * PI);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) M.a[a][b] *= c;
    return M;
}

// panel::velocityGradientPointDoublet (panel.cpp:813-840)
M3 vg_point_doublet(const PanelGeo& g, double mu, V3 poi) {
    M3 M;
    double x = g.center.x, y = g.center.y, z = g.center.z, x0 = poi.x, y0 = poi.y;
    double ddd = std::pow(std::pow(x - x0, 2) + std::pow(y - y0, 2) + z * z, 3.5);
    double uv = 3 * mu * g.area / (4 * PI);
    double w = -mu * g.area / (4 * PI);
    double dx2 = std::pow(x - x0, 2), dy2 = std::pow(y - y0, 2);
    M.a[0][0] = uv * z * (-4 * dx2 + dy2 + z * z) / ddd;
    M.a[1][0] = uv * (-5) * z * (y - y0) * (x - x0) / ddd;
    M.a[2][0] = uv * (x - x0) * (dx2 + dy2 - 4 * z * z) / ddd;
    M.a[0][1] = uv * (-5) * (x - x0) * (y - y0) * z / ddd;
    M.a[1][1] = uv * z * (dx2 - 4 * dy2 + z * z) / ddd;
    M.a[2][1] = uv * (y - y0) * (dx2 + dy2 - 4 * z * z) / ddd;
    M.a[0][2] = w * (-3) * (x - x0) * (dx2 + dy2 - 4 * z * z) / ddd;
    M.a[1][2] = w * (-3) * (y - y0) * (dx2 + dy2 - 4 * z * z) / ddd;
    M.a[2][2] = w * (-3) * z * (3 * dx2 + 3 * dy2 - 2 * z * z) / ddd;
    return M;
}

// panel::gradDoub (panel.cpp:876-923): d/dx_i of the vortex-segment
// velocity (a x b)(|a|+|b|) / (|a||b|(|a||b| + a.b) + (core |s|)^2)
M3 grad_doub(V3 a, V3 b, V3 s) {
    M3 M;
    double na = norm(a), nb = norm(b);
    V3 top = (na + nb) * cross(a, b);
    double cs = CORE * norm(s);
    double bot = na * nb * (na * nb + dot(a, b)) + cs * cs;
    V3 E[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double dan[3] = {a.x / na, a.y / na, a.z / na}, dbn[3] = {b.x / nb, b.y / nb, b.z / nb};
    for (int k = 0; k < 3; k++) {
        V3 dtop = (na + nb) * (cross(E[k], b) + cross(a, E[k])) + (dan[k] + dbn[k]) * cross(a, b);
        double dbot = dan[k] * nb * (na * nb + dot(a, b)) + na * dbn[k] * (na * nb + dot(a, b)) +
                      na * nb * (dan[k] * nb + na * dbn[k] + dot(E[k], b) + dot(a, E[k]));
        V3 row = (1.0 / (bot * bot)) * (bot * dtop - dbot * top);
        M.a[k][0] = row.x;
        M.a[k][1] = row.y;
        M.a[k][2] = row.z;
    }
    return M;
}

// panel::nearFilamentCheck (panel.cpp:925-962)
bool near_filament(V3 p1, V3 p2, V3 poi) {
    double L = norm(p2 - p1), d1 = norm(p1 - poi), d2 = norm(p2 - poi);
    if (d1 < CORE * L || d2 < CORE * L) return true;
    bool outside = d1 > L || d2 > L;
    double dist = norm(cross(p1 - poi, p2 - poi)) / L;
    return dist < CORE * L && !outside;
}

// panel::onPanelCheck (panel.cpp:140-153): the point's projection lies
// strictly inside the panel (the reference decides this with a convex hull
// of the local vertices plus the point, boundary points kept, which is this
// test for a convex panel) and its local z is below 5% of the longest side
bool on_panel(const PanelGeo& g, V3 poi) {
    V3 P = to_local(g, poi);
    for (int k = 0; k < g.nv; k++) {
        V3 a = to_local(g, g.p[k]), b = to_local(g, g.p[(k + 1) % g.nv]);
        double cz = (b.x - a.x) * (P.y - a.y) - (b.y - a.y) * (P.x - a.x);
        if (!(cz > 0)) return false;
    }
    return P.z < 0.05 * g.longSide;
}

// panel::velocityGradientDoublet (panel.cpp:842-874)
M3 vg_panel_doublet(const PanelGeo& g, double mu, V3 poi) {
    M3 M = m3_zero();
    for (int i = 0; i < g.nv; i++) {
        V3 p1 = g.p[i], p2 = g.p[(i + 1) % g.nv];
        V3 a = poi - p1, b = poi - p2, s = p2 - p1;
        bool onp = on_panel(g, poi);
        bool nearf = near_filament(p1, p2, poi);
        if (!nearf) {
            M3 d = grad_doub(a, b, s);
            for (int r = 0; r < 3; r++)
                for (int c = onp ? 2 : 0; c < 3; c++) M.a[r][c] += d.a[r][c];
        }
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M.a[r][c] *= mu / (4 * PI);
    return M;
}

M3 add3(M3 x, const M3& y) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) x.a[r][c] += y.a[r][c];
    return x;
}

// bodyPanel / wakePanel ::partStretching velocity-gradient matrix
M3 part_stretch_matrix(const PanelGeo& g, bool body, double sigma, double mu, V3 pos) {
    bool far = norm(pos - g.center) / g.longSide > 5;
    if (body) {
        if (far) return add3(vg_point_source(g, sigma, pos), vg_point_doublet(g, mu, pos));
        return add3(vg_panel_source(g, sigma, pos), vg_panel_doublet(g, mu, pos));
    }
    return far ? vg_point_doublet(g, mu, pos) : vg_panel_doublet(g, mu, pos);
}

}  // namespace

extern "C" {

/* velocity-gradient matrix M (row-major 3x3) of one panel at pos:
 * kind 0: partStretching's (body: source sigma + doublet mu; wake: doublet
 * mu), 1: point source, 2: panel source (local frame), 3: point doublet,
 * 4: panel doublet, 5: gradDoub of edge 0 */
void or_velgrad(int nv, const double* v, int body, int kind, double sigma, double mu, const double* pos,
                double* out) {
    PanelGeo g = make_panel(nv, v);
    V3 p{pos[0], pos[1], pos[2]};
    M3 M;
    switch (kind) {
        case 1: M = vg_point_source(g, sigma, p); break;
        case 2: M = vg_panel_source(g, sigma, p); break;
        case 3: M = vg_point_doublet(g, mu, p); break;
        case 4: M = vg_panel_doublet(g, mu, p); break;
        case 5: M = grad_doub(p - g.p[0], p - g.p[1], g.p[1] - g.p[0]); break;
        default: M = part_stretch_matrix(g, body != 0, sigma, mu, p); break;
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) out[3 * r + c] = M.a[r][c];
}

/* sum over panels of partStretching on m particles: out[3t..] = sum_j M_j alpha_t */
void or_panels_stretching(int nb, const int* nvb, const double* vb, const double* sigma, const double* mu, int nw,
                          const int* nvw, const double* vw, const double* mu_w, int m, const double* pos,
                          const double* alpha, double* out) {
    std::vector<PanelGeo> B(nb), W(nw);
    for (int j = 0; j < nb; j++) B[j] = make_panel(nvb[j], vb + 12 * j);
    for (int j = 0; j < nw; j++) W[j] = make_panel(nvw[j], vw + 12 * j);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < m; t++) {
        V3 p{pos[3 * t], pos[3 * t + 1], pos[3 * t + 2]};
        double al[3] = {alpha[3 * t], alpha[3 * t + 1], alpha[3 * t + 2]};
        double acc[3] = {0, 0, 0};
        auto apply = [&](const M3& M) {
            for (int r = 0; r < 3; r++) acc[r] += M.a[r][0] * al[0] + M.a[r][1] * al[1] + M.a[r][2] * al[2];
        };
        for (int j = 0; j < nb; j++) apply(part_stretch_matrix(B[j], true, sigma[j], mu[j], p));
        for (int j = 0; j < nw; j++) apply(part_stretch_matrix(W[j], false, 0.0, mu_w[j], p));
        out[3 * t] = acc[0];
        out[3 * t + 1] = acc[1];
        out[3 * t + 2] = acc[2];
    }
}

}  // extern "C"
