// Host precompute of the post-processing tables (see post.h).
#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>

#include "panelrec.h"
#include "post.h"

// Host precompute follows the reference operation order without FMA
// contraction: CHTLS fits on near-flat clusters are ill conditioned and
// amplify last-bit differences.
#pragma clang fp contract(off)

namespace cpgpu {

namespace {

// Monomial exponent lists of a Taylor series of order q in N variables, in
// the order of chtlsnd::derivSequence (chtlsnd.cpp:212-249): the last
// variable's power is the outer loop, the first variable's the inner loop.
void monomials(int q, int N, std::vector<std::vector<int>>& out) {
    out.clear();
    if (q == 0) {
        out.push_back(std::vector<int>(N, 0));
        return;
    }
    for (int i = 0; i <= q; i++) {
        if (N > 1) {
            std::vector<std::vector<int>> sub;
            monomials(q - i, N - 1, sub);
            for (auto& r : sub) {
                r.push_back(i);
                out.push_back(r);
            }
        } else {
            out.push_back({i});
        }
    }
}

// The monomial lists of chtls_weights, sorted by total degree with the
// constant term removed, built once per (order, N) and kept.
const std::vector<std::vector<int>>& sorted_monomials(int order, int N) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::vector<std::vector<int>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({order, N});
    if (it != cache.end()) return it->second;
    std::vector<std::vector<int>> ms;
    monomials(order, N, ms);
    std::stable_sort(ms.begin(), ms.end(), [](const std::vector<int>& a, const std::vector<int>& b) {
        int sa = 0, sb = 0;
        for (int v : a) sa += v;
        for (int v : b) sb += v;
        return sa < sb;
    });
    ms.erase(ms.begin());
    return cache.emplace(std::make_pair(order, N), std::move(ms)).first->second;
}

// x^e for the small exponents of the Taylor monomials.  pow returns x^0 = 1
// and x^1 = x exactly, so answering those two without the call changes no
// bit of the fit; larger exponents go through pow as before.
inline double ipow(double x, int e) { return e == 0 ? 1.0 : (e == 1 ? x : std::pow(x, e)); }

int factorial_i(int i) {
    int r = 1;
    for (int k = 2; k <= i; k++) r *= k;
    return r;
}

// Derivative weights of the CHTLS fit (chtlsnd.cpp:20-209, unconstrained):
// rows = monomials sorted by total degree (stable, constant term removed),
// columns = function observations.  Solved by Householder QR of the normal
// matrix, without pivoting or rank check, as Eigen's HouseholderQR::solve.
void chtls_weights(int N, const double* X0, int nf, const double* Xf, int order, int nb, const double* Vb,
                   std::vector<double>& F, int& t) {
    std::vector<double> dXf((size_t)nf * N), dXb((size_t)nb * N);
    double rmax = 0;
    for (int i = 0; i < nf; i++) {
        double s = 0;
        for (int k = 0; k < N; k++) {
            double d = Xf[i * N + k] - X0[k];
            dXf[i * N + k] = d;
            s += d * d;
        }
        rmax = std::max(rmax, std::sqrt(s));
    }
    for (auto& v : dXf) v /= rmax;
    // derivative observation points coincide with the function points here
    for (int i = 0; i < nb; i++)
        for (int k = 0; k < N; k++) dXb[i * N + k] = (Xf[i * N + k] - X0[k]) / rmax;
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
    const std::vector<std::vector<int>>& ms = sorted_monomials(order, N);
    t = (int)ms.size();
    std::vector<int> deg(t);
    std::vector<double> Am((size_t)nf * t), Bm((size_t)nb * t);
    for (int k = 0; k < t; k++) {
        double fct = 1;
        int dsum = 0;
        for (int i = 0; i < N; i++) {
            fct /= factorial_i(ms[k][i]);
            dsum += ms[k][i];
        }
        deg[k] = dsum;
        for (int i = 0; i < nf; i++) {
            double prod = 1;
            for (int a = 0; a < N; a++) prod *= ipow(dXf[i * N + a], ms[k][a]);
            Am[i + (size_t)k * nf] = fct * prod;
        }
        for (int i = 0; i < nb; i++) {
            double acc = 0;
            for (int a = 0; a < N; a++) {
                if (ms[k][a] == 0) continue;
                double prod = 1;
                for (int b = 0; b < N; b++) prod *= ipow(dXb[i * N + b], ms[k][b] - (b == a ? 1 : 0));
                acc += Vb[i * N + a] * ms[k][a] * prod;
            }
            Bm[i + (size_t)k * nb] = fct * acc;
        }
    }
    std::vector<double> M((size_t)t * t), R((size_t)t * nf);
    for (int c = 0; c < t; c++)
        for (int r = 0; r < t; r++) {
            double s1 = 0, s2 = 0;
            for (int i = 0; i < nf; i++) s1 += Am[i + (size_t)r * nf] * Wf[i] * Am[i + (size_t)c * nf];
            for (int i = 0; i < nb; i++) s2 += Bm[i + (size_t)r * nb] * Wb[i] * Bm[i + (size_t)c * nb];
            M[r + (size_t)c * t] = s1 + s2;
        }
    for (int j = 0; j < nf; j++)
        for (int r = 0; r < t; r++) R[r + (size_t)j * t] = Am[j + (size_t)r * nf] * Wf[j];
    // Householder QR of M applied to R, then back substitution
    std::vector<double> tau(t);
    const double tiny = std::numeric_limits<double>::min();
    for (int k = 0; k < t; k++) {
        double tail = 0;
        for (int i = k + 1; i < t; i++) tail += M[i + (size_t)k * t] * M[i + (size_t)k * t];
        double c0 = M[k + (size_t)k * t], beta;
        if (k == t - 1 || tail <= tiny) {
            tau[k] = 0;
            beta = c0;
            for (int i = k + 1; i < t; i++) M[i + (size_t)k * t] = 0;
        } else {
            beta = std::sqrt(c0 * c0 + tail);
            if (c0 >= 0) beta = -beta;
            for (int i = k + 1; i < t; i++) M[i + (size_t)k * t] /= (c0 - beta);
            tau[k] = (beta - c0) / beta;
        }
        M[k + (size_t)k * t] = beta;
        // Eigen's applyHouseholderOnTheLeft: tmp = essential^T * bottom, then
        // tmp += row(0) (this order matters for ill-conditioned fits)
        auto reflect = [&](double* col) {
            double s = 0;
            for (int i = k + 1; i < t; i++) s += M[i + (size_t)k * t] * col[i];
            s += col[k];
            col[k] -= tau[k] * s;
            for (int i = k + 1; i < t; i++) col[i] -= tau[k] * M[i + (size_t)k * t] * s;
        };
        if (tau[k] != 0) {
            for (int j = k + 1; j < t; j++) reflect(&M[(size_t)j * t]);
        }
    }
    for (int j = 0; j < nf; j++) {
        double* col = &R[(size_t)j * t];
        for (int k = 0; k < t; k++) {
            if (tau[k] == 0) continue;
            double s = 0;
            for (int i = k + 1; i < t; i++) s += M[i + (size_t)k * t] * col[i];
            s += col[k];
            col[k] -= tau[k] * s;
            for (int i = k + 1; i < t; i++) col[i] -= tau[k] * M[i + (size_t)k * t] * s;
        }
        for (int i = t - 1; i >= 0; i--) {
            double xi = col[i] / M[i + (size_t)i * t];
            col[i] = xi;
            for (int r = 0; r < i; r++) col[r] -= xi * M[r + (size_t)i * t];
        }
    }
    for (int i = 0; i < t; i++) {
        double sc = std::pow(rmax, deg[i]);
        for (int j = 0; j < nf; j++) R[i + (size_t)j * t] /= sc;
    }
    F.swap(R);
}

bool finite_all(const double* p, int n) {
    for (int i = 0; i < n; i++)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// A cluster without neighbours or with a zero-order fit has no derivative
// weights (the reference indexes an empty matrix there); NaN weights route the
// panel to the last-resort fallback velocity.
bool degenerate_fit(const Geometry& G, int i, int nf, double* w) {
    if (nf > 0 && G.ts_order[i] >= 1) return false;
    for (int k = 0; k < 3 * std::max(nf, 1); k++) w[k] = std::numeric_limits<double>::quiet_NaN();
    return true;
}

// 2D weights in the panel's local frame mapped back to global (velocity2D)
void weights_2d(const Geometry& G, int i, const int* cl, int nf, double* w) {
    if (degenerate_fit(G, i, nf, w)) return;
    const PanelGeom& p = G.body[i];
    PanelRec r = make_panel_rec(p);
    std::vector<double> Xf(2 * nf);
    for (int k = 0; k < nf; k++) {
        const Vec3 c = G.body[cl[k]].center;
        double d0 = c.x - p.center.x, d1 = c.y - p.center.y, d2 = c.z - p.center.z;
        Xf[2 * k] = r.l[0] * d0 + r.l[1] * d1 + r.l[2] * d2;
        Xf[2 * k + 1] = r.m[0] * d0 + r.m[1] * d1 + r.m[2] * d2;
    }
    double X0[2] = {0, 0};
    std::vector<double> F;
    int t;
    chtls_weights(2, X0, nf, Xf.data(), G.ts_order[i], 0, nullptr, F, t);
    for (int k = 0; k < nf; k++) {
        double f0 = F[0 + (size_t)k * t], f1 = F[1 + (size_t)k * t];
        for (int a = 0; a < 3; a++) w[3 * k + a] = f0 * r.l[a] + f1 * r.m[a];
    }
}

void weights_3d(const Geometry& G, int i, const int* cl, int nf, double* w) {
    if (degenerate_fit(G, i, nf, w)) return;
    const PanelGeom& p = G.body[i];
    std::vector<double> Xf(3 * nf), Vb(3 * nf);
    for (int k = 0; k < nf; k++) {
        const PanelGeom& q = G.body[cl[k]];
        Xf[3 * k] = q.center.x;
        Xf[3 * k + 1] = q.center.y;
        Xf[3 * k + 2] = q.center.z;
        Vb[3 * k] = q.bez.x;
        Vb[3 * k + 1] = q.bez.y;
        Vb[3 * k + 2] = q.bez.z;
    }
    double X0[3] = {p.center.x, p.center.y, p.center.z};
    std::vector<double> F;
    int t;
    chtls_weights(3, X0, nf, Xf.data(), G.ts_order[i], nf, Vb.data(), F, t);
    for (int k = 0; k < nf; k++)
        for (int a = 0; a < 3; a++) w[3 * k + a] = F[a + (size_t)k * t];
}

// wake::pntInWake (wake.cpp:601-640) over the wake's TE panels
Vec3 point_in_wake(const Geometry& G, const WakeInfo& W, double x, double y) {
    for (int wj : W.panels) {
        if (!G.wake_te[wj]) continue;
        int up = G.wake_upper[wj];
        for (int k = 0; k < 3; k++) {
            int n1 = G.body_te_edges[6 * up + 2 * k], n2 = G.body_te_edges[6 * up + 2 * k + 1];
            if (n1 < 0) continue;
            Vec3 p1 = G.nodes[n1], p2 = G.nodes[n2];
            if ((p1.y <= y && p2.y >= y) || (p1.y >= y && p2.y <= y)) {
                double tt = (y - p1.y) / (p2.y - p1.y);
                Vec3 pnt = {p1.x + tt * (p2.x - p1.x), p1.y + tt * (p2.y - p1.y), p1.z + tt * (p2.z - p1.z)};
                Vec3 nn = G.wake[wj].normal;
                // out = -(n x yhat)
                Vec3 out = {nn.z, 0.0, -nn.x};
                if (out.x < 0) out = {-out.x, -out.y, -out.z};
                double sc = (x - pnt.x) / out.x;
                return {pnt.x + sc * out.x, pnt.y + sc * out.y, pnt.z + sc * out.z};
            }
        }
    }
    return {0, 0, 0};
}

}  // namespace

void build_post_tables(const Geometry& G, PostHost& P) {
    const int n = (int)G.body.size();
    P.n = n;
    P.nbr.assign(16 * (size_t)n, 0);
    P.w1.assign(45 * (size_t)n, 0.0);
    P.w2.assign(45 * (size_t)n, 0.0);
    P.mode.assign(n, 0);
    parallel_for(n, 0, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            const int* cl = &G.cluster[16 * (size_t)i];
            int cnt = 0, idx[16];
            // velocity2D/3D drop the panel itself (the point is its centre)
            for (int k = 1; k < 16 && cl[k] >= 0; k++) idx[cnt++] = cl[k];
            P.nbr[16 * (size_t)i] = cnt;
            for (int k = 0; k < cnt; k++) P.nbr[16 * (size_t)i + 1 + k] = idx[k];
            double* w1 = &P.w1[45 * (size_t)i];
            double* w2 = &P.w2[45 * (size_t)i];
            if (G.tip[i]) {
                weights_2d(G, i, idx, cnt, w1);
            } else {
                weights_3d(G, i, idx, cnt, w1);
                if (!finite_all(w1, 3 * cnt)) {
                    weights_2d(G, i, idx, cnt, w2);
                    P.mode[i] = 1;
                }
            }
        }
    });

    // ---- Trefftz plane functionals
    const int nw = (int)G.wake.size();
    P.n_wakes = (int)G.wakes.size();
    P.n_st = 165;
    std::map<int, int> teMap;
    for (const WakeInfo& W : G.wakes)
        for (const WakeLineInfo& wl : W.lines) {
            teMap.emplace(wl.upper, 0);
            teMap.emplace(wl.lower, 0);
        }
    for (int j = 0; j < nw; j++)
        for (int s = 0; s < 4; s++) teMap.emplace(G.fold_idx[4 * j + s], 0);
    P.te_idx.clear();
    for (auto& kv : teMap) {
        kv.second = (int)P.te_idx.size();
        P.te_idx.push_back(kv.first);
    }
    P.n_te = (int)P.te_idx.size();
    const int nst = P.n_st, nte = P.n_te, nwk = (int)G.wakes.size();
    P.rphi.assign((size_t)nwk * nst * nte, 0.0);
    P.rv.assign((size_t)nwk * nst * nte, 0.0);
    P.rw.assign((size_t)nwk * nst * nte, 0.0);
    P.simpson.assign((size_t)nwk * nst, 0.0);
    P.yloc.assign((size_t)nwk * nst, 0.0);
    if (nw == 0) return;
    std::vector<PanelRec> wrec(nw);
    for (int j = 0; j < nw; j++) wrec[j] = make_panel_rec(G.wake[j]);
    // coefficient of body mu_b in wake mu_w (wakePanel::setMu, wakePanel.cpp:94-100)
    std::vector<int> foldTe(4 * (size_t)nw);
    std::vector<double> foldC(4 * (size_t)nw);
    for (int j = 0; j < nw; j++) {
        double t = G.fold_coef[j];
        double c[4] = {1 - t, t - 1, t, -t};
        for (int s = 0; s < 4; s++) {
            foldTe[4 * j + s] = teMap[G.fold_idx[4 * j + s]];
            foldC[4 * j + s] = c[s];
        }
    }
    for (int wk = 0; wk < nwk; wk++) {
        const WakeInfo& W = G.wakes[wk];
        const int nP = 164;
        double step = (W.yMax - W.yMin) / nP;
        double xT = W.x0 + 2 * (W.xf - W.x0) / 3;
        double* yl = &P.yloc[(size_t)wk * nst];
        double* sw = &P.simpson[(size_t)wk * nst];
        yl[0] = W.yMin;
        yl[nP] = W.yMax;
        for (int i = 0; i + 2 <= nP; i += 2) {
            sw[i] += step / 3;
            sw[i + 1] += 4 * step / 3;
            sw[i + 2] += step / 3;
        }
        parallel_for(nP - 1, 0, [&](int a, int b) {
            std::vector<double> gy(nw), gz(nw);
            for (int ii = a; ii < b; ii++) {
                int i = ii + 1;
                double y = W.yMin + i * step;
                yl[i] = y;
                double* rp = &P.rphi[((size_t)wk * nst + i) * nte];
                double* rv = &P.rv[((size_t)wk * nst + i) * nte];
                double* rw = &P.rw[((size_t)wk * nst + i) * nte];
                // dPhi = -wakeStrength(y) (wake.cpp:478-506)
                const std::vector<WakeLineInfo>& L = W.lines;
                size_t k1 = 0;
                if (y < L[1].y) k1 = 0;
                else if (y >= L.back().y) k1 = L.size() - 2;
                else
                    for (size_t q = 1; q + 1 < L.size(); q++)
                        if (L[q].y <= y && L[q + 1].y > y) k1 = q;
                double it = (y - L[k1].y) / (L[k1 + 1].y - L[k1].y);
                rp[teMap[L[k1].upper]] += -(1 - it);
                rp[teMap[L[k1].lower]] += (1 - it);
                rp[teMap[L[k1 + 1].upper]] += -it;
                rp[teMap[L[k1 + 1].lower]] += it;
                // Vradial (wake.cpp:508-597): geometry of the survey points
                Vec3 pw = point_in_wake(G, W, xT, y);
                double r, th = kPI / 4, dzmax = 0.3, delZ;
                Vec3 POI;
                POI.x = pw.x;
                r = pw.y >= 0 ? W.yMax - pw.y : pw.y - W.yMin;
                delZ = r * std::sin(th);
                if (delZ > dzmax) {
                    delZ = dzmax;
                    th = std::asin(dzmax / r);
                }
                POI.y = pw.y >= 0 ? W.yMax - r * std::cos(th) : W.yMin + r * std::cos(th);
                POI.z = pw.z + r * std::sin(th);
                const int m = 6;
                double dz = 0.5 * delZ, stp = 2 * dz / (m - 1);
                double dY[m], dZ[m];
                Vec3 p1[m], p2[m];
                for (int q = 0; q < m; q++) {
                    double delta = -dz + q * stp;
                    p1[q] = {POI.x + delta * 0, POI.y + delta * 1, POI.z + delta * 0};
                    p2[q] = {POI.x + delta * 0, POI.y + delta * 0, POI.z + delta * 1};
                    dY[q] = p1[q].y - POI.y;
                    dZ[q] = p2[q].z - POI.z;
                }
                double x0 = 0;
                std::vector<double> Fy, Fz;
                int t;
                chtls_weights(1, &x0, m, dY, 3, 0, nullptr, Fy, t);
                chtls_weights(1, &x0, m, dZ, 3, 0, nullptr, Fz, t);
                // phi(P) = -sum_w mu_w D_w(P)  (geometry::wakePotential)
                for (int w = 0; w < nw; w++) {
                    double d0 = dub_phi_inf_f64(wrec[w], POI.x, POI.y, POI.z);
                    double sy = 0, sz = 0;
                    for (int q = 0; q < m; q++) {
                        sy += Fy[0 + (size_t)q * t] * (dub_phi_inf_f64(wrec[w], p1[q].x, p1[q].y, p1[q].z) - d0);
                        sz += Fz[0 + (size_t)q * t] * (dub_phi_inf_f64(wrec[w], p2[q].x, p2[q].y, p2[q].z) - d0);
                    }
                    gy[w] = -sy;
                    gz[w] = -sz;
                }
                for (int w = 0; w < nw; w++)
                    for (int s = 0; s < 4; s++) {
                        rv[foldTe[4 * w + s]] += gy[w] * foldC[4 * w + s];
                        rw[foldTe[4 * w + s]] += gz[w] * foldC[4 * w + s];
                    }
            }
        });
    }
}

}  // namespace cpgpu
