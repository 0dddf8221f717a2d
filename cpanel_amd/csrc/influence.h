// Panel influence kernels shared by the gfx950 kernels and the host
// precompute (Trefftz functionals, CHTLS).  FP64 near-field functions keep the
// reference's operation order with FMA contraction disabled so that GPU and
// CPU agree to the last bits where the formula is well conditioned.
//
// Reference formulas: panel::vortexPhi (panel.cpp:713-735),
// bodyPanel::srcSidePhi (bodyPanel.cpp:361-373), bodyPanel::panelPhiInf
// (bodyPanel.cpp:132-190), panel::dubPhiInf (panel.cpp:208-252),
// panel::vortexV (panel.cpp:707-711), panel::pntDubV (panel.cpp:773-776),
// bodyPanel::panelVInf / srcSideV (bodyPanel.cpp:305-348, 375-390).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define CP_HD __host__ __device__ __forceinline__

namespace cpgpu {

constexpr double kPI = 3.14159265358979323846;

// One panel as the kernels read it: a 256-byte record.  Within a wavefront
// every lane reads the same record, so the loads go through the scalar cache.
struct alignas(256) PanelRec {
    double v[12];        // vertices x,y,z (nv of them)
    double c[3];         // centre
    double n[3];         // unit normal
    double l[3];         // local x: unit(vertex0 - centre) (panel::getLocalSys)
    double m[3];         // local y: n x l
    double area;
    double long_side;
    double far2_lo;      // 25*L^2*(1-1e-13): below -> near field for sure
    double far2_hi;      // 25*L^2*(1+1e-13): above -> far field for sure
    double area_4pi;     // area / (4 pi)
    int32_t nv;
    int32_t pad0;
};
static_assert(sizeof(PanelRec) == 256, "PanelRec must be 256 bytes");

// far-field test of the reference: |pjk| / longSide > 5, decided exactly
// from r^2 except inside a relative band of 1e-13 around the threshold (R: a
// panel record, or any record with far2_lo, far2_hi and long_side)
template <class R>
CP_HD bool is_far(const R& P, double r2) {
#pragma clang fp contract(off)
    if (r2 > P.far2_hi) return true;
    if (r2 < P.far2_lo) return false;
    return sqrt(r2) / P.long_side > 5;
}

CP_HD double sq_norm3(double x, double y, double z) {
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
}

// panel::vortexPhi, reference order of operations
CP_HD double vortex_phi(double PN, double Al, const double* a, const double* s, const double* l, const double* m,
                        double na, double nb) {
#pragma clang fp contract(off)
    double axs0 = a[1] * s[2] - a[2] * s[1];
    double axs1 = a[2] * s[0] - a[0] * s[2];
    double axs2 = a[0] * s[1] - a[1] * s[0];
    double q0 = l[1] * axs2 - l[2] * axs1;
    double q1 = l[2] * axs0 - l[0] * axs2;
    double q2 = l[0] * axs1 - l[1] * axs0;
    double PA = a[0] * q0 + a[1] * q1 + a[2] * q2;
    double sm = s[0] * m[0] + s[1] * m[1] + s[2] * m[2];
    double PB = PA - Al * sm;
    double num = sm * PN * (nb * PA - na * PB);
    double den = PA * PB + PN * PN * na * nb * (sm * sm);
    if (den == 0 && fabs(PN) < 1e-15) return PN >= 0 ? 0.5 * kPI : -0.5 * kPI;
    return atan2(num, den);
}

// Near-field (edge formula) source and doublet potential of a panel at a
// point; the caller has already excluded the far field.  itself: the point
// is the panel centre (|pjk| < 1e-10), doublet -> -0.5.  NV: the vertex
// count when it is known at compile time (3: a body triangle, so the
// per-vertex arrays are plain registers), 0 to read it from the record.
template <int NV>
CP_HD void near_phi_nv(const PanelRec& P, double px, double py, double pz, double& src, double& dub) {
#pragma clang fp contract(off)
    double pjk0 = px - P.c[0], pjk1 = py - P.c[1], pjk2 = pz - P.c[2];
    double r = sqrt(sq_norm3(pjk0, pjk1, pjk2));
    bool itself = r < 1e-10;
    double PN = pjk0 * P.n[0] + pjk1 * P.n[1] + pjk2 * P.n[2];
    const int nv = NV > 0 ? NV : P.nv;
    double d[4][3], nd[4];
    for (int k = 0; k < nv; k++) {
        d[k][0] = px - P.v[3 * k];
        d[k][1] = py - P.v[3 * k + 1];
        d[k][2] = pz - P.v[3 * k + 2];
        nd[k] = sqrt(sq_norm3(d[k][0], d[k][1], d[k][2]));
    }
    double s_src = 0, s_dub = 0;
    for (int k = 0; k < nv; k++) {
        int k2 = (k + 1 == nv) ? 0 : k + 1;
        const double* a = d[k];
        double s[3] = {P.v[3 * k2] - P.v[3 * k], P.v[3 * k2 + 1] - P.v[3 * k + 1], P.v[3 * k2 + 2] - P.v[3 * k + 2]};
        double sxa0 = s[1] * a[2] - s[2] * a[1];
        double sxa1 = s[2] * a[0] - s[0] * a[2];
        double sxa2 = s[0] * a[1] - s[1] * a[0];
        double Al = P.n[0] * sxa0 + P.n[1] * sxa1 + P.n[2] * sxa2;
        double phiV = 0;
        if (!itself) {
            phiV = vortex_phi(PN, Al, a, s, P.l, P.m, nd[k], nd[k2]);
            s_dub += phiV;
        }
        double A = nd[k], B = nd[k2], S = sqrt(sq_norm3(s[0], s[1], s[2]));
        double GL = 0;
        if (fabs(A + B - S) > 0 && S > 0) GL = 1 / S * log(fabs((A + B + S) / (A + B - S)));
        s_src += Al * GL - PN * phiV;
    }
    src = s_src / (4 * kPI);
    dub = itself ? -0.5 : s_dub / (4 * kPI);
}

CP_HD void near_phi_f64(const PanelRec& P, double px, double py, double pz, double& src, double& dub) {
    near_phi_nv<0>(P, px, py, pz, src, dub);
}

// a body triangle (every body panel: the .tri reader's panels are
// triangles), bit-identical to near_phi_f64 on it
CP_HD void near_phi_tri_f64(const PanelRec& P, double px, double py, double pz, double& src, double& dub) {
    near_phi_nv<3>(P, px, py, pz, src, dub);
}

// bodyPanel::panelPhiInf in FP64 (far field via the point formulas).
CP_HD void panel_phi_inf_f64(const PanelRec& P, double px, double py, double pz, double& src, double& dub) {
#pragma clang fp contract(off)
    double d0 = px - P.c[0], d1 = py - P.c[1], d2 = pz - P.c[2];
    double r2 = sq_norm3(d0, d1, d2);
    if (is_far(P, r2)) {
        double r = sqrt(r2);
        double PN = d0 * P.n[0] + d1 * P.n[1] + d2 * P.n[2];
        src = P.area / (4 * kPI * r);
        dub = PN * P.area / (4 * kPI * (r * r * r));
        return;
    }
    near_phi_f64(P, px, py, pz, src, dub);
}

// panel::dubPhiInf in FP64: self (|pjk| < 1e-7) -> -0.5; far -> point
// doublet; else the edge sum of vortexPhi.
CP_HD double dub_phi_inf_f64(const PanelRec& P, double px, double py, double pz) {
#pragma clang fp contract(off)
    double d0 = px - P.c[0], d1 = py - P.c[1], d2 = pz - P.c[2];
    double r2 = sq_norm3(d0, d1, d2);
    double r = sqrt(r2);
    double PN = d0 * P.n[0] + d1 * P.n[1] + d2 * P.n[2];
    if (r < 0.0000001) return -0.5;
    if (is_far(P, r2)) return PN * P.area / (4 * kPI * (r * r * r));
    const int nv = P.nv;
    double d[4][3], nd[4];
    for (int k = 0; k < nv; k++) {
        d[k][0] = px - P.v[3 * k];
        d[k][1] = py - P.v[3 * k + 1];
        d[k][2] = pz - P.v[3 * k + 2];
        nd[k] = sqrt(sq_norm3(d[k][0], d[k][1], d[k][2]));
    }
    double phi = 0;
    for (int k = 0; k < nv; k++) {
        int k2 = (k + 1 == nv) ? 0 : k + 1;
        double s[3] = {P.v[3 * k2] - P.v[3 * k], P.v[3 * k2 + 1] - P.v[3 * k + 1], P.v[3 * k2 + 2] - P.v[3 * k + 2]};
        const double* a = d[k];
        double sxa0 = s[1] * a[2] - s[2] * a[1];
        double sxa1 = s[2] * a[0] - s[0] * a[2];
        double sxa2 = s[0] * a[1] - s[1] * a[0];
        double Al = P.n[0] * sxa0 + P.n[1] * sxa1 + P.n[2] * sxa2;
        phi += vortex_phi(PN, Al, a, s, P.l, P.m, nd[k], nd[k2]);
    }
    return phi / (4 * kPI);
}

// ---- velocities (panel -> point) ------------------------------------------
constexpr double kCore = 0.05;  // panel::core (panel.h:67)

// panel::vortexV (panel.cpp:707-711): (a x b)(|a|+|b|) / (|a||b|(|a||b| + a.b) + (core |s|)^2)
CP_HD void vortex_v(const double* a, const double* b, const double* s, double* out) {
#pragma clang fp contract(off)
    double na = sqrt(sq_norm3(a[0], a[1], a[2]));
    double nb = sqrt(sq_norm3(b[0], b[1], b[2]));
    double cs = kCore * sqrt(sq_norm3(s[0], s[1], s[2]));
    double den = na * nb * (na * nb + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2])) + cs * cs;
    const double f = (na + nb) / den;  // one division for the three components
    out[0] = (a[1] * b[2] - a[2] * b[1]) * f;
    out[1] = (a[2] * b[0] - a[0] * b[2]) * f;
    out[2] = (a[0] * b[1] - a[1] * b[0]) * f;
}

// panel::pntDubV (panel.cpp:773-776): area (3 (pjk.n) pjk - r^2 n) / (4 pi r^5)
template <class R>
CP_HD void pnt_dub_v(const R& P, const double* pjk, double r, double* out) {
#pragma clang fp contract(off)
    double pn3 = 3 * (pjk[0] * P.n[0] + pjk[1] * P.n[1] + pjk[2] * P.n[2]);
    double r2 = r * r;
    double den = 4 * kPI * (r2 * r2 * r);  // r^5 (the reference's pow(r, 5), to a few ulp)
    const double f = P.area / den;         // one division for the three components
    for (int k = 0; k < 3; k++) out[k] = (pn3 * pjk[k] - r2 * P.n[k]) * f;
}

// the edge sums of panel_v_inf_f64 (a point the far test has excluded); NV:
// the vertex count when known at compile time (3: a body triangle), 0 to read
// it from the record
template <int NV>
CP_HD void panel_v_near_nv(const PanelRec& P, const double* p, double* vs, double* vd) {
#pragma clang fp contract(off)
    double pjk[3] = {p[0] - P.c[0], p[1] - P.c[1], p[2] - P.c[2]};
    double PN = pjk[0] * P.n[0] + pjk[1] * P.n[1] + pjk[2] * P.n[2];
    const int nv = NV > 0 ? NV : P.nv;
    double S[3] = {0, 0, 0}, D[3] = {0, 0, 0};
    for (int k = 0; k < nv; k++) {
        int k2 = (k + 1 == nv) ? 0 : k + 1;
        double a[3] = {p[0] - P.v[3 * k], p[1] - P.v[3 * k + 1], p[2] - P.v[3 * k + 2]};
        double b[3] = {p[0] - P.v[3 * k2], p[1] - P.v[3 * k2 + 1], p[2] - P.v[3 * k2 + 2]};
        double s[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
        double sxa0 = s[1] * a[2] - s[2] * a[1];
        double sxa1 = s[2] * a[0] - s[0] * a[2];
        double sxa2 = s[0] * a[1] - s[1] * a[0];
        double Al = P.n[0] * sxa0 + P.n[1] * sxa1 + P.n[2] * sxa2;
        double w[3];
        vortex_v(a, b, s, w);
        D[0] += w[0];
        D[1] += w[1];
        D[2] += w[2];
        // bodyPanel::srcSideV (bodyPanel.cpp:375-390)
        double A = sqrt(sq_norm3(a[0], a[1], a[2])), B = sqrt(sq_norm3(b[0], b[1], b[2]));
        double Sl = sqrt(sq_norm3(s[0], s[1], s[2]));
        double GL = 0;
        if (fabs(A + B - Sl) > 0 && Sl > 0) GL = 1 / Sl * log(fabs((A + B + Sl) / (A + B - Sl)));
        double CJK = vortex_phi(PN, Al, a, s, P.l, P.m, A, B);
        double sm = s[0] * P.m[0] + s[1] * P.m[1] + s[2] * P.m[2];
        double sl = s[0] * P.l[0] + s[1] * P.l[1] + s[2] * P.l[2];
        for (int q = 0; q < 3; q++) S[q] += GL * (sm * P.l[q] - sl * P.m[q]) + CJK * P.n[q];
    }
    for (int q = 0; q < 3; q++) {
        vd[q] = D[q] / (4 * kPI);
        vs[q] = S[q] / (4 * kPI);
    }
}

// bodyPanel::panelVInf (bodyPanel.cpp:305-348) in FP64: source and doublet
// velocity of a unit-strength panel at p.  Far field: point doublet and the
// point source with the reference's sign (negated, bodyPanel.cpp:314).
CP_HD void panel_v_inf_f64(const PanelRec& P, const double* p, double* vs, double* vd) {
#pragma clang fp contract(off)
    double pjk[3] = {p[0] - P.c[0], p[1] - P.c[1], p[2] - P.c[2]};
    double r2 = sq_norm3(pjk[0], pjk[1], pjk[2]);
    if (is_far(P, r2)) {
        double r = sqrt(r2);
        double den = 4 * kPI * (r * r * r);  // r^3 (the reference's pow(r, 3), to a few ulp)
        const double f = P.area / den;
        for (int k = 0; k < 3; k++) vs[k] = -1.0 * (pjk[k] * f);
        pnt_dub_v(P, pjk, r, vd);
        return;
    }
    panel_v_near_nv<0>(P, p, vs, vd);
}

// the edge sums of dub_v_inf_f64 (a point the far test has excluded)
template <int NV>
CP_HD void dub_v_near_nv(const PanelRec& P, const double* p, double* vd) {
#pragma clang fp contract(off)
    double D[3] = {0, 0, 0};
    const int nv = NV > 0 ? NV : P.nv;
    for (int k = 0; k < nv; k++) {
        int k2 = (k + 1 == nv) ? 0 : k + 1;
        double a[3] = {p[0] - P.v[3 * k], p[1] - P.v[3 * k + 1], p[2] - P.v[3 * k + 2]};
        double b[3] = {p[0] - P.v[3 * k2], p[1] - P.v[3 * k2 + 1], p[2] - P.v[3 * k2 + 2]};
        double s[3] = {P.v[3 * k2] - P.v[3 * k], P.v[3 * k2 + 1] - P.v[3 * k + 1], P.v[3 * k2 + 2] - P.v[3 * k + 2]};
        double w[3];
        vortex_v(a, b, s, w);
        D[0] += w[0];
        D[1] += w[1];
        D[2] += w[2];
    }
    for (int q = 0; q < 3; q++) vd[q] = D[q] / (4 * kPI);
}

// panel::dubVInf (panel.cpp:666-705) in FP64 (wake panels)
CP_HD void dub_v_inf_f64(const PanelRec& P, const double* p, double* vd) {
#pragma clang fp contract(off)
    double pjk[3] = {p[0] - P.c[0], p[1] - P.c[1], p[2] - P.c[2]};
    double r2 = sq_norm3(pjk[0], pjk[1], pjk[2]);
    if (is_far(P, r2)) {
        pnt_dub_v(P, pjk, sqrt(r2), vd);
        return;
    }
    dub_v_near_nv<0>(P, p, vd);
}

}  // namespace cpgpu
