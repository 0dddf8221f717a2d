// Panel -> particle vortex stretching (bodyPanel::partStretching,
// bodyPanel.cpp:697-1063; wakePanel::partStretching, wakePanel.cpp:266-289;
// panel.cpp:140-153, 813-962), FP64, for the gfx950 kernel and the host.
//
// G(i, j) = d u_j / d x_i of one panel at the particle (the reference's
// "dudx dvdx dwdx / ..." layout) and the stretching is G alpha.  Formulas as
// the reference writes them, including its frames: the near-field source
// gradient is taken in the panel's local frame and applied to the global
// strength; the far-field point source/doublet gradients use the panel
// centre's z; the source's d v / d y far-field entry is kept as written.
#pragma once

#include "influence.h"

namespace cpgpu {

struct Mat3 {
    double a[3][3];
};

CP_HD void mat3_zero(Mat3& M) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M.a[i][j] = 0;
}

// global -> local (l, m, n rows) of p - centre
CP_HD void to_local(const PanelRec& P, const double* p, double* q) {
#pragma clang fp contract(off)
    double d0 = p[0] - P.c[0], d1 = p[1] - P.c[1], d2 = p[2] - P.c[2];
    q[0] = P.l[0] * d0 + P.l[1] * d1 + P.l[2] * d2;
    q[1] = P.m[0] * d0 + P.m[1] * d1 + P.m[2] * d2;
    q[2] = P.n[0] * d0 + P.n[1] * d1 + P.n[2] * d2;
}

// bodyPanel::velocityGradientPointSource (bodyPanel.cpp:733-763), added to G
// (R: a panel record, or any record with c and area)
template <class R>
CP_HD void vg_point_source_add(const R& P, double sigma, const double* poi, Mat3& G) {
#pragma clang fp contract(off)
    double x = P.c[0], y = P.c[1], z = P.c[2], x0 = poi[0], y0 = poi[1], z0 = poi[2];
    double dx = x - x0, dy = y - y0;
    const double q2 = dx * dx + dy * dy + z * z;
    double sdd = q2 * q2 * sqrt(q2);  // q^2.5 (the reference's pow(q, 2.5), to a few ulp)
    const double isdd = 1 / sdd;  // one division for the nine entries
    double c = sigma * P.area / (4 * kPI);
    G.a[0][0] += c * ((-2 * dx * dx + dy * dy + z * z) * isdd);
    G.a[1][0] += c * (-3 * dx * dy * isdd);
    G.a[2][0] += c * (-3 * dx * z * isdd);
    G.a[0][1] += c * (-3 * dx * dy * isdd);
    G.a[1][1] += c * ((dx * dx - dy * dy + z * z) * isdd);
    G.a[2][1] += c * (-3 * z * dy * isdd);
    G.a[0][2] += c * (-3 * dx * (z - z0) * isdd);
    G.a[1][2] += c * (-3 * dy * (z - z0) * isdd);
    G.a[2][2] += c * ((-3 * z * (z - z0) + (dx * dx + dy * dy + z * z)) * isdd);
}

// bodyPanel::velocityGradientTriSource / QuadSource (bodyPanel.cpp:766-1063):
// per edge i -> i+1 in local coordinates, added to G.  NV: the vertex count
// when known at compile time (3: a body triangle), 0 to read it from the
// record (here and below)
template <int NV = 0>
CP_HD void vg_panel_source_add(const PanelRec& P, double sigma, const double* poi, Mat3& G) {
#pragma clang fp contract(off)
    const int nv = NV > 0 ? NV : P.nv;
    double X[4], Y[4];
    for (int k = 0; k < nv; k++) {
        double q[3];
        to_local(P, &P.v[3 * k], q);
        X[k] = q[0];
        Y[k] = q[1];
    }
    double L[3];
    to_local(P, poi, L);
    const double x = L[0], y = L[1], z = L[2];
    double r[4], drx[4], dry[4], drz[4], e[4], h[4];
    for (int k = 0; k < nv; k++) {
        double q = (x - X[k]) * (x - X[k]) + (y - Y[k]) * (y - Y[k]) + z * z;
        double sq = sqrt(q);
        const double isq = 1 / sq;  // one division for the three derivatives
        r[k] = sq;
        drx[k] = (x - X[k]) * isq;
        dry[k] = (y - Y[k]) * isq;
        drz[k] = z * isq;
        e[k] = (x - X[k]) * (x - X[k]) + z * z;
        h[k] = (x - X[k]) * (y - Y[k]);
    }
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < nv; i++) {
        const int j = (i + 1 == nv) ? 0 : i + 1;
        double d = sqrt((X[j] - X[i]) * (X[j] - X[i]) + (Y[j] - Y[i]) * (Y[j] - Y[i]));
        double S = r[i] + r[j];
        double Gk[3] = {drx[i] + drx[j], dry[i] + dry[j], drz[i] + drz[j]};
        // (S + d) 2d / d / (S - d) / (S + d)^2 of the reference, as
        // (2d / d) / ((S - d)(S + d)): two divisions for the six entries, the
        // same value to rounding and the same inf / NaN where d or S - d is 0
        const double cc = ((2 * d) / d) / ((S - d) * (S + d));
        for (int k = 0; k < 3; k++) {
            M[k][0] += (Y[j] - Y[i]) * Gk[k] * cc;
            M[k][1] += (X[i] - X[j]) * Gk[k] * cc;
        }
        const double mm = (Y[j] - Y[i]) / (X[j] - X[i]);
        for (int t = 0; t < 2; t++) {
            const int q = t == 0 ? i : j;
            const double sg = t == 0 ? 1.0 : -1.0;
            const double me = mm * e[q] - h[q];
            const double zr = z * r[q];
            const double mz = me / zr;
            // f / (zr zr) once for the three entries
            const double fz = (1 / (1 + mz * mz)) / (zr * zr);
            M[0][2] += sg * ((zr * (mm * 2 * (x - X[q]) - (y - Y[q])) - me * (z * drx[q])) * fz);
            M[1][2] += sg * ((-z * r[q] * (x - X[q]) - me * (z * dry[q])) * fz);
            M[2][2] += sg * ((zr * mm * (2 * z) - me * (r[q] + z * drz[q])) * fz);
        }
    }
    const double c = sigma / (4 * kPI);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) G.a[a][b] += M[a][b] * c;
}

// panel::velocityGradientPointDoublet (panel.cpp:813-840), added to G
template <class R>
CP_HD void vg_point_doublet_add(const R& P, double mu, const double* poi, Mat3& G) {
#pragma clang fp contract(off)
    double x = P.c[0], y = P.c[1], z = P.c[2], x0 = poi[0], y0 = poi[1];
    double dx2 = (x - x0) * (x - x0), dy2 = (y - y0) * (y - y0);
    const double q2 = dx2 + dy2 + z * z;
    double ddd = q2 * q2 * q2 * sqrt(q2);  // q^3.5 (the reference's pow(q, 3.5), to a few ulp)
    const double iddd = 1 / ddd;  // one division for the nine entries
    double uv = 3 * mu * P.area / (4 * kPI);
    double w = -mu * P.area / (4 * kPI);
    G.a[0][0] += uv * z * (-4 * dx2 + dy2 + z * z) * iddd;
    G.a[1][0] += uv * (-5) * z * (y - y0) * (x - x0) * iddd;
    G.a[2][0] += uv * (x - x0) * (dx2 + dy2 - 4 * z * z) * iddd;
    G.a[0][1] += uv * (-5) * (x - x0) * (y - y0) * z * iddd;
    G.a[1][1] += uv * z * (dx2 - 4 * dy2 + z * z) * iddd;
    G.a[2][1] += uv * (y - y0) * (dx2 + dy2 - 4 * z * z) * iddd;
    G.a[0][2] += w * (-3) * (x - x0) * (dx2 + dy2 - 4 * z * z) * iddd;
    G.a[1][2] += w * (-3) * (y - y0) * (dx2 + dy2 - 4 * z * z) * iddd;
    G.a[2][2] += w * (-3) * z * (3 * dx2 + 3 * dy2 - 2 * z * z) * iddd;
}

// panel::gradDoub (panel.cpp:876-923) scaled by s and added to G; only
// column j0.. of each row when `wcol_only` (on-panel points)
CP_HD void grad_doub_add(const double* a, const double* b, const double* sv, double scale, bool wcol_only, Mat3& G) {
#pragma clang fp contract(off)
    double na = sqrt(sq_norm3(a[0], a[1], a[2])), nb = sqrt(sq_norm3(b[0], b[1], b[2]));
    double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double cs = kCore * sqrt(sq_norm3(sv[0], sv[1], sv[2]));
    double top[3] = {(na + nb) * axb[0], (na + nb) * axb[1], (na + nb) * axb[2]};
    double bot = na * nb * (na * nb + ab) + cs * cs;
    // one reciprocal each for |a|, |b| and bot^2 in place of the reference's
    // fifteen divisions (the same values to rounding, the same inf / NaN
    // where a norm or bot is 0)
    const double ina = 1 / na, inb = 1 / nb, ibot2 = 1 / (bot * bot);
    for (int k = 0; k < 3; k++) {
        // e_k x b + a x e_k
        double ek[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
        double c1[3] = {ek[1] * b[2] - ek[2] * b[1], ek[2] * b[0] - ek[0] * b[2], ek[0] * b[1] - ek[1] * b[0]};
        double c2[3] = {a[1] * ek[2] - a[2] * ek[1], a[2] * ek[0] - a[0] * ek[2], a[0] * ek[1] - a[1] * ek[0]};
        double dan = a[k] * ina, dbn = b[k] * inb;
        double dbot = dan * nb * (na * nb + ab) + na * dbn * (na * nb + ab) + na * nb * (dan * nb + na * dbn + b[k] + a[k]);
        for (int j = wcol_only ? 2 : 0; j < 3; j++) {
            double dtop = (na + nb) * (c1[j] + c2[j]) + (dan + dbn) * axb[j];
            G.a[k][j] += scale * ((bot * dtop - dbot * top[j]) * ibot2);
        }
    }
}

// panel::nearFilamentCheck (panel.cpp:925-962)
CP_HD bool near_filament(const double* p1, const double* p2, const double* poi) {
#pragma clang fp contract(off)
    double s[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double u[3] = {p1[0] - poi[0], p1[1] - poi[1], p1[2] - poi[2]};
    double v[3] = {p2[0] - poi[0], p2[1] - poi[1], p2[2] - poi[2]};
    double L = sqrt(sq_norm3(s[0], s[1], s[2]));
    double d1 = sqrt(sq_norm3(u[0], u[1], u[2])), d2 = sqrt(sq_norm3(v[0], v[1], v[2]));
    if (d1 < kCore * L || d2 < kCore * L) return true;
    bool outside = d1 > L || d2 > L;
    double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
    double dist = sqrt(sq_norm3(cx, cy, cz)) / L;
    return dist < kCore * L && !outside;
}

// panel::onPanelCheck (panel.cpp:140-153): projection strictly inside the
// panel and local z below 5 % of the longest side
template <int NV = 0>
CP_HD bool on_panel(const PanelRec& P, const double* poi) {
#pragma clang fp contract(off)
    double q[3];
    to_local(P, poi, q);
    const int nv = NV > 0 ? NV : P.nv;
    for (int k = 0; k < nv; k++) {
        const int k2 = (k + 1 == nv) ? 0 : k + 1;
        double a[3], b[3];
        to_local(P, &P.v[3 * k], a);
        to_local(P, &P.v[3 * k2], b);
        double cz = (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0]);
        if (!(cz > 0)) return false;
    }
    return q[2] < 0.05 * P.long_side;
}

// panel::velocityGradientDoublet (panel.cpp:842-874), added to G
template <int NV = 0>
CP_HD void vg_panel_doublet_add(const PanelRec& P, double mu, const double* poi, Mat3& G) {
#pragma clang fp contract(off)
    const bool onp = on_panel<NV>(P, poi);
    const int nv = NV > 0 ? NV : P.nv;
    const double scale = mu / (4 * kPI);
    for (int i = 0; i < nv; i++) {
        const int j = (i + 1 == nv) ? 0 : i + 1;
        const double* p1 = &P.v[3 * i];
        const double* p2 = &P.v[3 * j];
        if (near_filament(p1, p2, poi)) continue;
        double a[3] = {poi[0] - p1[0], poi[1] - p1[1], poi[2] - p1[2]};
        double b[3] = {poi[0] - p2[0], poi[1] - p2[1], poi[2] - p2[2]};
        double s[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        grad_doub_add(a, b, s, scale, onp, G);
    }
}

// part_stretch_add for a point the far test has excluded (the near pass of
// the panel sums queues only such pairs)
template <int NV>
CP_HD void part_stretch_near_add(const PanelRec& P, bool body, double sigma, double mu, const double* pos, Mat3& G) {
    if (body) vg_panel_source_add<NV>(P, sigma, pos, G);
    vg_panel_doublet_add<NV>(P, mu, pos, G);
}

// partStretching gradient of one panel at pos, added to G
CP_HD void part_stretch_add(const PanelRec& P, bool body, double sigma, double mu, const double* pos, Mat3& G) {
#pragma clang fp contract(off)
    double d0 = pos[0] - P.c[0], d1 = pos[1] - P.c[1], d2 = pos[2] - P.c[2];
    const bool far = is_far(P, sq_norm3(d0, d1, d2));
    if (body) {
        if (far) {
            vg_point_source_add(P, sigma, pos, G);
            vg_point_doublet_add(P, mu, pos, G);
        } else {
            vg_panel_source_add(P, sigma, pos, G);
            vg_panel_doublet_add(P, mu, pos, G);
        }
    } else if (far) {
        vg_point_doublet_add(P, mu, pos, G);
    } else {
        vg_panel_doublet_add(P, mu, pos, G);
    }
}

}  // namespace cpgpu
