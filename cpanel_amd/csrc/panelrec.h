// Conversion of host panel geometry into the 256-byte device record.
#pragma once

#include <cmath>
#include <cstring>

#include "geometry.h"
#include "influence.h"

namespace cpgpu {

// Local frame of panel::getLocalSys (panel.cpp:156-168): l = unit(v0 - c),
// m = n x l, n = normal.
inline PanelRec make_panel_rec(const PanelGeom& p) {
    PanelRec r;
    std::memset(&r, 0, sizeof(r));
    r.nv = p.nv;
    for (int k = 0; k < p.nv; k++) {
        r.v[3 * k] = p.v[k].x;
        r.v[3 * k + 1] = p.v[k].y;
        r.v[3 * k + 2] = p.v[k].z;
    }
    r.c[0] = p.center.x;
    r.c[1] = p.center.y;
    r.c[2] = p.center.z;
    r.n[0] = p.normal.x;
    r.n[1] = p.normal.y;
    r.n[2] = p.normal.z;
    double lx = p.v[0].x - p.center.x, ly = p.v[0].y - p.center.y, lz = p.v[0].z - p.center.z;
    double s = lx * lx + ly * ly + lz * lz;
    if (s > 0) {
        double q = std::sqrt(s);
        lx /= q;
        ly /= q;
        lz /= q;
    }
    r.l[0] = lx;
    r.l[1] = ly;
    r.l[2] = lz;
    r.m[0] = r.n[1] * lz - r.n[2] * ly;
    r.m[1] = r.n[2] * lx - r.n[0] * lz;
    r.m[2] = r.n[0] * ly - r.n[1] * lx;
    r.area = p.area;
    r.long_side = p.long_side;
    double f2 = 25.0 * p.long_side * p.long_side;
    r.far2_lo = f2 * (1 - 1e-13);
    r.far2_hi = f2 * (1 + 1e-13);
    r.area_4pi = p.area / (4 * kPI);
    return r;
}

}  // namespace cpgpu
