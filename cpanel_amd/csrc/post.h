// Post-solve outputs: per-panel velocity / Cp / forces and the Trefftz plane.
//
// Everything that does not depend on the freestream is precomputed once per
// geometry on the host (FP64):
//   * CHTLS derivative weights of every panel's cluster
//     (bodyPanel::velocity3D/velocity2D, bodyPanel.cpp:526-589; chtlsnd.cpp),
//     so the surface velocity is a 16-point weighted sum of potentials;
//   * the Trefftz-plane functionals (wake::trefftzPlane, wake.cpp:146-188),
//     which are linear in the trailing-edge doublet strengths: per station,
//     wakeStrength (wake.cpp:478-506) and the two CHTLS potential
//     derivatives of Vradial (wake.cpp:508-597) become fixed rows over the
//     trailing-edge panels.
// The per-condition work then runs on the device for any number of
// conditions at once (post.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "geometry.h"

namespace cpgpu {

constexpr int kClusterMax = 15;  // cluster size minus the panel itself

struct PostHost {
    int n = 0;
    // velocity weights: per panel up to 15 neighbours, 3 weights each
    std::vector<int32_t> nbr;      // 16*n: [count, idx0..idx14]
    std::vector<double> w1;        // 45*n primary weights (x,y,z per neighbour)
    std::vector<double> w2;        // 45*n secondary (2D) weights, used when w1 gives NaN
    std::vector<int32_t> mode;     // 0: w1 only; 1: w1 then w2 on NaN
    // Trefftz rows over the trailing-edge panel set
    int n_te = 0;
    std::vector<int32_t> te_idx;   // body indices of the TE set
    int n_st = 0;                  // stations per wake (165 points incl. ends)
    std::vector<double> rphi, rv, rw;  // n_st x n_te each (row major)
    std::vector<double> simpson;   // n_st Simpson weights (step/3 * {1,4,2,...,4,1})
    std::vector<double> yloc;
    int n_wakes = 0;
};

// Builds the post-processing tables (CPU threads).
void build_post_tables(const Geometry& G, PostHost& out);

// Device copy of PostHost plus per-panel geometry needed by the kernel.
struct PostDev {
    int n = 0, n_te = 0, n_st = 0, n_wakes = 0;
    int32_t* nbr = nullptr;
    double* w1 = nullptr;
    double* w2 = nullptr;
    int32_t* mode = nullptr;
    int32_t* te_idx = nullptr;
    double* rphi = nullptr;
    double* rv = nullptr;
    double* rw = nullptr;
    double* simpson = nullptr;
    double* ctr = nullptr;   // 3*n centres (SoA: x[n], y[n], z[n])
    double* bez = nullptr;   // 3*n bezier normals (SoA)
    double* area = nullptr;  // n
};

struct PostParams {
    double PG, Sref, bref, cref, cg[3];
};

// Optional outputs of launch_post (device pointers, all may be null):
//   vel   3 per panel (x, y, z; V_x divided by the PG factor), condition c at
//         vel + c*ldvel;
//   phi   potential Vinf . c_i - mu_i, condition c at phi + c*ldphi;
//   span  per condition, wake and Trefftz station: (Cl / PG, Cdi / PG^2),
//         at span + ((c*n_wakes + w)*n_st + st)*2.
//   vpot  input: the velocity each panel's potential is formed with (3 per
//         panel, condition c at vpot + 3*c*n) instead of the freestream.
struct PostFields {
    const double* vpot = nullptr;
    double* vel = nullptr;
    int64_t ldvel = 0;
    double* phi = nullptr;
    int64_t ldphi = 0;
    double* span = nullptr;
};

// mu: device, column c at mu + c*ldmu (n_body rows); Vinf3: device 4*k
// (Vx, Vy, Vz, |V|).  Outputs: cp (device, n*k, may be null), res (device,
// 16*k: CL, CDi, Fbody[3], Fwind[3] filled on host, CM[3]).
void launch_post(const PostDev& P, const PostParams& pp, int k, const double* mu, int64_t ldmu,
                 const double* vinf4, double* cp, int64_t ldcp, double* partials, int npart_blocks,
                 double* res, hipStream_t s, const PostFields* fields = nullptr);
// Wake doublet strengths (Kutta fold of the body mu: fidx 4 body indices and
// ft the interpolation coefficient t per wake panel) and wake potentials
// (wctr: wake centres SoA, 3*nw) for k conditions; outputs nw per condition.
void launch_post_wake(int nw, const int32_t* fidx, const double* ft, const double* wctr, int k, const double* mu,
                      int64_t ldmu, const double* vinf4, double* wmu, double* wphi, hipStream_t s);
int post_partial_blocks(int n);

}  // namespace cpgpu
