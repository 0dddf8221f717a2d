// K4 / K5 device tables and launchers (vp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "cpgpu.h"
#include "device.h"
#include "vp_tree.h"

namespace cpgpu {

// device copy of a flattened particle octree (vp_tree.h)
struct DevBhTree {
    int n = 0;            // particles
    int n_nodes = 0;
    BhNode* nodes = nullptr;
    double* mpos = nullptr;   // members in leaf order, 3 per slot
    double* mstr = nullptr;
    double* mrad = nullptr;
    int* order = nullptr;     // slot -> particle index
};

// mode CP_BH_VEL_POINTS: m points pts[3m] -> out[3m]
void launch_bh_points(const DevBhTree& T, double theta, int m, const double* pts, double* out, hipStream_t s);
// modes CP_BH_VEL / CP_BH_STRETCH / CP_BH_DIFF on the tree's particles -> out[3n]
void launch_bh_particles(const DevBhTree& T, int mode, double theta, double* out, hipStream_t s);

// K slices of the panel -> point evaluation for m points and np panels
int p2p_slices(int m, int npanels);
// out[3m] = velocity at pts of the body (sigma, mu) and wake (muw) panels;
// part: KS * 3m scratch
void launch_panel_to_points(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                            const double* pts, int KS, double* part, double* out, hipStream_t s);

// out[3m] = sum over body and wake panels of partStretching at the particles
// (pos, alpha: 3 per particle); part: KS * 3m scratch
void launch_panel_stretching(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                             const double* pos, const double* alpha, int KS, double* part, double* out,
                             hipStream_t s);
// y[0..n) += a[0..n)
void launch_add_vec(int n, const double* a, double* y, hipStream_t s);
// out (column-major, ld) = -C mu for k columns: C is lda x m (FP64), mu m x k
void launch_buffer_rhs(int n, int64_t lda, int m, const double* Cm, const double* mu, int k, double* out, int64_t ld,
                       hipStream_t s);

}  // namespace cpgpu
