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
// caller-given targets up to which the tree walks run a wavefront per
// target (above: a lane per target)
constexpr int kBhWaveTargets = 16384;
void launch_bh_particles(const DevBhTree& T, int mode, double theta, double* out, hipStream_t s);

// modes CP_BH_VEL / CP_BH_STRETCH / CP_BH_DIFF on m caller-given targets
// (tpos, tstr: 3 per target; trad; tslot: the target's slot in T or -1)
void launch_bh_targets(const DevBhTree& T, int mode, double theta, int m, const double* tpos, const double* tstr,
                       const double* trad, const int* tslot, double* out, hipStream_t s);
// T's members take the current state of their particles (particle-indexed
// pos, str: 3 each; rad), keeping the tree's nodes and pseudo-particles
void launch_bh_refresh(const DevBhTree& T, const double* pos, const double* str, const double* rad, hipStream_t s);
// direct pairwise stretching (CP_BH_STRETCH) or diffusion (CP_BH_DIFF) of n
// particles, each particle the called object of the pair functions
void launch_pairs_direct(int mode, int n, const double* pos, const double* str, const double* rad, double* out,
                         hipStream_t s);
// Direct sums over n particles at m points split the particles into
// direct_slices(m, n) slices when m alone leaves the device idle (a few
// hundred Trefftz points against 80,000 particles), adding the slices'
// partial sums in order; part: direct_part_doubles(m, n) doubles of scratch.
int direct_slices(int m, int n);
size_t direct_part_doubles(int m, int n);
// direct Gaussian point velocity of n particles at m points
void launch_gauss_points(int n, const double* pos, const double* str, const double* rad, int m, const double* pts,
                         double* part, double* out, hipStream_t s);
// vortex-filament velocity (ends: 6 per filament, fstr) at m points
void launch_filament_points(int nf, const double* ends, const double* fstr, int m, const double* pts, double* out,
                            hipStream_t s);

// K slices of the far pass of the panel sums for m points and np panels
int p2p_slices(int m, int npanels);
// doubles of the `part` scratch the two launchers below need for m points
// and npanels panels (far and near slices, the point tiles, the spatial
// order and its sort)
size_t p2p_part_doubles(int m, int npanels);
// out[3m] = velocity at pts of the body (sigma, mu) and wake (muw) panels;
// f32far: far-field pairs in FP32 arithmetic (near field FP64)
void launch_panel_to_points(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                            const double* pts, double* part, double* out, hipStream_t s, bool f32far = false);
// out[3m] = sum over body and wake panels of partStretching at the particles
// (pos, alpha: 3 per particle)
void launch_panel_stretching(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                             const double* pos, const double* alpha, double* part, double* out, hipStream_t s,
                             bool f32far = false);
// Spatial order of m points (vp_build.hip): perm[k] = the point at position
// k of the order of their 30-bit Morton keys over the points' bounding box
// (stable); scratch: morton_sort_bytes(m) bytes.
size_t morton_sort_bytes(int m);
void sort_points_morton(int m, const double* pts, void* scratch, int* perm, hipStream_t s);

// The algebraic particle kernel (velInflAlgSmooth with one core radius for
// every particle): mode CP_BH_VEL (velocity; tstr unused) or CP_BH_STRETCH
// ((alpha . grad) u), in FP32 (f32) or FP64 arithmetic.
// Barnes-Hut on tree T at m targets: caller-given (tpos, tstr; tslot[t] the
// target's slot in T, -1 none, or tslot null) with out[3m] in target order,
// or T's own members (tpos = T.mpos, tstr = T.mstr, m = T.n, own_order =
// T.order) with out[3n] in particle order.
void launch_core_bh(const DevBhTree& T, int mode, double theta, double core, bool f32, int m, const double* tpos,
                    const double* tstr, const int* tslot, const int* own_order, double* out, hipStream_t s);
// direct sums over n particles at m targets; own: the targets are the
// particles themselves (m = n), each skipping itself
void launch_core_direct(int mode, double core, bool f32, int n, const double* pos, const double* str, int m,
                        const double* tpos, const double* tstr, bool own, double* part, double* out, hipStream_t s);
// y[0..n) += a[0..n)
void launch_add_vec(int n, const double* a, double* y, hipStream_t s);
// out (column-major, ld) = -C mu for k columns: C is lda x m (FP64), mu m x k
void launch_buffer_rhs(int n, int64_t lda, int m, const double* Cm, const double* mu, int k, double* out, int64_t ld,
                       hipStream_t s);

}  // namespace cpgpu
