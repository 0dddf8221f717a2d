// K4 / K5: vortex-particle evaluations on gfx950.
//
// K4, Barnes-Hut: one lane per target walks the flattened octree (pre-order
// with skip links, so the walk is a forward scan without a stack).  Particle
// targets are taken in leaf order, so the lanes of a wavefront are spatial
// neighbours and mostly walk the same nodes.  Per node: a leaf sums its
// members directly (self excluded by slot), an internal node is replaced by
// its pseudo-particle when s/d < theta, else the walk descends.  This is the
// reference's recursion (node.h:236-421) in the same visiting order.  FP64,
// FMA contraction off.
//
// K5, panel -> point velocity and stretching (bodyPanel::panelV,
// wakePanel::panelV, partStretching): a far pass over LDS-staged panel
// chunks and a near pass over compacted (panel, point) pairs, summed in a
// fixed order (see "panel -> point sums in two passes" below).
#include <hip/hip_runtime.h>

#include "device.h"
#include "stretch.h"
#include "vp.h"
#include "wavequeue.h"

namespace cpgpu {

namespace {

#pragma clang fp contract(off)

constexpr double kCoreOverlap = 1.5;  // particle::coreOverlap (particle.h:33)
constexpr double kNu = 1.983e-5;      // particle.cpp:126
constexpr int kT = 256;

struct Src {
    double x, y, z, ax, ay, az, r;
};

__device__ inline double nrm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// particle::velInflAlgSmooth (particle.cpp:31-38)
__device__ inline void k_alg(const Src& s, const double* p, double* acc) {
    double sigma = kCoreOverlap * s.r;
    double dx = p[0] - s.x, dy = p[1] - s.y, dz = p[2] - s.z;
    double r = nrm3(dx, dy, dz);
    double r2 = r * r, s2 = sigma * sigma;
    const double q = r2 + s2;
    double f = -1 / (4 * kPI) * (r2 + 2.5 * s2) / (q * q * sqrt(q));  // (r^2 + s^2)^2.5
    acc[0] += f * (dy * s.az - dz * s.ay);
    acc[1] += f * (dz * s.ax - dx * s.az);
    acc[2] += f * (dx * s.ay - dy * s.ax);
}

__device__ inline double sym_sigma(double rs, double rt) {
    double a = kCoreOverlap * rs, b = kCoreOverlap * rt;
    return sqrt(a * a + b * b) / sqrt(2.0);
}

// particle::velInfl(particle*) (particle.cpp:42-58): source s on target t
__device__ inline void k_gauss(const Src& s, const Src& t, double* acc) {
    double sigma = sym_sigma(s.r, t.r);
    double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
    double rho = nrm3(dx, dy, dz) / sigma;
    double K = (1 / (4 * kPI * rho) * erf(rho / sqrt(2.0)) - 1 / pow(2 * kPI, 1.5) * exp(-0.5 * rho * rho)) /
               (rho * rho);
    double f = -1 / (sigma * sigma * sigma) * K;
    acc[0] += f * (dy * s.az - dz * s.ay);
    acc[1] += f * (dz * s.ax - dx * s.az);
    acc[2] += f * (dx * s.ay - dy * s.ax);
}

// source->vortexStretching(target) (particle.cpp:80-118):
// (1/sigma^3) alpha_t x ((K I - F d d^T / sigma^2) alpha_s), d = x_s - x_t,
// zero beyond 5 source radii
__device__ inline void k_stretch(const Src& s, const Src& t, double* acc) {
    double dx = s.x - t.x, dy = s.y - t.y, dz = s.z - t.z;
    double dist = nrm3(dx, dy, dz);
    if (dist > 5 * s.r) return;
    double sig = sym_sigma(s.r, t.r);
    double rho = dist / sig;
    double xi = 1 / pow(2 * kPI, 1.5) * exp(-rho * rho / 2);
    double G = 1 / (4 * kPI * rho) * erf(rho / sqrt(2.0));
    double K = (G - xi) / (rho * rho);
    double F = (3 * K - xi) / (rho * rho);
    double c = 1 / (sig * sig) * F;
    // M alpha_s
    double da = dx * s.ax + dy * s.ay + dz * s.az;
    double mx = K * s.ax - c * dx * da;
    double my = K * s.ay - c * dy * da;
    double mz = K * s.az - c * dz * da;
    double s3 = 1 / (sig * sig * sig);
    acc[0] += s3 * (t.ay * mz - t.az * my);
    acc[1] += s3 * (t.az * mx - t.ax * mz);
    acc[2] += s3 * (t.ax * my - t.ay * mx);
}

// source->viscousDiffusion(target) (particle.cpp:122-140)
__device__ inline void k_diff(const Src& s, const Src& t, double* acc) {
    double dx = s.x - t.x, dy = s.y - t.y, dz = s.z - t.z;
    double sig = sym_sigma(s.r, t.r);
    double rho = nrm3(dx, dy, dz) / sig;
    double Vs = 4 * kPI / 3 * (s.r * s.r * s.r);
    double Vt = 4 * kPI / 3 * (t.r * t.r * t.r);
    double xi = 1 / (pow(2 * kPI, 1.5) * (rho * rho * rho)) * exp(-rho * rho / 2);
    double f = 2 * kNu / (sig * sig) * xi;
    acc[0] += f * (Vs * t.ax - Vt * s.ax);
    acc[1] += f * (Vs * t.ay - Vt * s.ay);
    acc[2] += f * (Vs * t.az - Vt * s.az);
}

__device__ inline Src member(const double* mpos, const double* mstr, const double* mrad, int i) {
    return Src{mpos[3 * i], mpos[3 * i + 1], mpos[3 * i + 2], mstr[3 * i], mstr[3 * i + 1], mstr[3 * i + 2], mrad[i]};
}

__device__ inline Src pseudo(const BhNode& nd) { return Src{nd.ex, nd.ey, nd.ez, nd.sx, nd.sy, nd.sz, nd.radius}; }

// node::isFarField (node.h:395-411): s/d < theta
__device__ inline bool far_cell(const BhNode& nd, double px, double py, double pz, double theta) {
    double d = nrm3(nd.ex - px, nd.ey - py, nd.ez - pz);
    return nd.width / d < theta;
}

// velocity at points, algebraic kernel (node::calcVel(POI))
__global__ __launch_bounds__(kT) void bh_points_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                       const double* __restrict__ mpos,
                                                       const double* __restrict__ mstr,
                                                       const double* __restrict__ mrad, double theta, int m,
                                                       const double* __restrict__ pts, double* __restrict__ out) {
    int t = blockIdx.x * kT + threadIdx.x;
    if (t >= m) return;
    double p[3] = {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]};
    double acc[3] = {0, 0, 0};
    int i = 1;  // root excluded: its children start at node 1
    while (i < n_nodes) {
        const BhNode nd = nodes[i];
        if (nd.count >= 0) {
            for (int k = 0; k < nd.count; k++) k_alg(member(mpos, mstr, mrad, nd.first + k), p, acc);
            i = nd.skip;
        } else if (far_cell(nd, p[0], p[1], p[2], theta)) {
            k_alg(pseudo(nd), p, acc);
            i = nd.skip;
        } else {
            i = i + 1;
        }
    }
    out[3 * t] = acc[0];
    out[3 * t + 1] = acc[1];
    out[3 * t + 2] = acc[2];
}

// velocity / stretching / diffusion on the tree's own particles
template <int MODE>
__global__ __launch_bounds__(kT) void bh_particles_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                          const double* __restrict__ mpos,
                                                          const double* __restrict__ mstr,
                                                          const double* __restrict__ mrad, int n, double theta,
                                                          const int* __restrict__ order, double* __restrict__ out) {
    int st = blockIdx.x * kT + threadIdx.x;  // target slot (leaf order)
    if (st >= n) return;
    const Src tg = member(mpos, mstr, mrad, st);
    double acc[3] = {0, 0, 0};
    // the wavefront's walk of bhc_kernel: uniform node index, a lane that
    // takes a cell's pseudo-particle sits out its subtree (same sums per
    // target, in the same order)
    int i = 1, skip_to = 0;
    while (i < n_nodes) {
        const BhNode nd = nodes[i];
        const bool act = skip_to <= i;
        if (nd.count >= 0) {
            if (act)
                for (int k = 0; k < nd.count; k++) {
                    int ms = nd.first + k;
                    if (ms == st) continue;  // pointer self-exclusion of node.h:272
                    const Src s = member(mpos, mstr, mrad, ms);
                    if (MODE == CP_BH_VEL)
                        k_gauss(s, tg, acc);
                    else if (MODE == CP_BH_STRETCH)
                        k_stretch(s, tg, acc);
                    else
                        k_diff(s, tg, acc);
                }
            i = nd.skip;
        } else {
            bool open = false;
            if (act) {
                if (far_cell(nd, tg.x, tg.y, tg.z, theta)) {
                    const Src s = pseudo(nd);
                    if (MODE == CP_BH_VEL)
                        k_gauss(s, tg, acc);
                    else if (MODE == CP_BH_STRETCH)
                        k_stretch(s, tg, acc);
                    else
                        k_diff(s, tg, acc);
                    skip_to = nd.skip;
                } else {
                    open = true;
                }
            }
            i = __ballot(open) ? i + 1 : nd.skip;
        }
    }
    const int o = order[st];
    out[3 * o] = acc[0];
    out[3 * o + 1] = acc[1];
    out[3 * o + 2] = acc[2];
}

// Barnes-Hut sums on caller-given particle targets.  The vortex-particle
// step loop evaluates the previous step's tree on the current particles,
// including those shed since it was built (cpCaseVP.cpp:58-83 rebuild after
// the strength update): tslot is the target's slot in this tree (its own
// member is skipped, node.h:272) or -1.
template <int MODE>
__global__ __launch_bounds__(kT) void bh_targets_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                        const double* __restrict__ mpos,
                                                        const double* __restrict__ mstr,
                                                        const double* __restrict__ mrad, double theta, int m,
                                                        const double* __restrict__ tpos,
                                                        const double* __restrict__ tstr,
                                                        const double* __restrict__ trad,
                                                        const int* __restrict__ tslot, double* __restrict__ out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= m) return;
    const Src tg{tpos[3 * t], tpos[3 * t + 1], tpos[3 * t + 2], tstr[3 * t], tstr[3 * t + 1], tstr[3 * t + 2], trad[t]};
    const int self = tslot[t];
    double acc[3] = {0, 0, 0};
    // the wavefront's walk of bhc_kernel (same sums per target)
    int i = 1, skip_to = 0;
    while (i < n_nodes) {
        const BhNode nd = nodes[i];
        const bool act = skip_to <= i;
        if (nd.count >= 0) {
            if (act)
                for (int k = 0; k < nd.count; k++) {
                    const int ms = nd.first + k;
                    if (ms == self) continue;
                    const Src sp = member(mpos, mstr, mrad, ms);
                    if (MODE == CP_BH_VEL)
                        k_gauss(sp, tg, acc);
                    else if (MODE == CP_BH_STRETCH)
                        k_stretch(sp, tg, acc);
                    else
                        k_diff(sp, tg, acc);
                }
            i = nd.skip;
        } else {
            bool open = false;
            if (act) {
                if (far_cell(nd, tg.x, tg.y, tg.z, theta)) {
                    const Src sp = pseudo(nd);
                    if (MODE == CP_BH_VEL)
                        k_gauss(sp, tg, acc);
                    else if (MODE == CP_BH_STRETCH)
                        k_stretch(sp, tg, acc);
                    else
                        k_diff(sp, tg, acc);
                    skip_to = nd.skip;
                } else {
                    open = true;
                }
            }
            i = __ballot(open) ? i + 1 : nd.skip;
        }
    }
    out[3 * t] = acc[0];
    out[3 * t + 1] = acc[1];
    out[3 * t + 2] = acc[2];
}

// caller-given targets up to which the tree walks run a wavefront per target
// (kBhWaveTargets in vp.h)
constexpr int kBhWaveMax = kBhWaveTargets;
constexpr int kBhcWaveMax = kBhWaveMax;

// The walk with one wavefront per target, for target sets too small to fill
// the device (the reference form's few thousand particles, a step's seed
// points): the node decisions are the wavefront's, a leaf's members are
// split over the lanes (member(ms) adds member ms to the lane's sums), a far
// cell is lane 0's (cell(nd)); the caller adds the lanes' sums.
template <class FM, class FC>
__device__ __forceinline__ void walk_wave(const BhNode* __restrict__ nodes, int n_nodes, double px, double py,
                                          double pz, double theta, int lane, FM&& member_fn, FC&& cell_fn) {
    int i = 1;
    while (i < n_nodes) {
        const BhNode nd = nodes[i];
        if (nd.count >= 0) {
            for (int k = lane; k < nd.count; k += 64) member_fn(nd.first + k);
            i = nd.skip;
        } else if (far_cell(nd, px, py, pz, theta)) {
            if (lane == 0) cell_fn(nd);
            i = nd.skip;
        } else {
            i = i + 1;
        }
    }
}

// sum of acc[3] over the wavefront's lanes in a fixed butterfly order (every
// lane ends with the total)
__device__ __forceinline__ void wave_sum3(double* acc) {
    for (int off = 32; off >= 1; off >>= 1)
        for (int q = 0; q < 3; q++) acc[q] += __shfl_xor(acc[q], off, 64);
}

__global__ __launch_bounds__(kT) void bh_points_wave_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                            const double* __restrict__ mpos,
                                                            const double* __restrict__ mstr,
                                                            const double* __restrict__ mrad, double theta, int m,
                                                            const double* __restrict__ pts,
                                                            double* __restrict__ out) {
    const int t = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= m) return;  // whole wavefronts
    double p[3] = {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]};
    double acc[3] = {0, 0, 0};
    walk_wave(nodes, n_nodes, p[0], p[1], p[2], theta, lane,
              [&](int ms) { k_alg(member(mpos, mstr, mrad, ms), p, acc); },
              [&](const BhNode& nd) { k_alg(pseudo(nd), p, acc); });
    wave_sum3(acc);
    if (lane == 0)
        for (int q = 0; q < 3; q++) out[3 * t + q] = acc[q];
}

template <int MODE>
__global__ __launch_bounds__(kT) void bh_targets_wave_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                             const double* __restrict__ mpos,
                                                             const double* __restrict__ mstr,
                                                             const double* __restrict__ mrad, double theta, int m,
                                                             const double* __restrict__ tpos,
                                                             const double* __restrict__ tstr,
                                                             const double* __restrict__ trad,
                                                             const int* __restrict__ tslot,
                                                             double* __restrict__ out) {
    const int t = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= m) return;  // whole wavefronts
    const Src tg{tpos[3 * t], tpos[3 * t + 1], tpos[3 * t + 2], tstr[3 * t], tstr[3 * t + 1], tstr[3 * t + 2], trad[t]};
    const int self = tslot[t];
    double acc[3] = {0, 0, 0};
    auto add = [&](const Src& sp) {
        if (MODE == CP_BH_VEL)
            k_gauss(sp, tg, acc);
        else if (MODE == CP_BH_STRETCH)
            k_stretch(sp, tg, acc);
        else
            k_diff(sp, tg, acc);
    };
    walk_wave(nodes, n_nodes, tg.x, tg.y, tg.z, theta, lane,
              [&](int ms) {
                  if (ms != self) add(member(mpos, mstr, mrad, ms));
              },
              [&](const BhNode& nd) { add(pseudo(nd)); });
    wave_sum3(acc);
    if (lane == 0)
        for (int q = 0; q < 3; q++) out[3 * t + q] = acc[q];
}

// direct pairwise stretching or diffusion on every particle i from every
// other particle j, in j order, as cpCaseVP::particleStrengthUpdate's
// direct branch calls them (cpCaseVP.cpp:489-496, 507-514): particle i is
// the called object (particles[i]->vortexStretching(particles[j])), the
// roles the tree walk gives to the other particle
template <int MODE>
__global__ __launch_bounds__(kT) void pairs_direct_kernel(int n, const double* __restrict__ pos,
                                                          const double* __restrict__ str,
                                                          const double* __restrict__ rad, double* __restrict__ out) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const Src me = member(pos, str, rad, i);
    double acc[3] = {0, 0, 0};
    for (int j = 0; j < n; j++) {
        if (j == i) continue;
        const Src o = member(pos, str, rad, j);
        if (MODE == CP_BH_STRETCH)
            k_stretch(me, o, acc);
        else
            k_diff(me, o, acc);
    }
    out[3 * i] = acc[0];
    out[3 * i + 1] = acc[1];
    out[3 * i + 2] = acc[2];
}

// the tree's members take the current state of the particles they hold
// (the reference's octree holds particle pointers: leaves see the particles
// as they are now, pseudo-particles stay as they were at the build)
__global__ __launch_bounds__(kT) void bh_refresh_kernel(int n, const int* __restrict__ order,
                                                        const double* __restrict__ pos,
                                                        const double* __restrict__ str,
                                                        const double* __restrict__ rad, double* __restrict__ mpos,
                                                        double* __restrict__ mstr, double* __restrict__ mrad) {
    const int sl = blockIdx.x * kT + threadIdx.x;
    if (sl >= n) return;
    const int i = order[sl];
    for (int q = 0; q < 3; q++) {
        mpos[3 * sl + q] = pos[3 * i + q];
        mstr[3 * sl + q] = str[3 * i + q];
    }
    mrad[sl] = rad[i];
}

// particle::velInfl(POI) (particle.cpp:60-77) summed over all particles in
// their order: Gaussian kernel, sigma = coreOverlap * radius, a particle at
// the point itself contributes nothing
// (the particles j0..j1 of slice blockIdx.y: slice q of out at out + 3 m q)
__global__ __launch_bounds__(kT) void gauss_points_kernel(int n, const double* __restrict__ ppos,
                                                          const double* __restrict__ pstr,
                                                          const double* __restrict__ prad, int m,
                                                          const double* __restrict__ pts, int chunk,
                                                          double* __restrict__ out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= m) return;
    const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    out += (int64_t)blockIdx.y * 3 * m;
    const double px = pts[3 * t], py = pts[3 * t + 1], pz = pts[3 * t + 2];
    double acc[3] = {0, 0, 0};
    for (int j = j0; j < j1; j++) {
        const double dx = px - ppos[3 * j], dy = py - ppos[3 * j + 1], dz = pz - ppos[3 * j + 2];
        const double d = nrm3(dx, dy, dz);
        if (d == 0) continue;
        const double sigma = kCoreOverlap * prad[j];
        const double rho = d / sigma;
        const double K = (1 / (4 * kPI * rho) * erf(rho / sqrt(2.0)) - 1 / pow(2 * kPI, 1.5) * exp(-0.5 * rho * rho)) /
                         (rho * rho);
        const double f = -1 / (sigma * sigma * sigma) * K;
        const double ax = pstr[3 * j], ay = pstr[3 * j + 1], az = pstr[3 * j + 2];
        acc[0] += f * (dy * az - dz * ay);
        acc[1] += f * (dz * ax - dx * az);
        acc[2] += f * (dx * ay - dy * ax);
    }
    out[3 * t] = acc[0];
    out[3 * t + 1] = acc[1];
    out[3 * t + 2] = acc[2];
}

// vortexFil::velInfl (vortexFil.cpp:26-40): strength / (4 pi) * vortexV over
// the filaments (ends p1, p2: 6 per filament), in order
__global__ __launch_bounds__(kT) void filament_points_kernel(int nf, const double* __restrict__ ends,
                                                             const double* __restrict__ fstr, int m,
                                                             const double* __restrict__ pts,
                                                             double* __restrict__ out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= m) return;
    const double p[3] = {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]};
    double acc[3] = {0, 0, 0};
    for (int f = 0; f < nf; f++) {
        const double* e = ends + 6 * (int64_t)f;
        const double a[3] = {p[0] - e[0], p[1] - e[1], p[2] - e[2]};
        const double b[3] = {p[0] - e[3], p[1] - e[4], p[2] - e[5]};
        const double sv[3] = {e[3] - e[0], e[4] - e[1], e[5] - e[2]};
        double w[3];
        vortex_v(a, b, sv, w);
        for (int q = 0; q < 3; q++) acc[q] += w[q] * fstr[f] / (4 * kPI);
    }
    out[3 * t] = acc[0];
    out[3 * t + 1] = acc[1];
    out[3 * t + 2] = acc[2];
}

// ---------------------------------------------------------------------------
// The configured unsteady case's particle kernel: velInflAlgSmooth's
// algebraic smoothing (particle.cpp:31-38) with one core radius for every
// particle (s2 = core^2), velocity and (alpha_t . grad) u stretching.  The
// pair offset is formed in FP64 (particles sit ~1e-2 apart at coordinates of
// order 10, below FP32's resolution of an absolute position); the rest runs
// in T (float: the FP32 path; double: the FP64 path the FP32 one is checked
// against).
template <class T>
__device__ inline void kc_vel(const double* sp, const T* sa, const double* tp, T s2, T* acc) {
    const T dx = (T)(tp[0] - sp[0]), dy = (T)(tp[1] - sp[1]), dz = (T)(tp[2] - sp[2]);
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T q = r2 + s2;
    const T f = (T)(-1.0 / (4 * kPI)) * (r2 + (T)2.5 * s2) / (q * q * sqrt(q));
    acc[0] += f * (dy * sa[2] - dz * sa[1]);
    acc[1] += f * (dz * sa[0] - dx * sa[2]);
    acc[2] += f * (dx * sa[1] - dy * sa[0]);
}

// u = f(r) x x a_s with x = target - source: (a_t . grad) u =
// f'(r)/r (a_t . x)(x x a_s) + f (a_t x a_s), f'/r = (3r^2 + 10.5 s^2) /
// (4 pi (r^2 + s^2)^3.5)
template <class T>
__device__ inline void kc_stretch(const double* sp, const T* sa, const double* tp, const T* ta, T s2, T* acc) {
    const T dx = (T)(tp[0] - sp[0]), dy = (T)(tp[1] - sp[1]), dz = (T)(tp[2] - sp[2]);
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T q = r2 + s2;
    const T sq = sqrt(q);
    const T f = (T)(-1.0 / (4 * kPI)) * (r2 + (T)2.5 * s2) / (q * q * sq);
    const T h = (T)(1.0 / (4 * kPI)) * ((T)3 * r2 + (T)10.5 * s2) / (q * q * q * sq);
    const T hd = h * (ta[0] * dx + ta[1] * dy + ta[2] * dz);
    acc[0] += hd * (dy * sa[2] - dz * sa[1]) + f * (ta[1] * sa[2] - ta[2] * sa[1]);
    acc[1] += hd * (dz * sa[0] - dx * sa[2]) + f * (ta[2] * sa[0] - ta[0] * sa[2]);
    acc[2] += hd * (dx * sa[1] - dy * sa[0]) + f * (ta[0] * sa[1] - ta[1] * sa[0]);
}

// Barnes-Hut with the configured kernel on m targets (the reference's walk:
// root excluded, far iff s/d < theta, decided in FP64 so both precisions
// visit the same cells; leaves skip the target's own slot).  KC_VEL at
// points / particles (tstr unused), KC_STRETCH on particles.  Targets are
// either caller-given (tslot[t]: the target's slot or -1; tslot null: no
// slot; result t to out[t]) or, with own_order, the tree's own members in
// slot order (target t is slot t; result to out[own_order[t]]).
enum { KC_VEL = 0, KC_STRETCH = 1 };
template <int MODE, class T>
__global__ __launch_bounds__(kT) void bhc_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                 const double* __restrict__ mpos, const double* __restrict__ mstr,
                                                 double theta, double s2d, int m, const double* __restrict__ tpos,
                                                 const double* __restrict__ tstr, const int* __restrict__ tslot,
                                                 const int* __restrict__ own_order, double* __restrict__ out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= m) return;
    const double tp[3] = {tpos[3 * t], tpos[3 * t + 1], tpos[3 * t + 2]};
    T ta[3] = {0, 0, 0};
    if (MODE == KC_STRETCH) {
        ta[0] = (T)tstr[3 * t];
        ta[1] = (T)tstr[3 * t + 1];
        ta[2] = (T)tstr[3 * t + 2];
    }
    const int self = own_order ? t : (tslot ? tslot[t] : -1);
    const int o = own_order ? own_order[t] : t;
    const T s2 = (T)s2d;
    T acc[3] = {0, 0, 0};
    // The wavefront walks the nodes together (uniform i: one load of each
    // node, one of each member, for all lanes).  A lane that takes a cell's
    // pseudo-particle sits out the cell's subtree (nodes below its skip
    // index); the wavefront descends when any active lane opens the cell and
    // jumps to the skip index when none does.  Each lane so meets exactly the
    // nodes of its own walk, in the same order: the same sums, bit for bit.
    int i = 1, skip_to = 0;
    while (i < n_nodes) {
        const BhNode nd = nodes[i];
        const bool act = skip_to <= i;
        if (nd.count >= 0) {
            if (act)
                for (int k = 0; k < nd.count; k++) {
                    const int ms = nd.first + k;
                    if (ms == self) continue;
                    const T sa[3] = {(T)mstr[3 * ms], (T)mstr[3 * ms + 1], (T)mstr[3 * ms + 2]};
                    if (MODE == KC_VEL)
                        kc_vel<T>(mpos + 3 * ms, sa, tp, s2, acc);
                    else
                        kc_stretch<T>(mpos + 3 * ms, sa, tp, ta, s2, acc);
                }
            i = nd.skip;
        } else {
            bool open = false;
            if (act) {
                if (far_cell(nd, tp[0], tp[1], tp[2], theta)) {
                    const double sp[3] = {nd.ex, nd.ey, nd.ez};
                    const T sa[3] = {(T)nd.sx, (T)nd.sy, (T)nd.sz};
                    if (MODE == KC_VEL)
                        kc_vel<T>(sp, sa, tp, s2, acc);
                    else
                        kc_stretch<T>(sp, sa, tp, ta, s2, acc);
                    skip_to = nd.skip;
                } else {
                    open = true;
                }
            }
            i = __ballot(open) ? i + 1 : nd.skip;
        }
    }
    out[3 * o] = acc[0];
    out[3 * o + 1] = acc[1];
    out[3 * o + 2] = acc[2];
}

// The same walk with one wavefront per target, for target sets too small to
// fill the device (a step's few hundred RK4 seeds, the collocation points):
// the node decisions are the wavefront's (uniform), a leaf's members are
// split over the lanes, a far cell's pseudo-particle is lane 0's; the lanes'
// sums are added in a fixed butterfly order.  Targets are caller-given
// points (tslot as in bhc_kernel).
template <int MODE, class T>
__global__ __launch_bounds__(kT) void bhc_wave_kernel(const BhNode* __restrict__ nodes, int n_nodes,
                                                      const double* __restrict__ mpos,
                                                      const double* __restrict__ mstr, double theta, double s2d,
                                                      int m, const double* __restrict__ tpos,
                                                      const double* __restrict__ tstr,
                                                      const int* __restrict__ tslot, double* __restrict__ out) {
    const int t = blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= m) return;  // whole wavefronts
    const double tp[3] = {tpos[3 * t], tpos[3 * t + 1], tpos[3 * t + 2]};
    T ta[3] = {0, 0, 0};
    if (MODE == KC_STRETCH) {
        ta[0] = (T)tstr[3 * t];
        ta[1] = (T)tstr[3 * t + 1];
        ta[2] = (T)tstr[3 * t + 2];
    }
    const int self = tslot ? tslot[t] : -1;
    const T s2 = (T)s2d;
    T acc[3] = {0, 0, 0};
    int i = 1;
    while (i < n_nodes) {
        const BhNode nd = nodes[i];
        if (nd.count >= 0) {
            for (int k = lane; k < nd.count; k += 64) {
                const int ms = nd.first + k;
                if (ms == self) continue;
                const T sa[3] = {(T)mstr[3 * ms], (T)mstr[3 * ms + 1], (T)mstr[3 * ms + 2]};
                if (MODE == KC_VEL)
                    kc_vel<T>(mpos + 3 * ms, sa, tp, s2, acc);
                else
                    kc_stretch<T>(mpos + 3 * ms, sa, tp, ta, s2, acc);
            }
            i = nd.skip;
        } else if (far_cell(nd, tp[0], tp[1], tp[2], theta)) {
            if (lane == 0) {
                const double sp[3] = {nd.ex, nd.ey, nd.ez};
                const T sa[3] = {(T)nd.sx, (T)nd.sy, (T)nd.sz};
                if (MODE == KC_VEL)
                    kc_vel<T>(sp, sa, tp, s2, acc);
                else
                    kc_stretch<T>(sp, sa, tp, ta, s2, acc);
            }
            i = nd.skip;
        } else {
            i = i + 1;
        }
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int q = 0; q < 3; q++) acc[q] += __shfl_xor(acc[q], off, 64);
    if (lane == 0) {
        out[3 * t] = acc[0];
        out[3 * t + 1] = acc[1];
        out[3 * t + 2] = acc[2];
    }
}

// direct sums of the configured kernel over n particles at m targets, in
// particle order; with own, target t is particle t and skips itself
template <int MODE, class T>
__global__ __launch_bounds__(kT) void dc_kernel(int n, const double* __restrict__ pos, const double* __restrict__ str,
                                                double s2d, int m, const double* __restrict__ tpos,
                                                const double* __restrict__ tstr, int own, int chunk,
                                                double* __restrict__ out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= m) return;
    const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    out += (int64_t)blockIdx.y * 3 * m;  // this slice's partial sums
    const double tp[3] = {tpos[3 * t], tpos[3 * t + 1], tpos[3 * t + 2]};
    T ta[3] = {0, 0, 0};
    if (MODE == KC_STRETCH) {
        ta[0] = (T)tstr[3 * t];
        ta[1] = (T)tstr[3 * t + 1];
        ta[2] = (T)tstr[3 * t + 2];
    }
    const int self = own ? t : -1;
    const T s2 = (T)s2d;
    T acc[3] = {0, 0, 0};
    for (int j = j0; j < j1; j++) {
        if (j == self) continue;
        const T sa[3] = {(T)str[3 * j], (T)str[3 * j + 1], (T)str[3 * j + 2]};
        if (MODE == KC_VEL)
            kc_vel<T>(pos + 3 * j, sa, tp, s2, acc);
        else
            kc_stretch<T>(pos + 3 * j, sa, tp, ta, s2, acc);
    }
    out[3 * t] = acc[0];
    out[3 * t + 1] = acc[1];
    out[3 * t + 2] = acc[2];
}

// ---- panel -> point sums in two passes ----------------------------------
// Far pass: every (point, panel) pair whose panel is far (is_far: the
// reference's |pjk| / longSide > 5) with the point source and point doublet
// forms, branch-free over panel chunks staged in LDS (a broadcast read per
// pair instead of a 256-byte record load), one lane per point, panels split
// over K slices.  Near pass: one wavefront per tile of 64 points; panels
// whose 5 longSide reach meets the tile's bounding sphere are candidates,
// the (panel, point) pairs of a candidate that are near join a queue
// (lane permutes), and every 64 queued pairs are evaluated by the whole
// wavefront, one pair per lane, with the reference's near-field formulas;
// each point then adds its pairs' results in queue order.  Both passes sum
// in a fixed order, so results do not depend on the launch.

// far-field data of one panel, staged in LDS by the far pass
struct FarRec {
    double c[3];
    double n[3];
    double area;
    double far2_lo, far2_hi, long_side;
    double sg, mj;  // source and doublet strength (sg = 0 on wake panels)
};

__device__ inline void load_far(const PanelRec* __restrict__ body, int nb, const PanelRec* __restrict__ wake,
                                const double* __restrict__ sigma, const double* __restrict__ mu,
                                const double* __restrict__ muw, int j, FarRec& F) {
    const bool isb = j < nb;
    const PanelRec& P = isb ? body[j] : wake[j - nb];
    for (int k = 0; k < 3; k++) {
        F.c[k] = P.c[k];
        F.n[k] = P.n[k];
    }
    F.area = P.area;
    F.far2_lo = P.far2_lo;
    F.far2_hi = P.far2_hi;
    F.long_side = P.long_side;
    F.sg = isb ? sigma[j] : 0.0;
    F.mj = isb ? mu[j] : muw[j - nb];
}

// far velocity of one panel (panel_v_inf_f64's far branch: point source
// with the reference's negated sign, bodyPanel.cpp:314, and point doublet),
// times its strengths, added to acc
__device__ inline void far_vel_f64(const FarRec& F, const double* p, double* acc) {
    const double pjk[3] = {p[0] - F.c[0], p[1] - F.c[1], p[2] - F.c[2]};
    const double r2 = sq_norm3(pjk[0], pjk[1], pjk[2]);
    if (!is_far(F, r2)) return;
    const double r = sqrt(r2);
    const double den = 4 * kPI * (r * r * r);
    const double f = F.area / den;
    double vs[3], vd[3];
    for (int k = 0; k < 3; k++) vs[k] = -1.0 * (pjk[k] * f);
    pnt_dub_v(F, pjk, r, vd);
    for (int q = 0; q < 3; q++) acc[q] += F.sg * vs[q] + F.mj * vd[q];
}

// the same in FP32 arithmetic (offset and far test in FP64)
__device__ inline void far_vel_f32(const FarRec& F, const double* p, float* acc) {
    const double d0 = p[0] - F.c[0], d1 = p[1] - F.c[1], d2 = p[2] - F.c[2];
    if (!is_far(F, sq_norm3(d0, d1, d2))) return;
    const float x = (float)d0, y = (float)d1, z = (float)d2;
    const float ri = rsqrtf(x * x + y * y + z * z);
    const float ri2 = ri * ri;
    const float ari3 = (float)(F.area * (1.0 / (4 * kPI))) * ri2 * ri;  // area / (4 pi r^3)
    const float nx = (float)F.n[0], ny = (float)F.n[1], nz = (float)F.n[2];
    const float t = 3.0f * (x * nx + y * ny + z * nz) * ri2;
    const float fd = (float)F.mj * ari3;
    const float fs = -(float)F.sg * ari3;
    acc[0] += fd * (t * x - nx) + fs * x;
    acc[1] += fd * (t * y - ny) + fs * y;
    acc[2] += fd * (t * z - nz) + fs * z;
}

// far velocity gradient (point source and point doublet gradients of
// bodyPanel::partStretching, with the reference's frames), added to G
__device__ inline void far_grad_f64(const FarRec& F, const double* p, Mat3& G) {
    const double d0 = p[0] - F.c[0], d1 = p[1] - F.c[1], d2 = p[2] - F.c[2];
    if (!is_far(F, sq_norm3(d0, d1, d2))) return;
    vg_point_source_add(F, F.sg, p, G);
    vg_point_doublet_add(F, F.mj, p, G);
}

// the same in FP32 arithmetic (far test in FP64; the forms use the panel
// centre's z as the reference writes them)
__device__ inline void far_grad_f32(const FarRec& F, const double* poi, float G[3][3]) {
    const double e0 = poi[0] - F.c[0], e1 = poi[1] - F.c[1], e2 = poi[2] - F.c[2];
    if (!is_far(F, sq_norm3(e0, e1, e2))) return;
    const float z = (float)F.c[2], zz0 = (float)(F.c[2] - poi[2]);
    const float dx = (float)(F.c[0] - poi[0]), dy = (float)(F.c[1] - poi[1]);
    const float dx2 = dx * dx, dy2 = dy * dy;
    const float q2 = dx2 + dy2 + z * z;
    const float sq = sqrtf(q2);
    const float a4 = (float)(F.area * (1.0 / (4 * kPI)));
    const float isdd = 1.0f / (q2 * q2 * sq);
    const float c = (float)F.sg * a4;
    G[0][0] += c * ((-2 * dx2 + dy2 + z * z) * isdd);
    G[1][0] += c * (-3 * dx * dy * isdd);
    G[2][0] += c * (-3 * dx * z * isdd);
    G[0][1] += c * (-3 * dx * dy * isdd);
    G[1][1] += c * ((dx2 - dy2 + z * z) * isdd);
    G[2][1] += c * (-3 * z * dy * isdd);
    G[0][2] += c * (-3 * dx * zz0 * isdd);
    G[1][2] += c * (-3 * dy * zz0 * isdd);
    G[2][2] += c * ((-3 * z * zz0 + (dx2 + dy2 + z * z)) * isdd);
    const float iddd = 1.0f / (q2 * q2 * q2 * sq);
    const float uv = 3 * (float)F.mj * a4;
    const float w = -(float)F.mj * a4;
    G[0][0] += uv * z * (-4 * dx2 + dy2 + z * z) * iddd;
    G[1][0] += uv * (-5) * z * dy * dx * iddd;
    G[2][0] += uv * dx * (dx2 + dy2 - 4 * z * z) * iddd;
    G[0][1] += uv * (-5) * dx * dy * z * iddd;
    G[1][1] += uv * z * (dx2 - 4 * dy2 + z * z) * iddd;
    G[2][1] += uv * dy * (dx2 + dy2 - 4 * z * z) * iddd;
    G[0][2] += w * (-3) * dx * (dx2 + dy2 - 4 * z * z) * iddd;
    G[1][2] += w * (-3) * dy * (dx2 + dy2 - 4 * z * z) * iddd;
    G[2][2] += w * (-3) * z * (3 * dx2 + 3 * dy2 - 2 * z * z) * iddd;
}

enum { PS_VEL = 0, PS_STRETCH = 1 };

// far pass over the panels of K slice blockIdx.y: part slice ks (3 per
// point: the velocity, or the gradient applied to the particle's alpha)
template <int MODE, bool F32>
__global__ __launch_bounds__(kT) void panels_far_kernel(const PanelRec* __restrict__ body, int nb,
                                                        const PanelRec* __restrict__ wake, int nw,
                                                        const double* __restrict__ sigma,
                                                        const double* __restrict__ mu,
                                                        const double* __restrict__ muw, int m,
                                                        const double* __restrict__ pts,
                                                        const double* __restrict__ alpha, int chunk,
                                                        double* __restrict__ part) {
    __shared__ FarRec sh[kT];
    const int t = blockIdx.x * kT + threadIdx.x;
    const bool live = t < m;
    const int ks = blockIdx.y;
    const int p0 = ks * chunk, p1 = min(nb + nw, p0 + chunk);
    double p[3] = {0, 0, 0};
    if (live)
        for (int k = 0; k < 3; k++) p[k] = pts[3 * t + k];
    double acc[3] = {0, 0, 0};
    float accf[3] = {0, 0, 0};
    Mat3 G;
    mat3_zero(G);
    float Gf[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int b0 = p0; b0 < p1; b0 += kT) {
        const int cnt = min(kT, p1 - b0);
        __syncthreads();
        if ((int)threadIdx.x < cnt) load_far(body, nb, wake, sigma, mu, muw, b0 + (int)threadIdx.x, sh[threadIdx.x]);
        __syncthreads();
        if (!live) continue;
        for (int k = 0; k < cnt; k++) {
            if (MODE == PS_VEL) {
                if (F32)
                    far_vel_f32(sh[k], p, accf);
                else
                    far_vel_f64(sh[k], p, acc);
            } else {
                if (F32)
                    far_grad_f32(sh[k], p, Gf);
                else
                    far_grad_f64(sh[k], p, G);
            }
        }
    }
    if (!live) return;
    double* o = part + ((int64_t)ks * m + t) * 3;
    if (MODE == PS_VEL) {
        for (int q = 0; q < 3; q++) o[q] = F32 ? (double)accf[q] : acc[q];
    } else {
        const double a0 = alpha[3 * t], a1 = alpha[3 * t + 1], a2 = alpha[3 * t + 2];
        for (int r = 0; r < 3; r++) {
            const double g0 = F32 ? (double)Gf[r][0] : G.a[r][0];
            const double g1 = F32 ? (double)Gf[r][1] : G.a[r][1];
            const double g2 = F32 ? (double)Gf[r][2] : G.a[r][2];
            o[r] = g0 * a0 + g1 * a1 + g2 * a2;
        }
    }
}

// bounding sphere of each tile of 64 points in the spatial order perm
// (centre, radius with a margin over its rounding): one wavefront per tile
__global__ __launch_bounds__(64) void point_tiles_kernel(int m, const double* __restrict__ pts,
                                                         const int* __restrict__ perm, double* __restrict__ tsph) {
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int t = tile * 64 + lane;
    const int src = perm[t < m ? t : tile * 64];  // the tile's first point stands in for missing ones
    const double x = pts[3 * src], y = pts[3 * src + 1], z = pts[3 * src + 2];
    double lo[3] = {x, y, z}, hi[3] = {x, y, z};
    for (int o = 32; o >= 1; o >>= 1)
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    double r = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    for (int o = 32; o >= 1; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    if (lane == 0) {
        tsph[4 * tile] = cx;
        tsph[4 * tile + 1] = cy;
        tsph[4 * tile + 2] = cz;
        tsph[4 * tile + 3] = r * (1 + 1e-12) + 1e-300;
    }
}

// near pass: one 64-lane workgroup per (tile of points in the spatial
// order perm, slice blockIdx.y of the panels); out: one slice of 3 per point
// per panel slice
template <int MODE>
__global__ __launch_bounds__(64) void panels_near_kernel(const PanelRec* __restrict__ body, int nb,
                                                         const PanelRec* __restrict__ wake, int nw,
                                                         const double* __restrict__ sigma,
                                                         const double* __restrict__ mu,
                                                         const double* __restrict__ muw, int m,
                                                         const double* __restrict__ pts,
                                                         const double* __restrict__ alpha,
                                                         const int* __restrict__ perm,
                                                         const double* __restrict__ tsph, int chunk,
                                                         double* __restrict__ out) {
    constexpr int NE = MODE == PS_VEL ? 3 : 9;  // values per pair
    __shared__ double sp[64][3];
    __shared__ double res[64][NE];
    __shared__ unsigned long long tmask[64];
    const int lane = threadIdx.x, tile = blockIdx.x;
    const int ts = tile * 64 + lane;
    const bool live = ts < m;
    const int t = live ? perm[ts] : 0;  // the point this lane holds
    double p[3] = {0, 0, 0};
    if (live)
        for (int k = 0; k < 3; k++) p[k] = pts[3 * t + k];
    for (int k = 0; k < 3; k++) sp[lane][k] = p[k];
    const double tc0 = tsph[4 * tile], tc1 = tsph[4 * tile + 1], tc2 = tsph[4 * tile + 2], tr = tsph[4 * tile + 3];
    double acc[NE];
    for (int e = 0; e < NE; e++) acc[e] = 0;
    __syncthreads();
    // one queued pair per lane: item = panel * 64 + point lane
    auto eval = [&](int item, bool ok) {
        double r[NE];
        for (int e = 0; e < NE; e++) r[e] = 0;
        int L = -1;
        if (ok) {
            const int j = item >> 6;
            L = item & 63;
            const double q[3] = {sp[L][0], sp[L][1], sp[L][2]};
            const bool isb = j < nb;
            const PanelRec& P = isb ? body[j] : wake[j - nb];
            // a queued pair is near (the far test excluded it): the edge
            // sums alone (for the velocity, body panels as triangles)
            if (MODE == PS_VEL) {
                if (isb) {
                    double vs[3], vd[3];
                    panel_v_near_nv<3>(P, q, vs, vd);
                    const double sg = sigma[j], mj = mu[j];
                    for (int k = 0; k < 3; k++) r[k] = sg * vs[k] + mj * vd[k];
                } else {
                    double vd[3];
                    dub_v_near_nv<0>(P, q, vd);
                    const double mw = muw[j - nb];
                    for (int k = 0; k < 3; k++) r[k] = mw * vd[k];
                }
            } else {
                // one call site for body and wake panels (the vertex count
                // from the record): one inlined copy of the gradient sums
                Mat3 Gq;
                mat3_zero(Gq);
                part_stretch_near_add<0>(P, isb, isb ? sigma[j] : 0.0, isb ? mu[j] : muw[j - nb], q, Gq);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) r[3 * a + b] = Gq.a[a][b];
            }
        }
        // each point lane adds the entries of its point in entry order: the
        // entries' bits are ORed into their point's mask, and a lane then
        // walks only its own bits (not all 64 entries)
        for (int e = 0; e < NE; e++) res[lane][e] = r[e];
        tmask[lane] = 0;
        __syncthreads();
        if (L >= 0) atomicOr(&tmask[L], 1ull << lane);
        __syncthreads();
        unsigned long long mine = tmask[lane];
        while (mine) {
            const int k = __ffsll(mine) - 1;
            mine &= mine - 1;
            for (int e = 0; e < NE; e++) acc[e] += res[k][e];
        }
        __syncthreads();
    };
    RowQueue q;
    const int jb = blockIdx.y * chunk, je = min(nb + nw, jb + chunk);
    for (int j0 = jb; j0 < je; j0 += 64) {
        const int jl = j0 + lane;
        bool cand = false;
        if (jl < je) {
            const PanelRec& P = jl < nb ? body[jl] : wake[jl - nb];
            // 5 longSide plus a margin above the is_far band and the
            // rounding of the sphere test: a panel it drops has no near point
            const double reach = 5.0 * P.long_side * (1.0 + 1e-9);
            const double dx = P.c[0] - tc0, dy = P.c[1] - tc1, dz = P.c[2] - tc2;
            const double lim = reach + tr;
            cand = dx * dx + dy * dy + dz * dz <= lim * lim;
        }
        uint64_t cm = __ballot(cand);
        while (cm) {
            const int b = __ffsll((unsigned long long)cm) - 1;
            cm &= cm - 1;
            const int j = j0 + b;
            const PanelRec& P = j < nb ? body[j] : wake[j - nb];
            bool near = false;
            if (live) {
                const double d0 = p[0] - P.c[0], d1 = p[1] - P.c[1], d2 = p[2] - P.c[2];
                near = !is_far(P, sq_norm3(d0, d1, d2));
            }
            queue_push(q, near, j * 64 + lane, lane, eval);
        }
    }
    eval(q.p, lane < q.qn);
    if (!live) return;
    double* o = out + ((int64_t)blockIdx.y * m + t) * 3;
    if (MODE == PS_VEL) {
        for (int k = 0; k < 3; k++) o[k] = acc[k];
    } else {
        const double a0 = alpha[3 * t], a1 = alpha[3 * t + 1], a2 = alpha[3 * t + 2];
        for (int r = 0; r < 3; r++) o[r] = acc[3 * r] * a0 + acc[3 * r + 1] * a1 + acc[3 * r + 2] * a2;
    }
}

__global__ void p2p_reduce_kernel(const double* __restrict__ part, int KS, int m, double* __restrict__ out) {
    int e = blockIdx.x * kT + threadIdx.x;
    if (e >= 3 * m) return;
    double s = 0;
    for (int q = 0; q < KS; q++) s += part[(int64_t)q * 3 * m + e];
    out[e] = s;
}

// out[i + c ld] = -sum_j C(i, j) mu[j + c m]: the second buffer-wake row's
// contribution to the VP right-hand side (cpCaseVP.cpp:272), FP64
__global__ __launch_bounds__(kT) void buffer_rhs_kernel(int n, int64_t lda, int m, const double* __restrict__ Cm,
                                                        const double* __restrict__ mu, int k, double* __restrict__ out,
                                                        int64_t ld) {
    const int i = blockIdx.x * kT + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= k) return;
    double acc = 0;
    for (int j = 0; j < m; j++) acc += Cm[i + (int64_t)j * lda] * mu[(int64_t)c * m + j];
    out[i + (int64_t)c * ld] = -acc;
}

__global__ void add_vec_kernel(int n, const double* __restrict__ a, double* __restrict__ y) {
    int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) y[i] += a[i];
}

}  // namespace

void launch_add_vec(int n, const double* a, double* y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(add_vec_kernel, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, s, n, a, y);
    CP_LAUNCH_CHECK();
}

void launch_buffer_rhs(int n, int64_t lda, int m, const double* Cm, const double* mu, int k, double* out, int64_t ld,
                       hipStream_t s) {
    if (n <= 0 || k <= 0) return;
    dim3 grid((unsigned)((n + kT - 1) / kT), (unsigned)k);
    hipLaunchKernelGGL(buffer_rhs_kernel, grid, dim3(kT), 0, s, n, lda, m, Cm, mu, k, out, ld);
    CP_LAUNCH_CHECK();
}

void launch_bh_points(const DevBhTree& T, double theta, int m, const double* pts, double* out, hipStream_t s) {
    if (m <= 0) return;
    if (m <= kBhWaveMax) {
        hipLaunchKernelGGL(bh_points_wave_kernel, dim3((unsigned)((m + kT / 64 - 1) / (kT / 64))), dim3(kT), 0, s,
                           T.nodes, T.n_nodes, T.mpos, T.mstr, T.mrad, theta, m, pts, out);
        CP_LAUNCH_CHECK();
        return;
    }
    hipLaunchKernelGGL(bh_points_kernel, dim3((unsigned)((m + kT - 1) / kT)), dim3(kT), 0, s, T.nodes, T.n_nodes,
                       T.mpos, T.mstr, T.mrad, theta, m, pts, out);
    CP_LAUNCH_CHECK();
}

void launch_bh_particles(const DevBhTree& T, int mode, double theta, double* out, hipStream_t s) {
    if (T.n <= 0) return;
    dim3 grid((unsigned)((T.n + kT - 1) / kT));
    switch (mode) {
        case CP_BH_VEL:
            hipLaunchKernelGGL(bh_particles_kernel<CP_BH_VEL>, grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos,
                               T.mstr, T.mrad, T.n, theta, T.order, out);
            break;
        case CP_BH_STRETCH:
            hipLaunchKernelGGL(bh_particles_kernel<CP_BH_STRETCH>, grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos,
                               T.mstr, T.mrad, T.n, theta, T.order, out);
            break;
        default:
            hipLaunchKernelGGL(bh_particles_kernel<CP_BH_DIFF>, grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos,
                               T.mstr, T.mrad, T.n, theta, T.order, out);
            break;
    }
    CP_LAUNCH_CHECK();
}

void launch_bh_targets(const DevBhTree& T, int mode, double theta, int m, const double* tpos, const double* tstr,
                       const double* trad, const int* tslot, double* out, hipStream_t s) {
    if (m <= 0) return;
    if (m <= kBhWaveMax) {
        const dim3 gw((unsigned)((m + kT / 64 - 1) / (kT / 64)));
        switch (mode) {
            case CP_BH_VEL:
                hipLaunchKernelGGL(bh_targets_wave_kernel<CP_BH_VEL>, gw, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos,
                                   T.mstr, T.mrad, theta, m, tpos, tstr, trad, tslot, out);
                break;
            case CP_BH_STRETCH:
                hipLaunchKernelGGL(bh_targets_wave_kernel<CP_BH_STRETCH>, gw, dim3(kT), 0, s, T.nodes, T.n_nodes,
                                   T.mpos, T.mstr, T.mrad, theta, m, tpos, tstr, trad, tslot, out);
                break;
            default:
                hipLaunchKernelGGL(bh_targets_wave_kernel<CP_BH_DIFF>, gw, dim3(kT), 0, s, T.nodes, T.n_nodes,
                                   T.mpos, T.mstr, T.mrad, theta, m, tpos, tstr, trad, tslot, out);
                break;
        }
        CP_LAUNCH_CHECK();
        return;
    }
    dim3 grid((unsigned)((m + kT - 1) / kT));
    switch (mode) {
        case CP_BH_VEL:
            hipLaunchKernelGGL(bh_targets_kernel<CP_BH_VEL>, grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos, T.mstr,
                               T.mrad, theta, m, tpos, tstr, trad, tslot, out);
            break;
        case CP_BH_STRETCH:
            hipLaunchKernelGGL(bh_targets_kernel<CP_BH_STRETCH>, grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos,
                               T.mstr, T.mrad, theta, m, tpos, tstr, trad, tslot, out);
            break;
        default:
            hipLaunchKernelGGL(bh_targets_kernel<CP_BH_DIFF>, grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos,
                               T.mstr, T.mrad, theta, m, tpos, tstr, trad, tslot, out);
            break;
    }
    CP_LAUNCH_CHECK();
}

void launch_bh_refresh(const DevBhTree& T, const double* pos, const double* str, const double* rad, hipStream_t s) {
    if (T.n <= 0) return;
    hipLaunchKernelGGL(bh_refresh_kernel, dim3((unsigned)((T.n + kT - 1) / kT)), dim3(kT), 0, s, T.n, T.order, pos,
                       str, rad, T.mpos, T.mstr, T.mrad);
    CP_LAUNCH_CHECK();
}

void launch_pairs_direct(int mode, int n, const double* pos, const double* str, const double* rad, double* out,
                         hipStream_t s) {
    if (n <= 0) return;
    dim3 grid((unsigned)((n + kT - 1) / kT));
    if (mode == CP_BH_STRETCH)
        hipLaunchKernelGGL(pairs_direct_kernel<CP_BH_STRETCH>, grid, dim3(kT), 0, s, n, pos, str, rad, out);
    else
        hipLaunchKernelGGL(pairs_direct_kernel<CP_BH_DIFF>, grid, dim3(kT), 0, s, n, pos, str, rad, out);
    CP_LAUNCH_CHECK();
}


int direct_slices(int m, int n) {
    const int tiles = (m + kT - 1) / kT;
    int KS = 1;
    while (KS < 64 && (int64_t)tiles * KS < 2048 && n / (2 * KS) >= 512) KS *= 2;
    return KS;
}

size_t direct_part_doubles(int m, int n) {
    const int KS = m > 0 ? direct_slices(m, n) : 1;
    return KS > 1 ? (size_t)KS * 3 * m : 1;
}

void launch_gauss_points(int n, const double* pos, const double* str, const double* rad, int m, const double* pts,
                         double* part, double* out, hipStream_t s) {
    if (m <= 0) return;
    const int KS = direct_slices(m, n);
    const int chunk = (n + KS - 1) / KS;
    hipLaunchKernelGGL(gauss_points_kernel, dim3((unsigned)((m + kT - 1) / kT), (unsigned)KS), dim3(kT), 0, s, n, pos,
                       str, rad, m, pts, chunk, KS > 1 ? part : out);
    CP_LAUNCH_CHECK();
    if (KS > 1) {
        hipLaunchKernelGGL(p2p_reduce_kernel, dim3((unsigned)((3 * m + kT - 1) / kT)), dim3(kT), 0, s,
                           (const double*)part, KS, m, out);
        CP_LAUNCH_CHECK();
    }
}

void launch_filament_points(int nf, const double* ends, const double* fstr, int m, const double* pts, double* out,
                            hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(filament_points_kernel, dim3((unsigned)((m + kT - 1) / kT)), dim3(kT), 0, s, nf, ends, fstr, m,
                       pts, out);
    CP_LAUNCH_CHECK();
}

int p2p_slices(int m, int npanels) {
    int blocks = (m + kT - 1) / kT;
    int ks = 1;
    while (ks < 64 && blocks * ks < 1024 && npanels / (2 * ks) >= 32) ks *= 2;
    return ks;
}

namespace {

constexpr int kSortMin = 4096;  // points below which the near pass keeps their order

__global__ void iota_kernel(int m, int* __restrict__ a) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < m) a[i] = i;
}

// the scratch of one panel-sum launch, in doubles from `part`: KS far and
// KN near slices of 3 per point, the point tiles' spheres, the spatial
// order, the sort's scratch
struct PanelSumPlan {
    int KS, KN, ntiles;
    size_t tsph, perm, sort, total;
};

PanelSumPlan plan_panel_sums(int m, int np) {
    PanelSumPlan P;
    P.KS = p2p_slices(m, np);
    P.ntiles = (m + 63) / 64;
    // the near work gathers in the few tiles close to the body: split the
    // panels finely so those tiles spread over many wavefronts (up to 32
    // slices of at least 48 panels; with few tiles, such as a step's few
    // hundred seed points, up to 256 slices of at least 12 panels until
    // about 1,024 workgroups)
    P.KN = 1;
    while ((P.KN < 32 && np / (2 * P.KN) >= 96) ||
           (P.KN < 256 && (int64_t)P.ntiles * P.KN < 1024 && np / (2 * P.KN) >= 24))
        P.KN *= 2;
    P.tsph = ((size_t)P.KS + P.KN) * 3 * (size_t)m;
    P.perm = P.tsph + 4 * (size_t)P.ntiles;
    P.sort = (P.perm + ((size_t)m + 1) / 2 + 31) / 32 * 32;  // 256-byte aligned
    P.total = P.sort + (m > kSortMin ? (morton_sort_bytes(m) + 7) / 8 : 0);
    return P;
}

// far pass into slices [0, KS), near pass into [KS, KS + KN), their sum
// into out
template <int MODE>
void launch_panel_sums(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                       const double* pts, const double* alpha, double* part, double* out, hipStream_t s,
                       bool f32far) {
    if (m <= 0) return;
    const int np = g.n + g.nw;
    const PanelSumPlan P = plan_panel_sums(m, np);
    const int chunk = (np + P.KS - 1) / P.KS;
    int* perm = reinterpret_cast<int*>(part + P.perm);
    double* tsph = part + P.tsph;
    dim3 grid((unsigned)((m + kT - 1) / kT), (unsigned)P.KS);
    if (f32far)
        hipLaunchKernelGGL((panels_far_kernel<MODE, true>), grid, dim3(kT), 0, s, g.body, g.n, g.wake, g.nw, sigma, mu,
                           muw, m, pts, alpha, chunk, part);
    else
        hipLaunchKernelGGL((panels_far_kernel<MODE, false>), grid, dim3(kT), 0, s, g.body, g.n, g.wake, g.nw, sigma,
                           mu, muw, m, pts, alpha, chunk, part);
    CP_LAUNCH_CHECK();
    // spatial order for the near pass's tiles; a few thousand points (the
    // seeding and collocation calls of the VP step) keep their own order,
    // which the sort's launches would cost more than they save
    if (m > kSortMin)
        sort_points_morton(m, pts, part + P.sort, perm, s);
    else
        hipLaunchKernelGGL(iota_kernel, dim3((unsigned)((m + kT - 1) / kT)), dim3(kT), 0, s, m, perm);
    hipLaunchKernelGGL(point_tiles_kernel, dim3((unsigned)P.ntiles), dim3(64), 0, s, m, pts, (const int*)perm, tsph);
    const int nchunk = (np + P.KN - 1) / P.KN;
    hipLaunchKernelGGL((panels_near_kernel<MODE>), dim3((unsigned)P.ntiles, (unsigned)P.KN), dim3(64), 0, s, g.body,
                       g.n, g.wake, g.nw, sigma, mu, muw, m, pts, alpha, (const int*)perm, (const double*)tsph, nchunk,
                       part + (size_t)P.KS * 3 * m);
    CP_LAUNCH_CHECK();
    hipLaunchKernelGGL(p2p_reduce_kernel, dim3((unsigned)((3 * m + kT - 1) / kT)), dim3(kT), 0, s,
                       (const double*)part, P.KS + P.KN, m, out);
    CP_LAUNCH_CHECK();
}

}  // namespace

size_t p2p_part_doubles(int m, int npanels) { return m > 0 ? plan_panel_sums(m, npanels).total : 1; }

void launch_panel_to_points(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                            const double* pts, double* part, double* out, hipStream_t s, bool f32far) {
    launch_panel_sums<PS_VEL>(g, sigma, mu, muw, m, pts, nullptr, part, out, s, f32far);
}

void launch_panel_stretching(const DevGeom& g, const double* sigma, const double* mu, const double* muw, int m,
                             const double* pos, const double* alpha, double* part, double* out, hipStream_t s,
                             bool f32far) {
    launch_panel_sums<PS_STRETCH>(g, sigma, mu, muw, m, pos, alpha, part, out, s, f32far);
}

void launch_core_bh(const DevBhTree& T, int mode, double theta, double core, bool f32, int m, const double* tpos,
                    const double* tstr, const int* tslot, const int* own_order, double* out, hipStream_t s) {
    if (m <= 0) return;
    const double s2 = core * core;
    // a wavefront per target below kBhcWaveMax caller-given targets (a lane
    // per target leaves the device idle there)
    if (!own_order && m <= kBhcWaveMax) {
        dim3 grid((unsigned)((m + kT / 64 - 1) / (kT / 64)));
#define CP_BHW(MODE, TY)                                                                                      \
    hipLaunchKernelGGL((bhc_wave_kernel<MODE, TY>), grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos, T.mstr, \
                       theta, s2, m, tpos, tstr, tslot, out)
        if (mode == CP_BH_STRETCH) {
            if (f32)
                CP_BHW(KC_STRETCH, float);
            else
                CP_BHW(KC_STRETCH, double);
        } else {
            if (f32)
                CP_BHW(KC_VEL, float);
            else
                CP_BHW(KC_VEL, double);
        }
#undef CP_BHW
        CP_LAUNCH_CHECK();
        return;
    }
    dim3 grid((unsigned)((m + kT - 1) / kT));
#define CP_BHC(MODE, TY)                                                                                      \
    hipLaunchKernelGGL((bhc_kernel<MODE, TY>), grid, dim3(kT), 0, s, T.nodes, T.n_nodes, T.mpos, T.mstr, theta, \
                       s2, m, tpos, tstr, tslot, own_order, out)
    if (mode == CP_BH_STRETCH) {
        if (f32)
            CP_BHC(KC_STRETCH, float);
        else
            CP_BHC(KC_STRETCH, double);
    } else {
        if (f32)
            CP_BHC(KC_VEL, float);
        else
            CP_BHC(KC_VEL, double);
    }
#undef CP_BHC
    CP_LAUNCH_CHECK();
}

void launch_core_direct(int mode, double core, bool f32, int n, const double* pos, const double* str, int m,
                        const double* tpos, const double* tstr, bool own, double* part, double* out, hipStream_t s) {
    if (m <= 0) return;
    const int KS = direct_slices(m, n);
    const int chunk = (n + KS - 1) / KS;
    dim3 grid((unsigned)((m + kT - 1) / kT), (unsigned)KS);
    const double s2 = core * core;
    double* dst = KS > 1 ? part : out;
#define CP_DC(MODE, TY)                                                                                        \
    hipLaunchKernelGGL((dc_kernel<MODE, TY>), grid, dim3(kT), 0, s, n, pos, str, s2, m, tpos, tstr, (int)own, \
                       chunk, dst)
    if (mode == CP_BH_STRETCH) {
        if (f32)
            CP_DC(KC_STRETCH, float);
        else
            CP_DC(KC_STRETCH, double);
    } else {
        if (f32)
            CP_DC(KC_VEL, float);
        else
            CP_DC(KC_VEL, double);
    }
#undef CP_DC
    CP_LAUNCH_CHECK();
    if (KS > 1) {  // the slices' partial sums in slice order
        hipLaunchKernelGGL(p2p_reduce_kernel, dim3((unsigned)((3 * m + kT - 1) / kT)), dim3(kT), 0, s,
                           (const double*)part, KS, m, out);
        CP_LAUNCH_CHECK();
    }
}

}  // namespace cpgpu
