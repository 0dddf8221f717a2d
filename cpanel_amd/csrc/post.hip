// Post-solve outputs on the device for k conditions at once:
//   phi_i = Vinf . c_i - mu_i               (panel::setPotential, panel.cpp:106-111)
//   V_i   = CHTLS weighted sum over the cluster, NaN fallbacks, V_x /= PG
//                                            (bodyPanel::pntVelocity, bodyPanel.cpp:490-524)
//   Cp_i  = 1 - (|V_i| / V)^2               (bodyPanel::computeCp, :591-594)
//   F, CM sums                               (cpCase::compVelocity, cpCase.cpp:342-365)
//   Trefftz CL, CDi                          (wake::trefftzPlane, wake.cpp:146-188)
// Sums use a fixed reduction tree, so a condition's outputs do not depend on
// which other conditions share the launch.
#include "device.h"
#include "post.h"

namespace cpgpu {

namespace {

constexpr int kT = 256;
constexpr int kNP = 6;  // partial sums per block: F[3], M[3]

__device__ double block_sum_post(double v, double* sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] += sh[t + o];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// sum over the 64 lanes of a wavefront with a fixed butterfly (every lane
// gets the total)
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int kPP = 64;   // panels per workgroup (one per lane)
constexpr int kPC = 16;   // conditions per workgroup (each wavefront every 4th)

// Per (panel, condition): potential, CHTLS velocity, Cp and the force and
// moment contributions.  A workgroup owns 64 consecutive panels and 16
// conditions: the panels' weights and neighbour lists (contiguous, 45 + 16
// words per panel) are staged in LDS once and serve every condition, so each
// weight is read from memory once per 16 conditions instead of once per
// condition.  Partial sums per (tile, condition) come from a wavefront sum.
__global__ __launch_bounds__(256) void post_panel_kernel(
    int n, int k, const int32_t* __restrict__ nbr, const double* __restrict__ w1, const double* __restrict__ w2,
    const int32_t* __restrict__ mode, const double* __restrict__ ctr, const double* __restrict__ bez,
    const double* __restrict__ area, PostParams pp, const double* __restrict__ mu, int64_t ldmu,
    const double* __restrict__ vinf4, double* __restrict__ cp, int64_t ldcp, double* __restrict__ partials,
    double* __restrict__ vel, int64_t ldvel, double* __restrict__ phi, int64_t ldphi,
    const double* __restrict__ vpot) {
    __shared__ double sw[kPP * 45];
    __shared__ int32_t snb[kPP * 16];
    const int i0 = blockIdx.x * kPP;
    const int np = min(kPP, n - i0);
    for (int e = threadIdx.x; e < np * 45; e += 256) sw[e] = w1[45 * (int64_t)i0 + e];
    for (int e = threadIdx.x; e < np * 16; e += 256) snb[e] = nbr[16 * (int64_t)i0 + e];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int i = i0 + lane;
    const bool live = lane < np;
    const double* cx = ctr;
    const double* cy = ctr + n;
    const double* cz = ctr + 2 * (int64_t)n;
    double bx = 0, by = 0, bz = 0, a = 0, r0 = 0, r1 = 0, r2 = 0;
    int cnt = 0, md = 0;
    if (live) {
        bx = bez[i];
        by = bez[n + i];
        bz = bez[2 * (int64_t)n + i];
        a = area[i];
        r0 = cx[i] - pp.cg[0];
        r1 = cy[i] - pp.cg[1];
        r2 = cz[i] - pp.cg[2];
        cnt = snb[16 * lane];
        md = mode[i];
    }
    const double* w = sw + 45 * lane;
    const int32_t* nb = snb + 16 * lane + 1;
    const int c1 = min(k, (int)(blockIdx.y + 1) * kPC);
    for (int c = blockIdx.y * kPC + (threadIdx.x >> 6); c < c1; c += 4) {
        const double Vx = vinf4[4 * c], Vy = vinf4[4 * c + 1], Vz = vinf4[4 * c + 2], Vm = vinf4[4 * c + 3];
        const double* m = mu + (int64_t)c * ldmu;
        // the velocity a panel's potential is formed with: the freestream,
        // or a per-panel one (the vortex-particle case adds the particle and
        // filament velocity at the centre, cpCaseVP::VinfPlusVecPot,
        // cpCaseVP.cpp:601-623)
        const double* vp = vpot ? vpot + (int64_t)c * 3 * n : nullptr;
        auto pot = [&](int j) {
            if (vp) return vp[3 * j] * cx[j] + vp[3 * j + 1] * cy[j] + vp[3 * j + 2] * cz[j] - m[j];
            return Vx * cx[j] + Vy * cy[j] + Vz * cz[j] - m[j];
        };
        double f[kNP] = {0, 0, 0, 0, 0, 0};
        if (live) {
            const double phi_i = pot(i);
            double df[kClusterMax];
            for (int q = 0; q < cnt; q++) df[q] = pot(nb[q]) - phi_i;
            double v0 = 0, v1 = 0, v2 = 0;
            for (int q = 0; q < cnt; q++) {
                v0 += w[3 * q] * df[q];
                v1 += w[3 * q + 1] * df[q];
                v2 += w[3 * q + 2] * df[q];
            }
            bool bad = cnt == 0 || (v0 != v0) || (v1 != v1) || (v2 != v2);
            if (bad && md == 1) {
                const double* ww = w2 + 45 * (int64_t)i;
                v0 = v1 = v2 = 0;
                for (int q = 0; q < cnt; q++) {
                    v0 += ww[3 * q] * df[q];
                    v1 += ww[3 * q + 1] * df[q];
                    v2 += ww[3 * q + 2] * df[q];
                }
                bad = cnt == 0 || (v0 != v0) || (v1 != v1) || (v2 != v2);
            }
            if (bad) {
                // (bez x Vinf) x bez   (bodyPanel.cpp:516-520)
                double q0 = by * Vz - bz * Vy, q1 = bz * Vx - bx * Vz, q2 = bx * Vy - by * Vx;
                v0 = q1 * bz - q2 * by;
                v1 = q2 * bx - q0 * bz;
                v2 = q0 * by - q1 * bx;
            }
            v0 /= pp.PG;
            if (vel) {
                double* vo = vel + (int64_t)c * ldvel + 3 * (int64_t)i;
                vo[0] = v0;
                vo[1] = v1;
                vo[2] = v2;
            }
            if (phi) phi[(int64_t)c * ldphi + i] = phi_i;
            double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2) / Vm;
            double Cp = 1 - vn * vn;
            if (cp) cp[(int64_t)c * ldcp + i] = Cp;
            double F0 = -Cp * a * bx, F1 = -Cp * a * by, F2 = -Cp * a * bz;
            f[0] = F0 / pp.Sref;
            f[1] = F1 / pp.Sref;
            f[2] = F2 / pp.Sref;
            f[3] = (r1 * F2 - r2 * F1) / (pp.Sref * pp.bref);
            f[4] = (r2 * F0 - r0 * F2) / (pp.Sref * pp.cref);
            f[5] = (r0 * F1 - r1 * F0) / (pp.Sref * pp.bref);
        }
        for (int q = 0; q < kNP; q++) {
            const double t = wave_sum(f[q]);
            if (lane == 0) partials[((int64_t)c * gridDim.x + blockIdx.x) * kNP + q] = t;
        }
    }
}

// one block per condition: force sums and the Trefftz plane
__global__ __launch_bounds__(256) void post_final_kernel(
    int nblk, const double* __restrict__ partials, int n_wakes, int n_st, int n_te,
    const int32_t* __restrict__ te_idx, const double* __restrict__ rphi, const double* __restrict__ rv,
    const double* __restrict__ rw, const double* __restrict__ simpson, PostParams pp,
    const double* __restrict__ mu, int64_t ldmu, const double* __restrict__ vinf4, double* __restrict__ res,
    double* __restrict__ span) {
    __shared__ double sh[kT];
    const int c = blockIdx.x;
    const double* m = mu + (int64_t)c * ldmu;
    const double Vm = vinf4[4 * c + 3];
    double out[kNP];
    for (int q = 0; q < kNP; q++) {
        double s = 0;
        for (int b = threadIdx.x; b < nblk; b += kT) s += partials[((int64_t)c * nblk + b) * kNP + q];
        out[q] = block_sum_post(s, sh);
    }
    double CL = 0, CD = 0;
    for (int w = 0; w < n_wakes; w++) {
        double cl = 0, cd = 0;
        for (int st = threadIdx.x; st < n_st; st += kT) {
            const int64_t row = ((int64_t)w * n_st + st) * n_te;
            double dphi = 0, v = 0, ww = 0;
            for (int k = 0; k < n_te; k++) {
                double mk = m[te_idx[k]];
                dphi += rphi[row + k] * mk;
                v += rv[row + k] * mk;
                ww += rw[row + k] * mk;
            }
            double sw = simpson[(int64_t)w * n_st + st];
            double Cl = 2 * dphi / (Vm * pp.Sref);
            double Cd = dphi * sqrt(v * v + ww * ww) / (Vm * Vm * pp.Sref);
            cl += sw * Cl;
            cd += sw * Cd;
            if (span) {
                double* so = span + (((int64_t)c * n_wakes + w) * n_st + st) * 2;
                so[0] = Cl / pp.PG;
                so[1] = Cd / (pp.PG * pp.PG);
            }
        }
        cl = block_sum_post(cl, sh);
        cd = block_sum_post(cd, sh);
        CL += cl / pp.PG;
        CD += cd / (pp.PG * pp.PG);
    }
    if (threadIdx.x == 0) {
        double* r = res + 16 * (int64_t)c;
        r[0] = CL;
        r[1] = CD;
        r[2] = out[0];
        r[3] = out[1];
        r[4] = out[2];
        r[8] = out[3];
        r[9] = out[4];
        r[10] = out[5];
    }
}

// wakePanel::setMu (wakePanel.cpp:94-100) and panel::setPotential
// (panel.cpp:106-111) for every wake panel and condition
__global__ __launch_bounds__(256) void post_wake_kernel(int nw, const int32_t* __restrict__ fidx,
                                                        const double* __restrict__ ft,
                                                        const double* __restrict__ wctr, const double* __restrict__ mu,
                                                        int64_t ldmu, const double* __restrict__ vinf4,
                                                        double* __restrict__ wmu, double* __restrict__ wphi) {
    const int c = blockIdx.y;
    const int j = blockIdx.x * kT + threadIdx.x;
    if (j >= nw) return;
    const double* m = mu + (int64_t)c * ldmu;
    const double t = ft[j];
    const int32_t* q = fidx + 4 * (int64_t)j;
    const double u = (1 - t) * m[q[0]] + (t - 1) * m[q[1]] + t * m[q[2]] - t * m[q[3]];
    const double Vx = vinf4[4 * c], Vy = vinf4[4 * c + 1], Vz = vinf4[4 * c + 2];
    if (wmu) wmu[(int64_t)c * nw + j] = u;
    if (wphi) wphi[(int64_t)c * nw + j] = Vx * wctr[j] + Vy * wctr[nw + j] + Vz * wctr[2 * (int64_t)nw + j] - u;
}

}  // namespace

void launch_post_wake(int nw, const int32_t* fidx, const double* ft, const double* wctr, int k, const double* mu,
                      int64_t ldmu, const double* vinf4, double* wmu, double* wphi, hipStream_t s) {
    if (nw <= 0 || k <= 0) return;
    dim3 g((unsigned)((nw + kT - 1) / kT), (unsigned)k);
    hipLaunchKernelGGL(post_wake_kernel, g, dim3(kT), 0, s, nw, fidx, ft, wctr, mu, ldmu, vinf4, wmu, wphi);
    CP_LAUNCH_CHECK();
}

int post_partial_blocks(int n) { return (n + kPP - 1) / kPP; }

void launch_post(const PostDev& P, const PostParams& pp, int k, const double* mu, int64_t ldmu,
                 const double* vinf4, double* cp, int64_t ldcp, double* partials, int nblk, double* res,
                 hipStream_t s, const PostFields* f) {
    const PostFields none;
    if (!f) f = &none;
    dim3 g1((unsigned)nblk, (unsigned)((k + kPC - 1) / kPC));
    hipLaunchKernelGGL(post_panel_kernel, g1, dim3(kT), 0, s, P.n, k, P.nbr, P.w1, P.w2, P.mode, P.ctr, P.bez,
                       P.area, pp, mu, ldmu, vinf4, cp, ldcp, partials, f->vel, f->ldvel, f->phi, f->ldphi, f->vpot);
    CP_LAUNCH_CHECK();
    hipLaunchKernelGGL(post_final_kernel, dim3((unsigned)k), dim3(kT), 0, s, nblk, partials, P.n_wakes, P.n_st,
                       P.n_te, P.te_idx, P.rphi, P.rv, P.rw, P.simpson, pp, mu, ldmu, vinf4, res, f->span);
    CP_LAUNCH_CHECK();
}

}  // namespace cpgpu
