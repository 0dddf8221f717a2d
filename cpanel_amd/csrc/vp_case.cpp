// Vortex-particle case: the time-step loop of cpCaseVP::run (cpCaseVP.cpp:
// 26-158) on the device kernels, behind cp_vp_* in cpgpu.h.
//
// State lives on the host between steps (particles, buffer-row strengths,
// filaments), the way the reference's cpCaseVP holds it; every sum over
// panels, particles or filaments runs on the geometry's stream:
//   * panel -> point velocity (body sources and doublets, first buffer row)
//     and the second buffer row's doublets: p2p_kernel;
//   * filaments: filament_points_kernel;
//   * particles: Barnes-Hut on the tree built after the previous strength
//     update, its leaves refreshed to the particles' current state (the
//     reference's octree holds particle pointers, its pseudo-particles are
//     fixed at the build), or the direct sums where the reference uses them
//     (time steps <= 2 or 3, or with acceleration off);
//   * panel -> particle stretching: stretch_kernel.
// The configured form (cp_vp_opts kernel / per_edge / core / fp32) swaps the
// particle kernel for the algebraic one with a fixed core (no diffusion
// term), sheds q particles per second-row panel, and evaluates the particle
// sums and the far-field panel sums in FP32; the step order is the same.
// The order of the steps, the far-field cut-offs (12 and 20 reference
// chords), the Adams-Bashforth updates and the RK4 seeding of new particles
// follow cpCaseVP.cpp:160-734; wake::trefftzPlaneVP (wake.cpp:191-325) gives
// the convergence test and the final CL, CDi.
#include <chrono>

#include "capi_internal.h"

using namespace cpgpu;
using namespace cpgpu::capi;

namespace {

constexpr double kPi = 3.14159265358979323846;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct V3 {
    double x = 0, y = 0, z = 0;
};

}  // namespace

struct cp_vp {
    cp_geom* g = nullptr;
    cp_vp_opts o{};
    int ts = 0;          // time steps completed
    int n = 0, m = 0;    // body panels; panels per buffer row
    double dt = 0;
    double Vb[3] = {0, 0, 0};
    double Vmag = 1, PG = 1;
    std::vector<double> mu, sigma, mu1, mu2, mu2prev;
    std::vector<double> vpot;  // velocity of each panel's potential, last solve
    double res = 0;
    int iters = 0;
    bool have_fil = false;
    std::vector<double> fil_ends, fil_str;
    // particles, in creation order
    std::vector<double> ppos, pstr, prad, pvel0, pdst0;
    std::vector<int> pshed, pparent;
    // where each particle was seeded: its second-row panel's centre, or the
    // point along the panel's span of a split panel (3 per particle)
    std::vector<double> pseed;
    // the configured form: algebraic kernel of core radius `core`, q
    // particles per panel and step, FP32 sums
    bool alg = false, f32 = false;
    int q = 1;
    double core = 0;
    // tree of the last build (its slot array maps the particles of that
    // build to their member slots)
    std::unique_ptr<cp_tree> tree;
    double CL = 0, CD = 0;
    bool converged = false;
    DevBuf w2rec, fil_dev;
    DevGeom row2;
    // the particles' positions, strengths and radii on the device: uploaded
    // when the host state (version ver, bumped by every update) is newer than
    // the copy (dev_ver); the tree's leaves hold version leaves_ver
    DevBuf dpos, dstr, drad;
    uint64_t ver = 1, dev_ver = 0, leaves_ver = 0;
};

namespace {

// ---------------------------------------------------------------- device sums
// each returns 3 per point on the host; all run on the geometry's stream

void to_dev(cp_ctx* c, DevBuf& b, const std::vector<double>& v, hipStream_t s) {
    dalloc(c, b, std::max<size_t>(v.size(), 1) * sizeof(double));
    if (!v.empty()) CP_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, s));
}

void from_dev(std::vector<double>& v, const DevBuf& b, size_t count, hipStream_t s) {
    v.resize(count);
    if (count) CP_HIP(hipMemcpyAsync(v.data(), b.p, count * sizeof(double), hipMemcpyDeviceToHost, s));
    CP_HIP(hipStreamSynchronize(s));
}

// The particle state on the device, uploaded once per change.  Every sum
// ends with a synchronisation of the stream before the host updates the
// state, so no copy is in flight when the host arrays change.
void sync_particles(cp_vp* v) {
    if (v->dev_ver == v->ver) return;
    cp_ctx* c = v->g->ctx;
    hipStream_t s = v->g->stream;
    auto put = [&](DevBuf& b, const std::vector<double>& h) {
        const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(double);
        if (b.bytes < bytes) dalloc(c, b, bytes + bytes / 8);  // headroom: the particles grow every step
        if (!h.empty()) CP_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    };
    put(v->dpos, v->ppos);
    put(v->dstr, v->pstr);
    put(v->drad, v->prad);
    v->dev_ver = v->ver;
}

// points on the device: the particles' own positions when pts is the
// particle array, else a copy in tmp
const double* dev_points(cp_vp* v, const std::vector<double>& pts, DevBuf& tmp) {
    if (&pts == &v->ppos) {
        sync_particles(v);
        return v->dpos.as<double>();
    }
    to_dev(v->g->ctx, tmp, pts, v->g->stream);
    return tmp.as<double>();
}

// body panels (sigma, mu) and the first buffer row (mu1): bodyPanel::panelV,
// wakePanel::panelV summed over the panels
std::vector<double> vel_panels(cp_vp* v, const std::vector<double>& pts) {
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    const int m = (int)(pts.size() / 3);
    std::vector<double> out;
    if (m == 0) return out;
    DevBuf dsig, dmu, dmw, dpts, dpart, dout;
    to_dev(c, dsig, v->sigma, s);
    to_dev(c, dmu, v->mu, s);
    to_dev(c, dmw, v->mu1, s);
    const double* P = dev_points(v, pts, dpts);
    dalloc(c, dpart, p2p_part_doubles(m, g->dev.n + g->dev.nw) * sizeof(double));
    dalloc(c, dout, (size_t)3 * m * sizeof(double));
    {
        TimedScope ts("panel_to_points", s);
        launch_panel_to_points(g->dev, dsig.as<double>(), dmu.as<double>(), dmw.as<double>(), m, P,
                               dpart.as<double>(), dout.as<double>(), s, v->f32);
    }
    from_dev(out, dout, (size_t)3 * m, s);
    return out;
}

// the second buffer row's doublets (mu2)
std::vector<double> vel_row2(cp_vp* v, const std::vector<double>& pts) {
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    const int m = (int)(pts.size() / 3);
    std::vector<double> out;
    if (m == 0) return out;
    DevBuf dmw, dpts, dpart, dout;
    to_dev(c, dmw, v->mu2, s);
    const double* P = dev_points(v, pts, dpts);
    dalloc(c, dpart, p2p_part_doubles(m, v->m) * sizeof(double));
    dalloc(c, dout, (size_t)3 * m * sizeof(double));
    launch_panel_to_points(v->row2, nullptr, nullptr, dmw.as<double>(), m, P, dpart.as<double>(), dout.as<double>(),
                           s, v->f32);
    from_dev(out, dout, (size_t)3 * m, s);
    return out;
}

std::vector<double> vel_filaments(cp_vp* v, const std::vector<double>& pts) {
    const int m = (int)(pts.size() / 3);
    std::vector<double> out((size_t)3 * m, 0.0);
    if (!v->have_fil || m == 0) return out;
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    DevBuf dstr, dpts, dout;
    to_dev(c, dstr, v->fil_str, s);
    const double* P = dev_points(v, pts, dpts);
    dalloc(c, dout, (size_t)3 * m * sizeof(double));
    launch_filament_points(v->m, v->fil_dev.as<double>(), dstr.as<double>(), m, P, dout.as<double>(), s);
    from_dev(out, dout, (size_t)3 * m, s);
    return out;
}

// direct point velocity of every particle: the Gaussian particle::velInfl(
// POI), or the algebraic kernel in the configured form
std::vector<double> vel_direct(cp_vp* v, const std::vector<double>& pts) {
    const int m = (int)(pts.size() / 3), np = (int)v->prad.size();
    std::vector<double> out((size_t)3 * m, 0.0);
    if (m == 0 || np == 0) return out;
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    DevBuf dpts, dpart, dout;
    const double* P = dev_points(v, pts, dpts);
    sync_particles(v);
    dalloc(c, dpart, direct_part_doubles(m, np) * sizeof(double));
    dalloc(c, dout, (size_t)3 * m * sizeof(double));
    if (v->alg)
        launch_core_direct(CP_BH_VEL, v->core, v->f32, np, v->dpos.as<double>(), v->dstr.as<double>(), m, P, nullptr,
                           false, dpart.as<double>(), dout.as<double>(), s);
    else
        launch_gauss_points(np, v->dpos.as<double>(), v->dstr.as<double>(), v->drad.as<double>(), m, P,
                            dpart.as<double>(), dout.as<double>(), s);
    from_dev(out, dout, (size_t)3 * m, s);
    return out;
}

// the tree's leaves take the particles' current state
void refresh_tree(cp_vp* v) {
    cp_tree* t = v->tree.get();
    if (!t || t->dev.n == 0) return;
    if (v->leaves_ver == v->ver) return;
    hipStream_t s = v->g->stream;
    sync_particles(v);
    launch_bh_refresh(t->dev, v->dpos.as<double>(), v->dstr.as<double>(), v->drad.as<double>(), s);
    CP_HIP(hipStreamSynchronize(s));
    v->leaves_ver = v->ver;
}

bool tree_usable(const cp_vp* v) { return v->tree && v->tree->dev.n_nodes > 1; }

// Barnes-Hut velocity at points (algebraic kernel, particleFMM::barnesHutVel)
std::vector<double> vel_bh_points(cp_vp* v, const std::vector<double>& pts) {
    const int m = (int)(pts.size() / 3);
    std::vector<double> out((size_t)3 * m, 0.0);
    if (m == 0 || !tree_usable(v)) return out;
    refresh_tree(v);
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    DevBuf dpts, dout;
    to_dev(c, dpts, pts, s);
    dalloc(c, dout, (size_t)3 * m * sizeof(double));
    {
        TimedScope ts("bh_eval", s);
        if (v->alg)
            launch_core_bh(v->tree->dev, CP_BH_VEL, v->tree->theta, v->core, v->f32, m, dpts.as<double>(), nullptr,
                           nullptr, nullptr, dout.as<double>(), s);
        else
            launch_bh_points(v->tree->dev, v->tree->theta, m, dpts.as<double>(), dout.as<double>(), s);
    }
    from_dev(out, dout, (size_t)3 * m, s);
    return out;
}

// Barnes-Hut sums on every current particle (particle-target modes)
std::vector<double> bh_on_particles(cp_vp* v, int mode) {
    const int np = (int)v->prad.size();
    std::vector<double> out((size_t)3 * np, 0.0);
    if (np == 0 || !tree_usable(v)) return out;
    refresh_tree(v);
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    DevBuf dsl, dout;
    sync_particles(v);
    const double* dp = v->dpos.as<double>();
    const double* ds = v->dstr.as<double>();
    const double* dr = v->drad.as<double>();
    // Above kBhWaveTargets particles (a lane per target), the tree's members
    // (the particles of its build, refreshed to their current state) walk as
    // the tree's own targets, in its spatial order (coherent wavefronts),
    // their sums written back through its order array, and only the particles
    // shed since the build (no slot: -1) as caller-given targets.  Each
    // target's walk and self-exclusion are those of a caller-given target
    // with its slot: the same sums.  Below it every particle is a
    // caller-given target (a wavefront each).
    const DevBhTree& T = v->tree->dev;
    const int nb = np > kBhWaveTargets ? std::min(np, T.n) : 0;
    dalloc(c, dout, (size_t)3 * np * sizeof(double));
    if (np > nb) {  // the caller-given targets' slots
        const int nm = std::min(np, T.n);
        dalloc(c, dsl, (size_t)(np - nb) * sizeof(int));
        if (nm > nb)
            CP_HIP(hipMemcpyAsync(dsl.p, v->tree->slot.as<int>() + nb, (size_t)(nm - nb) * sizeof(int),
                                  hipMemcpyDeviceToDevice, s));
        if (np > nm) CP_HIP(hipMemsetAsync(dsl.as<int>() + (nm - nb), 0xFF, (size_t)(np - nm) * sizeof(int), s));
    }
    {
        TimedScope ts("bh_eval", s);
        if (nb > 0) {
            if (v->alg)
                launch_core_bh(T, mode, v->tree->theta, v->core, v->f32, nb, T.mpos, T.mstr, nullptr, T.order,
                               dout.as<double>(), s);
            else
                launch_bh_particles(T, mode, v->tree->theta, dout.as<double>(), s);
        }
        if (np > nb) {
            double* o = dout.as<double>() + 3 * (size_t)nb;
            if (v->alg)
                launch_core_bh(T, mode, v->tree->theta, v->core, v->f32, np - nb, dp + 3 * (size_t)nb,
                               ds + 3 * (size_t)nb, dsl.as<int>(), nullptr, o, s);
            else
                launch_bh_targets(T, mode, v->tree->theta, np - nb, dp + 3 * (size_t)nb, ds + 3 * (size_t)nb,
                                  dr + nb, dsl.as<int>(), o, s);
        }
    }
    from_dev(out, dout, (size_t)3 * np, s);
    return out;
}

std::vector<double> pairs_direct(cp_vp* v, int mode) {
    const int np = (int)v->prad.size();
    std::vector<double> out((size_t)3 * np, 0.0);
    if (np == 0) return out;
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    DevBuf dpart, dout;
    sync_particles(v);
    const double* dp = v->dpos.as<double>();
    const double* ds = v->dstr.as<double>();
    dalloc(c, dout, (size_t)3 * np * sizeof(double));
    if (v->alg) {
        dalloc(c, dpart, direct_part_doubles(np, np) * sizeof(double));
        launch_core_direct(mode, v->core, v->f32, np, dp, ds, np, dp, ds, true, dpart.as<double>(), dout.as<double>(),
                           s);
    } else {
        launch_pairs_direct(mode, np, dp, ds, v->drad.as<double>(), dout.as<double>(), s);
    }
    from_dev(out, dout, (size_t)3 * np, s);
    return out;
}

// body and first-row panel stretching on the particles (partStretching)
std::vector<double> panel_stretch(cp_vp* v) {
    const int np = (int)v->prad.size();
    std::vector<double> out((size_t)3 * np, 0.0);
    if (np == 0) return out;
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    hipStream_t s = g->stream;
    DevBuf dsig, dmu, dmw, dpart, dout;
    to_dev(c, dsig, v->sigma, s);
    to_dev(c, dmu, v->mu, s);
    to_dev(c, dmw, v->mu1, s);
    sync_particles(v);
    dalloc(c, dpart, p2p_part_doubles(np, g->dev.n + g->dev.nw) * sizeof(double));
    dalloc(c, dout, (size_t)3 * np * sizeof(double));
    {
        TimedScope ts("panel_stretching", s);
        launch_panel_stretching(g->dev, dsig.as<double>(), dmu.as<double>(), dmw.as<double>(), np,
                                v->dpos.as<double>(), v->dstr.as<double>(), dpart.as<double>(), dout.as<double>(), s,
                                v->f32);
    }
    from_dev(out, dout, (size_t)3 * np, s);
    return out;
}

void build_tree(cp_vp* v) {
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    const int np = (int)v->prad.size();
    std::unique_ptr<cp_tree> t(new cp_tree);
    t->ctx = c;
    t->theta = v->o.theta >= 0 ? v->o.theta : 0.5;
    if (v->alg) {
        t->kernel = CP_VP_ALGEBRAIC;
        t->core = v->core;
        t->fp32 = v->f32;
    }
    const int maxm = v->o.max_members > 0 ? v->o.max_members : 10;
    bool built = false;
    if (np > 0) {  // on the device (the host only when the keys are too shallow)
        hipStream_t s = g->stream;
        sync_particles(v);
        TimedScope ts("bh_build", s);
        built = build_tree_device(c, np, v->dpos.as<double>(), v->dstr.as<double>(), v->drad.as<double>(), maxm, s,
                                  t.get());
        CP_HIP(hipStreamSynchronize(s));
    }
    if (!built) build_tree_host(c, np, v->ppos.data(), v->pstr.data(), v->prad.data(), maxm, t.get());
    v->tree = std::move(t);
    v->leaves_ver = 0;  // refreshed before its first use, as every tree is
}

// ---------------------------------------------------------------- helpers

// wakePanel::setMu over the first buffer row (wakePanel.cpp:94-100)
void fold_row1(cp_vp* v) {
    const Geometry& G = v->g->host;
    for (int j = 0; j < v->m; j++) {
        const double t = G.fold_coef[j];
        const int* q = &G.fold_idx[4 * (size_t)j];
        v->mu1[j] = (1 - t) * v->mu[q[0]] + (t - 1) * v->mu[q[1]] + t * v->mu[q[2]] - t * v->mu[q[3]];
    }
}

// cpCaseVP::convectBufferWake (cpCaseVP.cpp:160-170)
void convect_buffer(cp_vp* v) {
    for (int i = 0; i < v->m; i++) {
        v->mu2prev[i] = v->mu2[i];
        v->mu2[i] = v->mu1[v->g->host.w2.parent[i]];
    }
}

// A mu = -B sigma - C mu2 (cpCaseVP::solveVPmatrixEq; with_c false: the
// first step's cpCase::solveMatrixEq)
cp_status solve(cp_vp* v, bool with_c) {
    cp_solve_opts so{};
    so.mach = v->o.mach;
    so.Sref = v->o.Sref;
    so.bref = v->o.bref;
    so.cref = v->o.cref;
    for (int a = 0; a < 3; a++) so.cg[a] = v->o.cg[a];
    so.tol = v->o.tol > 0 ? v->o.tol : 1e-12;
    std::vector<double> x0;
    const bool warm = v->o.warm_start && v->ts > 1;
    cp_solve_info info{};
    std::vector<double> mu(v->n);
    cp_status st = cp_solve_rhs(v->g, (cp_prec)v->o.prec, 1, v->sigma.data(), with_c ? v->mu2.data() : nullptr,
                                warm ? v->mu.data() : nullptr, &so, mu.data(), &info);
    if (st != CP_OK && st != CP_NOT_CONVERGED) return st;
    v->mu = mu;
    v->res = info.rel_residual;
    v->iters = info.iters;
    fold_row1(v);
    return st;
}

// velocities of the particles for convection (velocityInflFromEverything(
// particle*), cpCaseVP.cpp:680-734)
std::vector<double> particle_velocities(cp_vp* v) {
    const int np = (int)v->prad.size();
    std::vector<double> vel((size_t)3 * np);
    if (np == 0) return vel;
    const double c20 = 20 * v->o.cref, c12 = 12 * v->o.cref;
    std::vector<double> part;
    if (v->o.accelerate && v->ts > 2)
        part = bh_on_particles(v, CP_BH_VEL);
    else
        part = vel_direct(v, v->ppos);
    const std::vector<double> pan = vel_panels(v, v->ppos);
    const std::vector<double> r2 = vel_row2(v, v->ppos);
    const std::vector<double> fil = vel_filaments(v, v->ppos);
    for (int i = 0; i < np; i++) {
        const double* p = &v->ppos[3 * (size_t)i];
        const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (int q = 0; q < 3; q++) {
            double u = v->Vb[q];
            if (!(r > c20)) {
                u += part[3 * i + q];
                if (!(r > c12)) u = u + pan[3 * i + q] + r2[3 * i + q] + fil[3 * i + q];
            }
            vel[3 * i + q] = u;
        }
    }
    return vel;
}

// velocityInflFromEverything(POI) (cpCaseVP.cpp:641-678) at points, with the
// particle term from the tree (algebraic) or the direct Gaussian sum
std::vector<double> everything_at(cp_vp* v, const std::vector<double>& pts, bool bh) {
    const int m = (int)(pts.size() / 3);
    const std::vector<double> part = bh ? vel_bh_points(v, pts) : vel_direct(v, pts);
    const std::vector<double> pan = vel_panels(v, pts);
    const std::vector<double> r2 = vel_row2(v, pts);
    const std::vector<double> fil = vel_filaments(v, pts);
    std::vector<double> out((size_t)3 * m);
    for (int i = 0; i < 3 * m; i++) out[i] = v->Vb[i % 3] + part[i] + pan[i] + r2[i] + fil[i];
    return out;
}

// cpCaseVP::rungeKuttaStepper for a set of seed points
std::vector<double> rk4(cp_vp* v, const std::vector<double>& P, bool bh) {
    const double dt = v->dt;
    const size_t N = P.size();
    std::vector<double> k1 = everything_at(v, P, bh), x(N);
    for (size_t i = 0; i < N; i++) x[i] = P[i] + k1[i] * dt / 2;
    std::vector<double> k2 = everything_at(v, x, bh);
    for (size_t i = 0; i < N; i++) x[i] = P[i] + k2[i] * dt / 2;
    std::vector<double> k3 = everything_at(v, x, bh);
    for (size_t i = 0; i < N; i++) x[i] = P[i] + k3[i] * dt;
    std::vector<double> k4 = everything_at(v, x, bh);
    std::vector<double> out(N);
    for (size_t i = 0; i < N; i++) out[i] = P[i] + dt * (k1[i] / 6 + k2[i] / 3 + k3[i] / 3 + k4[i] / 6);
    return out;
}

// cpCaseVP::convectParticles (cpCaseVP.cpp:544-570), Adams-Bashforth
void convect_particles(cp_vp* v) {
    const int np = (int)v->prad.size();
    if (np == 0) return;
    const std::vector<double> vel = particle_velocities(v);
    const double dt = v->dt;
    for (int i = 0; i < np; i++) {
        double* p0 = &v->pvel0[3 * (size_t)i];
        const bool first = p0[0] == 0 && p0[1] == 0 && p0[2] == 0;
        for (int q = 0; q < 3; q++) {
            const double u = vel[3 * i + q];
            double& x = v->ppos[3 * (size_t)i + q];
            x = first ? x + dt * u : x + dt * (1.5 * u - 0.5 * p0[q]);
        }
        for (int q = 0; q < 3; q++) p0[q] = vel[3 * i + q];
    }
    v->ver++;
}

// cpCaseVP::collapseBufferWake (cpCaseVP.cpp:319-367): one particle per
// second-row panel (strength from the panel's edges, RK4 seed position,
// radius from the neighbours' spacing), then the filaments
void collapse_buffer(cp_vp* v) {
    const VpBuffer& W = v->g->host.w2;
    const int m = v->m;
    std::vector<V3> str(m);
    for (int i = 0; i < m; i++) {  // wakePanel::edgeStrength (wakePanel.cpp:405-451)
        const double mu = v->mu2[i];
        V3 s;
        for (int j = 1; j < 4; j++) {
            if (!W.eio_first[i][j]) continue;
            Vec3 Rj, Ri;
            double k;
            if (j == 2) {
                Rj = W.pio[i][2];
                Ri = W.pio[i][3];
                k = mu - v->mu2prev[i];
            } else {
                Rj = j == 1 ? W.pio[i][1] : W.pio[i][3];
                Ri = j == 1 ? W.pio[i][2] : W.pio[i][0];
                const int o = W.eio_nbr[i][j];
                k = o >= 0 ? mu - v->mu2[o] : mu;
            }
            s.x += k * (Ri.x - Rj.x);
            s.y += k * (Ri.y - Rj.y);
            s.z += k * (Ri.z - Rj.z);
        }
        str[i] = s;
    }
    // seeds: the panel's centre, or with q > 1 particles per panel the
    // points (k + 0.5) / q along its span (the mean of its two spanwise
    // edges, pointsInOrder 0 -> 1 and 3 -> 2), each carrying 1/q of the
    // panel's edge strength
    const int q = v->q;
    std::vector<double> seed((size_t)3 * m * q);
    for (int i = 0; i < m; i++) {
        const Vec3 c = W.panels[i].center;
        const std::array<Vec3, 4>& p = W.pio[i];
        const double sx = 0.5 * ((p[1].x - p[0].x) + (p[2].x - p[3].x));
        const double sy = 0.5 * ((p[1].y - p[0].y) + (p[2].y - p[3].y));
        const double sz = 0.5 * ((p[1].z - p[0].z) + (p[2].z - p[3].z));
        for (int k = 0; k < q; k++) {
            double* o = &seed[3 * ((size_t)i * q + k)];
            const double u = q == 1 ? 0.0 : (k + 0.5) / q - 0.5;
            o[0] = q == 1 ? c.x : c.x + u * sx;
            o[1] = q == 1 ? c.y : c.y + u * sy;
            o[2] = q == 1 ? c.z : c.z + u * sz;
        }
    }
    const bool bh = v->o.accelerate && v->ts > 3;
    std::vector<double> pos;
    if (bh) pos = rk4(v, seed, true);  // the tree does not hold this step's new particles: one batch
    const double dt = v->dt;
    for (int i = 0; i < m; i++) {
        double radius = v->core;
        if (!v->alg) {
            // wakePanel::getPartRadius (wakePanel.cpp:228-249)
            const Vec3 c = W.panels[i].center;
            const double cx = c.x + dt * v->Vb[0], cy = c.y + dt * v->Vb[1], cz = c.z + dt * v->Vb[2];
            double sum = 0;
            int cnt = 0;
            for (int k = 0; k < 2; k++) {
                const int o = W.rad_nbr[i][k];
                if (o < 0) continue;
                const Vec3 d = W.panels[o].center;
                const double dx = cx - (d.x + dt * v->Vb[0]), dy = cy - (d.y + dt * v->Vb[1]),
                             dz = cz - (d.z + dt * v->Vb[2]);
                sum += std::fabs(std::sqrt(dx * dx + dy * dy + dz * dz));
                cnt++;
            }
            radius = sum / (double)cnt;
        }
        for (int k = 0; k < q; k++) {
            const double* sd = &seed[3 * ((size_t)i * q + k)];
            double x[3];
            if (bh) {
                for (int a = 0; a < 3; a++) x[a] = pos[3 * ((size_t)i * q + k) + a];
            } else {  // direct sums see the particles added so far
                const std::vector<double> r = rk4(v, std::vector<double>{sd[0], sd[1], sd[2]}, false);
                for (int a = 0; a < 3; a++) x[a] = r[a];
            }
            v->ppos.insert(v->ppos.end(), {x[0], x[1], x[2]});
            if (q == 1)
                v->pstr.insert(v->pstr.end(), {str[i].x, str[i].y, str[i].z});
            else
                v->pstr.insert(v->pstr.end(), {str[i].x / q, str[i].y / q, str[i].z / q});
            v->prad.push_back(radius);
            v->pvel0.insert(v->pvel0.end(), {0.0, 0.0, 0.0});
            v->pdst0.insert(v->pdst0.end(), {0.0, 0.0, 0.0});
            v->pshed.push_back(v->ts);
            v->pparent.push_back(i);
            v->pseed.insert(v->pseed.end(), {sd[0], sd[1], sd[2]});
            v->ver++;
        }
    }
    if (!v->have_fil) {  // vortexFil from pointsInOrder 2 -> 3
        v->fil_ends.resize((size_t)6 * m);
        for (int i = 0; i < m; i++) {
            const Vec3 a = W.pio[i][2], b = W.pio[i][3];
            double* e = &v->fil_ends[6 * (size_t)i];
            e[0] = a.x;
            e[1] = a.y;
            e[2] = a.z;
            e[3] = b.x;
            e[4] = b.y;
            e[5] = b.z;
        }
        upload(v->g->ctx, v->fil_dev, v->fil_ends);
        v->have_fil = true;
    }
    v->fil_str.resize(m);
    for (int i = 0; i < m; i++) v->fil_str[i] = -v->mu2[i];
}

// cpCaseVP::particleStrengthUpdate (cpCaseVP.cpp:463-541)
void strength_update(cp_vp* v) {
    const int np = (int)v->prad.size();
    if (np == 0) return;
    const double c20 = 20 * v->o.cref, c12 = 12 * v->o.cref;
    const bool accel = v->o.accelerate != 0;
    const bool bh_normal = accel && v->ts > 3;
    // the particle sums each class of particle needs
    std::vector<double> bh_st, bh_df, d_st, d_df;
    bool need_bh = false, need_direct = false;
    for (int i = 0; i < np; i++) {
        const double ax = std::fabs(v->ppos[3 * (size_t)i]);
        if (ax > c20) continue;
        if (ax > c12) (accel ? need_bh : need_direct) = true;
        else (bh_normal ? need_bh : need_direct) = true;
    }
    const bool diffusion = !v->alg;  // the algebraic kernel carries no diffusion term
    if (need_bh) {
        bh_st = bh_on_particles(v, CP_BH_STRETCH);
        if (diffusion) bh_df = bh_on_particles(v, CP_BH_DIFF);
    }
    if (need_direct) {
        d_st = pairs_direct(v, CP_BH_STRETCH);
        if (diffusion) d_df = pairs_direct(v, CP_BH_DIFF);
    }
    const std::vector<double> pst = panel_stretch(v);
    std::vector<double> upd((size_t)3 * np, 0.0);
    for (int i = 0; i < np; i++) {
        const double ax = std::fabs(v->ppos[3 * (size_t)i]);
        if (ax > c20) continue;
        const bool far1 = ax > c12;
        const bool use_bh = far1 ? accel : bh_normal;
        const std::vector<double>& st = use_bh ? bh_st : d_st;
        const std::vector<double>& df = use_bh ? bh_df : d_df;
        for (int q = 0; q < 3; q++) {
            double diff = diffusion ? df[3 * i + q] : 0.0;
            if (!far1) diff += pst[3 * i + q];
            upd[3 * i + q] = diff + st[3 * i + q];
        }
    }
    const double dt = v->dt;
    for (int i = 0; i < np; i++) {
        double* d0 = &v->pdst0[3 * (size_t)i];
        const bool first = d0[0] == 0 && d0[1] == 0 && d0[2] == 0;
        for (int q = 0; q < 3; q++) {
            double& a = v->pstr[3 * (size_t)i + q];
            const double u = upd[3 * i + q];
            a = first ? a + u * dt : a + dt * (1.5 * u - 0.5 * d0[q]);
        }
        for (int q = 0; q < 3; q++) d0[q] = upd[3 * i + q];
    }
    v->ver++;
}

// sigma_i = -(V_i . n_i) with V_i = Vinf + particle and filament velocity at
// the centre (cpCaseVP::setSourceStrengthsVP, cpCaseVP.cpp:173-202); V_i is
// also the velocity of the panel's potential (VinfPlusVecPot)
void sources_from_particles(cp_vp* v) {
    const Geometry& G = v->g->host;
    const int n = v->n;
    std::vector<double> ctr((size_t)3 * n);
    for (int i = 0; i < n; i++) {
        ctr[3 * i] = G.body[i].center.x;
        ctr[3 * i + 1] = G.body[i].center.y;
        ctr[3 * i + 2] = G.body[i].center.z;
    }
    const std::vector<double> part = v->o.accelerate ? vel_bh_points(v, ctr) : vel_direct(v, ctr);
    const std::vector<double> fil = vel_filaments(v, ctr);
    v->vpot.resize((size_t)3 * n);
    for (int i = 0; i < n; i++) {
        double V[3];
        for (int q = 0; q < 3; q++) {
            V[q] = v->Vb[q] + part[3 * i + q] + fil[3 * i + q];
            v->vpot[3 * i + q] = V[q];
        }
        const Vec3 nn = G.body[i].normal;
        v->sigma[i] = -(V[0] * nn.x + V[1] * nn.y + V[2] * nn.z) + 0;
    }
}

// wake::wakeStrength (wake.cpp:478-506) over the body doublets
double wake_strength(const cp_vp* v, const WakeInfo& W, double y) {
    const std::vector<WakeLineInfo>& L = W.lines;
    size_t k1 = 0;
    if (y < L[1].y) k1 = 0;
    else if (y >= L.back().y) k1 = L.size() - 2;
    else
        for (size_t i = 1; i + 1 < L.size(); i++)
            if (L[i].y <= y && L[i + 1].y > y) k1 = i;
    const WakeLineInfo &a = L[k1], &b = L[k1 + 1];
    const double interp = (y - a.y) / (b.y - a.y);
    const double sa = v->mu[a.upper] - v->mu[a.lower], sb = v->mu[b.upper] - v->mu[b.lower];
    return sa + interp * (sb - sa);
}

// wake::trefftzPlaneVP (wake.cpp:191-325) over every wake, CL / PG and
// CDi / PG^2 (cpCaseVP::trefftzPlaneAnalysisVP, cpCaseVP.cpp:418-429)
void trefftz_vp(cp_vp* v) {
    const Geometry& G = v->g->host;
    int steps = v->ts;
    if (steps % 2 != 0) steps++;
    const int row = steps / 2 + 2;
    double CLs = 0, CDs = 0;
    for (const WakeInfo& wk : G.wakes) {
        std::vector<int> un;
        for (int i = 0; i < (int)v->pshed.size(); i++)
            if (v->pshed[i] == row) un.push_back(i);
        // the shedding position of particle k (the reference's shedding
        // panel centre)
        auto seed = [&](int k) {
            const double* p = &v->pseed[3 * (size_t)k];
            return V3{p[0], p[1], p[2]};
        };
        std::vector<int> sp;  // by the shedding position's y, first minimum
        while (!un.empty()) {
            size_t lo = 0;
            double ly = seed(un[0]).y;
            for (size_t i = 0; i < un.size(); i++) {
                const double y = seed(un[i]).y;
                if (y < ly) {
                    ly = y;
                    lo = i;
                }
            }
            sp.push_back(un[lo]);
            un.erase(un.begin() + (std::ptrdiff_t)lo);
        }
        double CL = 0, CD = 0;
        if (sp.size() >= 2) {
            auto P = [&](int k) {
                const double* p = &v->ppos[3 * (size_t)sp[k]];
                return V3{p[0], p[1], p[2]};
            };
            auto dist = [](V3 a, V3 b) {
                const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                return std::sqrt(dx * dx + dy * dy + dz * dz);
            };
            double Slen = 0;
            for (size_t i = 0; i + 1 < sp.size(); i++) Slen += dist(P((int)i), P((int)i + 1));
            const int nP = (int)sp.size() * 3;
            const double step = Slen / nP;
            std::vector<V3> S;
            std::vector<int> S1, S2;
            S.push_back(P(0));
            S1.push_back(0);
            S2.push_back(1);
            V3 next = P(1), pt = P(0);
            int ni = 1;
            for (int i = 1; i < nP; i++) {
                const double d = dist(pt, next);
                if (step < d) {
                    const V3 dir{(next.x - pt.x) / d, (next.y - pt.y) / d, (next.z - pt.z) / d};
                    pt = V3{pt.x + step * dir.x, pt.y + step * dir.y, pt.z + step * dir.z};
                } else {
                    pt = next;
                    if (ni + 1 >= (int)sp.size()) break;  // the curve's end (rounding)
                    next = P(ni + 1);
                    ni++;
                    const double dn = dist(pt, next);
                    const V3 dir{(next.x - pt.x) / dn, (next.y - pt.y) / dn, (next.z - pt.z) / dn};
                    pt = V3{pt.x + (step - d) * dir.x, pt.y + (step - d) * dir.y, pt.z + (step - d) * dir.z};
                }
                S.push_back(pt);
                S1.push_back(ni - 1);
                S2.push_back(ni);
            }
            std::vector<double> pts((size_t)3 * S.size());
            for (size_t i = 0; i < S.size(); i++) {
                pts[3 * i] = S[i].x;
                pts[3 * i + 1] = S[i].y;
                pts[3 * i + 2] = S[i].z;
            }
            const std::vector<double> pv = vel_direct(v, pts);
            std::vector<double> Cl(nP + 1, 0.0), Cd(nP + 1, 0.0);
            for (size_t i = 1; i < S.size(); i++) {
                const int a = sp[S1[i]], b = sp[S2[i]];
                const V3 pa = P(S1[i]), pb = P(S2[i]);
                const V3 ca = seed(a), cb = seed(b);
                const double ya = ca.y, yb = cb.y;
                // wake::particlePntInWakeY
                const double dd = dist(pa, pb), da = dist(pa, S[i]), db = dist(S[i], pb);
                double y;
                if (da < 1e-10) y = ya;
                else if (db < 1e-10) y = yb;
                else y = ya + da / dd * (yb - ya);
                // wake::stretchFactor
                const double dshed = std::sqrt((cb.x - ca.x) * (cb.x - ca.x) + (cb.y - ca.y) * (cb.y - ca.y) +
                                               (cb.z - ca.z) * (cb.z - ca.z));
                const double st = dshed / dd;
                const double w = std::fabs(pv[3 * i + 2]);
                const double dphi = -wake_strength(v, wk, y) * st;
                Cl[i] = 2 * dphi / (v->Vmag * v->o.Sref);
                Cd[i] = dphi * w / (v->Vmag * v->Vmag * v->o.Sref);
            }
            for (int i = 0; i < (int)Cl.size() - 2; i += 2) {
                CL += 1.0 / 3 * step * (Cl[i] + 4 * Cl[i + 1] + Cl[i + 2]);
                CD += 1.0 / 3 * step * (Cd[i] + 4 * Cd[i + 1] + Cd[i + 2]);
            }
        }
        CLs += CL / v->PG;
        CDs += CD / (v->PG * v->PG);
    }
    v->CL = CLs;
    v->CD = CDs;
}

}  // namespace

extern "C" {

cp_status cp_geometry_upload_vp(cp_ctx* c, const cp_mesh* m, double V, double dt, uint64_t key, cp_geom** out) {
    if (!c || !m || !out || !m->xyz || !m->tri || !m->ids || m->n_nodes <= 0 || m->n_tris <= 0 || !(V > 0))
        return CP_BAD_ARG;
    VpParams p;
    p.V = V;
    p.dt = dt;
    return guarded_impl(c, nullptr, [&] { return build_geom(c, m, key, out, &p); });
}

cp_status cp_vp_buffer_rows(const cp_geom* g, int32_t* m, double* dt, double* row1, double* row2) {
    if (!g || !g->host.vp) return CP_BAD_ARG;
    const Geometry& G = g->host;
    const int mm = (int)G.w2.panels.size();
    if (m) *m = mm;
    if (dt) *dt = G.w2.dt;
    for (int i = 0; i < mm; i++)
        for (int k = 0; k < 4; k++) {
            if (row1) {
                const Vec3 a = G.wake[i].v[k];
                double* o = row1 + 12 * (size_t)i + 3 * k;
                o[0] = a.x;
                o[1] = a.y;
                o[2] = a.z;
            }
            if (row2) {
                const Vec3 a = G.w2.panels[i].v[k];
                double* o = row2 + 12 * (size_t)i + 3 * k;
                o[0] = a.x;
                o[1] = a.y;
                o[2] = a.z;
            }
        }
    return CP_OK;
}

cp_status cp_vp_create(cp_geom* g, const cp_vp_opts* o, cp_vp** out) {
    if (!g || !o || !out || !g->host.vp || !(o->fs.V > 0)) return CP_BAD_ARG;
    if (o->prec != CP_FP64 && o->prec != CP_FP32_REFINE && o->prec != CP_FP32_REFINE64) return CP_BAD_ARG;
    if (o->kernel != CP_VP_REFERENCE && o->kernel != CP_VP_ALGEBRAIC) return CP_BAD_ARG;
    *out = nullptr;
    cp_ctx* c = g->ctx;
    if (c->device < 0) return fail(c, CP_NO_DEVICE, "host-only context");
    std::unique_ptr<cp_vp> v(new cp_vp);
    v->g = g;
    v->o = *o;
    v->n = g->dev.n;
    v->m = (int)g->host.w2.panels.size();
    v->dt = g->host.w2.dt;
    v->Vmag = o->fs.V;
    v->PG = std::sqrt(1 - o->mach * o->mach);
    v->alg = o->kernel == CP_VP_ALGEBRAIC;
    v->f32 = o->fp32 != 0;
    v->q = std::max(1, (int)o->per_edge);
    if (v->alg) {
        v->core = o->core;
        if (!(v->core > 0)) {  // 0.1 x the mean panel size
            double sum = 0;
            for (const PanelGeom& P : g->host.body) sum += std::sqrt(P.area);
            v->core = 0.1 * sum / (double)std::max<size_t>(1, g->host.body.size());
        }
    }
    double T[9];
    wind_to_body(o->fs, T, v->Vb);
    v->mu.assign(v->n, 0.0);
    v->sigma.assign(v->n, 0.0);
    v->mu1.assign(v->m, 0.0);
    v->mu2.assign(v->m, 0.0);
    v->mu2prev.assign(v->m, 0.0);
    // the second row's doublet influence C on the collocation points
    std::vector<int32_t> nv(v->m, 4);
    std::vector<double> verts((size_t)12 * v->m);
    cp_vp_buffer_rows(g, nullptr, nullptr, nullptr, verts.data());
    cp_status st = cp_set_buffer_wake(g, v->m, nv.data(), verts.data());
    if (st != CP_OK) return st;
    if (o->prec == CP_FP64 || o->prec == CP_FP32_REFINE64) {
        if ((st = cp_assemble(g, CP_FP64)) != CP_OK) return st;
    }
    if (o->prec != CP_FP64) {
        if ((st = cp_assemble(g, CP_FP32_REFINE)) != CP_OK) return st;
    }
    st = guarded_g(g, [&]() -> cp_status {
        std::vector<PanelRec> rec(v->m);
        for (int i = 0; i < v->m; i++) rec[i] = make_panel_rec(g->host.w2.panels[i]);
        upload(c, v->w2rec, rec);
        v->row2 = DevGeom();
        v->row2.n = 0;
        v->row2.nw = v->m;
        v->row2.wake = v->w2rec.as<PanelRec>();
        return CP_OK;
    });
    if (st != CP_OK) return st;
    geom_ref(g);  // the case keeps its geometry (and through it the context) alive
    *out = v.release();
    return CP_OK;
}

cp_status cp_vp_step(cp_vp* v, cp_vp_step_info* info) {
    if (!v) return CP_BAD_ARG;
    cp_geom* g = v->g;
    const double t0 = now_s();
    cp_status st = CP_OK;
    v->ts++;
    if (v->ts == 1) {
        // setSourceStrengths + solveMatrixEq (cpCaseVP.cpp:38-40, 233-263)
        st = guarded_g(g, [&]() -> cp_status {
            for (int i = 0; i < v->n; i++) {
                const Vec3 nn = g->host.body[i].normal;
                v->sigma[i] = -(v->Vb[0] * nn.x + v->Vb[1] * nn.y + v->Vb[2] * nn.z) + 0;
            }
            v->vpot.assign((size_t)3 * v->n, 0.0);
            for (int i = 0; i < v->n; i++)
                for (int q = 0; q < 3; q++) v->vpot[3 * i + q] = v->Vb[q];
            return CP_OK;
        });
        if (st == CP_OK) st = solve(v, false);
    } else if (v->ts == 2) {
        // convectBufferWake + solveVPmatrixEq (cpCaseVP.cpp:47-51)
        convect_buffer(v);
        st = solve(v, true);
    } else {
        st = guarded_g(g, [&]() -> cp_status {
            convect_particles(v);
            collapse_buffer(v);
            convect_buffer(v);
            strength_update(v);
            if (v->o.accelerate) build_tree(v);
            sources_from_particles(v);
            return CP_OK;
        });
        if (st == CP_OK) st = solve(v, true);
        if (st == CP_OK && v->ts >= 5) {
            // cpCaseVP::solutionConvergence (cpCaseVP.cpp:204-230)
            const double pCL = v->CL, pCD = v->CD;
            st = guarded_g(g, [&]() -> cp_status {
                trefftz_vp(v);
                return CP_OK;
            });
            const double dCD = std::fabs((v->CD - pCD) / pCD), dCL = std::fabs((v->CL - pCL) / pCL);
            v->converged = dCL < 1e-4 && dCD < 1e-4;
        }
    }
    if (info) {
        info->step = v->ts;
        info->n_particles = (int32_t)v->prad.size();
        info->converged = v->converged ? 1 : 0;
        info->solve_iters = v->iters;
        info->CL_trefftz = v->CL;
        info->CDi_trefftz = v->CD;
        info->rel_residual = v->res;
        info->ms = 1e3 * (now_s() - t0);
    }
    return st;
}

cp_status cp_vp_finish(cp_vp* v, cp_solve_info* info, double* cp) {
    if (!v || !info) return CP_BAD_ARG;
    cp_geom* g = v->g;
    cp_ctx* c = g->ctx;
    return guarded_g(g, [&]() -> cp_status {
        ensure_post(g);
        // cpCaseVP::compVelocity (steady) with the potentials of the last
        // solve, then trefftzPlaneAnalysisVP
        const int n = v->n;
        const int64_t lda = g->dev.lda;
        hipStream_t s = g->stream;
        std::vector<double> v4{v->Vb[0], v->Vb[1], v->Vb[2], v->Vmag};
        DevBuf dv4, dmu, dvp, dpart, dres, dcp;
        upload(c, dv4, v4);
        std::vector<double> mu_l((size_t)lda, 0.0);
        std::copy(v->mu.begin(), v->mu.end(), mu_l.begin());
        upload(c, dmu, mu_l);
        upload(c, dvp, v->vpot);
        const int nblk = post_partial_blocks(n);
        dalloc(c, dpart, (size_t)nblk * 6 * sizeof(double));
        dalloc(c, dres, 16 * sizeof(double));
        if (cp) dalloc(c, dcp, (size_t)n * sizeof(double));
        PostParams pp;
        pp.PG = v->PG;
        pp.Sref = v->o.Sref;
        pp.bref = v->o.bref;
        pp.cref = v->o.cref;
        for (int a = 0; a < 3; a++) pp.cg[a] = v->o.cg[a];
        PostFields F;
        F.vpot = dvp.as<double>();
        launch_post(g->post, pp, 1, dmu.as<double>(), lda, dv4.as<double>(), cp ? dcp.as<double>() : nullptr, n,
                    dpart.as<double>(), nblk, dres.as<double>(), s, &F);
        std::vector<double> r(16);
        CP_HIP(hipMemcpyAsync(r.data(), dres.p, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
        if (cp) CP_HIP(hipMemcpyAsync(cp, dcp.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        double T[9], Vb[3];
        wind_to_body(v->o.fs, T, Vb);
        *info = cp_solve_info{};
        fill_info(r.data(), T, *info);
        trefftz_vp(v);
        info->CL = v->CL;
        info->CDi = v->CD;
        info->iters = v->iters;
        info->rel_residual = v->res;
        return CP_OK;
    });
}

cp_status cp_vp_particles(const cp_vp* v, int32_t* n, double* pos, double* strength, double* radius, int32_t* shed) {
    if (!v) return CP_BAD_ARG;
    const size_t np = v->prad.size();
    if (n) *n = (int32_t)np;
    if (pos) std::copy(v->ppos.begin(), v->ppos.end(), pos);
    if (strength) std::copy(v->pstr.begin(), v->pstr.end(), strength);
    if (radius) std::copy(v->prad.begin(), v->prad.end(), radius);
    if (shed) std::copy(v->pshed.begin(), v->pshed.end(), shed);
    return CP_OK;
}

cp_status cp_vp_state(const cp_vp* v, double* mu, double* sigma, double* mu_row1, double* mu_row2) {
    if (!v) return CP_BAD_ARG;
    if (mu) std::copy(v->mu.begin(), v->mu.end(), mu);
    if (sigma) std::copy(v->sigma.begin(), v->sigma.end(), sigma);
    if (mu_row1) std::copy(v->mu1.begin(), v->mu1.end(), mu_row1);
    if (mu_row2) std::copy(v->mu2.begin(), v->mu2.end(), mu_row2);
    return CP_OK;
}

void cp_vp_release(cp_vp* v) {
    if (!v) return;
    cp_geom* g = v->g;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    delete v;
    geom_unref(g);
}

}  // extern "C"
