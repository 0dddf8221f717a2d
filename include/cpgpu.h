/*
 * cpgpu — C ABI of the MI355X (gfx950) CPanel solve path.
 *
 * The reference (jakeddavis/CPanel) has no plugin API; the seams it exposes
 * are C++ class methods.  Each entry point below replaces one of them:
 *
 *   cp_geometry_upload / cp_geometry_from_tri
 *       geometry::geometry -> readTri topology     (CPanel/geometry.cpp:132-484)
 *   cp_assemble
 *       geometry::setInfCoeff, getA/getB           (CPanel/geometry.cpp:809-891,
 *                                                   CPanel/geometry.h:310-312)
 *   cp_solve
 *       cpCase::run = setSourceStrengths + solveMatrixEq + compVelocity +
 *       trefftzPlaneAnalysis                       (CPanel/cpCase.cpp:31-139,
 *                                                   174-214, 342-365, 541-552)
 *   cp_bh_* / cp_panel_to_points
 *       particleFMM::build / barnesHutVel / barnesHutStretch,
 *       bodyPanel::panelV / partStretching         (CPanel/particleFMM.h:52-56,
 *                                                   CPanel/bodyPanel.cpp:122-130,697-729)
 *
 * Conventions: every call returns cp_status; nothing exits or prompts; host
 * buffers are caller owned, device buffers context owned; matrices are
 * column-major.  One context per GPU.  Each geometry has its own stream and
 * lock: calls on one geometry serialise, calls on different geometries may
 * run concurrently from different host threads.  Barnes-Hut trees use the
 * context's stream and lock.
 */
#ifndef CPGPU_H
#define CPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CP_OK = 0,
    CP_NOT_CONVERGED = 1,   /* solve ran but ||r||/||b|| > the requested tol (1e-10 by
                               default, the reference's flag, cpCase.cpp:199-202) */
    CP_OUT_OF_BUDGET = 2,   /* request exceeds the context's HBM budget */
    CP_BAD_MESH = 3,        /* topology the reference cannot handle */
    CP_HIP_ERROR = 4,       /* a HIP runtime call failed; see cp_last_hip_error */
    CP_BAD_ARG = 5,
    CP_NO_DEVICE = 6,
    CP_NOT_ASSEMBLED = 7
} cp_status;

typedef enum {
    CP_FP64 = 0,            /* plain FP64 assembly and FP64 BiCGSTAB */
    CP_FP32_REFINE = 1,     /* FP32 A,B (FP64 near field), FP32 Krylov, FP64 refinement */
    CP_FP32_REFINE64 = 2    /* FP32 Krylov refined against the resident FP64 A,B: the
                               refined solution is the FP64 system's (both sets assembled) */
} cp_prec;

typedef struct cp_ctx cp_ctx;
typedef struct cp_geom cp_geom;

/* ---- context ------------------------------------------------------------ */
/* device = -1 creates a host-only context: geometry/topology queries work,
 * device work returns CP_NO_DEVICE.  hbm_budget_bytes = 0: no budget. */
cp_status cp_create(int device, uint64_t hbm_budget_bytes, cp_ctx** out);
void cp_destroy(cp_ctx* ctx);
const char* cp_last_error(const cp_ctx* ctx);
int cp_last_hip_error(const cp_ctx* ctx);
const char* cp_status_string(cp_status s);
/* bytes of device memory currently held by the context */
uint64_t cp_bytes_in_use(const cp_ctx* ctx);

/* ---- geometry (replaces geometry::readTri, geometry.cpp:132-484) --------- */
typedef struct {
    int32_t n_nodes;
    const double* xyz;      /* 3*n_nodes */
    int32_t n_tris;
    const int32_t* tri;     /* 3*n_tris, 0-based node indices */
    const int32_t* ids;     /* surface ID per triangle; >1000 = wake */
} cp_mesh;

typedef struct {
    int32_t n_body;         /* body panels (rows/cols of A, B) */
    int32_t n_wake;         /* prescribed wake panels */
    int32_t n_wakelines;
    int32_t n_nodes;
    int64_t lda;            /* device leading dimension of A, B */
    int32_t n_wakes;        /* wakes (one Trefftz plane and spanwise file each) */
    int32_t n_stations;     /* Trefftz stations per wake (165) */
} cp_geom_info;

cp_status cp_geometry_upload(cp_ctx* ctx, const cp_mesh* mesh, uint64_t key, cp_geom** out);
cp_status cp_geometry_from_tri(cp_ctx* ctx, const char* path, uint64_t key, cp_geom** out);
void cp_geometry_release(cp_geom* g);
cp_status cp_geometry_info(const cp_geom* g, cp_geom_info* info);
/* body panels in solve order: file triangle index, centre[3], normal[3], area */
cp_status cp_geometry_body(const cp_geom* g, int32_t* tri_index, double* center, double* normal, double* area);
/* Kutta folding table (wakePanel::interpPanels, wakePanel.cpp:48-81) */
cp_status cp_geometry_wake(const cp_geom* g, int32_t* tri_index, int32_t* idx4, double* coef);
/* CHTLS clusters (bodyPanel::setCluster, bodyPanel.cpp:402-487), 16 per panel */
cp_status cp_geometry_clusters(const cp_geom* g, int32_t* cluster16, int32_t* ts_order, int32_t* tip);

/* ---- assembly (geometry::setInfCoeff, geometry.cpp:809-891) -------------- */
cp_status cp_assemble(cp_geom* g, cp_prec prec);
/* copy A and B (as FP64) to host arrays with leading dimension ld */
cp_status cp_read_matrices(cp_geom* g, cp_prec prec, double* A, double* B, int64_t ld);
/* Sets the FP64 matrices A and B (n x n, column-major, leading dimension
   ld >= n) of a geometry to caller-given ones, in place of
   cp_assemble(CP_FP64): replaces geometry::readInfCoeff (geometry.cpp:1125-1166),
   which loads A and B from a .infCoeff file instead of computing them
   (cpanel_amd/infcoeff.py reads and writes that format). The FP32 matrices of
   the mixed paths are not touched (cp_assemble(CP_FP32_REFINE) builds them). */
cp_status cp_load_matrices(cp_geom* g, const double* A, const double* B, int64_t ld);

/* ---- steady solve (cpCase::run, cpCase.cpp:31-139) ----------------------- */
typedef struct {
    double V, alpha_deg, beta_deg;     /* cpCase::windToBody (cpCase.cpp:142-166) */
} cp_freestream;

typedef struct {
    double mach;            /* Prandtl-Glauert factor sqrt(1-M^2) (cpCase.h:116) */
    double Sref, bref, cref;
    double cg[3];
    double tol;             /* target ||r||/||b|| (<=0: 1e-10) */
    int32_t max_iter;       /* Krylov iterations per solve (<=0: 2N) */
    int32_t max_refine;     /* FP64 refinement sweeps, CP_FP32_REFINE (<=0: 20) */
    double inner_tol;       /* FP32 inner solve reduction per sweep (<=0: 2e-4, relaxed on the
                               last sweep to what tol still needs) */
    int32_t precond;        /* CP_FP32_REFINE inner solve: 0 = strip block Jacobi
                               (default), 1 = Jacobi (the reference's); the FP64
                               path always uses Jacobi */
    int32_t pad;
} cp_solve_opts;

typedef struct {
    double CL, CDi;         /* Trefftz plane (cpCase.cpp:541-552) */
    double Fbody[3], Fwind[3], CM[3];   /* surface integration (cpCase.cpp:342-365) */
    double rel_residual;    /* final ||b - A mu|| / ||b|| in FP64 */
    int32_t iters;          /* Krylov iterations */
    int32_t refinements;    /* FP64 refinement sweeps */
    int32_t converged;      /* rel_residual <= the requested tol (default 1e-10) */
    int32_t pad;
} cp_solve_info;

/* Solves k freestream conditions on one assembled geometry.  mu, cp: host
 * arrays n_body*k (column c = condition c) or NULL; info: k entries. */
cp_status cp_solve(cp_geom* g, cp_prec prec, int32_t k, const cp_freestream* fs, const cp_solve_opts* opts,
                   double* mu, double* cp, cp_solve_info* info);

/* Several geometries' mixed-precision steady solves in one pass, replacing
 * the reference's sequential loop of cases (CPanelMgr.cpp:49-88, one
 * cpCase::run at a time; a caller of the reference runs one geometry per
 * process).  geoms: G distinct geometries of one context (G <= 16), each
 * assembled for prec (CP_FP32_REFINE or CP_FP32_REFINE64); k[i] <= 64
 * conditions fs[i] and options opts[i] (an array of G) for geometry i;
 * mu[i], cp[i] (host, n_body_i * k[i], or NULL; the arrays themselves may be
 * NULL) and info[i] (k[i] entries) as in cp_solve.  Every Krylov iteration stage runs as one launch over all
 * geometries still iterating, geometries of uneven panel count included;
 * each geometry's results are bit-identical to its own cp_solve with the
 * same options.  Returns CP_NOT_CONVERGED if any condition missed tol. */
cp_status cp_solve_group(int32_t G, cp_geom* const* geoms, cp_prec prec, const int32_t* k,
                         const cp_freestream* const* fs, const cp_solve_opts* opts, double* const* mu,
                         double* const* cp, cp_solve_info* const* info);

/* Y = A X through exactly one device kernel (tests, kernel timing); the
 * product A*x that cpCase::solveMatrixEq hands to Eigen (cpCase.cpp:191-198).
 *   kernel 0: FP64 GEMV on the FP64 A (k launches)
 *   kernel 1: FP32 MFMA GEMM on the FP32 A, X rounded to FP32 (k <= 64)
 *   kernel 2: FP64-accumulated MFMA GEMM on the FP32 A (k <= 64)
 *   kernel 3: kernel 2 plus the FP64 near-field correction (entries with
 *             |c_i - c_j| <= longSide_j carry their FP64 value)
 *   kernel 4: FP64 MFMA GEMM on the FP64 A (k <= 64; needs both sets)
 * X, Y: host, n_body x k column-major. */
cp_status cp_matvec(cp_geom* g, int kernel, int32_t k, const double* X, double* Y);

/* y = A x (FP64 A, the product of cp_matvec kernel 0) through caller-given
 * DEVICE buffers, allocating nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  X: k columns of
 * n_body entries (column stride ldx >= n_body); Y: k columns of lda entries
 * (cp_geometry_info's lda, stride ldy >= lda; rows n_body..lda-1 are set to
 * zero).  The split-K partial sums live in a static 20 MiB device array per
 * GPU (CP_BAD_ARG if a geometry's split would exceed it, which needs a
 * leading dimension above 524,288); calls that overlap on one GPU race on
 * that array.  Results equal cp_matvec kernel 0 bit for bit.  Synchronous. */
cp_status cp_matvec_f64_noalloc(cp_geom* g, int32_t k, const double* X, int64_t ldx, double* Y, int64_t ldy);

/* ---- vortex particles: Barnes-Hut (particleFMM::build / barnesHutVel /
 *      barnesHutStretch / barnesHutDiff, CPanel/particleFMM.h:52-56) ------- */
typedef struct cp_tree cp_tree;

typedef enum {
    CP_BH_VEL_POINTS = 0,   /* velocity at points, algebraic kernel
                               (particleFMM::barnesHutVel(Vector3d), node.h:236-262) */
    CP_BH_VEL = 1,          /* velocity on the particles, Gaussian kernel
                               (barnesHutVel(particle*), node.h:264-290) */
    CP_BH_STRETCH = 2,      /* vortex stretching (barnesHutStretch, node.h:292-318) */
    CP_BH_DIFF = 3          /* viscous diffusion (barnesHutDiff, node.h:333-359) */
} cp_bh_mode;

/* particle kernel of a tree or a vortex-particle case */
typedef enum {
    CP_VP_REFERENCE = 0,    /* the reference's kernels: Gaussian particle-particle
                               velocity, stretching and diffusion with per-particle
                               radii, algebraic velocity at points (particle.cpp:31-140) */
    CP_VP_ALGEBRAIC = 1     /* one fixed core for everything: velInflAlgSmooth's
                               algebraic kernel (particle.cpp:31-38) with core radius
                               `core` for velocity at points and on particles, and its
                               (alpha . grad) u for stretching; no diffusion term */
} cp_vp_kernel;

typedef struct {
    int32_t max_members;    /* leaf capacity (<=0: 10, cpCaseVP.cpp:79) */
    int32_t kernel;         /* cp_vp_kernel (0: the reference's) */
    double theta;           /* opening angle s/d (<0: 0.5, octree.h:37) */
    double core;            /* CP_VP_ALGEBRAIC: core radius (> 0) */
    int32_t fp32;           /* CP_VP_ALGEBRAIC: 1 = pair arithmetic in FP32 (pair
                               offsets formed in FP64), 0 = FP64 */
    int32_t host_build;     /* 1: build the tree on the host (the same tree; for
                               checking the device build); 0: on the device */
} cp_bh_opts;

/* Builds the particle octree and its pseudo-particles (octree::addData,
 * octree.h:151-186; particleFMM::computeMultExp, particleFMM.cpp:29-58) on the
 * device: octant keys, a radix sort, the nodes level by level.  xyz, alpha: 3
 * per particle (interleaved), radius: 1 per particle.  Particles closer than
 * 2^-21 of the root width in a cell that must still split exceed the keys'
 * depth: that tree is built on the host (the same tree).  On a host-only
 * context the tree is built on the host for cp_bh_info / cp_bh_read only. */
cp_status cp_bh_build(cp_ctx* ctx, int32_t n, const double* xyz, const double* alpha, const double* radius,
                      const cp_bh_opts* opts, cp_tree** out);
/* node count, deepest level (root = 1), root origin[3] and half width */
cp_status cp_bh_info(const cp_tree* t, int32_t* n_nodes, int32_t* levels, double* origin, double* half);
/* Barnes-Hut sums from the root's children (root excluded,
 * particleFMM.cpp:65-116).  CP_BH_VEL_POINTS: m points pts[3m] -> out[3m];
 * the other modes: the tree's own n particles -> out[3n] (pts ignored).
 * A CP_VP_ALGEBRAIC tree has no CP_BH_DIFF (CP_BAD_ARG). */
cp_status cp_bh_eval(cp_tree* t, cp_bh_mode mode, int32_t m, const double* pts, double* out);
/* Diagnostics (no counterpart in the reference): the flattened tree in
 * pre-order, nodes: 8 per node (pseudo-particle x, y, z, cell width, strength
 * x, y, z, radius), links: 3 per node (next node after the subtree, first
 * member slot, member count or -1 for an internal node), order: member slot
 * -> particle index (n); any output may be NULL. */
cp_status cp_bh_read(const cp_tree* t, double* nodes, int32_t* links, int32_t* order);
void cp_bh_release(cp_tree* t);

/* ---- panel -> point velocity (bodyPanel::panelV, bodyPanel.cpp:122-130;
 *      wakePanel::panelV, wakePanel.cpp:89-92) ------------------------------
 * vel[3m] = sum over body panels of sigma_j vSrc_j + mu_j vDub_j plus wake
 * panels mu_w vDub_w at the points pts[3m]; sigma, mu: n_body (solve order),
 * mu_wake: n_wake or NULL (zero).  fp32 = 1: the far-field pairs (the
 * point source and doublet of bodyPanel.cpp:305-348 beyond 5 longSide) in
 * FP32 arithmetic, the near field in FP64; 0: everything FP64. */
cp_status cp_panel_to_points(cp_geom* g, const double* sigma, const double* mu, const double* mu_wake, int32_t m,
                             const double* pts, double* vel, int32_t fp32);

/* ---- panel -> particle stretching (bodyPanel::partStretching,
 *      bodyPanel.cpp:697-729; wakePanel::partStretching, wakePanel.cpp:266-289)
 * out[3m] = sum over body panels (source sigma, doublet mu) and wake panels
 * (doublet mu_wake, NULL = zero) of the velocity-gradient matrix at each
 * particle applied to its strength: pos, alpha: 3 per particle.  fp32 as for
 * cp_panel_to_points (far-field point source and doublet gradients). */
cp_status cp_panel_stretching(cp_geom* g, const double* sigma, const double* mu, const double* mu_wake, int32_t m,
                              const double* pos, const double* alpha, double* out, int32_t fp32);

/* ---- vortex-particle solve (cpCaseVP::solveVPmatrixEq, cpCaseVP.cpp:265-301)
 * The VP case's second buffer-wake row: m panels (nv[j] = 3 or 4 vertices,
 * verts[12 j + 3 q + d]) whose doublet influence on the body collocation
 * points forms C (n_body x m), as geometry::setInfCoeff builds it
 * (geometry.cpp:876-882).  m = 0 clears it. */
cp_status cp_set_buffer_wake(cp_geom* g, int32_t m, const int32_t* nv, const double* verts);
/* Solves A mu = -B sigma - C mu_w2 for k columns with caller-given source
 * strengths (sigma: n_body x k, column-major; the VP case's sigma includes the
 * particle-induced velocity, cpCaseVP.cpp:173-200), optional buffer-wake
 * strengths mu_w2 (m x k) and optional initial guess x0 (n_body x k; NULL =
 * zero, the reference's choice; a previous step's mu = warm start).  mu:
 * n_body x k.  info: iterations, refinements, residual (no post-processing). */
cp_status cp_solve_rhs(cp_geom* g, cp_prec prec, int32_t k, const double* sigma, const double* mu_w2,
                       const double* x0, const cp_solve_opts* opts, double* mu, cp_solve_info* info);

/* ---- vortex-particle case (cpCaseVP::run, cpCaseVP.cpp:26-158) -----------
 * A geometry in vortex-particle mode has no prescribed wake: every
 * trailing-edge edge gets a first buffer-row quad (the wake that folds into A)
 * and a second-row quad behind it (geometry::readTri with the VP flag,
 * geometry.cpp:267-397; createVPWakeSurfaces :1307-1342).  The rows start at
 * the trailing edge and extend along the trailing-edge bisector by 1.3 V dt
 * and V dt; dt <= 0 takes the mean trailing-edge edge length / V
 * (geometry::calcTimeStep, :1276-1304). */
cp_status cp_geometry_upload_vp(cp_ctx* ctx, const cp_mesh* mesh, double V, double dt, uint64_t key, cp_geom** out);
/* m quads per row, the time step, and the quads' vertices (m x 4 x 3 per
 * row, in node order); any output may be NULL. */
cp_status cp_vp_buffer_rows(const cp_geom* g, int32_t* m, double* dt, double* row1, double* row2);

typedef struct cp_vp cp_vp;

typedef struct {
    cp_freestream fs;       /* V, alpha, beta (steady freestream) */
    double mach, Sref, bref, cref, cg[3];
    int32_t prec;           /* cp_prec of the matrix solves */
    int32_t accelerate;     /* 1: Barnes-Hut sums ("Accelerate Code"), 0: direct */
    int32_t warm_start;     /* 1: each solve starts from the previous mu (the
                               reference starts from zero) */
    int32_t max_members;    /* tree leaf capacity (<= 0: 10, cpCaseVP.cpp:79) */
    double theta;           /* opening angle (< 0: 0.5) */
    double tol;             /* solve tolerance (<= 0: 1e-12) */
    /* the configured form of the case (0 everywhere: the reference's) */
    int32_t kernel;         /* cp_vp_kernel of the particle sums */
    int32_t per_edge;       /* particles shed per second-row panel and step
                               (<= 1: one, cpCaseVP.cpp:319-341); q > 1 splits the
                               panel's edge strength into q equal particles seeded
                               at (k + 0.5) / q along its span */
    double core;            /* CP_VP_ALGEBRAIC core radius (<= 0: 0.1 x the mean
                               sqrt(area) of the body panels) */
    int32_t fp32;           /* 1: far-field panel -> particle sums (velocity,
                               stretching) and, with CP_VP_ALGEBRAIC, the particle
                               sums in FP32 arithmetic; solves, near field and
                               forces stay FP64 */
    int32_t pad;
} cp_vp_opts;

typedef struct {
    int32_t step;           /* time step just completed (1-based) */
    int32_t n_particles;
    int32_t converged;      /* cpCaseVP::solutionConvergence: relative change
                               of the Trefftz CL and CDi < 1e-4 (steps >= 5) */
    int32_t solve_iters;
    double CL_trefftz, CDi_trefftz;  /* latest wake::trefftzPlaneVP values */
    double rel_residual;
    double ms;              /* wall time of the step */
} cp_vp_step_info;

/* Starts a case on a vortex-particle geometry (assembles what opts->prec
 * needs and the second row's influence C). */
cp_status cp_vp_create(cp_geom* g, const cp_vp_opts* opts, cp_vp** out);
/* One time step of cpCaseVP::run: step 1 solves the first-row system, step 2
 * convects the buffer row and re-solves, later steps convect the particles
 * (Adams-Bashforth), shed one particle per second-row panel (RK4 seed),
 * update the strengths (stretching, diffusion, panel stretching), rebuild
 * the tree, set the sources from the particle and filament velocities and
 * re-solve; from step 5 the Trefftz convergence test runs. */
cp_status cp_vp_step(cp_vp* v, cp_vp_step_info* info);
/* After the last step: surface velocities, Cp and forces with the
 * particle-induced potentials (cpCaseVP::compVelocity) and the final
 * wake::trefftzPlaneVP CL, CDi in info; cp: n_body or NULL. */
cp_status cp_vp_finish(cp_vp* v, cp_solve_info* info, double* cp);
/* particles in creation order: pos, strength (3 each), radius, shed step */
cp_status cp_vp_particles(const cp_vp* v, int32_t* n, double* pos, double* strength, double* radius, int32_t* shed);
/* body mu and sigma (n_body), first- and second-row doublet strengths (m) */
cp_status cp_vp_state(const cp_vp* v, double* mu, double* sigma, double* mu_row1, double* mu_row2);
void cp_vp_release(cp_vp* v);

/* ---- serving: continuous batching of (geometry, conditions) requests -----
 * Replaces the reference's sequential case loop (CPanelMgr.cpp:49-88).
 * Worker threads (slots) take requests in arrival order; each geometry is
 * uploaded and assembled once (CP_FP32_REFINE) and kept on the device under
 * the caller's key, with its own stream, so requests on different
 * geometries overlap on the GPU; queued requests for the same geometry and
 * options are merged into one batched solve of up to 64 conditions.  New
 * geometries are admitted against hbm_budget (0: no limit), evicting idle
 * geometries least recently used first.  When a worker takes a request for
 * a resident geometry of at most 4,096 body panels, queued requests for other
 * idle resident geometries of that size join it (up to 8 geometries) and the
 * worker solves them together with cp_solve_group; each result is the same
 * as in a solve of its own (environment: CPGPU_SERVE_GROUP = geometries per
 * group, 1 turns grouping off; CPGPU_SERVE_GROUP_N = panel limit). */
typedef struct cp_server cp_server;

typedef struct {
    double queue_s;         /* submit -> a worker took the request */
    double prepare_s;       /* geometry upload + assembly (0 on a cache hit) */
    double solve_s;         /* batched solve + post-processing */
    double latency_s;       /* submit -> result */
    int32_t cache_hit;      /* geometry was resident */
    int32_t batch_conditions;  /* columns of the solve it ran in */
    int32_t n_body;
    int32_t group_geometries;  /* geometries of the grouped solve it ran in (1: alone) */
} cp_request_stats;

typedef struct {
    int64_t submitted, completed, conditions, batches;
    int64_t cache_hits, cache_misses, evictions;
    int64_t resident_geometries;
    uint64_t resident_bytes_est, device_bytes;
    int64_t grouped_solves;  /* cp_solve_group calls of two or more geometries */
} cp_server_stats;

cp_status cp_server_create(cp_ctx* ctx, int32_t slots, uint64_t hbm_budget, cp_server** out);
/* k <= 64 conditions; the mesh is copied (used if the key is not resident);
 * want_mu: keep mu (n_body x k) for cp_server_wait */
cp_status cp_server_submit(cp_server* s, uint64_t key, const cp_mesh* mesh, int32_t k, const cp_freestream* fs,
                           const cp_solve_opts* opts, int32_t want_mu, uint64_t* id);
/* blocks until request id is done; info: k entries, mu: n_body*k or NULL.
 * Returns the request's cp_solve status. */
cp_status cp_server_wait(cp_server* s, uint64_t id, cp_solve_info* info, double* mu, cp_request_stats* st);
cp_status cp_server_get_stats(cp_server* s, cp_server_stats* out);
/* finishes running requests, releases every resident geometry */
void cp_server_destroy(cp_server* s);

/* JSON text describing the batched solver of a geometry: tuned GEMM shape
 * and the autotuner's timings, block-Jacobi blocks, near-field entries. */
cp_status cp_solver_report(cp_geom* g, char* buf, int32_t len);

/* ---- per-case output fields ----------------------------------------------
 * The arrays the reference writes for every case (cpCase::writeFiles,
 * cpCase.cpp:667-709): surfaceData.vtu cell data (writeBodyData, :711-758),
 * wakeData.vtu (writeWakeData, :883-908) and spanwiseData_Wake<i>.csv
 * (writeSpanwiseData, :910-935).  Every pointer may be NULL; arrays are
 * condition-major (condition q's block follows condition q-1's). */
typedef struct {
    double* sigma;          /* k x n_body: source strengths -Vinf . n_i */
    double* velocity;       /* k x n_body x 3: panel velocity (V_x / PG) */
    double* potential;      /* k x n_body: Vinf . c_i - mu_i */
    double* cp;             /* k x n_body */
    double* wake_mu;        /* k x n_wake: wakePanel::setMu */
    double* wake_potential; /* k x n_wake */
    double* spanwise;       /* k x n_wakes x n_stations x 3: y, Cl / PG,
                               Cdi / PG^2 (wake::trefftzPlane, wake.cpp:146-188) */
} cp_case_fields;

/* Evaluates the fields of k conditions on the device from their body doublet
 * strengths mu (k x n_body, as cp_solve returns them); opts supplies the
 * Mach number and reference quantities (tolerances are not used). */
cp_status cp_case_fields_eval(cp_geom* g, int32_t k, const cp_freestream* fs, const cp_solve_opts* opts,
                              const double* mu, cp_case_fields* out);

/* Diagnostic, no counterpart in the reference: the strip block-Jacobi
 * preconditioner of the mixed-precision solve (built on first use after
 * cp_assemble(CP_FP32_REFINE)).  sizes[0] = blocks, sizes[1] = panels
 * covered, sizes[2] = packed inverse entries (all 0 when the point-Jacobi
 * fallback is in use).  ptr (blocks+1), perm (panels) and inv (FP32 inverse of
 * block q, column-major, at sum of earlier b^2) may be NULL. */
cp_status cp_block_jacobi_read(cp_geom* g, int64_t* sizes, int32_t* ptr, int32_t* perm, float* inv);

/* ---- timing: HIP-event totals of the named kernel on the context stream -- */
cp_status cp_timing_reset(cp_ctx* ctx);
/* kernel: "assemble_f64", "assemble_f32", "gemv_f64", "gemm_f32", "resid_f64", ...
 * returns total milliseconds and launch count since the last reset */
cp_status cp_timing_get(cp_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches);
cp_status cp_timing_enable(cp_ctx* ctx, int on);
/* records only the named kernel's intervals while timing is on (NULL or "":
 * every kernel); no counterpart in the reference: for measuring one kernel
 * without the event records of the others in the timed region */
cp_status cp_timing_only(cp_ctx* ctx, const char* kernel);

#ifdef __cplusplus
}
#endif
#endif /* CPGPU_H */
