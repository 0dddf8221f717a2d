/*
 * FP64 CPU oracle for the CPanel solve hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a line-by-line restatement of the
 * reference CPanel algorithms (jakeddavis/CPanel, read as text) and exists to
 * CHECK the HIP product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (cpanel_amd/) never
 * links, imports or calls it.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each part (what pins it) is listed in oracle/README.md.
 */
#ifndef CPANEL_ORACLE_H
#define CPANEL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_geom or_geom;

/* ---- single-panel influence kernels (KATs) ------------------------------ */
/* bodyPanel::panelPhiInf (bodyPanel.cpp:132-190) for a triangle (nv=3) or
 * quad (nv=4) with vertices v[3*nv]; returns source and doublet potential. */
void or_panel_phi_inf(int nv, const double* v, const double* poi,
                      double* phi_src, double* phi_dub);
/* panel::dubPhiInf (panel.cpp:208-252) */
double or_dub_phi_inf(int nv, const double* v, const double* poi);
/* bodyPanel::panelVInf (bodyPanel.cpp:305-348): vsrc[3], vdub[3] */
void or_panel_v_inf(int nv, const double* v, const double* poi,
                    double* vsrc, double* vdub);
/* panel::dubVInf (panel.cpp:666-705) */
void or_dub_v_inf(int nv, const double* v, const double* poi, double* vdub);

/* ---- geometry ----------------------------------------------------------- */
/* Build the panel geometry and topology from 0-based connectivity as
 * geometry::readTri does after "connectivity.array()-1"
 * (geometry.cpp:132-484).  ids >1000 are wake surfaces. */
or_geom* or_geom_create(int n_nodes, const double* xyz, int n_tris,
                        const int32_t* tri, const int32_t* ids);
/* Parse a CPanel .tri file (geometry.cpp:155-239). NULL on error. */
or_geom* or_geom_read_tri(const char* path);
void or_geom_destroy(or_geom* g);

int or_geom_n_body(const or_geom* g);
int or_geom_n_wake(const or_geom* g);
int or_geom_n_wakelines(const or_geom* g);
/* Body panels in the reference order (geometry.cpp:452-458): per panel the
 * file triangle index, centre[3], normal[3], area, longSide, tip flag. */
void or_geom_body_info(const or_geom* g, int32_t* tri_index, double* center,
                       double* normal, double* area, double* long_side,
                       int32_t* tip_flag);
/* Wake panels (geometry.cpp:459-469): file triangle index and the Kutta
 * folding table of wakePanel::interpPanels (wakePanel.cpp:48-81):
 * idx4[4*nw] = body indices (U1, L1, U2, L2), coef[nw] = interpCoeff. */
void or_geom_wake_info(const or_geom* g, int32_t* tri_index, int32_t* idx4,
                       double* coef);
/* Cluster used for the CHTLS velocity (bodyPanel.cpp:402-487):
 * out[16*nb] body indices (-1 padded), order[nb] = TSorder. */
void or_geom_clusters(const or_geom* g, int32_t* out16, int32_t* ts_order);

/* ---- assembly (geometry::setInfCoeff, geometry.cpp:809-891) ------------- */
/* A, B column-major nb x nb with leading dimension lda. */
void or_assemble(const or_geom* g, double* A, double* B, int64_t lda,
                 int nthreads);
/* Columns j0..j1-1 of the body-body loop of or_assemble (no wake folding):
 * a bounded sample of the assembly for single-thread timing. */
void or_assemble_cols(const or_geom* g, double* A, double* B, int64_t lda, int j0, int j1,
                      int nthreads);

/* ---- solve -------------------------------------------------------------- */
/* Eigen::BiCGSTAB<MatrixXd> with its default Jacobi preconditioner and x0=0
 * (cpCase.cpp:196-198).  tol<=0 means Eigen's default (machine epsilon);
 * maxit<=0 means Eigen's default (2n).  Returns iterations; *err is
 * ||r||/||b|| as Eigen's error(). */
int or_bicgstab(int n, const double* A, int64_t lda, const double* b,
                double* x, double tol, int maxit, int nthreads, double* err);

/* ---- full steady case (cpCase::run without file output) ----------------- */
typedef struct {
    double V, alpha_deg, beta_deg, mach;
    double Sref, bref, cref, cg[3];
} or_case_params;

typedef struct {
    double CL, CDi;            /* Trefftz plane (cpCase.cpp:541-552) */
    double Fbody[3], Fwind[3]; /* surface integrated (cpCase.cpp:342-365) */
    double CM[3];
    double solve_err;          /* Eigen error() of the solve */
    int    solve_iters;
} or_case_result;

/* Runs setSourceStrengths, solveMatrixEq, compVelocity, trefftzPlaneAnalysis
 * on assembled A, B (cpCase.cpp:31-139).  Per-body-panel outputs (any may be
 * NULL): mu, sigma, phi, vel[3*nb], cp.  Wake panel mu[nw].  Spanwise
 * Trefftz arrays (165 each): yloc, cl, cd. */
int or_case_run(or_geom* g, const double* A, const double* B, int64_t lda,
                const or_case_params* p, double tol, int maxit, int nthreads,
                or_case_result* res, double* mu, double* sigma, double* phi,
                double* vel, double* cp, double* wake_mu, double* yloc,
                double* span_cl, double* span_cd);

/* Post-processing only, from given body doublet strengths (used to check the
 * GPU solve's outputs on the same mu). */
void or_case_post(or_geom* g, const or_case_params* p, const double* mu_in,
                  or_case_result* res, double* phi, double* vel, double* cp,
                  double* wake_mu);

/* Wake induced potential at a point, geometry::wakePotential
 * (geometry.cpp:1229-1237), with the wake mu left by the last case run. */
double or_wake_potential(const or_geom* g, const double* pnt);

/* ---- vortex-particle kernels (vp_oracle.cpp) ---------------------------- */
/* particle octree (octree.h:39-195, node.h:30-214), max members per leaf and
 * Barnes-Hut opening angle theta (octree.h:37); pseudo-particles as
 * particleFMM::computeMultExp (particleFMM.cpp:29-58) */
void* or_bh_build(int n, const double* xyz, const double* alpha, const double* radius,
                  int max_members, double theta);
void or_bh_destroy(void* tree);
void or_bh_info(void* tree, int* n_nodes, int* levels, double* origin, double* half);
/* mode 0: velocity at points (algebraic kernel), 1: velocity on particles
 * (Gaussian), 2: stretching, 3: diffusion (particleFMM.cpp:65-116) */
void or_bh_eval(void* tree, int mode, int m, const double* pts, double* out);
/* direct sums with the same kernels; mode 4: Gaussian velocity at points */
void or_vp_direct(int mode, int n, const double* xyz, const double* alpha, const double* radius,
                  int m, const double* pts, double* out);
void or_vp_pair(int mode, const double* xs, const double* as, double rs, const double* xt,
                const double* at, double rt, double* out);
/* panel -> point velocity of body (sigma, mu) and wake (mu_w) panels
 * (bodyPanel::panelV bodyPanel.cpp:122-130, wakePanel::panelV wakePanel.cpp:89-92) */
void or_panels_to_points(int nb, const int* nvb, const double* vb, const double* sigma,
                         const double* mu, int nw, const int* nvw, const double* vw,
                         const double* mu_w, int m, const double* pts, double* out);

/* ---- panel -> particle stretching (bodyPanel::partStretching, bodyPanel.cpp:697-1063;
 * wakePanel::partStretching, wakePanel.cpp:266-289; panel.cpp:813-962) ------- */
void or_velgrad(int nv, const double* v, int body, int kind, double sigma, double mu, const double* pos,
                double* out);
void or_panels_stretching(int nb, const int* nvb, const double* vb, const double* sigma, const double* mu, int nw,
                          const int* nvw, const double* vw, const double* mu_w, int m, const double* pos,
                          const double* alpha, double* out);

#ifdef __cplusplus
}
#endif
#endif
