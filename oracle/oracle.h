/*
 * oracle.h -- CPU FP64 restatement of DNDS's 2D VR finite-volume hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP implementation and the CPU baseline of the benchmark; no product code
 * links, imports or calls it.  It is written independently of the product
 * sources: it restates the reference's loops (face loop with in-order
 * scatter, cell loop in cell2face order) and its expression order, compiled
 * with -ffp-contract=off.
 *
 * Array layouts are the same logical layouts the product's C ABI documents
 * (reference layouts): u [n_cell][4]; uRec [n_cell][36] column-major 9x4 per
 * cell; setup arrays row-major.
 */
#ifndef DNDS_ORACLE_H
#define DNDS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_mesh {
    int64_t n_cell, n_face;
    const int64_t *cell_ptr;     /* [n_cell+1]              */
    const int64_t *cell2face;    /* local face order        */
    const int64_t *face2cell;    /* [n_face][2], INT64_MIN  */
    const int32_t *face_bc;      /* BoundaryType values     */
} oracle_mesh;

typedef struct oracle_setup {
    const double *volume;           /* [n_cell]            */
    const double *face_area;        /* [n_face]            */
    const double *face_center;      /* [n_face][2]         */
    const double *face_norm_center; /* [n_face][2]         */
    const double *face_norm;        /* [n_face][3][2]      */
    const double *face_weight;      /* [n_face][10]        */
    const double *face_d0;          /* [n_face][3][2][9]   */
    const double *face_dibj_g0;     /* [n_face][2][10][9]  */
    const double *matrix_secondary; /* [n_face][2][9][9]   */
    const double *ainv;             /* [n_cell][9][9]      */
    const double *ainv_b;           /* [n_c2f][9][9]       */
    const double *ainv_bvec;        /* [n_c2f][9]          */
    const double *face_gauss_xy;    /* [n_face][3][2][2] scaled (X, Y) of each LINE_3 point per side (viscous terms) */
    const double *cell_scale;       /* [n_cell][2]         */
} oracle_setup;

typedef struct oracle_setting {
    int32_t riemann_solver;        /* 1 Roe, 2 HLLC, 3 HLLEP, 11-15 Roe_M1-M5 */
    double gamma;
    double far_field[4];
    double jacobi_relax;
    double tang_weight;            /* for the WBAP functional */
    int32_t tang_weight_direction; /* 0 Norm, 1 Bary          */
    double wbap_smooth_indicator_scale;
    double wbap_n_std;
    int32_t gmres_code;            /* fsolve: 0 LU-SGS, 1 GMRES, 2 LU-SGS then GMRES */
    int32_t n_gmres_space;         /* Krylov subspace size      */
    int32_t n_gmres_iter;          /* restarts                  */
    int32_t limiter_procedure;     /* 0 WBAP V2, 1 WBAP V3 (limiterProcedure) */
    double mu_gas;                 /* idealGasProperty: 0 = Euler            */
    double pr_gas, r_gas, t_ref, c_sutherland;
    const int64_t *sor_scan;       /* non-null: the steps' reconstruction sweeps are
                                      SOR_Instead sweeps in this scan order      */
    int32_t sor_inverse;           /* SOR_InverseScanning: then the reverse order */
} oracle_setting;

/* VRFiniteVolume2D::ReconstructionJacobiStep, SOR_Instead = false
 * (DNDS_FV_VR.hpp:1474-1713): urec_new receives the sweep, urec is then
 * overwritten with it.  Returns 0 or a nonzero error code. */
int oracle_jacobi_step(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                       double *urec, double *urec_new, double t);
/* SOR_Instead: one in-place sweep in the scan order `scan` (n_cell entries) */
int oracle_sor_step(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                    double *urec, const int64_t *scan, double t);

/* EulerEvaluator::EvaluateDt (DNDS_FV_EulerEvaluator.cpp:14-183) for the
 * Euler model (muGas = 0): fills dt, *dt_min, lambda_face, delta_lambda_face. */
int oracle_evaluate_dt(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                       double cfl, double max_dt, int32_t use_local_dt, double *dt, double *dt_min,
                       double *lambda_face, double *delta_lambda_face);

/* EulerEvaluator::EvaluateRHS (DNDS_FV_EulerEvaluator.cpp:197-640) for the
 * Euler model: face loop, Gauss sum, in-order scatter. */
int oracle_evaluate_rhs(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                        const double *delta_lambda_face, const double *u, const double *urec, double t, double *rhs);

/* VRFiniteVolume2D::ReconstructionWBAPLimitFacialV3 (DNDS_FV_VR.hpp:2053-2329)
 * with the solver's characteristic matrices (DNDS_FV_EulerSolver.hpp:749-795):
 * urec is read only, urec_new receives the limited reconstruction,
 * if_use_limiter the smoothness indicator. */
int oracle_wbap_v3(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                   const double *urec, double *urec_new, double *if_use_limiter, int32_t if_all);

/* VRFiniteVolume2D::ReconstructionWBAPLimitFacialV2 (DNDS_FV_VR.hpp:1749-2047),
 * same arguments as oracle_wbap_v3. */
int oracle_wbap_v2(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                   const double *urec, double *urec_new, double *if_use_limiter, int32_t if_all);

/* n_steps of ExplicitSSPRK3LocalDt::Step (DNDS_ODE.hpp:131-178) with frhs =
 * n_sweeps Jacobi sweeps [+ WBAP V2 or V3 by limiter_procedure] + EvaluateRHS (on the limited
 * reconstruction when use_limiter) and fdt = EvaluateDt.  *t advances by the
 * global minimum dt of each step. */
int oracle_ssprk3_steps(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, int32_t n_steps,
                        int32_t n_sweeps, int32_t use_limiter, double cfl, double max_dt, int32_t use_local_dt,
                        double *u, double *urec, double *t);

/* Gas-dynamics kernels exposed for unit tests (rotated frame, 4 variables). */
int oracle_roe_flux(const double *UL, const double *UR, double gamma, double dlambda, double *F);
int oracle_roe_m_flux(const double *UL, const double *UR, double gamma, double dlambda, int scheme, double *F);
int oracle_hllc_flux(const double *UL, const double *UR, double gamma, double dlambda, double *F);
int oracle_hllep_flux(const double *UL, const double *UR, double gamma, double dlambda, double *F);
/* Left / right eigenvector matrices of the x-direction Euler Jacobian
 * (EulerGasLeftEigenVector / EulerGasRightEigenVector), row-major 4x4. */
int oracle_eigenvectors(const double *U, double gamma, double *L, double *R);
/* generateBoundaryValue (DNDS_FV_EulerEvaluator.hpp:958-1275). */
int oracle_boundary_value(const oracle_setting *cfg, double *UL, const double *unit_norm, const double *p, double t,
                          int32_t bc, int32_t fix_ul, double *UR);

/* LU-SGS, jacobianTypeCode 0, Euler model (DNDS_FV_EulerEvaluator.cpp:995-1357).
 * scan2cell (may be NULL: identity, the reference's default order) is the
 * order of the sequential sweeps.  jinv[i] = 1 / D_i from
 * oracle_lusgs_matrix_init; lambda_face from oracle_evaluate_dt.  uInc and
 * uIncNew may alias (the solver passes the same array). */
int oracle_lusgs_matrix_init(const oracle_mesh *m, const oracle_setup *s, const double *dTau, double dt,
                             double alphaDiag, const double *lambda_face, double *jinv);
int oracle_lusgs_forward(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                         const int64_t *scan2cell, const double *jinv, const double *lambda_face, double alphaDiag,
                         const double *rhs, const double *u, const double *uInc, double *uIncNew);
int oracle_lusgs_backward(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                          const int64_t *scan2cell, const double *jinv, const double *lambda_face, double alphaDiag,
                          const double *rhs, const double *u, const double *uInc, double *uIncNew);

/* Implicit-Euler dual-time iterations with LU-SGS (DNDS_ODE.hpp:201-265;
 * DNDS_FV_EulerSolver.hpp:713-1000): up to max_iter iterations, stopping when
 * the L1 density increment falls below threshold x its running maximum. */
int oracle_implicit_euler(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                          const int64_t *scan2cell, int32_t max_iter, int32_t n_sweeps, double cfl, double dt,
                          int32_t use_local_dt, double threshold, double t, double *u, double *urec,
                          double *res_hist, int32_t *n_done);
int oracle_bdf_steps(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const int64_t *scan2cell,
                     int32_t k, int32_t n_steps, int32_t max_iter, int32_t n_sweeps, double cfl, double dt,
                     int32_t use_local_dt, double threshold, double t, double *u, double *urec, int32_t *n_iter);
int oracle_sdirk4_step(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const int64_t *scan2cell,
                       int32_t max_iter, int32_t n_sweeps, double cfl, double dt, int32_t use_local_dt,
                       double threshold, double t, double *u, double *urec, int32_t *n_iter);

/* LUSGSMatrixVec (DNDS_FV_EulerEvaluator.cpp:1062-1145): Ax_i = D_i x_i -
 * sum_f (alpha/2) S_f / V_i dF(u_j, x_j) over all neighbours; diag = D. */
int oracle_lusgs_matrix_vec(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                            const double *diag, const double *lambda_face, double alphaDiag, const double *u,
                            const double *x, double *Ax);
/* the solver's fsolve (DNDS_FV_EulerSolver.hpp:852-975) for cfg->gmres_code:
 * LUSGSMatrixInit, the compression of rhs, LU-SGS and / or GMRES. xinc out. */
int oracle_fsolve(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const int64_t *scan2cell,
                  const double *dTau, double dt, double alphaDiag, const double *lambda_face, const double *u,
                  double *rhs, double *xinc);

/* fluxJacobian0_Right_Times_du (DNDS_FV_EulerEvaluator.hpp:928-956), Euler */
int oracle_flux_jacobian_times_du(const double *U, const double *unit_norm, const double *dU, double lambda_main,
                                  double gamma, double *dF);

/* worker threads for the per-cell / per-face loops (results do not depend on it) */
void oracle_set_threads(int32_t n);

#ifdef __cplusplus
}
#endif

#endif

