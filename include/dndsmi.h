/*
 * dndsmi.h -- C ABI of the MI355X (gfx950) implementation of DNDS's 2D
 * variational-reconstruction finite-volume hot path.
 *
 * Everything here is plain C: pointers, sizes and status codes.  No HIP,
 * torch or C++ type crosses this boundary.  Host arrays are caller-owned and
 * are only read or written during the call that receives them; device state
 * lives in objects owned by the library (dndsmi_ctx, dndsmi_array).
 *
 * Reference interfaces replaced (paths relative to the DNDS source tree):
 *   mesh topology       CompactFacedMeshSerialRW, InterpolateTopology
 *                       DNDS_Mesh.hpp:238-350, 634-709, 1330-1361
 *   geometry + VR setup ImplicitFiniteVolume2D ctor DNDS_FV_VR.hpp:29-102,
 *                       VRFiniteVolume2D::Initialization DNDS_FV_VR.cpp:3-13
 *   Jacobi sweep        VRFiniteVolume2D::ReconstructionJacobiStep
 *                       DNDS_FV_VR.hpp:1474-1713
 *   time step           EulerEvaluator::EvaluateDt DNDS_FV_EulerEvaluator.cpp:14-183
 *   right-hand side     EulerEvaluator::EvaluateRHS DNDS_FV_EulerEvaluator.cpp:197-640
 *   SSP-RK3 step        ODE::ExplicitSSPRK3LocalDt::Step DNDS_ODE.hpp:131-178
 *
 * Every function returns DNDSMI_OK (0) or one of the error codes below; the
 * message of the last failure of a context is returned by dndsmi_last_error.
 * The reference aborts on the same conditions (DNDS_assert, DNDS_Defines.h:31-44);
 * the C++ facade in dndsmi.hpp restores that behaviour.
 */
#ifndef DNDSMI_H
#define DNDSMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    DNDSMI_OK = 0,
    DNDSMI_ERR_INVALID_ARGUMENT = 1,
    DNDSMI_ERR_HIP = 2,
    DNDSMI_ERR_NAN = 3,              /* non-finite value detected (DNDS_FV_EulerEvaluator.hpp:576-588) */
    DNDSMI_ERR_NONPOSITIVE_STATE = 4,/* rho <= 0 or negative sound speed squared in a flux (DNDS_Gas.hpp:322, 438) */
    DNDSMI_ERR_UNSUPPORTED = 5,      /* option outside the implemented set */
    DNDSMI_ERR_OUT_OF_MEMORY = 6,
    DNDSMI_ERR_SETUP = 7,            /* degenerate geometry during setup (DNDS_FV_VR.hpp:67-69, DNDS_FV_VR.cpp:579) */
    DNDSMI_ERR_IO = 8                /* a file could not be opened, read or written */
};

/* ---- boundary kinds: the values of BoundaryType, DNDS_Mesh.hpp:21-35 ----- */
enum {
    DNDSMI_BC_INNER = -1,
    DNDSMI_BC_WALL = 1,
    DNDSMI_BC_FARFIELD = 2,
    DNDSMI_BC_WALL_EULER = 3,
    DNDSMI_BC_WALL_NOSLIP = 4,
    DNDSMI_BC_SPECIAL_DMRFAR = -2,
    DNDSMI_BC_SPECIAL_RTFAR = -3,
    DNDSMI_BC_SPECIAL_IVFAR = -4,
    DNDSMI_BC_SPECIAL_2DRIEMANNFAR = -5
};

/* second entry of face2cell for a boundary face (FACE_2_VOL_EMPTY, DNDS_Mesh.hpp:19) */
#define DNDSMI_NO_CELL INT64_MIN

/* ======================================================================== */
/* Mesh                                                                     */
/* ======================================================================== */

/*
 * A 2D mesh of linear triangles (3 nodes) and quadrilaterals (4 nodes),
 * counter-clockwise, with the reference's conventions (DNDS_Mesh.hpp:238-350):
 * faces are numbered in order of first appearance when walking the cells in
 * order and each cell's faces in local order (tri 01,12,20; quad 01,12,23,30),
 * the face's nodes and its first cell f2c[0] come from that first walk, so the
 * face normal (dy, -dx) points out of f2c[0].  A periodic face is an inner
 * face whose second cell lies across the domain; face_shift[f] is the
 * translation that moves f2c[1]'s nodes next to the face (zero otherwise).
 */
typedef struct dndsmi_mesh_view {
    int64_t n_node, n_cell, n_face;
    const double *node_xy;         /* [n_node][2]                            */
    const int64_t *cell_ptr;       /* [n_cell+1] offsets into cell2node and cell2face */
    const int64_t *cell2node;      /* CCW node lists                          */
    const int64_t *cell2face;      /* local face order                        */
    const int64_t *face2node;      /* [n_face][2]                            */
    const int64_t *face2cell;      /* [n_face][2], DNDSMI_NO_CELL if boundary */
    const int32_t *face2cell_ref;  /* [n_face][2] local face index in each cell (face2cellRefLocal) */
    const int32_t *face_bc;        /* [n_face] DNDSMI_BC_*                    */
    const double *face_shift;      /* [n_face][2]                            */
} dndsmi_mesh_view;

enum {
    DNDSMI_MESH_TRI_SPLIT = 0, /* each quad cut along a diagonal chosen by a seeded coin flip */
    DNDSMI_MESH_QUAD = 1,      /* quads only */
    DNDSMI_MESH_MIXED = 2      /* each quad kept (p = 1/2, seeded) or cut as in TRI_SPLIT */
};

typedef struct dndsmi_mesh_spec {
    int32_t kind;                  /* DNDSMI_MESH_*                           */
    int32_t periodic_x, periodic_y;
    int64_t nx, ny;                /* quads per direction                      */
    double x0, x1, y0, y1;         /* domain                                   */
    double jitter;                 /* interior nodes moved by U(-j, j) * local spacing, per axis */
    double y_ratio;                /* geometric ratio of successive y spacings (1 = uniform) */
    uint64_t seed;
    int32_t bc_xmin, bc_xmax, bc_ymin, bc_ymax; /* DNDSMI_BC_* of non-periodic sides */
} dndsmi_mesh_spec;

typedef struct dndsmi_mesh dndsmi_mesh;

/* Deterministic seeded generator of structured-as-unstructured meshes. */
int dndsmi_mesh_generate(const dndsmi_mesh_spec *spec, dndsmi_mesh **out);
/* Topology from raw cells (the Gmsh-reader path): nodes, CCW cells, and the
 * boundary edges with their DNDSMI_BC_* kind; faces are built as the
 * reference's InterpolateTopology does (DNDS_Mesh.hpp:238-350). */
int dndsmi_mesh_from_cells(int64_t n_node, const double *node_xy, int64_t n_cell, const int64_t *cell_ptr,
                           const int64_t *cell2node, int64_t n_bnd, const int64_t *bnd2node, const int32_t *bnd_bc,
                           dndsmi_mesh **out);
/* GmshReader::FileRead + InterpolateTopology as CompactFacedMeshSerialRWBuild
 * calls them (DNDS_Mesh.hpp:87-350, 760-835): a Gmsh 2.2 ASCII file of
 * first-order triangles and quadrangles (physical groups of dimension 2) and
 * boundary lines (dimension 1) named bc-1 (Wall), bc-2 (Farfield), bc-3
 * (Wall_Euler), bc-4 (Wall_NoSlip), bc-DMRFar, bc-RTFar, bc-IVFar or
 * bc-2DRiemannFar.  Cells in element-id order, counter-clockwise; point
 * elements (dimension 0) are skipped.  err (may be NULL) receives a
 * description of a failure. */
int dndsmi_mesh_read_gmsh(const char *path, dndsmi_mesh **out, char *err, size_t err_len);
/* CompactFacedMeshSerialRW::PrintSerialPartPltBinaryDataArray(fname, 0,
 * n_data, names, data, t) on one rank (DNDS_Mesh.hpp:1634-1814): a Tecplot
 * binary file (#!TDV112), one FEQUADRILATERAL zone, X/Y/Z at the nodes, the
 * n_data arrays data[k][n_cell] and iPart at the cell centres; triangles
 * repeat their third node. */
int dndsmi_mesh_write_plt(const dndsmi_mesh *mesh, const char *path, int32_t n_data, const char *const *names,
                          const double *data, double t, char *err, size_t err_len);
int dndsmi_mesh_get_view(const dndsmi_mesh *mesh, dndsmi_mesh_view *view);
/* Checks a mesh view before use (no counterpart in the reference, whose
 * meshes come from its own reader): every node, face and cell index in
 * range, each cell's cell2face entries and the faces' face2cell /
 * face2cell_ref naming each other, 3 or 4 faces per cell, and a
 * DNDSMI_BC_* kind on every boundary face.  dndsmi_vr_create,
 * dndsmi_ctx_create and dndsmi_warmstart_create run the same check and
 * return DNDSMI_ERR_INVALID_ARGUMENT on a view that fails it; this entry
 * also describes the failure in err (may be NULL). */
int dndsmi_mesh_view_check(const dndsmi_mesh_view *view, char *err, size_t err_len);
void dndsmi_mesh_destroy(dndsmi_mesh *mesh);

/* ======================================================================== */
/* VR setup (host, FP64)                                                    */
/* ======================================================================== */

/* VRFiniteVolume2D::Setting, DNDS_FV_VR.hpp:170-230 (names kept) */
typedef struct dndsmi_vr_setting {
    int32_t rec_order;              /* P_ORDER: 3, or 1 (no limiter data)    */
    double jacobi_relax;            /* JacobiRelax                            */
    double wall_weight, far_weight; /* wallWeight, farWeight                  */
    double tang_weight;             /* tangWeight                             */
    double tang_weight_mod_min;     /* tangWeightModMin (only 1 implemented)   */
    int32_t tang_weight_direction;  /* 0 = TWD_Norm, 1 = TWD_Bary             */
    int32_t anisotropic_lengths;    /* anisotropicLengths                     */
    int32_t weight_scheme_geom;     /* 0 None, 1 D, 2 S, 3 SDHQM              */
    int32_t weight_scheme_dir;      /* 0 None, 1 optHQM                       */
    double weight_len_scale;        /* weightLenScale                         */
    double wbap_smooth_indicator_scale; /* WBAP_SmoothIndicatorScale          */
    double wbap_n_std;              /* WBAP_nStd                              */
    int32_t use_local_coord;        /* useLocalCoord (no limiter data)        */
    int32_t orthogonalize_base;     /* must be 0                              */
    int32_t norm_wbap;              /* must be 0                              */
    int32_t build_limiter_data;     /* also build the WBAP face data          */
    int32_t sor_instead;            /* SOR_Instead: in-place (Gauss-Seidel) sweeps,
                                       over the cells colour by colour on the device */
    int32_t sor_inverse_scanning;   /* SOR_InverseScanning: each SOR sweep is a forward
                                       and a backward pass (colours reversed)   */
} dndsmi_vr_setting;

/* The reference's defaults, DNDS_FV_VR.hpp:172-228. */
void dndsmi_vr_setting_default(dndsmi_vr_setting *s);

/*
 * Setup products, host FP64, flattened in the reference's logical layout.
 * Row-major small matrices; 9 = NDOF-1 rec dofs at p=3, 10 = NDIFF.
 */
typedef struct dndsmi_vr_view {
    int64_t n_cell, n_face, n_c2f;    /* n_c2f = cell_ptr[n_cell]                  */
    const double *volume;             /* [n_cell] volumeLocal (TRI_7 / QUAD_9)     */
    const double *face_area;          /* [n_face] faceArea (LINE_5)                */
    const double *cell_bary;          /* [n_cell][2] cellBaries (TRI_4 / QUAD_9)   */
    const double *base_moment;        /* [n_cell][10] baseMoments                  */
    const double *face_center;        /* [n_face][2] faceCenters                   */
    const double *face_norm_center;   /* [n_face][2] faceNormCenter                */
    const double *face_norm;          /* [n_face][3][2] faceNorms at LINE_3 points */
    const double *face_weight;        /* [n_face][10] faceWeights                  */
    const double *face_d0;            /* [n_face][3][2][9] row 0 of faceDiBjGaussBatch, cols 1..9 */
    const double *face_dibj_g0;       /* [n_face][2][10][9] all rows at LINE_3 point 0 (NULL unless limiter data) */
    const double *matrix_secondary;   /* [n_face][2][9][9] matrixSecondaryBatch (NULL unless limiter data) */
    const double *ainv;               /* [n_cell][9][9] matrixBatch.m(0)          */
    const double *ainv_b;             /* [n_c2f][9][9] matrixBatch.m(ic2f+1)      */
    const double *ainv_bvec;          /* [n_c2f][9] vectorBatch rows              */
    /* [n_face][3][2][2] (X, Y) = ((px - cx) / sx, (py - cy) / sy) of LINE_3
     * point g seen from side s (local coordinates: xi = A (p - c), A the
     * side's cell_frame): the arguments of the monomials whose values (minus
     * the cell's moments) and derivatives form face_d0 and face_dibj_g0; the
     * device evaluates the bases from these. */
    const double *face_gauss_xy;
    const double *cell_scale;         /* [n_cell][2] (sx, sy) of FDiffBaseValue (CoordMinMaxScale) */
    /* [n_cell][2][2] A = invPJacobi of the local inertia coordinates, row-major
     * (DNDS_FV_VR.hpp:358-412; DNDS_FV_VR.cpp:118-139); NULL in global
     * coordinates.  With local coordinates the device evaluates the bases'
     * values only: contexts on such a setup refuse mu_gas > 0 (the viscous
     * terms need derivative rows) with DNDSMI_ERR_UNSUPPORTED, and the setup
     * refuses build_limiter_data. */
    const double *cell_frame;
} dndsmi_vr_view;

typedef struct dndsmi_vr dndsmi_vr;

/* ImplicitFiniteVolume2D ctor + VRFiniteVolume2D::Initialization on the host,
 * nthreads worker threads (0 = hardware concurrency). */
int dndsmi_vr_create(const dndsmi_mesh_view *mesh, const dndsmi_vr_setting *setting, int32_t nthreads,
                     dndsmi_vr **out);
int dndsmi_vr_get_view(const dndsmi_vr *vr, dndsmi_vr_view *view);
void dndsmi_vr_destroy(dndsmi_vr *vr);

/* ======================================================================== */
/* Device context and arrays                                                */
/* ======================================================================== */

/* EulerEvaluator<NS_2D>::Setting, DNDS_FV_EulerEvaluator.hpp:139-196: the
 * inviscid subset, then idealGasProperty's viscous constants (:151-160).
 * mu_gas = 0 is the Euler model (the viscous terms are skipped; with the
 * reference's formulas they are exactly zero then); mu_gas > 0 adds the
 * laminar Navier-Stokes terms of NS_2D with Sutherland's law, and then
 * pr_gas, r_gas and t_ref must be positive.  CpGas = r_gas gamma / (gamma - 1)
 * as the solver sets it (DNDS_FV_EulerSolver.hpp:334-339). */
typedef struct dndsmi_euler_setting {
    int32_t riemann_solver;           /* RiemannSolverType: 1 Roe, 2 HLLC, 3 HLLEP, 11-15 Roe_M1-M5 */
    double gamma;                     /* idealGasProperty.gamma                 */
    double far_field[4];              /* farFieldStaticValue (rho, rhou, rhov, E) */
    double mu_gas;                    /* idealGasProperty.muGas (reference default 1) */
    double pr_gas;                    /* prGas (0.72)                           */
    double r_gas;                     /* Rgas (1)                               */
    double t_ref;                     /* TRef (273.15)                          */
    double c_sutherland;              /* CSutherland (110.4)                    */
} dndsmi_euler_setting;

/* The reference's RiemannSolverType codes (DNDS_FV_EulerEvaluator.hpp:140-150). */
enum {
    DNDSMI_RS_ROE = 1,
    DNDSMI_RS_HLLC = 2,
    DNDSMI_RS_HLLEP = 3,
    DNDSMI_RS_ROE_M1 = 11,
    DNDSMI_RS_ROE_M2 = 12,
    DNDSMI_RS_ROE_M3 = 13,
    DNDSMI_RS_ROE_M4 = 14,
    DNDSMI_RS_ROE_M5 = 15
};

typedef struct dndsmi_ctx dndsmi_ctx;
typedef struct dndsmi_array dndsmi_array;

/* Array kinds: DOF = nvar doubles per cell (ArrayDOFV<4>), REC = 9 x nvar per
 * cell (ArrayRecV), SCALAR = one double per cell (dt, ifUseLimiter). */
enum { DNDSMI_ARRAY_DOF = 0, DNDSMI_ARRAY_REC = 1, DNDSMI_ARRAY_SCALAR = 2 };

/*
 * Uploads the mesh and the setup to the current HIP device.  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream); every call on the
 * context is asynchronous on it unless it returns host data.
 */
int dndsmi_ctx_create(const dndsmi_mesh_view *mesh, const dndsmi_vr_view *vr, const dndsmi_vr_setting *vsetting,
                      const dndsmi_euler_setting *esetting, void *stream, dndsmi_ctx **out);
void dndsmi_ctx_destroy(dndsmi_ctx *ctx);
int dndsmi_last_error(const dndsmi_ctx *ctx, char *buf, size_t n);
int dndsmi_sync(dndsmi_ctx *ctx);
/* Bytes of device memory held by the context (setup + scratch, not arrays). */
int64_t dndsmi_ctx_device_bytes(const dndsmi_ctx *ctx);

int dndsmi_array_create(dndsmi_ctx *ctx, int32_t kind, dndsmi_array **out);
void dndsmi_array_destroy(dndsmi_array *a);
/* Host data in the reference layout: DOF [n_cell][4]; REC per cell column-major
 * 9x4, element (r, v) at [cell*36 + r + 9*v] (VarVector::m_by_ncol); SCALAR [n_cell]. */
int dndsmi_array_upload(dndsmi_array *a, const double *host);
int dndsmi_array_download(const dndsmi_array *a, double *host);
int dndsmi_array_copy(dndsmi_array *dst, const dndsmi_array *src);
int dndsmi_array_set_constant(dndsmi_array *a, double value);

/* ReconstructionJacobiStep(u, uRec, uRecNewBuf, FBoundary) with SOR_Instead =
 * false: uRecNewBuf receives the sweep, then uRec = uRecNewBuf.  t is the time
 * handed to the boundary function of Wall_NoSlip faces. */
int dndsmi_reconstruction_jacobi_step(dndsmi_ctx *ctx, const dndsmi_array *u, dndsmi_array *uRec,
                                      dndsmi_array *uRecNewBuf, double t);

/* EvaluateDt(dt, u, CFL, dtMinall, MaxDt, UseLocaldt); also refreshes the
 * evaluator's lambdaFace / deltaLambdaFace.  dtMinall may be NULL (no host sync). */
int dndsmi_evaluate_dt(dndsmi_ctx *ctx, dndsmi_array *dt, const dndsmi_array *u, double CFL, double *dtMinall,
                       double MaxDt, int32_t UseLocaldt);

/* EvaluateRHS(rhs, u, uRec, t). */
int dndsmi_evaluate_rhs(dndsmi_ctx *ctx, dndsmi_array *rhs, const dndsmi_array *u, const dndsmi_array *uRec,
                        double t);

/* ReconstructionWBAPLimitFacialV3(u, uRec, uRecNewBuf, uRecNewBuf1,
 * ifUseLimiter, ifAll, fML, fMR) (DNDS_FV_VR.hpp:2053-2329) with the Euler
 * characteristic matrices of the solver (DNDS_FV_EulerSolver.hpp:749-795):
 * uRecNewBuf receives the limited reconstruction (uRec for cells below
 * WBAP_SmoothIndicatorScale), uRec is read only.  ifUseLimiter (SCALAR, may
 * be NULL) receives the smoothness indicator; uRecNewBuf1 (may be NULL)
 * receives a copy of the result.  Needs a setup built with build_limiter_data. */
int dndsmi_wbap_limit_facial_v3(dndsmi_ctx *ctx, const dndsmi_array *u, const dndsmi_array *uRec,
                                dndsmi_array *uRecNewBuf, dndsmi_array *uRecNewBuf1, dndsmi_array *ifUseLimiter,
                                int32_t ifAll);

/* ReconstructionWBAPLimitFacialV2(u, uRec, uRecNewBuf, uRecNewBuf1,
 * ifUseLimiter, ifAll, fML, fMR) (DNDS_FV_VR.hpp:1749-2047; the solver's
 * limiterProcedure 0, its default, DNDS_FV_EulerSolver.hpp:108 and 796-806):
 * the same indicator as V3; per order 3, 2, 1 each limited cell limits its
 * rows against every neighbour's buffer (higher orders already limited)
 * mapped to its basis.  Arguments and outputs as for V3. */
int dndsmi_wbap_limit_facial_v2(dndsmi_ctx *ctx, const dndsmi_array *u, const dndsmi_array *uRec,
                                dndsmi_array *uRecNewBuf, dndsmi_array *uRecNewBuf1, dndsmi_array *ifUseLimiter,
                                int32_t ifAll);

/* The solver's limiterProcedure (DNDS_FV_EulerSolver.hpp:108): the WBAP the
 * steps below apply when use_limiter is set, 0 for V2 (the default) or 1 for
 * V3.  Any other value is DNDSMI_ERR_INVALID_ARGUMENT. */
int dndsmi_set_limiter_procedure(dndsmi_ctx *ctx, int32_t limiterProcedure);
/* The ifAll argument the SSP-RK3 steps pass to the limiter: every cell is
 * limited, whatever its smoothness indicator.  The reference's frhs passes
 * iter < nPartialLimiterStartLocal && step < nPartialLimiterStart
 * (DNDS_FV_EulerSolver.hpp:796-806); the caller sets it per step (0 after
 * the context is created). */
int dndsmi_set_limiter_if_all(dndsmi_ctx *ctx, int32_t ifAll);

/* One ExplicitSSPRK3LocalDt::Step over frhs = n_sweeps Jacobi sweeps, then
 * (use_limiter) WBAP V2 or V3 (dndsmi_set_limiter_procedure) into a context
 * buffer, then EvaluateRHS on the
 * limited or unlimited reconstruction; fdt = EvaluateDt(CFL, MaxDt,
 * UseLocaldt).  Stage times follow DNDS_ODE.hpp:138-151 (t + dt, t + dt/4,
 * t + 2dt/3) with the global minimum dt; *dt_out (may be NULL) receives it. */
int dndsmi_ssprk3_step(dndsmi_ctx *ctx, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                       int32_t n_sweeps, int32_t use_limiter, double CFL, double MaxDt, int32_t UseLocaldt, double t,
                       double *dt_out);

/* Mixed precision: the same SSP-RK3 step with FP32 sweeps and flux storage.
 * No reference counterpart (a new form of ExplicitSSPRK3LocalDt::Step).
 * Per stage, the first Jacobi sweep runs in FP64 and yields the increment
 * r = J(x) - x; the other n_sweeps - 1 sweeps iterate the correction
 * d <- G d + r in FP32 with A^-1 B rounded to FP32 (G the linear part of the
 * sweep), and x += d in FP64: in real arithmetic exactly n_sweeps FP64
 * sweeps, with rounding confined to the correction (iterative refinement).
 * EvaluateRHS reconstructs the face values in FP32 and stores the face
 * fluxes in FP32; the Riemann solver, the boundary states, WBAP, the time
 * step and the stage update stay FP64.  Wall_Euler / Wall_NoSlip faces
 * enter the correction sweeps through the linear part of their boundary
 * correction.  n_sweeps >= 2. */
int dndsmi_ssprk3_step_mixed(dndsmi_ctx *ctx, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                             int32_t n_sweeps, int32_t use_limiter, double CFL, double MaxDt, int32_t UseLocaldt,
                             double t, double *dt_out);

/* ======================================================================== */
/* Implicit path: LU-SGS and the implicit-Euler dual-time step              */
/* ======================================================================== */

/* The device sweeps LU-SGS in colour-major order: cells are greedily
 * coloured in index order (no two neighbours share a colour) and scanned
 * colour by colour, cells ascending within a colour.  This is one of the
 * reference's sequential scan orders (SOR_iScan2iCell, DNDS_FV_VR.hpp:1376-1463),
 * so a sweep equals the reference's sweep in that order.  colour may be NULL;
 * otherwise it receives n_cell colours. */
int dndsmi_lusgs_colouring(const dndsmi_ctx *ctx, int32_t *colour, int32_t *n_colours);

/* EulerEvaluator::LUSGSMatrixInit(dTau, dt, alphaDiag, u, uRec, jacobianCode, t)
 * (DNDS_FV_EulerEvaluator.cpp:995-1047; declared DNDS_FV_EulerEvaluator.hpp:1429-1433):
 * the diagonal from lambdaFace of the last dndsmi_evaluate_dt on this
 * context.  u and uRec (either may be NULL) and t enter the reference only
 * through the source-term Jacobian, which the NS_2D model lacks; jacobianCode
 * other than 0 (the analytic diagonal) is DNDSMI_ERR_UNSUPPORTED. */
int dndsmi_lusgs_matrix_init(dndsmi_ctx *ctx, const dndsmi_array *dTau, double dt, double alphaDiag,
                             const dndsmi_array *u, const dndsmi_array *uRec, int32_t jacobianCode, double t);

/* EulerEvaluator::UpdateLUSGSForward / UpdateLUSGSBackward(alphaDiag, rhs, u,
 * uInc, uIncNew) (DNDS_FV_EulerEvaluator.cpp:1149-1357; declared
 * DNDS_FV_EulerEvaluator.hpp:1440-1454).  uInc and uIncNew may be the same
 * array (the solver's in-place call); distinct arrays read the old uInc. */
int dndsmi_update_lusgs_forward(dndsmi_ctx *ctx, double alphaDiag, const dndsmi_array *rhs, const dndsmi_array *u,
                                const dndsmi_array *uInc, dndsmi_array *uIncNew);
int dndsmi_update_lusgs_backward(dndsmi_ctx *ctx, double alphaDiag, const dndsmi_array *rhs, const dndsmi_array *u,
                                 const dndsmi_array *uInc, dndsmi_array *uIncNew);

/* EulerEvaluator::EvaluateResidual(res, rhs, P, volWise)
 * (DNDS_FV_EulerEvaluator.cpp:1557-1592): per variable, (sum |r|^P)^(1/P)
 * for P = 1, 2, max |r| for P >= 3; res receives 4 doubles (synchronises). */
int dndsmi_evaluate_residual(dndsmi_ctx *ctx, const dndsmi_array *rhs, int32_t P, int32_t volWise, double *res);

/* No counterpart in the reference.  dndsmi_evaluate_residual for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same result from the same
 * kernels, with the block partials kept in a static scratch array of the
 * library (2 MiB per GPU, held until the process ends) instead of buffers the
 * context allocates on first use.  It calls no hipMalloc or hipFree.  n_cell
 * at most 16,777,216 (DNDSMI_ERR_INVALID_ARGUMENT above that).  All calls on
 * one GPU share the scratch: calls that overlap in time on one GPU (other
 * streams or host threads) race. */
int dndsmi_evaluate_residual_noalloc(dndsmi_ctx *ctx, const dndsmi_array *rhs, int32_t P, int32_t volWise,
                                     double *res);

/* ODE::ImplicitEulerDualTimeStep::Step (DNDS_ODE.hpp:201-265) with the
 * solver's lambdas (DNDS_FV_EulerSolver.hpp:713-1000) for jacobianTypeCode 0,
 * gmresCode 0, nSGSIterationInternal 0, no limiter: per iteration
 * EvaluateDt (pseudo time step dTau), n_sweeps Jacobi sweeps, EvaluateRHS at
 * t + dt, the dual-time right-hand side, LUSGSMatrixInit, forward and
 * backward LU-SGS in place, u += uInc; stops after max_iter iterations or
 * when the L1 density increment falls below threshold times its running
 * maximum.  *n_iter receives the iterations done; res_hist (may be NULL)
 * the L1 increment norms, 4 per iteration. */
int dndsmi_implicit_euler_step(dndsmi_ctx *ctx, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                               int32_t n_sweeps, double CFL, double dt, int32_t UseLocaldt, int32_t max_iter,
                               double threshold, double t, int32_t *n_iter, double *res_hist);

/* EulerEvaluator::LUSGSMatrixVec (DNDS_FV_EulerEvaluator.cpp:1062-1145):
 * AuInc_i = D_i uInc_i - sum_f (alphaDiag/2) S_f / V_i dF(u_j, uInc_j) over
 * every neighbour, D from the last LUSGSMatrixInit. */
int dndsmi_lusgs_matrix_vec(dndsmi_ctx *ctx, double alphaDiag, const dndsmi_array *u, const dndsmi_array *uInc,
                            dndsmi_array *AuInc);

/* The linear solver of the dual-time steps' fsolve (the solver's gmresCode,
 * nGmresSpace, nGmresIter; DNDS_FV_EulerSolver.hpp:117-119, 852-975):
 * 0 LU-SGS forward + backward, 1 GMRES_LeftPreconditioned (DNDS_Linear.hpp:
 * 13-108) with LU-SGS as the preconditioner from a zero start, 2 LU-SGS then
 * GMRES from its result.  n_space >= 1, n_iter >= 0 restarts. */
int dndsmi_set_linear_solver(dndsmi_ctx *ctx, int32_t gmres_code, int32_t n_space, int32_t n_iter);

/* ODE::ImplicitSDIRK4DualTimeStep::Step (DNDS_ODE.hpp:270-372): three
 * stages i, each started from the step's initial state and iterated with the
 * same frhs / fdtau / fsolve / fstop as dndsmi_implicit_euler_step, with
 * alphaDiag = a_ii, the boundary time t + c_i dt and the stage right-hand side
 * (xLast - x) / dt + sum_{j<=i} a_ij rhs_j; then u = xLast + sum_j b_j dt rhs_j.
 * n_iter (may be NULL) receives the iterations of each stage (3 values);
 * res_hist (may be NULL) the L1 increment norms, max_iter x 4 per stage. */
int dndsmi_sdirk4_step(dndsmi_ctx *ctx, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew, int32_t n_sweeps,
                       double CFL, double dt, int32_t UseLocaldt, int32_t max_iter, double threshold, double t,
                       int32_t *n_iter, double *res_hist);

/* ODE::ImplicitBDFDualTimeStep (DNDS_ODE.hpp:392-512): a BDF integrator of
 * order k (1..4) holding the previous k-1 states on the device.  Step i uses
 * order min(i, k) (the reference's start-up).  A handle outlives nothing:
 * after dndsmi_ctx_destroy its steps fail and only destroy is valid. */
typedef struct dndsmi_bdf dndsmi_bdf;
int dndsmi_bdf_create(dndsmi_ctx *ctx, int32_t k, dndsmi_bdf **out);
void dndsmi_bdf_destroy(dndsmi_bdf *bdf);

/* ImplicitBDFDualTimeStep::Step (DNDS_ODE.hpp:441-500) with the same
 * frhs / fdtau / fsolve / fstop as dndsmi_implicit_euler_step: per iteration
 * dTau = EvaluateDt / alpha0, the sweeps and EvaluateRHS at t + dt,
 * rhs = alpha0 rhs - x / dt + (c1 / dt) xLast + sum_j (c_j / dt) xPrev_j,
 * LUSGSMatrixInit and LU-SGS with alphaDiag = alpha0, u += uInc; then the
 * step's initial state joins the history.  dt must equal the previous
 * steps' dt (the reference asserts it). */
int dndsmi_bdf_step(dndsmi_bdf *bdf, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew, int32_t n_sweeps,
                    double CFL, double dt, int32_t UseLocaldt, int32_t max_iter, double threshold, double t,
                    int32_t *n_iter, double *res_hist);

/* ======================================================================== */
/* Learned warm start of the reconstruction (no counterpart in the reference) */
/* ======================================================================== */

/* An MLP (58 -> 128 -> 128 -> 36, GELU) that maps a cell's stencil of cell
 * means and geometry moments to an initial guess of its rec coefficients,
 * trained on converged reconstructions the method itself produced.  Inputs:
 * the differences u_j - u_i over 12 stencil cells (breadth-first over faces),
 * per variable divided by s_v, their largest magnitude but at least
 * 1e-3 max_v |u_i,v|; the 9 base moments;
 * 0.5 log(volume).  Outputs: rec coefficients [r + 9 v] divided by s_v.
 * Training and inference run on the device with BF16 matrix-core GEMMs (FP32
 * accumulation) and Adam.  Plain Jacobi sweeps continue from the guess to
 * the caller's tolerance, so the method decides the result and remains the
 * fallback. */
typedef struct dndsmi_warmstart dndsmi_warmstart;

/* capacity = training samples (cells) the set can hold; seed for the weights
 * and the batch order */
int dndsmi_warmstart_create(dndsmi_ctx *ctx, const dndsmi_mesh_view *mesh, const dndsmi_vr_view *vr, uint64_t seed,
                            int64_t capacity, dndsmi_warmstart **out);
void dndsmi_warmstart_destroy(dndsmi_warmstart *ws);
int64_t dndsmi_warmstart_num_params(void);
/* parameters as one FP32 block: W1 [58][128], b1 [128], W2 [128][128], b2 [128], W3 [128][36], b3 [36] */
int dndsmi_warmstart_get_params(dndsmi_warmstart *ws, float *host);
/* sets the parameters and clears the Adam state */
int dndsmi_warmstart_set_params(dndsmi_warmstart *ws, const float *host);
/* features [n_cell][58] and scales [n_cell][4] of a state (either may be NULL) */
int dndsmi_warmstart_features(dndsmi_warmstart *ws, const dndsmi_array *u, float *X, float *S);
/* appends (features of u, converged uRec / s_v) for every cell */
int dndsmi_warmstart_add_samples(dndsmi_warmstart *ws, const dndsmi_array *u, const dndsmi_array *uRec);
int64_t dndsmi_warmstart_num_samples(const dndsmi_warmstart *ws);
/* appends host rows (X [rows][58], T [rows][36]), e.g. harvested on other meshes */
int dndsmi_warmstart_add_rows(dndsmi_warmstart *ws, const float *X, const float *T, int64_t rows);
/* the training targets of (u, uRec): T [n_cell][36] = uRec / s_v */
int dndsmi_warmstart_targets(dndsmi_warmstart *ws, const dndsmi_array *u, const dndsmi_array *uRec, float *T);
/* mean-squared-error loss and its gradient (same block layout as the
 * parameters) on host rows X [rows][58], T [rows][36]; grad may be NULL */
int dndsmi_warmstart_loss_grad(dndsmi_warmstart *ws, const float *X, const float *T, int32_t rows, float *grad,
                               double *loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) with the given gradient */
int dndsmi_warmstart_adam_step(dndsmi_warmstart *ws, const float *grad, double lr);
/* epochs over the set in shuffled batches, Adam with a cosine learning rate
 * from lr to lr_final; loss_hist (may be NULL) gets the last batch's loss of
 * each epoch */
int dndsmi_warmstart_train(dndsmi_warmstart *ws, int32_t epochs, int32_t batch, double lr, double lr_final,
                           double *loss_hist);
/* uRec = the network's guess for the state u */
int dndsmi_warmstart_predict(dndsmi_warmstart *ws, const dndsmi_array *u, dndsmi_array *uRec);

/* No counterpart in the reference: the learned warm start inside
 * dndsmi_ssprk3_step_mixed, whose stages otherwise start their sweeps from
 * the uRec the previous stage left (the reference's own warm start,
 * DNDS_ODE.hpp:131-178, DNDS_FV_EulerSolver.hpp:713-837).  With ws attached
 * (ws on this context; NULL detaches), every stage first adds to uRec the
 * network's prediction on the change of the cell means since the previous
 * stage.  Plain fallback, judged by the plain method itself: the first
 * check_steps steps after attaching, and again check_steps steps every
 * recheck_interval steps after a passed check (0: no re-checks), run twice
 * -- plain on shadow copies of u and uRec, then with the correction, both
 * with the plain run's time step (returned in dt_out).  At the end of a
 * check, if the corrected state differs from the plain one by more than
 * tolerance in the relative L2 norm over all cells, or if the corrected run
 * met a non-positive state, u and uRec become the plain run's and the
 * correction stays off until the next attach.  Check steps cost two steps
 * and synchronise once at their end; the shadow copies (u, uRec) are
 * allocated on the first check. */
int dndsmi_ssprk3_set_warm_start(dndsmi_ctx *ctx, dndsmi_warmstart *ws, int32_t check_steps,
                                 int32_t recheck_interval, double tolerance);
/* stats[0] stages that added the correction, stats[1] checks passed,
 * stats[2] checks failed (fallbacks to the plain run), counted since
 * dndsmi_ssprk3_set_warm_start; last_deviation (may be NULL) the relative
 * L2 difference found by the last check (-1 before the first) */
int dndsmi_ssprk3_warm_start_stats(const dndsmi_ctx *ctx, int64_t stats[3], double *last_deviation);

/* Jacobi sweeps (as dndsmi_reconstruction_jacobi_step) until the reference's
 * rec increment sum_i |uRec_i - uRecNew_i|^2 / n_cell (DNDS_FV_VR.hpp:1684-1705)
 * falls below tol, or max_sweeps; *sweeps and *res receive the count and the
 * last increment.  No direct counterpart: the reference runs a fixed count. */
int dndsmi_reconstruct_to_tolerance(dndsmi_ctx *ctx, const dndsmi_array *u, dndsmi_array *uRec,
                                    dndsmi_array *uRecNewBuf, double tol, int32_t max_sweeps, double t,
                                    int32_t *sweeps, double *res);

/* ======================================================================== */
/* Measurement                                                              */
/* ======================================================================== */

/* Kernels of the hot path, for per-kernel timing and byte counts. */
enum {
    DNDSMI_K_JACOBI = 0,     /* Jacobi sweep                                 */
    DNDSMI_K_WALL_CORR = 1,  /* Wall_Euler / Wall_NoSlip correction of a sweep */
    DNDSMI_K_DT_FACE = 2,    /* EvaluateDt, face part                         */
    DNDSMI_K_DT_CELL = 3,    /* EvaluateDt, cell part + block minima          */
    DNDSMI_K_RHS_FACE = 4,   /* EvaluateRHS, inner-face fluxes                */
    DNDSMI_K_RHS_CELL = 5,   /* EvaluateRHS gather (+ SSP-RK3 stage update)   */
    DNDSMI_K_WBAP_FACE = 6,  /* WBAP smoothness-indicator face terms          */
    DNDSMI_K_WBAP_CELL = 7,  /* WBAP indicator, copy of unlimited cells       */
    DNDSMI_K_WBAP_LIM = 8,   /* WBAP characteristic limiting of flagged cells */
    DNDSMI_K_LUSGS_SWEEP = 9,/* one colour of an LU-SGS forward/backward sweep */
    DNDSMI_K_DELTA32 = 10,   /* mixed precision: FP32 correction sweep       */
    DNDSMI_K_RHS_FACE32 = 11,/* mixed precision: inner-face fluxes, FP32 reconstruction */
    DNDSMI_K_RHS_CELL32 = 12,/* mixed precision: FP32 gather + stage update  */
    DNDSMI_K_LUSGS_MATVEC = 13,/* LUSGSMatrixVec (GMRES operator)            */
    DNDSMI_K_RHS_BFACE = 14, /* EvaluateRHS: boundary-face fluxes (FP64)     */
    DNDSMI_K_READ_PROBE = 15,/* read-bandwidth probe of the sweep's A^-1 B stream */
    DNDSMI_K_COUNT = 16
};

/* Turns HIP-event brackets around every launch on or off and clears the
 * records (synchronises the stream).  No counterpart in the reference, which
 * times with MPI_Wtime (DNDS_Profiling.h:7-37). */
int dndsmi_profile_enable(dndsmi_ctx *ctx, int32_t on);
/* Number of recorded launches of `kernel` and their summed duration (ms). */
int dndsmi_profile_query(dndsmi_ctx *ctx, int32_t kernel, int64_t *count, double *total_ms);
/* Algorithmic HBM bytes of one launch of `kernel` on this context's mesh
 * (DNDSMI_K_LUSGS_SWEEP: of one whole sweep, i.e. all its colour launches). */
int64_t dndsmi_kernel_bytes(const dndsmi_ctx *ctx, int32_t kernel);

/* No counterpart in the reference (measurement): reps launches of a kernel
 * that streams the sweep's A^-1 B pairs exactly as the Jacobi sweep does,
 * without its gathers (DNDSMI_K_READ_PROBE in the profile): the read
 * bandwidth the sweep's streaming part can reach on this device. */
int dndsmi_read_probe(dndsmi_ctx *ctx, int32_t reps);

#ifdef __cplusplus
}
#endif

#endif /* DNDSMI_H */
