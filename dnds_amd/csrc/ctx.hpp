// Device context shared by the HIP translation units (ctx.hip: mesh/setup
// upload, explicit path; lusgs.hip: LU-SGS and the implicit driver).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"

namespace dndsmi_host {

constexpr int kBlock = 256;

inline int64_t pad64(int64_t n) { return (n + 63) / 64 * 64; }
inline unsigned grid_for(int64_t n, int block = kBlock) { return unsigned((n + block - 1) / block); }

} // namespace dndsmi_host

using dndsmi_dev::GasParams;

struct dndsmi_ctx {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    std::string err;
    int64_t nc = 0, nf = 0, ldc = 0, ldf = 0;
    // device numbering of the mesh's cells and faces (host arrays keep the
    // mesh numbering)
    std::vector<int32_t> c2d, f2d;
    int nfmax = 3;
    GasParams gas{};
    double relax = 1.0;
    // SOR_Instead / SOR_InverseScanning: in-place sweeps over the colours of
    // the multicolour scan (col_cells), forward (then backward)
    bool sor = false, sor_inverse = false;
    // SOR_Instead (colour-major device numbering): colour k holds the device
    // cells [sor_start[k], sor_start[k + 1]), those with a far-field,
    // Wall_Euler or Wall_NoSlip face from sor_wall[k] on
    std::vector<int64_t> sor_start, sor_wall;
    int64_t bytes = 0;
    std::vector<void *> allocs;
    // live arrays and other objects bound to this context, detached when it
    // is destroyed first
    std::unordered_set<struct dndsmi_array *> arrays;
    std::unordered_set<struct dndsmi_warmstart *> warmstarts;
    std::unordered_set<struct dndsmi_bdf *> bdfs;

    // cell side
    int32_t *nbr = nullptr;    // [nfmax][ldc] neighbour cell, -1 = boundary slot, -2 = absent slot
    int32_t *sbc = nullptr;    // [nfmax][ldc] boundary kind of the slot's face
    int32_t *sface = nullptr;  // [nfmax][ldc] face index of the slot
    int32_t *gat = nullptr;    // [nfmax][ldc] (face << 1 | side) sorted by face, -1 = none
    double *M = nullptr;       // [nfmax][41][ldc][2] A^-1 B entry pairs
    double *bv = nullptr;      // [nfmax*9][ldc]
    double *vol = nullptr;     // [ldc]
    int32_t *wall_idx = nullptr; // [nfmax][ldc] index into the wall-face arrays, -1 = none
    int64_t n_wall = 0;
    double *wall_ainv = nullptr; // [81][n_wall]
    double *wall_d0 = nullptr;   // [27][n_wall]
    double *wall_nrm = nullptr;  // [6][n_wall]
    double *wall_w0 = nullptr;   // [n_wall]
    int32_t *wall_cell = nullptr; // [n_wall]
    int32_t *wall_bc = nullptr;   // [n_wall]
    double *wall_corr = nullptr;  // [36][n_wall]
    int64_t n_wall_faces = 0;     // real count (n_wall >= 1 for allocation)
    // face side
    int32_t *f2c = nullptr;    // [2][ldf]
    int32_t *fbc = nullptr;    // [ldf]
    double *fxy = nullptr;     // [6][ldf][2] (X, Y) of LINE_3 point g seen from side s, row 3 s + g
    double *cmom = nullptr;    // [9][ldc] base moments 1..9 of the cell
    double *cscale = nullptr;  // [ldc][2] (sx, sy) of the cell's scaled monomials
    double *fn = nullptr;      // [6][ldf] normals at the LINE_3 points
    double *fnc = nullptr;     // [2][ldf] normal at the face centre
    double *farea = nullptr;   // [ldf]
    double *fctr = nullptr;    // [2][ldf]
    int32_t *bfaces = nullptr; // boundary face indices
    int64_t n_bface = 0;
    // WBAP data (present when the setup carries limiter data)
    bool has_limiter = false;
    double *fw = nullptr;      // [10][ldf] face weights
    double *msec = nullptr;    // [162][ldf] matrixSecondary (R2L, L2R)
    double *wslot = nullptr;   // [nfmax][4][ldc] indicator terms of the cell's faces, cell2face slot order
    uint8_t *wpos = nullptr;   // [ldf] the face's cell2face slot in cell L (low 4 bits) and in cell R (high)
    double *uF = nullptr;      // [nfmax*28][ldc] WBAP per-cell face scratch
    double *ind = nullptr;     // [ldc] indicator when the caller passes none
    double *xlim = nullptr;    // [36][ldc] limited reconstruction inside SSP-RK3 steps (limited cells only)
    uint8_t *limf = nullptr;   // [ldc] 1 where the steps' limiter wrote the cell into xlim
    bool wbap_full_copy = true; // the last WBAP launch copied the unlimited cells (byte model)
    int32_t *lim_list = nullptr;  // [ldc] cells to limit
    int32_t *lim_count = nullptr; // [1]
    double tang_weight = 1.0;
    int twd_norm = 1;
    double wbap_scale = 1e-10, wbap_nstd = 10.0;
    int32_t limiter_procedure = 0; // the steps' limiter: 0 WBAP V2, 1 WBAP V3
    int32_t limiter_if_all = 0;    // the steps' limiter limits every cell (the solver's ifAll)
    // evaluator scratch (EulerEvaluator members lambdaFace, deltaLambdaFace)
    double *lambda_face = nullptr;
    double *dlambda_face = nullptr;
    // face fluxes integrated over the face, stored at the slot of each side's
    // cell: [nfmax][8][ldc], per slot F then the side's flux-balance term,
    // signed as the cell adds them (side 0: F, -FLfix; side 1: -F, FRfix)
    double *fslot = nullptr;
    uint8_t *fpos = nullptr;   // [ldf] slot of the face in cell L (low 4 bits) and in cell R (high 4 bits)
    bool straight_faces = false; // every face's normal is bitwise the same at its three LINE_3 points
    double *dt_part = nullptr; // [grid] partial minima
    double *dt_min = nullptr;  // [1]
    int32_t *errw = nullptr;   // [1] error bits
    double *xlast = nullptr;   // [ldc][4] SSP-RK3 / dual-time xLast
    double *dt_tmp = nullptr;  // [ldc]
    // multicolour LU-SGS: greedy colouring in cell order, cells of a colour
    // contiguous in col_cells (ascending), so that the colour-major scan
    // order is a valid sequential LU-SGS order
    int32_t n_colour = 0;
    std::vector<int64_t> col_start;   // [n_colour + 1]
    std::vector<int32_t> colour_host; // [nc]
    int32_t *colour = nullptr;        // [ldc]
    int32_t *col_cells = nullptr;     // [nc]
    double *jinv = nullptr;           // [ldc] 1 / LU-SGS diagonal, scan order
    double *jdiag = nullptr;          // [ldc] the LU-SGS diagonal, scan order (first use)
    // fsolve's linear solver: 0 LU-SGS, 1 GMRES, 2 LU-SGS then GMRES
    // (left-preconditioned by LU-SGS), with its subspace size and restarts
    int32_t gmres_code = 0, n_gmres_space = 10, n_gmres_iter = 2;
    double *gm_v = nullptr;           // [n_gmres_space + 3][ldc][4] Krylov vectors, A v, ML b (first use)
    double *gm_h = nullptr;           // [(n_gmres_space + 1) * n_gmres_space] Hessenberg matrix
    double *gm_part = nullptr;        // [n_gmres_space + 1][grid] dot partial sums
    double *gm_y = nullptr;           // [n_gmres_space] least-squares solution
    int32_t gm_alloc_space = 0;
    // per scan position p (col_cells order), built on first LU-SGS use:
    // per slot [s][ldc] the neighbour used by the forward / backward sweep
    // (-1 if none), the face (-1 if no slot), the unit normal out of the
    // cell and the face area; per position the volume
    int32_t *sc_jf = nullptr, *sc_jb = nullptr, *sc_face = nullptr;
    double *sc_n0 = nullptr, *sc_n1 = nullptr, *sc_area = nullptr, *sc_vol = nullptr;
    // mixed precision (allocated on first use): FP32 copy of A^-1 B, FP32
    // correction vectors, FP32 face fluxes
    float *M32 = nullptr;             // [nfmax][21][ldc][4]
    float *r32 = nullptr, *dA32 = nullptr, *dB32 = nullptr; // [36][ldc]
    float *fslot32 = nullptr;         // [nfmax][8][ldc]
    float *wcorr32 = nullptr;         // [36][n_wall] linear wall correction of a correction sweep
    // implicit dual-time scratch (allocated on first use)
    double *imp_rhs = nullptr;        // [ldc][4]
    double *imp_xinc = nullptr;       // [ldc][4]
    double *imp_dtau = nullptr;       // [ldc]
    double *imp_stage = nullptr;      // [3][ldc][4] SDIRK4 stage right-hand sides (first use)
    double *res_part = nullptr;       // [4][grid]
    double *res = nullptr;            // [4]
    // the learned warm start inside the mixed-precision step
    // (dndsmi_ssprk3_set_warm_start): the previous stage's cell means and
    // their change [ldc][4], the plain shadow run of a check window
    // ([ldc][4], [36][ldc]) and its reduction partials (all first use)
    struct dndsmi_warmstart *step_ws = nullptr;
    double *ws_uprev = nullptr, *ws_du = nullptr;
    double *ws_sh_u = nullptr, *ws_sh_x = nullptr, *ws_part = nullptr;
    bool ws_have_prev = false, ws_rejected = false, ws_first = true;
    int32_t ws_check_steps = 1, ws_recheck = 0, ws_win_left = 0, ws_since_check = 0;
    double ws_tol = 0.0, ws_last_dev = -1.0;
    int64_t ws_stats[3] = {0, 0, 0}; // stages corrected, windows passed, windows failed (fallbacks)
    double *probe_out = nullptr; // [ldc] output of the read-bandwidth probe (first use)
    // device address of this GPU's static scratch of the allocation-free
    // residual (set at context creation; never freed: it is not an allocation)
    double *noalloc_scratch = nullptr;
    // per-kernel HIP-event timing (dndsmi_profile_*)
    bool prof_on = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> prof_recs; // (kernel id, index of the start event)
    int prof_open = -1;
    size_t prof_open_ev = 0;

    template <class T>
    int alloc(T *&p, int64_t n) {
        size_t b = sizeof(T) * size_t(std::max<int64_t>(n, 1));
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) {
            err = std::string("hipMalloc: ") + hipGetErrorString(e);
            return DNDSMI_ERR_OUT_OF_MEMORY;
        }
        allocs.push_back(q);
        bytes += int64_t(b);
        p = static_cast<T *>(q);
        return DNDSMI_OK;
    }
    int hip(hipError_t e, const char *what) {
        if (e != hipSuccess) {
            err = std::string(what) + ": " + hipGetErrorString(e);
            return DNDSMI_ERR_HIP;
        }
        return DNDSMI_OK;
    }
    template <class T>
    int upload(T *dst, const std::vector<T> &src) {
        int rc = hip(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, stream),
                     "hipMemcpyAsync");
        if (rc)
            return rc;
        return hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    int launch_check(const char *what) { return hip(hipGetLastError(), what); }
};

struct dndsmi_array {
    dndsmi_ctx *ctx = nullptr;
    int32_t kind = 0;
    double *d = nullptr;
    int64_t elems = 0; // doubles allocated
};

#define DNDSMI_TRY(expr)                                                                                               \
    do {                                                                                                               \
        int rc_ = (expr);                                                                                              \
        if (rc_ != DNDSMI_OK)                                                                                          \
            return rc_;                                                                                                \
    } while (0)

namespace dndsmi_host {

// profiling brackets around one launch (no-ops unless profiling is on)
void prof_begin(dndsmi_ctx *c, int id);
void prof_end(dndsmi_ctx *c);
// reads and clears the device error word; a status for the first bit set
int check_errw(dndsmi_ctx *c);
// the learned warm start inside the mixed-precision step (warmstart.hip):
// before a stage's sweeps, the change of the means since the previous stage
// and, when `use`, x += the network's prediction on it; a check window's
// shadow copies, and its verdict at the end (restores the plain run's u and
// x when the corrected run failed or strayed beyond the tolerance)
int ws_stage_begin(dndsmi_ctx *c, const double *u, double *x, bool use);
int ws_window_begin(dndsmi_ctx *c, const double *u, const double *x);
int ws_window_end(dndsmi_ctx *c, double *u, double *x, bool failed);
// looks up this GPU's static scratch of dndsmi_evaluate_residual_noalloc
// (once per device; the device must be current) into c->noalloc_scratch
int noalloc_scratch_init(dndsmi_ctx *c);
bool same_ctx(const dndsmi_ctx *c, const dndsmi_array *a, int32_t kind);
// the explicit path's launches (ctx.hip)
int launch_jacobi(dndsmi_ctx *c, const double *u, const double *x, double *xn, double t);
// one SOR sweep (SOR_Instead) in place on x, colour by colour
int launch_sor(dndsmi_ctx *c, const double *u, double *x, double t);
// one reconstruction sweep of a step: Jacobi into xn, then the two swap; with
// SOR_Instead in place on x (xn untouched)
int step_sweep(dndsmi_ctx *c, const double *u, double *&x, double *&xn, double t);
int launch_dt(dndsmi_ctx *c, const double *u, double cfl, double max_dt, double *dt_out);
// xalt/limf (may be null): cells with limf set are read from xalt instead of x
int launch_rhs(dndsmi_ctx *c, const double *u, const double *x, double t, double *rhs, int stage, double *xs,
               const double *dt, double coefA, double coefB, const double *xalt = nullptr,
               const uint8_t *limf = nullptr);
// full_copy: xo receives every cell (x for those not limited); otherwise only
// the limited cells, flagged in c->limf
int launch_wbap(dndsmi_ctx *c, const double *u, const double *x, double *xo, double *ind, int if_all,
                int procedure, bool full_copy = true);
// the mixed-precision pieces (ctx.hip)
int ensure_mixed(dndsmi_ctx *c);
int launch_jacobi_mixed(dndsmi_ctx *c, const double *u, double *x, int n_sweeps, double t);
int launch_rhs_mixed(dndsmi_ctx *c, const double *u, const double *x, double t, double *rhs, int stage, double *xs,
                     const double *dt, double coefA, double coefB, const double *xalt = nullptr,
                     const uint8_t *limf = nullptr);

} // namespace dndsmi_host
