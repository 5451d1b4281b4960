// Multicolour LU-SGS and the implicit-Euler dual-time driver (gfx950).
//
// The reference's LU-SGS (DNDS_FV_EulerEvaluator.cpp:995-1357) is a pair of
// sequential sweeps over a scan order: forward uses the neighbours scanned
// earlier, backward those scanned later.  Here the scan order is
// colour-major: the cells are coloured so that no two neighbours share a
// colour, and a colour's cells are scanned together.  Within one colour no
// cell reads another of the same colour, so one launch per colour computes
// exactly the sequential sweep in that order: the same operations in the
// same order per cell as the oracle run with the same scan order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace dndsmi_dev;
using namespace dndsmi_host;

namespace {

// fluxJacobian0_Right_Times_du (DNDS_FV_EulerEvaluator.hpp:928-956) with
// IdealGasUIncrement and GasInviscidFluxFacialIncrement (DNDS_Gas.hpp:122-146):
// dF = d(F(U).n)/dU dU - lambda dU
__device__ inline void flux_jacobian_times_du(double gamma, const double U[4], double n0, double n1,
                                              const double dU[4], double lambdaMain, double dF[4]) {
    const double velo0 = U[1] / U[0], velo1 = U[2] / U[0];
    double p, asqr, H;
    ideal_gas_thermal(U[3], U[0], velo0 * velo0 + velo1 * velo1, gamma, p, asqr, H);
    const double dVelo0 = (dU[1] - U[1] / U[0] * dU[0]) / U[0];
    const double dVelo1 = (dU[2] - U[2] / U[0] * dU[0]) / U[0];
    const double dp = (gamma - 1) * (dU[3] - 0.5 * ((dU[1] * velo0 + dU[2] * velo1) + (U[1] * dVelo0 + U[2] * dVelo1)));
    const double vn = velo0 * n0 + velo1 * n1;
    const double dvn = dVelo0 * n0 + dVelo1 * n1;
    dF[0] = dU[1] * n0 + dU[2] * n1;
    dF[1] = dU[1] * vn + U[1] * dvn + n0 * dp;
    dF[2] = dU[2] * vn + U[2] * dvn + n1 * dp;
    dF[3] = (dU[3] + dp) * vn + (U[3] + p) * dvn;
#pragma unroll
    for (int k = 0; k < 4; k++)
        dF[k] -= lambdaMain * dU[k];
}

// Scan-order geometry of the sweeps, one thread per scan position: which
// neighbours each sweep uses (a lower colour is earlier, a higher one later),
// the face, its unit normal pointed out of the cell, its area, the volume.
template <int NF>
__global__ __launch_bounds__(kBlock) void k_lusgs_scan(int32_t nc, int64_t ldc, int64_t ldf,
                                                       const int32_t *__restrict__ cells,
                                                       const int32_t *__restrict__ nbr,
                                                       const int32_t *__restrict__ sface,
                                                       const int32_t *__restrict__ f2c,
                                                       const int32_t *__restrict__ colour,
                                                       const double *__restrict__ fnc,
                                                       const double *__restrict__ farea,
                                                       const double *__restrict__ vol, int32_t *__restrict__ jf,
                                                       int32_t *__restrict__ jb, int32_t *__restrict__ sf,
                                                       double *__restrict__ sn0, double *__restrict__ sn1,
                                                       double *__restrict__ sS, double *__restrict__ sV) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nc)
        return;
    const int32_t i = cells[p], col = colour[i];
    sV[p] = vol[i];
#pragma unroll
    for (int s = 0; s < NF; s++) {
        const int32_t j = nbr[s * ldc + i], f = sface[s * ldc + i];
        const int32_t cj = j >= 0 ? colour[j] : col;
        jf[s * ldc + p] = j >= 0 && cj < col ? j : -1;
        jb[s * ldc + p] = j >= 0 && cj > col ? j : -1;
        sf[s * ldc + p] = f;
        double n0 = 0.0, n1 = 0.0, S = 0.0;
        if (f >= 0) {
            normalized(fnc[f], fnc[ldf + f], n0, n1);
            if (f2c[f] != i) {
                n0 *= -1.0;
                n1 *= -1.0;
            }
            S = farea[f];
        }
        sn0[s * ldc + p] = n0, sn1[s * ldc + p] = n1, sS[s * ldc + p] = S;
    }
}

// LUSGSMatrixInit (DNDS_FV_EulerEvaluator.cpp:995-1047), jacobianCode 0, zero
// source Jacobian, per scan position:
// jinv = 1 / (1/dTau_i + 1/dt + sum_f (alpha/2) S_f lambda_f / V_i)
template <int NF>
__global__ __launch_bounds__(kBlock) void k_lusgs_init(int32_t nc, int64_t ldc, const int32_t *__restrict__ cells,
                                                       const int32_t *__restrict__ sf, const double *__restrict__ sS,
                                                       const double *__restrict__ sV,
                                                       const double *__restrict__ lambda_face,
                                                       const double *__restrict__ dtau, double dt, double alpha,
                                                       double *__restrict__ jinv, double *__restrict__ jdiag) {
    const int32_t p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= nc)
        return;
    double fp = 1.0 / dtau[cells[p]] + 1.0 / dt;
    const double v = sV[p];
#pragma unroll
    for (int s = 0; s < NF; s++) {
        const int32_t f = sf[s * ldc + p];
        if (f < 0)
            continue;
        fp += (0.5 * alpha) * sS[s * ldc + p] * lambda_face[f] / v;
    }
    jdiag[p] = fp + alpha * 0.0; // jacobianCell = fpDivisor I + alphaDiag 0 (source Jacobian)
    jinv[p] = 1.0 / (fp + alpha * 0.0);
}

// LUSGSMatrixVec (DNDS_FV_EulerEvaluator.cpp:1062-1145), per scan position:
// Ax_i = D_i x_i - sum_f (alpha/2) S_f / V_i dF(u_j, x_j) over every
// neighbour (the earlier ones and the later ones), faces in cell2face order
template <int NF>
__global__ __launch_bounds__(kBlock) void k_lusgs_matvec(int32_t nc, int64_t ldc, const int32_t *__restrict__ cells,
                                                         const int32_t *__restrict__ jf,
                                                         const int32_t *__restrict__ jb,
                                                         const int32_t *__restrict__ sf,
                                                         const double *__restrict__ sn0,
                                                         const double *__restrict__ sn1,
                                                         const double *__restrict__ sS,
                                                         const double *__restrict__ sV,
                                                         const double *__restrict__ lambda_face,
                                                         const double *__restrict__ jdiag, double gamma, double alpha,
                                                         const double *__restrict__ u, const double *__restrict__ x,
                                                         double *__restrict__ Ax) {
    const int32_t p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= nc)
        return;
    const int32_t i = cells[p];
    const double vi = sV[p];
    double buf[4] = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < NF; s++) {
        const int32_t a = jf[s * ldc + p], b = jb[s * ldc + p];
        const int32_t j = a >= 0 ? a : b;
        if (j < 0)
            continue;
        const int32_t f = sf[s * ldc + p];
        double uj[4], xj[4], dF[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
            uj[v] = u[4 * int64_t(j) + v], xj[v] = x[4 * int64_t(j) + v];
        flux_jacobian_times_du(gamma, uj, sn0[s * ldc + p], sn1[s * ldc + p], xj, lambda_face[f], dF);
        const double cf = (0.5 * alpha) * sS[s * ldc + p] / vi;
#pragma unroll
        for (int v = 0; v < 4; v++)
            buf[v] -= cf * dF[v];
    }
    const double d = jdiag[p];
#pragma unroll
    for (int v = 0; v < 4; v++)
        Ax[4 * int64_t(i) + v] = d * x[4 * int64_t(i) + v] - buf[v];
}

// ------------------------------------------------------- GMRES vectors ----
// Per-block partial sums of nq dot products w . V_q (V_q = V + q ldv), each
// cell's 4 products summed in order, then a fixed tree over the block.
constexpr int kMaxDots = 33;
__global__ __launch_bounds__(kBlock) void k_gm_dots(int32_t nc, int32_t nq, const double *__restrict__ w,
                                                    const double *__restrict__ V, int64_t ldv,
                                                    double *__restrict__ part) {
    __shared__ double red[kBlock];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double wi[4] = {0, 0, 0, 0};
    if (i < nc)
#pragma unroll
        for (int v = 0; v < 4; v++)
            wi[v] = w[4 * int64_t(i) + v];
    for (int32_t q = 0; q < nq; q++) {
        double a = 0.0;
        if (i < nc)
#pragma unroll
            for (int v = 0; v < 4; v++)
                a += wi[v] * V[q * ldv + 4 * int64_t(i) + v];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int st = kBlock / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st)
                red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            part[int64_t(q) * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// folds the partials of each dot product in a fixed order; out[q * ostride]
// = the sum, or its square root (norms)
__global__ __launch_bounds__(64) void k_gm_dots_final(int32_t nblk, int32_t nq, const double *__restrict__ part,
                                                      int32_t take_sqrt, double *__restrict__ out, int32_t ostride) {
    const int lane = threadIdx.x;
    for (int32_t q = 0; q < nq; q++) {
        double a = 0.0;
        for (int32_t b = lane; b < nblk; b += 64)
            a += part[int64_t(q) * nblk + b];
        for (int off = 32; off > 0; off >>= 1)
            a += __shfl_down(a, off, 64);
        if (lane == 0)
            out[q * ostride] = take_sqrt ? sqrt(a) : a;
    }
}

// y += x * a (ArrayDOFV::addTo, DNDS_DerivedArray.hpp:372-377)
__global__ __launch_bounds__(kBlock) void k_gm_axpy(int64_t n, double *__restrict__ y, const double *__restrict__ x,
                                                    double a) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n)
        y[k] += x[k] * a;
}

// v *= a, with a given, or a = 1 / h[0] read from the device
__global__ __launch_bounds__(kBlock) void k_gm_scale(int64_t n, double *__restrict__ v, double a,
                                                     const double *__restrict__ h) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const double s = h ? 1.0 / h[0] : a;
    v[k] *= s;
}

// classical Gram-Schmidt subtraction: v += V_i * (-h_i) for i = 0 .. nq-1 in
// turn, h_i = h[i * hstride] on the device; or, with sign +1, x += V_j y_j
__global__ __launch_bounds__(kBlock) void k_gm_combine(int64_t n, int32_t nq, double *__restrict__ v,
                                                       const double *__restrict__ V, int64_t ldv,
                                                       const double *__restrict__ h, int32_t hstride, int32_t negate) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    double a = v[k];
    for (int32_t q = 0; q < nq; q++) {
        const double c = h[q * hstride];
        a += V[q * ldv + k] * (negate ? -c : c);
    }
    v[k] = a;
}

// One colour of an LU-SGS sweep (UpdateLUSGSForward / UpdateLUSGSBackward,
// DNDS_FV_EulerEvaluator.cpp:1149-1357), scan positions [b, b + n): forward
//   uIncNew_i = jinv_i (rhs_i - sum_{earlier j} c_f dF(u_j, uInc_j))
// backward
//   uIncNew_i += jinv_i (- sum_{later j} c_f dF(u_j, uInc_j)),
// c_f = (alpha/2) S_f / V_i, the normal pointing out of i, faces in cell2face
// order.  "Earlier" is a lower colour, "later" a higher one.
template <int NF, bool FORWARD>
__global__ __launch_bounds__(kBlock) void k_lusgs_sweep(int32_t n, int32_t b, int64_t ldc,
                                                        const int32_t *__restrict__ cells,
                                                        const int32_t *__restrict__ sj,
                                                        const int32_t *__restrict__ sf,
                                                        const double *__restrict__ sn0,
                                                        const double *__restrict__ sn1,
                                                        const double *__restrict__ sS,
                                                        const double *__restrict__ sV,
                                                        const double *__restrict__ lambda_face,
                                                        const double *__restrict__ jinv, double gamma, double alpha,
                                                        const double *__restrict__ rhs, const double *__restrict__ u,
                                                        const double *uInc, double *uIncNew, int32_t *errw) {
    const int32_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int32_t p = b + t, i = cells[p];
    double buf[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        buf[v] = FORWARD ? rhs[4 * int64_t(i) + v] : 0.0;
    const double vi = sV[p];
#pragma unroll
    for (int s = 0; s < NF; s++) {
        const int32_t j = sj[s * ldc + p];
        if (j < 0)
            continue;
        const int32_t f = sf[s * ldc + p];
        const double n0 = sn0[s * ldc + p], n1 = sn1[s * ldc + p];
        double uj[4], dj[4], dF[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
            uj[v] = u[4 * int64_t(j) + v], dj[v] = uInc[4 * int64_t(j) + v];
        flux_jacobian_times_du(gamma, uj, n0, n1, dj, lambda_face[f], dF);
        const double cf = (0.5 * alpha) * sS[s * ldc + p] / vi;
#pragma unroll
        for (int v = 0; v < 4; v++)
            buf[v] -= cf * dF[v];
    }
    const double d = jinv[p];
    bool finite = true;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        double o = d * buf[v];
        if (!FORWARD)
            o = uIncNew[4 * int64_t(i) + v] + o;
        finite &= isfinite(o);
        uIncNew[4 * int64_t(i) + v] = o;
    }
    if (!finite)
        atomicOr(errw, ERRBIT_NAN);
}

// Dual-time right-hand side of ImplicitEulerDualTimeStep::Step
// (DNDS_ODE.hpp:253-258) and fsolve's compression (DNDS_FV_EulerSolver.hpp:870-873):
//   rhs = ((xLast - x) * (1/dt) + rhs); rhs = (rhs * dTau) / dTau; xinc = 0
__global__ __launch_bounds__(kBlock) void k_dual_rhs(int32_t nc, const double *__restrict__ xlast,
                                                     const double *__restrict__ x, double rdt,
                                                     const double *__restrict__ dtau, double *__restrict__ rhs,
                                                     double *__restrict__ xinc) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    const double dti = dtau[i];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int64_t k = 4 * int64_t(i) + v;
        double r = xlast[k];
        r -= x[k];
        r *= rdt;
        r = r + rhs[k];
        rhs[k] = (r * dti) / dti;
        xinc[k] = 0.0;
    }
}

// EvaluateResidual partial sums (DNDS_FV_EulerEvaluator.cpp:1557-1592) of
// one block, optionally after x += xinc: P = 1 sums |r|, P = 2 sums r^2,
// P >= 3 takes max |r|; volWise multiplies by the volume (P < 3).
__global__ __launch_bounds__(kBlock) void k_residual(int32_t nc, const double *__restrict__ r,
                                                     double *__restrict__ x, const double *__restrict__ vol,
                                                     int32_t P, int32_t vol_wise, double *__restrict__ part) {
    __shared__ double red[4][kBlock];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (i < nc) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double rv = r[4 * int64_t(i) + v];
            if (x)
                x[4 * int64_t(i) + v] += rv;
            const double a = fabs(rv);
            double q = P >= 3 ? a : (P == 2 ? a * a : a);
            if (P < 3 && vol_wise)
                q *= vol[i];
            acc[v] = q;
        }
    }
#pragma unroll
    for (int v = 0; v < 4; v++)
        red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
#pragma unroll
            for (int v = 0; v < 4; v++)
                red[v][threadIdx.x] = P >= 3 ? fmax(red[v][threadIdx.x], red[v][threadIdx.x + st])
                                             : red[v][threadIdx.x] + red[v][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 4)
        part[threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

// one wavefront per variable folds the block partials in a fixed order
__global__ __launch_bounds__(kBlock) void k_residual_final(int32_t nblk, int32_t P, const double *__restrict__ part,
                                                           double *__restrict__ res) {
    const int v = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (v >= 4)
        return;
    double a = 0.0;
    for (int32_t b = lane; b < nblk; b += 64)
        a = P >= 3 ? fmax(a, part[v * nblk + b]) : a + part[v * nblk + b];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(a, off, 64);
        a = P >= 3 ? fmax(a, o) : a + o;
    }
    if (lane == 0)
        res[v] = a;
}

// fdtau's division by alphaDiag (DNDS_FV_EulerSolver.hpp:846-847)
__global__ __launch_bounds__(kBlock) void k_div_scalar(int32_t n, double *__restrict__ p, double a) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] /= a;
}

// BDF dual-time right-hand side (DNDS_ODE.hpp:459-471), in the reference's
// order: rhs *= alpha0; rhs += x * (-1/dt); rhs += xLast * (c1/dt);
// rhs += xPrev_j * (c_j/dt) for the previous states in the history's order;
// then fsolve's compression rhs = (rhs dTau) / dTau and xinc = 0.
__global__ __launch_bounds__(kBlock) void k_bdf_rhs(int32_t nc, double alpha0, const double *__restrict__ x,
                                                    double mdt, const double *__restrict__ xlast, double c1,
                                                    int32_t nprev, const double *__restrict__ p0, double c2,
                                                    const double *__restrict__ p1, double c3,
                                                    const double *__restrict__ p2, double c4,
                                                    const double *__restrict__ dtau, double *__restrict__ rhs,
                                                    double *__restrict__ xinc) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    const double dti = dtau[i];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int64_t k = 4 * int64_t(i) + v;
        double r = rhs[k];
        r *= alpha0;
        r += x[k] * mdt;
        r += xlast[k] * c1;
        if (nprev > 0)
            r += p0[k] * c2;
        if (nprev > 1)
            r += p1[k] * c3;
        if (nprev > 2)
            r += p2[k] * c4;
        rhs[k] = (r * dti) / dti;
        xinc[k] = 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_scalar(int32_t n, const double *__restrict__ src,
                                                        double *__restrict__ dst) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[0];
}

int ensure_implicit_scratch(dndsmi_ctx *c) {
    if (c->imp_rhs)
        return DNDSMI_OK;
    DNDSMI_TRY(c->alloc(c->imp_rhs, 4 * c->ldc));
    DNDSMI_TRY(c->alloc(c->imp_xinc, 4 * c->ldc));
    DNDSMI_TRY(c->alloc(c->imp_dtau, c->ldc));
    DNDSMI_TRY(c->alloc(c->res_part, 4 * int64_t(grid_for(c->nc))));
    DNDSMI_TRY(c->alloc(c->res, 4));
    return DNDSMI_OK;
}

int ensure_lusgs_scan(dndsmi_ctx *c) {
    if (c->sc_jf)
        return DNDSMI_OK;
    const int64_t n = int64_t(c->nfmax) * c->ldc;
    DNDSMI_TRY(c->alloc(c->sc_jf, n));
    DNDSMI_TRY(c->alloc(c->sc_jb, n));
    DNDSMI_TRY(c->alloc(c->sc_face, n));
    DNDSMI_TRY(c->alloc(c->sc_n0, n));
    DNDSMI_TRY(c->alloc(c->sc_n1, n));
    DNDSMI_TRY(c->alloc(c->sc_area, n));
    DNDSMI_TRY(c->alloc(c->sc_vol, c->ldc));
    const int nc = int(c->nc);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_lusgs_scan<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->ldf,
                           c->col_cells, c->nbr, c->sface, c->f2c, c->colour, c->fnc, c->farea, c->vol, c->sc_jf,
                           c->sc_jb, c->sc_face, c->sc_n0, c->sc_n1, c->sc_area, c->sc_vol);
    else
        hipLaunchKernelGGL(k_lusgs_scan<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->ldf,
                           c->col_cells, c->nbr, c->sface, c->f2c, c->colour, c->fnc, c->farea, c->vol, c->sc_jf,
                           c->sc_jb, c->sc_face, c->sc_n0, c->sc_n1, c->sc_area, c->sc_vol);
    return c->launch_check("k_lusgs_scan");
}

int launch_lusgs_init(dndsmi_ctx *c, const double *dtau, double dt, double alpha) {
    DNDSMI_TRY(ensure_lusgs_scan(c));
    if (!c->jdiag)
        DNDSMI_TRY(c->alloc(c->jdiag, c->ldc));
    const int nc = int(c->nc);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_lusgs_init<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->col_cells,
                           c->sc_face, c->sc_area, c->sc_vol, c->lambda_face, dtau, dt, alpha, c->jinv, c->jdiag);
    else
        hipLaunchKernelGGL(k_lusgs_init<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->col_cells,
                           c->sc_face, c->sc_area, c->sc_vol, c->lambda_face, dtau, dt, alpha, c->jinv, c->jdiag);
    return c->launch_check("k_lusgs_init");
}

template <bool FORWARD>
int launch_lusgs_sweep(dndsmi_ctx *c, double alpha, const double *rhs, const double *u, const double *uInc,
                       double *uIncNew) {
    DNDSMI_TRY(ensure_lusgs_scan(c));
    const int32_t *sj = FORWARD ? c->sc_jf : c->sc_jb;
    for (int32_t kk = 0; kk < c->n_colour; kk++) {
        const int32_t col = FORWARD ? kk : c->n_colour - 1 - kk;
        const int64_t b = c->col_start[col], n = c->col_start[col + 1] - b;
        if (n == 0)
            continue;
        prof_begin(c, DNDSMI_K_LUSGS_SWEEP);
        if (c->nfmax == 3)
            hipLaunchKernelGGL((k_lusgs_sweep<3, FORWARD>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                               int32_t(n), int32_t(b), c->ldc, c->col_cells, sj, c->sc_face, c->sc_n0, c->sc_n1,
                               c->sc_area, c->sc_vol, c->lambda_face, c->jinv, c->gas.gamma, alpha, rhs, u, uInc,
                               uIncNew, c->errw);
        else
            hipLaunchKernelGGL((k_lusgs_sweep<4, FORWARD>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                               int32_t(n), int32_t(b), c->ldc, c->col_cells, sj, c->sc_face, c->sc_n0, c->sc_n1,
                               c->sc_area, c->sc_vol, c->lambda_face, c->jinv, c->gas.gamma, alpha, rhs, u, uInc,
                               uIncNew, c->errw);
        prof_end(c);
        DNDSMI_TRY(c->launch_check("k_lusgs_sweep"));
    }
    return DNDSMI_OK;
}

int launch_lusgs_matvec(dndsmi_ctx *c, double alpha, const double *u, const double *x, double *Ax) {
    DNDSMI_TRY(ensure_lusgs_scan(c));
    const int nc = int(c->nc);
    prof_begin(c, DNDSMI_K_LUSGS_MATVEC);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_lusgs_matvec<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->col_cells,
                           c->sc_jf, c->sc_jb, c->sc_face, c->sc_n0, c->sc_n1, c->sc_area, c->sc_vol, c->lambda_face,
                           c->jdiag, c->gas.gamma, alpha, u, x, Ax);
    else
        hipLaunchKernelGGL(k_lusgs_matvec<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->col_cells,
                           c->sc_jf, c->sc_jb, c->sc_face, c->sc_n0, c->sc_n1, c->sc_area, c->sc_vol, c->lambda_face,
                           c->jdiag, c->gas.gamma, alpha, u, x, Ax);
    prof_end(c);
    return c->launch_check("k_lusgs_matvec");
}

// least squares min |H y - beta e1| for the (k+1) x k Hessenberg H
// (row-major) by Householder reflections
std::vector<double> hessenberg_lsq_householder(std::vector<double> H, int k, double beta) {
    const int m = k + 1;
    std::vector<double> g(static_cast<size_t>(m), 0.0), w(static_cast<size_t>(m), 0.0);
    g[0] = beta;
    for (int j = 0; j < k; j++) {
        double nrm = 0.0;
        for (int r = j; r < m; r++)
            nrm += H[size_t(r) * k + j] * H[size_t(r) * k + j];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0)
            continue;
        const double a = H[size_t(j) * k + j], alpha = a > 0 ? -nrm : nrm;
        for (int r = 0; r < m; r++)
            w[size_t(r)] = r < j ? 0.0 : H[size_t(r) * k + j];
        w[size_t(j)] -= alpha;
        double ww = 0.0;
        for (int r = j; r < m; r++)
            ww += w[size_t(r)] * w[size_t(r)];
        if (ww == 0.0)
            continue;
        for (int col = j; col < k; col++) {
            double d = 0.0;
            for (int r = j; r < m; r++)
                d += w[size_t(r)] * H[size_t(r) * k + col];
            const double f = 2.0 * d / ww;
            for (int r = j; r < m; r++)
                H[size_t(r) * k + col] -= f * w[size_t(r)];
        }
        double d = 0.0;
        for (int r = j; r < m; r++)
            d += w[size_t(r)] * g[size_t(r)];
        const double f = 2.0 * d / ww;
        for (int r = j; r < m; r++)
            g[size_t(r)] -= f * w[size_t(r)];
    }
    std::vector<double> y(size_t(k), 0.0);
    for (int j = k - 1; j >= 0; j--) {
        double acc = g[size_t(j)];
        for (int col = j + 1; col < k; col++)
            acc -= H[size_t(j) * k + col] * y[size_t(col)];
        y[size_t(j)] = acc / H[size_t(j) * k + j];
    }
    return y;
}

int ensure_gmres(dndsmi_ctx *c) {
    const int k = c->n_gmres_space;
    if (c->gm_v && c->gm_alloc_space >= k)
        return DNDSMI_OK;
    auto release = [&](double *&p) {
        if (p) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(p);
            c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), static_cast<void *>(p)), c->allocs.end());
            p = nullptr;
        }
    };
    release(c->gm_v), release(c->gm_h), release(c->gm_part), release(c->gm_y);
    DNDSMI_TRY(c->alloc(c->gm_v, int64_t(k + 3) * 4 * c->ldc));
    DNDSMI_TRY(c->alloc(c->gm_h, int64_t(k + 1) * k));
    DNDSMI_TRY(c->alloc(c->gm_part, int64_t(k + 1) * grid_for(c->nc)));
    DNDSMI_TRY(c->alloc(c->gm_y, k));
    c->gm_alloc_space = k;
    return DNDSMI_OK;
}

// GMRES_LeftPreconditioned::solve (DNDS_Linear.hpp:48-104) with FA =
// LUSGSMatrixVec, FML = LU-SGS forward + backward, FStop never true (the
// solver's, DNDS_FV_EulerSolver.hpp:959-969); x is updated in place.
int gmres_solve(dndsmi_ctx *c, double alpha, const double *b, const double *u, double *x) {
    DNDSMI_TRY(ensure_gmres(c));
    const int k = c->n_gmres_space;
    const int64_t ldv = 4 * c->ldc, n = 4 * c->nc;
    const int nc = int(c->nc);
    const unsigned gn = grid_for(n), gc = grid_for(nc);
    double *V = c->gm_v, *Vt = V + int64_t(k + 1) * ldv, *MLb = V + int64_t(k + 2) * ldv;
    auto FML = [&](const double *in, double *out) -> int {
        DNDSMI_TRY(launch_lusgs_sweep<true>(c, alpha, in, u, out, out));
        return launch_lusgs_sweep<false>(c, alpha, in, u, out, out);
    };
    auto dots = [&](const double *w, const double *Vq, int nq, int take_sqrt, double *out, int ostride) -> int {
        hipLaunchKernelGGL(k_gm_dots, dim3(gc), dim3(kBlock), 0, c->stream, nc, nq, w, Vq, ldv, c->gm_part);
        hipLaunchKernelGGL(k_gm_dots_final, dim3(1), dim3(64), 0, c->stream, int32_t(gc), nq, c->gm_part, take_sqrt,
                           out, ostride);
        return c->launch_check("GMRES dot products");
    };
    DNDSMI_TRY(FML(b, MLb));
    std::vector<double> H(size_t(k + 1) * size_t(k));
    for (int32_t iRestart = 0; iRestart <= c->n_gmres_iter; iRestart++) {
        DNDSMI_TRY(launch_lusgs_matvec(c, alpha, u, x, Vt));
        DNDSMI_TRY(FML(Vt, V));
        hipLaunchKernelGGL(k_gm_axpy, dim3(gn), dim3(kBlock), 0, c->stream, n, V, MLb, -1.0);
        DNDSMI_TRY(dots(V, V, 1, 1, c->gm_y, 1));
        double beta = 0.0;
        DNDSMI_TRY(c->hip(hipMemcpyAsync(&beta, c->gm_y, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy"));
        DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
        if (std::fabs(beta) < 1e-200 || iRestart == c->n_gmres_iter)
            break;
        hipLaunchKernelGGL(k_gm_scale, dim3(gn), dim3(kBlock), 0, c->stream, n, V, -1.0 / beta, nullptr);
        for (int j = 0; j < k; j++) {
            double *Vj = V + int64_t(j) * ldv, *Vn = Vj + ldv;
            DNDSMI_TRY(launch_lusgs_matvec(c, alpha, u, Vj, Vt));
            DNDSMI_TRY(FML(Vt, Vn));
            DNDSMI_TRY(dots(Vn, V, j + 1, 0, c->gm_h + j, k)); // h(i, j), i <= j
            hipLaunchKernelGGL(k_gm_combine, dim3(gn), dim3(kBlock), 0, c->stream, n, j + 1, Vn, V, ldv, c->gm_h + j,
                               k, 1);
            DNDSMI_TRY(dots(Vn, Vn, 1, 1, c->gm_h + int64_t(j + 1) * k + j, 1)); // h(j + 1, j)
            hipLaunchKernelGGL(k_gm_scale, dim3(gn), dim3(kBlock), 0, c->stream, n, Vn, 0.0,
                               c->gm_h + int64_t(j + 1) * k + j);
            DNDSMI_TRY(c->launch_check("GMRES Arnoldi"));
        }
        DNDSMI_TRY(c->hip(hipMemcpyAsync(H.data(), c->gm_h, sizeof(double) * H.size(), hipMemcpyDeviceToHost,
                                         c->stream),
                          "copy"));
        DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
        for (int i = 0; i <= k; i++) // below the subdiagonal: zero (never written on the device)
            for (int j = 0; j < i - 1; j++)
                H[size_t(i) * k + j] = 0.0;
        const std::vector<double> y = hessenberg_lsq_householder(H, k, beta);
        DNDSMI_TRY(c->hip(hipMemcpyAsync(c->gm_y, y.data(), sizeof(double) * size_t(k), hipMemcpyHostToDevice,
                                         c->stream),
                          "copy"));
        hipLaunchKernelGGL(k_gm_combine, dim3(gn), dim3(kBlock), 0, c->stream, n, k, x, V, ldv, c->gm_y, 1, 0);
        DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize")); // y lives on the host stack
        DNDSMI_TRY(c->launch_check("GMRES update"));
    }
    return DNDSMI_OK;
}

// the solver's fsolve after the right-hand side and its compression (the
// drivers' rhs kernels also zero xinc): LUSGSMatrixInit, then LU-SGS and / or
// GMRES as c->gmres_code says (DNDS_FV_EulerSolver.hpp:852-975)
int fsolve(dndsmi_ctx *c, const double *dtau, double dt, double alpha, const double *rhs, const double *u,
           double *xinc) {
    DNDSMI_TRY(launch_lusgs_init(c, dtau, dt, alpha));
    if (c->gmres_code == 0 || c->gmres_code == 2) {
        DNDSMI_TRY(launch_lusgs_sweep<true>(c, alpha, rhs, u, xinc, xinc));
        DNDSMI_TRY(launch_lusgs_sweep<false>(c, alpha, rhs, u, xinc, xinc));
    }
    if (c->gmres_code != 0)
        DNDSMI_TRY(gmres_solve(c, alpha, rhs, u, xinc));
    return DNDSMI_OK;
}

// residual into c->res (device) and then into host res[4]; x += r first when x
// part: [4][grid_for(nc)] block partials, res_dev: [4]; both written before read
int residual(dndsmi_ctx *c, const double *r, double *x, int32_t P, int32_t vol_wise, double *res_host, double *part,
             double *res_dev) {
    const unsigned g = grid_for(c->nc);
    hipLaunchKernelGGL(k_residual, dim3(g), dim3(kBlock), 0, c->stream, int32_t(c->nc), r, x, c->vol, P, vol_wise,
                       part);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kBlock), 0, c->stream, int32_t(g), P, part, res_dev);
    DNDSMI_TRY(c->launch_check("residual kernels"));
    double h[4];
    DNDSMI_TRY(c->hip(hipMemcpyAsync(h, res_dev, sizeof(h), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync"));
    DNDSMI_TRY(check_errw(c));
    for (int v = 0; v < 4; v++)
        res_host[v] = P >= 3 ? h[v] : (P == 2 ? std::sqrt(h[v]) : h[v]);
    return DNDSMI_OK;
}

} // namespace

extern "C" int dndsmi_lusgs_colouring(const dndsmi_ctx *c, int32_t *colour, int32_t *n_colours) {
    if (!c || !n_colours)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    *n_colours = c->n_colour;
    if (colour)
        std::copy(c->colour_host.begin(), c->colour_host.end(), colour);
    return DNDSMI_OK;
}

extern "C" int dndsmi_lusgs_matrix_init(dndsmi_ctx *c, const dndsmi_array *dTau, double dt, double alphaDiag,
                                        const dndsmi_array *u, const dndsmi_array *uRec, int32_t jacobianCode,
                                        double t) {
    (void)t; // the NS_2D diagonal has no time-dependent source term
    if (!c || !same_ctx(c, dTau, DNDSMI_ARRAY_SCALAR) || (u && !same_ctx(c, u, DNDSMI_ARRAY_DOF)) ||
        (uRec && !same_ctx(c, uRec, DNDSMI_ARRAY_REC)))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (jacobianCode != 0) {
        c->err = "LUSGSMatrixInit: only jacobianCode 0 (the analytic diagonal) is implemented";
        return DNDSMI_ERR_UNSUPPORTED;
    }
    return launch_lusgs_init(c, dTau->d, dt, alphaDiag);
}

extern "C" int dndsmi_update_lusgs_forward(dndsmi_ctx *c, double alphaDiag, const dndsmi_array *rhs,
                                           const dndsmi_array *u, const dndsmi_array *uInc, dndsmi_array *uIncNew) {
    if (!c || !same_ctx(c, rhs, DNDSMI_ARRAY_DOF) || !same_ctx(c, u, DNDSMI_ARRAY_DOF) ||
        !same_ctx(c, uInc, DNDSMI_ARRAY_DOF) || !same_ctx(c, uIncNew, DNDSMI_ARRAY_DOF) || uIncNew == u ||
        uIncNew == rhs)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    return launch_lusgs_sweep<true>(c, alphaDiag, rhs->d, u->d, uInc->d, uIncNew->d);
}

extern "C" int dndsmi_update_lusgs_backward(dndsmi_ctx *c, double alphaDiag, const dndsmi_array *rhs,
                                            const dndsmi_array *u, const dndsmi_array *uInc, dndsmi_array *uIncNew) {
    if (!c || !same_ctx(c, rhs, DNDSMI_ARRAY_DOF) || !same_ctx(c, u, DNDSMI_ARRAY_DOF) ||
        !same_ctx(c, uInc, DNDSMI_ARRAY_DOF) || !same_ctx(c, uIncNew, DNDSMI_ARRAY_DOF) || uIncNew == u ||
        uIncNew == rhs)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    return launch_lusgs_sweep<false>(c, alphaDiag, rhs->d, u->d, uInc->d, uIncNew->d);
}

extern "C" int dndsmi_lusgs_matrix_vec(dndsmi_ctx *c, double alphaDiag, const dndsmi_array *u,
                                       const dndsmi_array *uInc, dndsmi_array *AuInc) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uInc, DNDSMI_ARRAY_DOF) ||
        !same_ctx(c, AuInc, DNDSMI_ARRAY_DOF) || AuInc == uInc || AuInc == u)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (!c->jdiag) {
        c->err = "LUSGSMatrixVec needs a LUSGSMatrixInit first";
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    return launch_lusgs_matvec(c, alphaDiag, u->d, uInc->d, AuInc->d);
}

extern "C" int dndsmi_set_linear_solver(dndsmi_ctx *c, int32_t gmres_code, int32_t n_space, int32_t n_iter) {
    if (!c || gmres_code < 0 || gmres_code > 2 || n_space < 1 || n_space > kMaxDots - 1 || n_iter < 0)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    c->gmres_code = gmres_code, c->n_gmres_space = n_space, c->n_gmres_iter = n_iter;
    return DNDSMI_OK;
}

extern "C" int dndsmi_evaluate_residual(dndsmi_ctx *c, const dndsmi_array *rhs, int32_t P, int32_t volWise,
                                        double *res) {
    if (!c || !same_ctx(c, rhs, DNDSMI_ARRAY_DOF) || !res || P < 1)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(ensure_implicit_scratch(c));
    return residual(c, rhs->d, nullptr, P, volWise, res, c->res_part, c->res);
}

// ---- the allocation-free residual ---------------------------------------
//
// dndsmi_evaluate_residual_noalloc computes what dndsmi_evaluate_residual
// does, with the same kernels, but keeps the block partials and the four
// results in `scratch`, a static array of the library on each GPU, instead
// of the context's lazily allocated buffers.  Host code must not pass
// `scratch` itself to a kernel (on the host the name is only a stand-in);
// its device address is looked up once per device and kept in
// g_scratch_dev.  Every k_residual block writes its partials before
// k_residual_final reads them, so nothing needs clearing between calls.
// All calls on one GPU share the array: calls that overlap in time on one
// GPU (from other streams or host threads) race.
constexpr int64_t kNoAllocMaxBlocks = 65536; // n_cell <= 65536 * 256 = 16,777,216
constexpr int64_t kNoAllocScratch = 4 * kNoAllocMaxBlocks + 4;
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};
} // namespace

namespace dndsmi_host {
int noalloc_scratch_init(dndsmi_ctx *c) {
    if (c->device < 0 || c->device >= kMaxDevices) {
        c->err = "device ordinal beyond the allocation-free scratch table";
        return DNDSMI_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            c->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return DNDSMI_ERR_HIP;
        }
        g_scratch_dev[c->device] = static_cast<double *>(p);
    }
    c->noalloc_scratch = g_scratch_dev[c->device];
    return DNDSMI_OK;
}
} // namespace dndsmi_host

extern "C" int dndsmi_evaluate_residual_noalloc(dndsmi_ctx *c, const dndsmi_array *rhs, int32_t P, int32_t volWise,
                                                double *res) {
    if (!c || !same_ctx(c, rhs, DNDSMI_ARRAY_DOF) || !res || P < 1)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    const int64_t g = grid_for(c->nc);
    if (g > kNoAllocMaxBlocks) {
        c->err = "dndsmi_evaluate_residual_noalloc: n_cell above the static scratch's maximum of " +
                 std::to_string(kNoAllocMaxBlocks * kBlock) + " cells";
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    if (!c->noalloc_scratch) {
        c->err = "dndsmi_evaluate_residual_noalloc: the static scratch was not looked up";
        return DNDSMI_ERR_HIP;
    }
    double *part = c->noalloc_scratch;
    return residual(c, rhs->d, nullptr, P, volWise, res, part, part + 4 * kNoAllocMaxBlocks);
}

extern "C" int dndsmi_implicit_euler_step(dndsmi_ctx *c, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                                          int32_t n_sweeps, double CFL, double dt, int32_t UseLocaldt,
                                          int32_t max_iter, double threshold, double t, int32_t *n_iter,
                                          double *res_hist) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNew, DNDSMI_ARRAY_REC) || uRec == uRecNew || n_sweeps < 0 || max_iter < 1 || !(dt > 0))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(ensure_implicit_scratch(c));
    const int nc = int(c->nc);
    const double alpha = 1.0;
    // xLast = x
    DNDSMI_TRY(c->hip(hipMemcpyAsync(c->xlast, u->d, sizeof(double) * 4 * c->nc, hipMemcpyDeviceToDevice, c->stream),
                      "copy"));
    double base[4] = {0, 0, 0, 0};
    if (n_iter)
        *n_iter = 0;
    for (int32_t iter = 1; iter <= max_iter; iter++) {
        // fdtau: EvaluateDt(dTau, u, CFL, dtMin, 1e100, UseLocaldt), dTau /= alpha (alpha = 1)
        DNDSMI_TRY(launch_dt(c, u->d, CFL, 1e100, c->imp_dtau));
        if (!UseLocaldt) {
            hipLaunchKernelGGL(k_fill_scalar, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->dt_min,
                               c->imp_dtau);
            DNDSMI_TRY(c->launch_check("k_fill_scalar"));
        }
        // frhs at t + dt: sweeps then EvaluateRHS
        for (int k = 0; k < n_sweeps; k++) {
            DNDSMI_TRY(step_sweep(c, u->d, uRec->d, uRecNew->d, t + dt));
        }
        DNDSMI_TRY(launch_rhs(c, u->d, uRec->d, t + dt, c->imp_rhs, 0, nullptr, nullptr, 0, 0));
        hipLaunchKernelGGL(k_dual_rhs, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->xlast, u->d, 1.0 / dt,
                           c->imp_dtau, c->imp_rhs, c->imp_xinc);
        DNDSMI_TRY(c->launch_check("k_dual_rhs"));
        // fsolve
        DNDSMI_TRY(fsolve(c, c->imp_dtau, dt, alpha, c->imp_rhs, u->d, c->imp_xinc));
        // x += xinc, fstop on the L1 norm of xinc
        double res[4];
        DNDSMI_TRY(residual(c, c->imp_xinc, u->d, 1, 0, res, c->res_part, c->res));
        if (res_hist)
            for (int v = 0; v < 4; v++)
                res_hist[4 * (iter - 1) + v] = res[v];
        for (int v = 0; v < 4; v++)
            base[v] = iter == 1 ? res[v] : std::max(base[v], res[v]);
        if (n_iter)
            *n_iter = iter;
        if (!std::isfinite(res[0])) {
            c->err = "non-finite increment in the implicit iteration";
            return DNDSMI_ERR_NAN;
        }
        if (res[0] / base[0] < threshold)
            break;
    }
    return DNDSMI_OK;
}

// ---------------------------------------------------------------- SDIRK4 ----

namespace {
// ImplicitSDIRK4DualTimeStep butcherA / butcherC / butcherB (DNDS_ODE.hpp:374-390)
constexpr double kZeta = 0.128886400515;
constexpr double kSA[3][3] = {{kZeta, 0, 0}, {0.5 - kZeta, kZeta, 0}, {2 * kZeta, 1 - 4 * kZeta, kZeta}};
constexpr double kSC[3] = {kZeta, 0.5, 1 - kZeta};
constexpr double kSB[3] = {1. / (6 * ((2 * kZeta - 1) * (2 * kZeta - 1))),
                           (4 * (kZeta * kZeta) - 4 * kZeta + 2. / 3.) / ((2 * kZeta - 1) * (2 * kZeta - 1)),
                           1. / (6 * ((2 * kZeta - 1) * (2 * kZeta - 1)))};

// stage right-hand side (DNDS_ODE.hpp:333-337): rhs = xLast; rhs -= x;
// rhs *= 1/dt; rhs += rhs_j a_ij for j <= i; then fsolve's compression and
// xinc = 0
__global__ __launch_bounds__(kBlock) void k_sdirk_rhs(int32_t nc, const double *__restrict__ xlast,
                                                      const double *__restrict__ x, double rdt,
                                                      const double *__restrict__ stage, int64_t lds, int32_t ns,
                                                      double a0, double a1, double a2,
                                                      const double *__restrict__ dtau, double *__restrict__ rhs,
                                                      double *__restrict__ xinc) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    const double dti = dtau[i];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int64_t k = 4 * int64_t(i) + v;
        double r = xlast[k];
        r -= x[k];
        r *= rdt;
        r += stage[k] * a0;
        if (ns > 1)
            r += stage[lds + k] * a1;
        if (ns > 2)
            r += stage[2 * lds + k] * a2;
        rhs[k] = (r * dti) / dti;
        xinc[k] = 0.0;
    }
}

// the step's result (DNDS_ODE.hpp:368-370): x = xLast; x += rhs_j (b_j dt)
__global__ __launch_bounds__(kBlock) void k_sdirk_final(int32_t nc, const double *__restrict__ xlast,
                                                        const double *__restrict__ stage, int64_t lds, double w0,
                                                        double w1, double w2, double *__restrict__ x) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int64_t k = 4 * int64_t(i) + v;
        double r = xlast[k];
        r += stage[k] * w0;
        r += stage[lds + k] * w1;
        r += stage[2 * lds + k] * w2;
        x[k] = r;
    }
}
} // namespace

extern "C" int dndsmi_sdirk4_step(dndsmi_ctx *c, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                                  int32_t n_sweeps, double CFL, double dt, int32_t UseLocaldt, int32_t max_iter,
                                  double threshold, double t, int32_t *n_iter, double *res_hist) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNew, DNDSMI_ARRAY_REC) || uRec == uRecNew || n_sweeps < 0 || max_iter < 1 || !(dt > 0))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(ensure_implicit_scratch(c));
    if (!c->imp_stage)
        DNDSMI_TRY(c->alloc(c->imp_stage, 3 * 4 * c->ldc));
    const int nc = int(c->nc);
    const int64_t lds = 4 * c->ldc;
    const size_t nbytes = sizeof(double) * 4 * c->nc;
    DNDSMI_TRY(c->hip(hipMemcpyAsync(c->xlast, u->d, nbytes, hipMemcpyDeviceToDevice, c->stream), "copy"));
    for (int ib = 0; ib < 3; ib++) {
        double *stage_rhs = c->imp_stage + ib * lds;
        const double aii = kSA[ib][ib];
        // x = xLast
        DNDSMI_TRY(c->hip(hipMemcpyAsync(u->d, c->xlast, nbytes, hipMemcpyDeviceToDevice, c->stream), "copy"));
        double base[4] = {0, 0, 0, 0};
        if (n_iter)
            n_iter[ib] = 0;
        for (int32_t iter = 1; iter <= max_iter; iter++) {
            DNDSMI_TRY(launch_dt(c, u->d, CFL, 1e100, c->imp_dtau));
            if (!UseLocaldt)
                hipLaunchKernelGGL(k_fill_scalar, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->dt_min,
                                   c->imp_dtau);
            hipLaunchKernelGGL(k_div_scalar, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->imp_dtau, aii);
            DNDSMI_TRY(c->launch_check("fdtau"));
            const double ts = t + kSC[ib] * dt;
            for (int q = 0; q < n_sweeps; q++) {
                DNDSMI_TRY(step_sweep(c, u->d, uRec->d, uRecNew->d, ts));
            }
            DNDSMI_TRY(launch_rhs(c, u->d, uRec->d, ts, stage_rhs, 0, nullptr, nullptr, 0, 0));
            hipLaunchKernelGGL(k_sdirk_rhs, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->xlast, u->d,
                               1.0 / dt, c->imp_stage, lds, ib + 1, kSA[ib][0], kSA[ib][1], kSA[ib][2], c->imp_dtau,
                               c->imp_rhs, c->imp_xinc);
            DNDSMI_TRY(c->launch_check("k_sdirk_rhs"));
            DNDSMI_TRY(fsolve(c, c->imp_dtau, dt, aii, c->imp_rhs, u->d, c->imp_xinc));
            double res[4];
            DNDSMI_TRY(residual(c, c->imp_xinc, u->d, 1, 0, res, c->res_part, c->res));
            if (res_hist)
                for (int v = 0; v < 4; v++)
                    res_hist[(int64_t(ib) * max_iter + (iter - 1)) * 4 + v] = res[v];
            for (int v = 0; v < 4; v++)
                base[v] = iter == 1 ? res[v] : std::max(base[v], res[v]);
            if (n_iter)
                n_iter[ib] = iter;
            if (!std::isfinite(res[0])) {
                c->err = "non-finite increment in the SDIRK4 iteration";
                return DNDSMI_ERR_NAN;
            }
            if (res[0] / base[0] < threshold)
                break;
        }
    }
    hipLaunchKernelGGL(k_sdirk_final, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->xlast, c->imp_stage, lds,
                       kSB[0] * dt, kSB[1] * dt, kSB[2] * dt, u->d);
    return c->launch_check("k_sdirk_final");
}

// ------------------------------------------------------------------ BDF ----

struct dndsmi_bdf {
    dndsmi_ctx *ctx = nullptr;
    int32_t k = 2;
    double *prev = nullptr; // [k - 1][ldc][4] previous steps' initial states
    double dt_prev[3] = {0, 0, 0};
    int32_t cn_prev = 0, prev_start = 0;
    ~dndsmi_bdf() {
        if (prev)
            (void)hipFree(prev);
    }
};

void dndsmi_bdf_detach(dndsmi_bdf *b) { b->ctx = nullptr; } // keeps its own allocation

namespace {
// ImplicitBDFDualTimeStep<TDATA>::BDFCoefs (DNDS_ODE.hpp:506-511)
constexpr double kBDFCoefs[4][5] = {{1. / 1., 1. / 1., 0., 0., 0.},
                                    {2. / 3., 4. / 3., -1. / 3., 0., 0.},
                                    {6. / 11., 18. / 11., -9. / 11., 2. / 11., 0.},
                                    {12. / 25., 48. / 25., -36. / 25., 16. / 25., -3. / 25.}};
inline int32_t pos_mod(int32_t a, int32_t m) { return ((a % m) + m) % m; }
} // namespace

extern "C" int dndsmi_bdf_create(dndsmi_ctx *c, int32_t k, dndsmi_bdf **out) {
    if (!c || !out || k < 1 || k > 4)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto *b = new (std::nothrow) dndsmi_bdf();
    if (!b)
        return DNDSMI_ERR_OUT_OF_MEMORY;
    b->ctx = c, b->k = k, b->prev_start = k - 2 >= 0 ? k - 2 : 0;
    if (k > 1) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(double) * 4 * size_t(c->ldc) * size_t(k - 1)) != hipSuccess) {
            delete b;
            c->err = "hipMalloc failed for the BDF history";
            return DNDSMI_ERR_OUT_OF_MEMORY;
        }
        b->prev = static_cast<double *>(q);
    }
    c->bdfs.insert(b);
    *out = b;
    return DNDSMI_OK;
}

extern "C" void dndsmi_bdf_destroy(dndsmi_bdf *b) {
    if (b && b->ctx) {
        (void)hipStreamSynchronize(b->ctx->stream);
        b->ctx->bdfs.erase(b);
    }
    delete b;
}

extern "C" int dndsmi_bdf_step(dndsmi_bdf *b, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                               int32_t n_sweeps, double CFL, double dt, int32_t UseLocaldt, int32_t max_iter,
                               double threshold, double t, int32_t *n_iter, double *res_hist) {
    if (!b || !b->ctx)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    dndsmi_ctx *c = b->ctx;
    if (!same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNew, DNDSMI_ARRAY_REC) || uRec == uRecNew || n_sweeps < 0 || max_iter < 1 || !(dt > 0))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    const int32_t prev_siz = b->k - 1, k_cur = b->cn_prev + 1;
    for (int32_t ip = 0; ip < b->cn_prev; ip++)
        if (!(std::abs(b->dt_prev[pos_mod(ip + b->prev_start, prev_siz)] - dt) < dt * 1e-8)) {
            c->err = "BDF step: dt differs from the previous steps' dt";
            return DNDSMI_ERR_INVALID_ARGUMENT;
        }
    DNDSMI_TRY(ensure_implicit_scratch(c));
    const int nc = int(c->nc);
    const double *co = kBDFCoefs[k_cur - 1];
    const double alpha0 = co[0];
    const double *ph[3] = {nullptr, nullptr, nullptr};
    double pc[3] = {0, 0, 0};
    for (int32_t ip = 0; ip < b->cn_prev; ip++) {
        ph[ip] = b->prev + 4 * c->ldc * pos_mod(ip + b->prev_start, prev_siz);
        pc[ip] = co[2 + ip] / dt;
    }
    // xLast = x
    DNDSMI_TRY(c->hip(hipMemcpyAsync(c->xlast, u->d, sizeof(double) * 4 * c->nc, hipMemcpyDeviceToDevice, c->stream),
                      "copy"));
    double base[4] = {0, 0, 0, 0};
    if (n_iter)
        *n_iter = 0;
    for (int32_t iter = 1; iter <= max_iter; iter++) {
        // fdtau(dTau, alpha0)
        DNDSMI_TRY(launch_dt(c, u->d, CFL, 1e100, c->imp_dtau));
        if (!UseLocaldt)
            hipLaunchKernelGGL(k_fill_scalar, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->dt_min,
                               c->imp_dtau);
        hipLaunchKernelGGL(k_div_scalar, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->imp_dtau, alpha0);
        DNDSMI_TRY(c->launch_check("fdtau"));
        // frhs at t + dt
        for (int q = 0; q < n_sweeps; q++) {
            DNDSMI_TRY(step_sweep(c, u->d, uRec->d, uRecNew->d, t + dt));
        }
        DNDSMI_TRY(launch_rhs(c, u->d, uRec->d, t + dt, c->imp_rhs, 0, nullptr, nullptr, 0, 0));
        hipLaunchKernelGGL(k_bdf_rhs, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, alpha0, u->d, -1. / dt,
                           c->xlast, co[1] / dt, b->cn_prev, ph[0], pc[0], ph[1], pc[1], ph[2], pc[2], c->imp_dtau,
                           c->imp_rhs, c->imp_xinc);
        DNDSMI_TRY(c->launch_check("k_bdf_rhs"));
        // fsolve with alphaDiag = alpha0
        DNDSMI_TRY(fsolve(c, c->imp_dtau, dt, alpha0, c->imp_rhs, u->d, c->imp_xinc));
        double res[4];
        DNDSMI_TRY(residual(c, c->imp_xinc, u->d, 1, 0, res, c->res_part, c->res));
        if (res_hist)
            for (int v = 0; v < 4; v++)
                res_hist[4 * (iter - 1) + v] = res[v];
        for (int v = 0; v < 4; v++)
            base[v] = iter == 1 ? res[v] : std::max(base[v], res[v]);
        if (n_iter)
            *n_iter = iter;
        if (!std::isfinite(res[0])) {
            c->err = "non-finite increment in the BDF iteration";
            return DNDSMI_ERR_NAN;
        }
        if (res[0] / base[0] < threshold)
            break;
    }
    // the step's initial state joins the history (DNDS_ODE.hpp:488-495)
    if (prev_siz) {
        b->prev_start = pos_mod(b->prev_start - 1, prev_siz);
        DNDSMI_TRY(c->hip(hipMemcpyAsync(b->prev + 4 * c->ldc * b->prev_start, c->xlast, sizeof(double) * 4 * c->nc,
                                         hipMemcpyDeviceToDevice, c->stream),
                          "copy"));
        b->dt_prev[b->prev_start] = dt;
        b->cn_prev = std::min(b->cn_prev + 1, prev_siz);
    }
    return DNDSMI_OK;
}
