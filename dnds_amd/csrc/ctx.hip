// Device context, arrays and the FP64 kernels of the hot path (gfx950).
//
// The context holds the mesh and the VR setup in device memory, repacked from
// the reference layouts into structure-of-arrays streams so that a wavefront
// of 64 cells (or faces) reads each matrix element as one contiguous 512-byte
// segment.  All kernels are launched on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "ctx.hpp"

using namespace dndsmi_dev;
using namespace dndsmi_host;

namespace dndsmi_host {

hipEvent_t prof_event(dndsmi_ctx *c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess)
            return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// brackets a launch with events when profiling is on
void prof_begin(dndsmi_ctx *c, int id) {
    if (!c->prof_on)
        return;
    c->prof_open = id;
    c->prof_open_ev = c->ev_used;
    hipEvent_t e = prof_event(c);
    if (e)
        (void)hipEventRecord(e, c->stream);
}

void prof_end(dndsmi_ctx *c) {
    if (!c->prof_on || c->prof_open < 0)
        return;
    hipEvent_t e = prof_event(c);
    if (e) {
        (void)hipEventRecord(e, c->stream);
        c->prof_recs.push_back({c->prof_open, c->prof_open_ev});
    }
    c->prof_open = -1;
}

} // namespace dndsmi_host

// ========================================================================
// kernels
// ========================================================================

// Boundary correction of a Wall_Euler / Wall_NoSlip face w to its cell's
// sweep (DNDS_FV_VR.hpp:1577-1657), from the cell's coefficients xi and mean
// ui: corr = relax A^-1 sum_g w_g |n_g| D0^T W0^2 (uBV - u_i).  ADD adds the
// correction to corr (xi may then be corr itself: xi is read in full before
// corr is written); otherwise it is stored there.
template <bool ADD>
__device__ inline void wall_correction(int32_t w, int32_t bc, const double *xi, const double ui[4],
                                       const double *__restrict__ wall_ainv, const double *__restrict__ wall_d0,
                                       const double *__restrict__ wall_nrm, const double *__restrict__ wall_w0,
                                       int64_t ldw, double relax, const GasParams &gp, double t, int32_t *errw,
                                       double *corr) {
    double BCc[36];
    for (int k = 0; k < 36; k++)
        BCc[k] = 0.0;
    const double w0 = wall_w0[w];
    for (int g = 0; g < 3; g++) {
        const double gw = g == 1 ? kLine3W1 : kLine3W0;
        const double nx = wall_nrm[(2 * g) * ldw + w], ny = wall_nrm[(2 * g + 1) * ldw + w];
        double uBL[4];
        for (int v = 0; v < 4; v++) {
            double acc = 0;
            for (int c = 0; c < 9; c++)
                acc += wall_d0[(g * 9 + c) * ldw + w] * xi[c + 9 * v];
            uBL[v] = acc;
        }
        for (int v = 0; v < 4; v++)
            uBL[v] += ui[v];
        double n0, n1;
        stable_normalized(nx, ny, n0, n1);
        double uBV[4];
        if (bc == BC_WALL_EULER) {
            for (int v = 0; v < 4; v++)
                uBV[v] = uBL[v];
            const double dd = n0 * uBV[1] + n1 * uBV[2];
            uBV[1] = uBL[1] - n0 * dd;
            uBV[2] = uBL[2] - n1 * dd;
        } else if (!boundary_value(gp, uBL, n0, n1, 0.0, 0.0, t, bc, uBV)) {
            atomicOr(errw, ERRBIT_NONPOSITIVE);
            for (int v = 0; v < 4; v++)
                uBV[v] = uBL[v];
        }
        const double nl = sqrt(nx * nx + ny * ny);
        for (int r = 0; r < 9; r++)
            for (int v = 0; v < 4; v++) {
                double inc = wall_d0[(g * 9 + r) * ldw + w] * w0 * w0 * (uBV[v] - ui[v]);
                inc *= nl;
                BCc[r + 9 * v] += inc * gw;
            }
    }
    for (int r = 0; r < 9; r++)
        for (int v = 0; v < 4; v++) {
            double acc = 0;
            for (int c = 0; c < 9; c++)
                acc += wall_ainv[(r * 9 + c) * ldw + w] * BCc[c + 9 * v];
            if (ADD)
                corr[r + 9 * v] += relax * acc;
            else
                corr[r + 9 * v] = relax * acc;
        }
}

// The Jacobi sweep's wall corrections, computed before the sweep from the
// current iterate; the sweep adds each in the face's slot, which keeps the
// reference's accumulation order.
__global__ __launch_bounds__(kBlock) void k_wall_corr(int32_t nw, int64_t ldc, const int32_t *__restrict__ wcell,
                                                      const int32_t *__restrict__ wbc,
                                                      const double *__restrict__ wall_ainv,
                                                      const double *__restrict__ wall_d0,
                                                      const double *__restrict__ wall_nrm,
                                                      const double *__restrict__ wall_w0, int64_t ldw,
                                                      const double *__restrict__ u, const double *__restrict__ x,
                                                      double relax, GasParams gp, double t,
                                                      double *__restrict__ corr, int32_t *errw) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw)
        return;
    const int32_t i = wcell[w];
    double xi[36], ui[4], cw[36];
    for (int k = 0; k < 36; k++)
        xi[k] = x[k * ldc + i];
    for (int v = 0; v < 4; v++)
        ui[v] = u[4 * int64_t(i) + v];
    wall_correction<false>(w, wbc[w], xi, ui, wall_ainv, wall_d0, wall_nrm, wall_w0, ldw, relax, gp, t, errw, cw);
    for (int k = 0; k < 36; k++)
        corr[k * ldw + w] = cw[k];
}

// One SOR pass over cells of one colour (SOR_Instead,
// DNDS_FV_VR.hpp:1495-1662), in place on x.  No two cells of a colour are
// neighbours, so the pass equals the reference's sequential in-place scan
// over those cells, and the colours in turn equal its scan in colour-major
// order.  As in the reference, every term is added to the cell's own
// coefficients: the far-field product M x_i and the Wall_Euler / Wall_NoSlip
// corrections take the partial sum after (1 - w) scaling and the faces before
// them (the product is formed from a copy, then added).
//
// The cells are the device range [base, base + n): with SOR_Instead the
// device numbering is colour-major.  SELF: the cells may have far-field,
// Wall_Euler or Wall_NoSlip faces, the terms that read the cell's own partial
// sum; the numbering puts those cells after the colour's others, so that the
// others (all but the boundary layer) run without the registers of the copy
// and of the wall correction.  x and xo are the
// same array: the new coefficients go out through xo so that the compiler
// does not keep the 36 addresses of the first reads live until the stores.
template <int NF, bool SELF>
__global__ __launch_bounds__(kBlock) void k_sor(int32_t n, int32_t base, int64_t ldc, const int32_t *__restrict__ nbr,
                                                const int32_t *__restrict__ sbc, const int32_t *__restrict__ wall_idx,
                                                const double *__restrict__ M, const double *__restrict__ bv,
                                                const double *__restrict__ u, const double *x, double *xo,
                                                double relax, const double *__restrict__ wall_ainv,
                                                const double *__restrict__ wall_d0,
                                                const double *__restrict__ wall_nrm,
                                                const double *__restrict__ wall_w0, int64_t ldw, GasParams gp,
                                                double t, int32_t *errw) {
    const int32_t tid = xcd_block() * blockDim.x + threadIdx.x;
    if (tid >= n)
        return;
    const int32_t i = base + tid;
    double nw[36];
#pragma unroll
    for (int k = 0; k < 36; k++)
        nw[k] = (1 - relax) * x[k * ldc + i];
    double ui[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        ui[v] = u[4 * int64_t(i) + v];
#pragma unroll 1
    for (int s = 0; s < NF; s++) {
        const int32_t j = nbr[s * ldc + i];
        if (j == -2)
            continue;
        const double *bs = bv + int64_t(s) * 9 * ldc + i;
        double xj[36], du[4], bmul = 1.0;
        if (j >= 0) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                du[v] = u[4 * int64_t(j) + v] - ui[v];
#pragma unroll
            for (int k = 0; k < 36; k++)
                xj[k] = x[k * ldc + j];
        } else {
            const int32_t bc = sbc[s * ldc + i];
            if (bc == BC_WALL) {
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    const double b = bs[r * ldc];
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        nw[r + 9 * v] += relax * ((0.0 - ui[v]) * b);
                }
                continue;
            }
            if (!SELF) { // not in this launch's lists
                atomicOr(errw, ERRBIT_BC);
                continue;
            }
            if (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) {
                wall_correction<true>(wall_idx[s * ldc + i], bc, nw, ui, wall_ainv, wall_d0, wall_nrm, wall_w0, ldw,
                                      relax, gp, t, errw, nw);
                continue;
            }
            if (!is_far(bc)) {
                atomicOr(errw, ERRBIT_BC);
                continue;
            }
            // far field: M_if times the partial sum, no mean-value term
#pragma unroll
            for (int v = 0; v < 4; v++)
                du[v] = 0.0;
#pragma unroll
            for (int k = 0; k < 36; k++)
                xj[k] = nw[k];
            bmul = 0.0;
        }
        const dbl2 *Ms = reinterpret_cast<const dbl2 *>(M) + m_vec_index(s, 0, i, ldc, kMPairs);
        dbl2 mp = {0.0, 0.0};
        double tt[4] = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 81; e++) {
            const int r = e / 9, c = e % 9;
            if (e % 2 == 0)
                mp = __builtin_nontemporal_load(Ms + (e / 2) * 64);
            const double m = e % 2 == 0 ? mp.x : mp.y;
            if (c == 0)
                tt[0] = tt[1] = tt[2] = tt[3] = 0.0;
#pragma unroll
            for (int v = 0; v < 4; v++)
                tt[v] += m * xj[c + 9 * v];
            if (c == 8) {
                const double b = bmul * __builtin_nontemporal_load(bs + r * ldc);
#pragma unroll
                for (int v = 0; v < 4; v++)
                    nw[r + 9 * v] += relax * (tt[v] + du[v] * b);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 36; k++)
        __builtin_nontemporal_store(nw[k], &xo[k * ldc + i]);
}

// One block-Jacobi VR sweep, SOR off (ReconstructionJacobiStep,
// DNDS_FV_VR.hpp:1495-1682): new_i = (1-w) x_i + w sum_f [M_if x_j + b_if (u_j - u_i)^T]
// in cell2face order, with the reference's boundary variants: far field
// couples the cell to itself (M_if x_i), a Wall face adds b_if (0 - u_i)^T,
// Wall_Euler / Wall_NoSlip faces add the correction of k_wall_corr.
//
// RESID (the FP64 step of the mixed-precision sweeps): instead of the new
// iterate, writes the increment r = J(x) - x rounded to FP32 into r32 and d32
// (the start of the FP32 correction sweeps, k_delta32).
template <int NF, bool RESID = false>
__global__ __launch_bounds__(kBlock) void k_jacobi(int32_t nc, int64_t ldc, const int32_t *__restrict__ nbr,
                                                   const int32_t *__restrict__ sbc,
                                                   const int32_t *__restrict__ wall_idx,
                                                   const double *__restrict__ M, const double *__restrict__ bv,
                                                   const double *__restrict__ u, const double *__restrict__ x,
                                                   double *__restrict__ xn, double relax,
                                                   const double *__restrict__ corr, int64_t ldw, int32_t *errw,
                                                   float *__restrict__ r32 = nullptr,
                                                   float *__restrict__ d32 = nullptr) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    double nw[36];
#pragma unroll
    for (int k = 0; k < 36; k++)
        nw[k] = (1 - relax) * x[k * ldc + i];
    double ui[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        ui[v] = u[4 * int64_t(i) + v];
#pragma unroll 1
    for (int s = 0; s < NF; s++) {
        int32_t j = nbr[s * ldc + i];
        if (j == -2)
            continue;
        const double *bs = bv + int64_t(s) * 9 * ldc + i;
        double bmul = 1.0;
        if (j < 0) {
            const int32_t bc = sbc[s * ldc + i];
            if (bc == BC_WALL) {
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    const double b = bs[r * ldc];
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        nw[r + 9 * v] += relax * ((0.0 - ui[v]) * b);
                }
                continue;
            }
            if (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) {
                const int64_t w = wall_idx[s * ldc + i];
#pragma unroll
                for (int k = 0; k < 36; k++)
                    nw[k] += corr[k * ldw + w];
                continue;
            }
            if (!is_far(bc)) {
                atomicOr(errw, ERRBIT_BC);
                continue;
            }
            j = i; // far field: M_if x_i, no mean-value term
            bmul = 0.0;
        }
        // A^-1 B entries e = 9 r + c, two per 16-byte load (entry 81 pads)
        const dbl2 *Ms = reinterpret_cast<const dbl2 *>(M) + m_vec_index(s, 0, i, ldc, kMPairs);
        double du[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
            du[v] = bmul * (u[4 * int64_t(j) + v] - ui[v]);
        double xj[36];
#pragma unroll
        for (int k = 0; k < 36; k++)
            xj[k] = x[k * ldc + j];
        dbl2 mp = {0.0, 0.0};
        double tt[4] = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 81; e++) {
            const int r = e / 9, c = e % 9;
            if (e % 2 == 0)
                mp = __builtin_nontemporal_load(Ms + (e / 2) * 64);
            const double m = e % 2 == 0 ? mp.x : mp.y;
            if (c == 0)
                tt[0] = tt[1] = tt[2] = tt[3] = 0.0;
#pragma unroll
            for (int v = 0; v < 4; v++)
                tt[v] += m * xj[c + 9 * v];
            if (c == 8) {
                const double b = bmul * __builtin_nontemporal_load(bs + r * ldc);
#pragma unroll
                for (int v = 0; v < 4; v++)
                    nw[r + 9 * v] += relax * (tt[v] + du[v] * b);
            }
        }
    }
    // the new iterate streams out with non-temporal stores: written once
    // and read by the next launch, it would only displace the neighbours'
    // reconstructions that this sweep still gathers (4 % faster on the 2.1M
    // triangles, 2 % on the 6.3M mixed cells)
    if (RESID) {
#pragma unroll
        for (int k = 0; k < 36; k++) {
            const float r = float(nw[k] - x[k * ldc + i]);
            __builtin_nontemporal_store(r, &r32[k * ldc + i]);
            __builtin_nontemporal_store(r, &d32[k * ldc + i]);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 36; k++)
        __builtin_nontemporal_store(nw[k], &xn[k * ldc + i]);
}

// Read-bandwidth probe of the sweep's dominant stream: every cell streams its
// slots' A^-1 B pairs exactly as k_jacobi does (16-byte non-temporal loads,
// the same slots skipped) with none of the sweep's gathers, and writes one
// double.  Its rate is the ceiling of the sweep's streaming part.
template <int NF>
__global__ __launch_bounds__(kBlock) void k_read_probe(int32_t nc, int64_t ldc, const int32_t *__restrict__ nbr,
                                                       const double *__restrict__ M, double *__restrict__ out) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    double a0 = 0.0, a1 = 0.0;
#pragma unroll 1
    for (int s = 0; s < NF; s++) {
        if (nbr[s * ldc + i] == -2)
            continue;
        const dbl2 *Ms = reinterpret_cast<const dbl2 *>(M) + m_vec_index(s, 0, i, ldc, kMPairs);
#pragma unroll
        for (int e = 0; e < kMPairs; e++) {
            const dbl2 m = __builtin_nontemporal_load(Ms + e * 64);
            a0 += m.x, a1 += m.y;
        }
    }
    out[i] = a0 + a1;
}

// Linear part of the Wall_Euler / Wall_NoSlip correction (k_wall_corr) for a
// correction d of the reconstruction: the wall state is the reconstructed
// value D d with the normal momentum removed (Wall_Euler) or the momentum
// reversed (Wall_NoSlip), both linear, and the cell mean drops out.
__global__ __launch_bounds__(kBlock) void k_wall_corr_lin(int32_t nw, int64_t ldc, const int32_t *__restrict__ wcell,
                                                          const int32_t *__restrict__ wbc,
                                                          const double *__restrict__ wall_ainv,
                                                          const double *__restrict__ wall_d0,
                                                          const double *__restrict__ wall_nrm,
                                                          const double *__restrict__ wall_w0, int64_t ldw,
                                                          const float *__restrict__ d, float relax,
                                                          float *__restrict__ corr) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw)
        return;
    const int32_t i = wcell[w];
    const int32_t bc = wbc[w];
    float di[36], BCc[36];
    for (int k = 0; k < 36; k++)
        di[k] = d[k * ldc + i], BCc[k] = 0.f;
    const float w0 = float(wall_w0[w]);
    for (int g = 0; g < 3; g++) {
        const float gw = float(g == 1 ? kLine3W1 : kLine3W0);
        const double nx = wall_nrm[(2 * g) * ldw + w], ny = wall_nrm[(2 * g + 1) * ldw + w];
        float uBL[4];
        for (int v = 0; v < 4; v++) {
            float acc = 0.f;
            for (int c = 0; c < 9; c++)
                acc = fmaf(float(wall_d0[(g * 9 + c) * ldw + w]), di[c + 9 * v], acc);
            uBL[v] = acc;
        }
        double n0, n1;
        stable_normalized(nx, ny, n0, n1);
        float uBV[4] = {uBL[0], uBL[1], uBL[2], uBL[3]};
        if (bc == BC_WALL_EULER) {
            const float dd = float(n0) * uBL[1] + float(n1) * uBL[2];
            uBV[1] = uBL[1] - float(n0) * dd;
            uBV[2] = uBL[2] - float(n1) * dd;
        } else {
            uBV[1] = -uBL[1];
            uBV[2] = -uBL[2];
        }
        const float nl = float(sqrt(nx * nx + ny * ny));
        for (int r = 0; r < 9; r++)
            for (int v = 0; v < 4; v++)
                BCc[r + 9 * v] += float(wall_d0[(g * 9 + r) * ldw + w]) * w0 * w0 * uBV[v] * nl * gw;
    }
    for (int r = 0; r < 9; r++)
        for (int v = 0; v < 4; v++) {
            float acc = 0.f;
            for (int c = 0; c < 9; c++)
                acc = fmaf(float(wall_ainv[(r * 9 + c) * ldw + w]), BCc[c + 9 * v], acc);
            corr[(r + 9 * v) * ldw + w] = relax * acc;
        }
}

// FP32 correction sweep of the mixed-precision reconstruction.  The Jacobi
// map is affine, J(x) = G x + c with G x = (1-w) x + w sum_f M_if x_j (far
// field faces couple the cell to itself), so with r = J(x) - x the sweeps
//   d_0 = r,  d_{k+1} = G d_k + r
// give x + d_k = J^{k+1}(x) exactly in real arithmetic: n - 1 FP32 sweeps
// after one FP64 step reproduce n FP64 sweeps, with FP32 rounding confined
// to the correction d (iterative refinement).  M is read as FP32.  LAST adds
// the correction to the FP64 iterate in place.  Wall_Euler / Wall_NoSlip
// faces contribute the linear part of their boundary correction, computed
// from d by k_wall_corr_lin (the wall state is linear in the reconstructed
// value there).
template <int NF, bool LAST>
__global__ __launch_bounds__(kBlock) void k_delta32(int32_t nc, int64_t ldc, const int32_t *__restrict__ nbr,
                                                    const int32_t *__restrict__ sbc, const float *__restrict__ M32,
                                                    const float *__restrict__ r32, const float *__restrict__ din,
                                                    float *__restrict__ dout, float relax,
                                                    double *__restrict__ x, const int32_t *__restrict__ wall_idx,
                                                    const float *__restrict__ wcorr, int64_t ldw) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    float nw[36];
#pragma unroll
    for (int k = 0; k < 36; k++)
        nw[k] = (1.0f - relax) * din[k * ldc + i];
#pragma unroll 1
    for (int s = 0; s < NF; s++) {
        int32_t j = nbr[s * ldc + i];
        if (j == -2)
            continue;
        if (j < 0) {
            const int32_t bc = sbc[s * ldc + i];
            if (bc == BC_WALL) // constant term: in r only
                continue;
            if (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) {
                const int64_t w = wall_idx[s * ldc + i];
#pragma unroll
                for (int k = 0; k < 36; k++)
                    nw[k] += wcorr[k * ldw + w];
                continue;
            }
            j = i; // far field
        }
        // FP32 A^-1 B entries, four per 16-byte load (81..83 pad)
        const flt4 *Ms = reinterpret_cast<const flt4 *>(M32) + m_vec_index(s, 0, i, ldc, kMQuads);
        float dj[36];
#pragma unroll
        for (int k = 0; k < 36; k++)
            dj[k] = din[k * ldc + j];
        flt4 mq = {0.f, 0.f, 0.f, 0.f};
        float tt[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int e = 0; e < 81; e++) {
            const int r = e / 9, c = e % 9;
            if (e % 4 == 0)
                mq = __builtin_nontemporal_load(Ms + (e / 4) * 64);
            const float m = mq[e % 4];
            if (c == 0)
                tt[0] = tt[1] = tt[2] = tt[3] = 0.f;
#pragma unroll
            for (int v = 0; v < 4; v++)
                tt[v] = fmaf(m, dj[c + 9 * v], tt[v]);
            if (c == 8)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    nw[r + 9 * v] = fmaf(relax, tt[v], nw[r + 9 * v]);
        }
    }
#pragma unroll
    for (int k = 0; k < 36; k++) {
        const float d = nw[k] + r32[k * ldc + i];
        if (LAST)
            x[k * ldc + i] += double(d);
        else
            __builtin_nontemporal_store(d, &dout[k * ldc + i]); // streamed, as in k_jacobi
    }
}

// A^-1 B from FP64 pairs to FP32 quads (both in the m_vec_index layout)
__global__ __launch_bounds__(kBlock) void k_m_to_f32(int32_t nfmax, int64_t ldc, const double *__restrict__ M,
                                                     float *__restrict__ M32) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(nfmax) * kMQuads * ldc)
        return;
    const int64_t i = t % ldc, sq = t / ldc;
    const int s = int(sq / kMQuads), q = int(sq % kMQuads);
    const int64_t o = m_vec_index(s, q, i, ldc, kMQuads) * 4;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const int k = 4 * q + w;
        M32[o + w] = k < 81 ? float(M[m_vec_index(s, k / 2, i, ldc, kMPairs) * 2 + k % 2]) : 0.f;
    }
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *__restrict__ src,
                                                   float *__restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = float(src[i]);
}

// EvaluateDt face loop (DNDS_FV_EulerEvaluator.cpp:25-137): lambdaFace and
// deltaLambdaFace per face (muGas = 0).
__global__ __launch_bounds__(kBlock) void k_dt_face(int32_t nf, int64_t ldf, const int32_t *__restrict__ f2c,
                                                    const double *__restrict__ fnc, const double *__restrict__ u,
                                                    GasParams gp, const double *__restrict__ farea,
                                                    const double *__restrict__ vol, double *__restrict__ lf,
                                                    double *__restrict__ dlf, int32_t *errw) {
    const double gamma = gp.gamma;
    const int32_t f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const int32_t L = f2c[f], R = f2c[ldf + f];
    double n0, n1;
    normalized(fnc[f], fnc[ldf + f], n0, n1);
    double UL[4];
    for (int k = 0; k < 4; k++)
        UL[k] = u[4 * int64_t(L) + k];
    double uMean[4] = {UL[0], UL[1], UL[2], UL[3]};
    double pL, asqrL, HL, pR, asqrR, HR;
    const double vL0 = UL[1] / UL[0], vL1 = UL[2] / UL[0];
    double vR0 = vL0, vR1 = vL1;
    ideal_gas_thermal(UL[3], UL[0], vL0 * vL0 + vL1 * vL1, gamma, pL, asqrL, HL);
    pR = pL, HR = HL, asqrR = asqrL;
    if (R >= 0) {
        double UR[4];
        for (int k = 0; k < 4; k++)
            UR[k] = u[4 * int64_t(R) + k];
        for (int k = 0; k < 4; k++)
            uMean[k] = (uMean[k] + UR[k]) * 0.5;
        vR0 = UR[1] / UR[0], vR1 = UR[2] / UR[0];
        ideal_gas_thermal(UR[3], UR[0], vR0 * vR0 + vR1 * vR1, gamma, pR, asqrR, HR);
    }
    if (!(uMean[0] > 0))
        atomicOr(errw, ERRBIT_NONPOSITIVE);
    const double veloNMean = 0.5 * ((vL0 + vR0) * n0 + (vL1 + vR1) * n1);
    const double pMean = (pL + pR) * 0.5;
    const double aMean = sqrt(gamma * pMean / uMean[0]);
    const double lambdaConvection = fabs(veloNMean) + aMean;
    if (gp.mu > 0) { // DNDS_FV_EulerEvaluator.cpp:79-119, NS_2D
        const double lamVis = sutherland_mu(gp, pMean, uMean[0]) / uMean[0] * fmax(4. / 3., gamma / gp.pr);
        const double volL = vol[L], volR = R >= 0 ? vol[R] : volL;
        lf[f] = lambdaConvection + lamVis * farea[f] * (1. / volL + 1. / volR);
    } else {
        lf[f] = lambdaConvection;
    }
    dlf[f] = fabs((vR0 - vL0) * n0 + (vR1 - vL1) * n1) + sqrt(fabs(asqrR - asqrL)) * 0.7071;
}

// EvaluateDt cell loop (DNDS_FV_EulerEvaluator.cpp:132-150): lambdaCell from
// the faces in ascending order, dt_i and the block minimum.
template <int NF>
__global__ __launch_bounds__(kBlock) void k_dt_cell(int32_t nc, int64_t ldc, const int32_t *__restrict__ gat,
                                                    const double *__restrict__ lf, const double *__restrict__ farea,
                                                    const double *__restrict__ vol, double cfl, double max_dt,
                                                    double *__restrict__ dt, double *__restrict__ part) {
    __shared__ double red[kBlock];
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    double d = 3e200;
    if (i < nc) {
        double lam = 0.0;
#pragma unroll
        for (int s = 0; s < NF; s++) {
            const int32_t g = gat[s * ldc + i];
            if (g < 0)
                continue;
            const int32_t f = g >> 1;
            lam += lf[f] * farea[f];
        }
        d = fmin(cfl * vol[i] / (lam + 1e-100), max_dt);
        dt[i] = d;
    }
    red[threadIdx.x] = d;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_min_final(int32_t n, const double *__restrict__ part,
                                                      double *__restrict__ out, int32_t *errw) {
    __shared__ double red[kBlock];
    double d = 3e200;
    for (int32_t k = threadIdx.x; k < n; k += kBlock)
        d = fmin(d, part[k]);
    red[threadIdx.x] = d;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
        if (!isfinite(red[0]))
            atomicOr(errw, ERRBIT_NAN);
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_from(int32_t n, const double *__restrict__ src,
                                                      double *__restrict__ dst) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[0];
}

// EvaluateRHS face loop (DNDS_FV_EulerEvaluator.cpp:215-492): left/right
// states at the three LINE_3 points, Riemann flux, flux-balance fluxes,
// Gauss sum.  Stores the integrated F, FLfix, FRfix per face.  BOUNDARY
// selects the boundary-face instance (run over the list of boundary faces);
// the inner instance skips boundary faces, so neither carries the other's
// registers.
//
// MIXED: the face reconstruction D0 . x runs in FP32 (FP32 bases, FP32
// accumulation) and the integrated fluxes are stored in FP32; the Riemann
// solver and the boundary states stay FP64.
//
// STRAIGHT (inner faces of a mesh whose face normals are bitwise equal at the
// three points, as on straight edges): the unit normal, |n| and the
// flux-balance fluxes are computed at point 0 and reused, which yields the
// same bits as recomputing them from the same inputs.
//
// VIS (muGas > 0, FP64 only): also the gradients of both sides at each point
// from rows 1, 2 of the DiBj blocks (DNDS_FV_EulerEvaluator.cpp:297-346), the
// face gradient with the jump term over distGRP = 2 min(V) / A, which walls
// push to 3e200 (:356-395, USE_DISABLE_DIST_GRP_FIX_AT_WALL), and fluxFace's
// viscous flux (DNDS_FV_EulerEvaluator.hpp:397-430, 574-577).
template <bool BOUNDARY, bool MIXED = false, bool STRAIGHT = false, bool VIS = false>
__global__ __launch_bounds__(kBlock, VIS && !BOUNDARY ? 2 : 1) void k_rhs_face(int32_t n, int64_t ldf, int64_t ldc,
                                                     const int32_t *__restrict__ list,
                                                     const int32_t *__restrict__ f2c,
                                                     const int32_t *__restrict__ fbc, const double *__restrict__ fxy,
                                                     const double *__restrict__ cmom,
                                                     const double *__restrict__ fn, const double *__restrict__ fctr,
                                                     const double *__restrict__ dlf, const double *__restrict__ u,
                                                     const double *__restrict__ x, GasParams gp, double t,
                                                     const uint8_t *__restrict__ fpos,
                                                     double *__restrict__ fslot, int32_t *errw,
                                                     float *__restrict__ fslot32,
                                                     const double *__restrict__ cscale = nullptr,
                                                     const double *__restrict__ vol = nullptr,
                                                     const double *__restrict__ farea = nullptr,
                                                     const double *__restrict__ xalt = nullptr,
                                                     const uint8_t *__restrict__ limf = nullptr) {
    static_assert(!(VIS && MIXED), "the viscous terms run in FP64 only");
    const int32_t tid = xcd_block() * blockDim.x + threadIdx.x;
    if (tid >= n)
        return;
    const int32_t f = BOUNDARY ? list[tid] : tid;
    const int32_t L = f2c[f], R = f2c[ldf + f];
    if (!BOUNDARY && R < 0)
        return;
    double uL[4], uR[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++)
        uL[k] = u[4 * int64_t(L) + k];
    // reconstructed increments D0 . x at the 3 points, both sides; the bases
    // D0 (row 0 of the DiBj blocks) are evaluated from the points' scaled
    // coordinates and the cell's moments, per side in c order as stored bases
    // would be read.  MIXED: the bases are rounded to FP32 and the products
    // accumulated in FP32.
    double incL[3][4], incR[3][4];
#pragma unroll
    for (int g = 0; g < 3; g++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            incL[g][v] = 0.0, incR[g][v] = 0.0;
    const dbl2 *xy = reinterpret_cast<const dbl2 *>(fxy) + f;
#pragma unroll 1
    for (int sd = 0; sd < (BOUNDARY ? 1 : 2); sd++) {
        const int32_t cell = sd ? R : L;
        // the steps' limiter leaves its unlimited cells in x and flags the others
        const double *xs = (limf && limf[cell]) ? xalt : x;
        double pX[3][4], pY[3][4];
#pragma unroll
        for (int g = 0; g < 3; g++) {
            const dbl2 q = __builtin_nontemporal_load(xy + (sd * 3 + g) * ldf);
            int_powers(q.x, pX[g]);
            int_powers(q.y, pY[g]);
        }
        double inc[3][4];
        float acc[3][4];
#pragma unroll
        for (int g = 0; g < 3; g++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                inc[g][v] = 0.0, acc[g][v] = 0.f;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const double m = cmom[c * ldc + cell];
            double xc[4];
#pragma unroll
            for (int v = 0; v < 4; v++)
                xc[v] = xs[(c + 9 * v) * ldc + cell];
#pragma unroll
            for (int g = 0; g < 3; g++) {
                const double d = pX[g][kBX[c + 1]] * pY[g][kBY[c + 1]] - m;
                if (MIXED) {
                    const float d32 = float(d);
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        acc[g][v] = fmaf(d32, float(xc[v]), acc[g][v]);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        inc[g][v] += d * xc[v];
                }
            }
        }
#pragma unroll
        for (int g = 0; g < 3; g++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double r = MIXED ? double(acc[g][v]) : inc[g][v];
                if (sd)
                    incR[g][v] = r;
                else
                    incL[g][v] = r;
            }
    }
    if (!BOUNDARY) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            uR[k] = u[4 * int64_t(R) + k];
    }
    const double dl = dlf[f];
    const int32_t bc = BOUNDARY ? fbc[f] : BC_INNER;
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
        acc[k] = 0.0;
    bool ok = true, okb = true;
    double n0, n1, nl, FL[4], FR[4];
    // the viscous variant keeps its points apart: unrolled, their gradient
    // work would be scheduled together and crowd the registers
#pragma unroll(VIS && !BOUNDARY ? 1 : 3)
    for (int g = 0; g < 3; g++) {
        double ULxy[4], URxy[4], URm[4];
        if (!STRAIGHT || g == 0) {
            const double nx = fn[(2 * g) * ldf + f], ny = fn[(2 * g + 1) * ldf + f];
            normalized(nx, ny, n0, n1);
            nl = sqrt(nx * nx + ny * ny);
        }
        compress_rec_part(uL, incL[g], ULxy);
        if (!BOUNDARY) {
            compress_rec_part(uR, incR[g], URxy);
#pragma unroll
            for (int k = 0; k < 4; k++)
                URm[k] = uR[k];
        } else {
            const double px = fctr[f], py = fctr[ldf + f];
            ok &= boundary_value(gp, ULxy, n0, n1, px, py, t, bc, URxy);
            ok &= boundary_value(gp, uL, n0, n1, px, py, t, bc, URm);
        }
        double F[4];
        if (!STRAIGHT || g == 0)
            okb = ok && balance_fluxes(gp, uL, URm, n0, n1, FL, FR);
        ok &= ok && okb && riemann_face(gp, ULxy, URxy, n0, n1, dl, F);
        if (VIS) {
            // the gradients of both sides at this point, formed here rather
            // than kept for all three points
            auto grad_at = [&](int32_t cell, int sd, double G[2][4]) {
                const dbl2 q = xy[(sd * 3 + g) * ldf];
                double px[4], py[4];
                int_powers(q.x, px);
                int_powers(q.y, py);
                const dbl2 sc = reinterpret_cast<const dbl2 *>(cscale)[cell];
                const double isx = 1.0 / sc.x, isy = 1.0 / sc.y; // the rows' scales as reciprocals
                const double *xs = (limf && limf[cell]) ? xalt : x;
#pragma unroll
                for (int d = 0; d < 2; d++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        G[d][v] = 0.0;
                // d/dx: bx X^(bx-1) Y^by / sx, d/dy: by X^bx Y^(by-1) / sy
#pragma unroll
                for (int c = 0; c < 9; c++) {
                    const int bx = kBX[c + 1], by = kBY[c + 1];
                    double xc[4];
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        xc[v] = xs[(c + 9 * v) * ldc + cell];
                    if (bx > 0) {
                        const double dx = double(bx) * px[bx - 1] * py[by] * isx;
#pragma unroll
                        for (int v = 0; v < 4; v++)
                            G[0][v] += dx * xc[v];
                    }
                    if (by > 0) {
                        const double dy = double(by) * px[bx] * py[by - 1] * isy;
#pragma unroll
                        for (int v = 0; v < 4; v++)
                            G[1][v] += dy * xc[v];
                    }
                }
            };
            double gL[2][4], gR[2][4];
            grad_at(L, 0, gL);
            if (!BOUNDARY)
                grad_at(R, 1, gR);
            double minVol = vol[L];
            if (!BOUNDARY)
                minVol = fmin(minVol, vol[R]);
            double distGRP = minVol / farea[f] * 2;
            distGRP += (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) ? 3e200 : 0.0;
            double GM[2][4], UM[4];
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const double nd = d ? n1 : n0;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const double gr = BOUNDARY ? gL[d][v] : gR[d][v];
                    GM[d][v] = (gr + gL[d][v]) * 0.5 + (1.0 / distGRP) * (nd * (URxy[v] - ULxy[v]));
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++)
                UM[v] = 0.5 * (ULxy[v] + URxy[v]);
            const double vm0 = UM[1] / UM[0], vm1 = UM[2] / UM[0];
            double pM, aM, HM;
            ideal_gas_thermal(UM[3], UM[0], vm0 * vm0 + vm1 * vm1, gp.gamma, pM, aM, HM);
            const double muf = sutherland_mu(gp, pM, UM[0]);
            const double k = gp.cp * (muf - muf) / 0.9 + gp.cp * muf / gp.pr;
            double vn[4];
            viscous_flux_n(UM, GM, n0, n1, bc == BC_WALL_NOSLIP, gp.gamma, muf, k, gp.cp, vn);
#pragma unroll
            for (int v = 0; v < 4; v++)
                F[v] += vn[v];
        }
        const double w = g == 1 ? kLine3W1 : kLine3W0;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double a = F[v] * nl, b = FL[v] * nl, c = FR[v] * nl;
            acc[v] += a * w;
            acc[4 + v] += b * w;
            acc[8 + v] += c * w;
        }
    }
    if (!ok)
        atomicOr(errw, ERRBIT_NONPOSITIVE);
    // each side's two terms go to its cell's slot for this face, signed as
    // the cell applies them (negation is exact, so the cell's sums are
    // unchanged bit for bit)
    const int32_t pos = fpos[f];
    const int64_t oL = int64_t(pos & 15) * 8 * ldc + L, oR = int64_t(pos >> 4) * 8 * ldc + R;
    if (MIXED) {
#pragma unroll
        for (int v = 0; v < 4; v++)
            fslot32[oL + v * ldc] = float(acc[v]), fslot32[oL + (4 + v) * ldc] = -float(acc[4 + v]);
        if (!BOUNDARY)
#pragma unroll
            for (int v = 0; v < 4; v++)
                fslot32[oR + v * ldc] = -float(acc[v]), fslot32[oR + (4 + v) * ldc] = float(acc[8 + v]);
        return;
    }
    // plain stores (with non-temporal ones this kernel ran 19 % slower)
#pragma unroll
    for (int v = 0; v < 4; v++)
        fslot[oL + v * ldc] = acc[v], fslot[oL + (4 + v) * ldc] = -acc[4 + v];
    if (!BOUNDARY)
#pragma unroll
        for (int v = 0; v < 4; v++)
            fslot[oR + v * ldc] = -acc[v], fslot[oR + (4 + v) * ldc] = acc[8 + v];
}

// EvaluateRHS scatter as a per-cell gather in ascending face order
// (DNDS_FV_EulerEvaluator.cpp:501-508): the face kernel has stored each
// face's two terms at the cell's slot for that face, slots in ascending face
// order, signed so that the cell only adds (F/V then -FLfix/V for side 0,
// -F/V then FRfix/V for side 1).  Optionally fused with the SSP-RK3 stage
// update of ExplicitSSPRK3LocalDt::Step (DNDS_ODE.hpp:136-157):
//   stage 0: rhs = gathered value
//   stage 1: xLast = x; x = x + rhs dt
//   stage 2/3: x = ((x + rhs dt) c/(1-c) + xLast)(1-c), c = 1/4, 2/3
template <int NF, typename FT = double>
__global__ __launch_bounds__(kBlock) void k_rhs_cell(int32_t nc, int64_t ldc, const int32_t *__restrict__ gat,
                                                     const FT *__restrict__ fslot,
                                                     const double *__restrict__ vol, double *__restrict__ rhs,
                                                     int stage, double *__restrict__ x,
                                                     double *__restrict__ xlast, const double *__restrict__ dt,
                                                     const double *__restrict__ dtmin, double coefA, double coefB,
                                                     int32_t *errw) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    const FT V = FT(vol[i]);
    // a cell has 3 or 4 faces: only the 4-slot layout has an absent last slot
    const int ns = NF == 4 && gat[3 * ldc + i] < 0 ? 3 : NF;
    FT r[4] = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < NF; s++) {
        if (s >= ns)
            break;
#pragma unroll
        for (int v = 0; v < 4; v++)
            r[v] += fslot[(s * 8 + v) * ldc + i] / V;
#pragma unroll
        for (int v = 0; v < 4; v++)
            r[v] += fslot[(s * 8 + 4 + v) * ldc + i] / V;
    }
    bool finite = true;
#pragma unroll
    for (int v = 0; v < 4; v++)
        finite &= isfinite(r[v]);
    if (!finite)
        atomicOr(errw, ERRBIT_NAN);
    if (stage == 0) {
#pragma unroll
        for (int v = 0; v < 4; v++)
            rhs[4 * int64_t(i) + v] = double(r[v]);
        return;
    }
    const double d = dt ? dt[i] : dtmin[0];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double xo = x[4 * int64_t(i) + v];
        double xv = xo + double(r[v]) * d;
        if (stage == 1) {
            xlast[4 * int64_t(i) + v] = xo;
        } else {
            xv *= coefA;
            xv += xlast[4 * int64_t(i) + v];
            xv *= coefB;
        }
        x[4 * int64_t(i) + v] = xv;
    }
}


// WBAP V3 step 0, face part (DNDS_FV_VR.hpp:2084-2126): for each inner face
// the (IJI, ISI) diagonals of the (rho, E) jump functional at the first
// LINE_3 point with the LINE_1 weight 2, for both adjacent cells.  As in the
// reference, a cell evaluates the face's side-0 basis with its own
// coefficients and the side-1 basis with its neighbour's.
__global__ __launch_bounds__(kBlock) void k_wbap_face(int32_t nf, int64_t ldf, int64_t ldc,
                                                      const int32_t *__restrict__ f2c,
                                                      const double *__restrict__ fxy, const double *__restrict__ cmom,
                                                      const double *__restrict__ cscale,
                                                      const double *__restrict__ fw,
                                                      const double *__restrict__ fn, const double *__restrict__ u,
                                                      const double *__restrict__ x, double tang_weight, int twd_norm,
                                                      const uint8_t *__restrict__ wpos, double *__restrict__ wslot) {
    const int32_t f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const int32_t L = f2c[f], R = f2c[ldf + f];
    if (R < 0)
        return;
    // The DiBj blocks of both sides at point 0 (side 0: cell L's basis,
    // side 1: cell R's), evaluated from the point's scaled coordinates and
    // the cells' scales and moments; zero entries are skipped (adding them
    // would not change any sum).  Cell side cs sees its own coefficients on
    // the side-0 basis and the neighbour's on the side-1 basis.  The rows are
    // taken in the functional's groups {0}, {1, 2}, {3, 4, 5}, {6..9}: one
    // pass over the coefficients per group, whose terms then join the
    // indicator sums in their order.
    const dbl2 *xy = reinterpret_cast<const dbl2 *>(fxy) + f;
    const dbl2 q0 = __builtin_nontemporal_load(xy), q1 = __builtin_nontemporal_load(xy + 3 * ldf);
    const dbl2 *sc = reinterpret_cast<const dbl2 *>(cscale);
    const dbl2 s0 = sc[L], s1 = sc[R];
    double pX0[4], pY0[4], pX1[4], pY1[4], sX0[4], sY0[4], sX1[4], sY1[4];
    int_powers(q0.x, pX0), int_powers(q0.y, pY0), int_powers(q1.x, pX1), int_powers(q1.y, pY1);
    int_powers(s0.x, sX0), int_powers(s0.y, sY0), int_powers(s1.x, sX1), int_powers(s1.y, sY1);
    double W[10];
#pragma unroll
    for (int k = 0; k < 10; k++)
        W[k] = fw[k * ldf + f];
    const double gx = fn[f], gy = fn[ldf + f];
    const FFParams P = ff_params(W, gx, gy, tang_weight, twd_norm);
    const double nl = sqrt(gx * gx + gy * gy);
    const double uL[2] = {u[4 * int64_t(L)], u[4 * int64_t(L) + 3]};
    const double uR[2] = {u[4 * int64_t(R)], u[4 * int64_t(R) + 3]};
    double sums[2][2][2]; // [cell side][rho, E][IJI, ISI]
#pragma unroll
    for (int k = 0; k < 8; k++)
        (&sums[0][0][0])[k] = 0.0;
    auto pass = [&](auto group) {
        constexpr int G = decltype(group)::value;
        constexpr int r0 = G == 0 ? 0 : G == 1 ? 1 : G == 2 ? 3 : 6, nr = G == 0 ? 1 : G == 1 ? 2 : G == 2 ? 3 : 4;
        // each pass reloads the coefficients (cache hits) rather than keeping
        // all of them in registers across the passes
        __asm__ volatile("" ::: "memory");
        // b[0]: side-0 basis x L, b[1]: side-0 x R, b[2]: side-1 x R, b[3]: side-1 x L
        double b[4][nr][2];
#pragma unroll
        for (int k = 0; k < 4 * nr * 2; k++)
            (&b[0][0][0])[k] = 0.0;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const double xL[2] = {x[c * ldc + L], x[(c + 27) * ldc + L]};
            const double xR[2] = {x[c * ldc + R], x[(c + 27) * ldc + R]};
            const double m0 = G == 0 ? cmom[c * ldc + L] : 0.0, m1 = G == 0 ? cmom[c * ldc + R] : 0.0;
#pragma unroll
            for (int rr = 0; rr < nr; rr++) {
                const int r = r0 + rr;
                const int coef = dibj_coef(r, c + 1);
                if (coef == 0)
                    continue;
                const int ex = kBX[c + 1] - kBX[r], ey = kBY[c + 1] - kBY[r];
                double d0 = double(coef) * pX0[ex] * pY0[ey] * 1.0;
                double d1 = double(coef) * pX1[ex] * pY1[ey] * 1.0;
                if (r == 0) {
                    d0 -= m0, d1 -= m1;
                } else {
                    d0 = d0 / (sX0[kBX[r]] * sY0[kBY[r]] * 1.0);
                    d1 = d1 / (sX1[kBX[r]] * sY1[kBY[r]] * 1.0);
                }
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    b[0][rr][q] += d0 * xL[q];
                    b[1][rr][q] += d0 * xR[q];
                    b[2][rr][q] += d1 * xR[q];
                    b[3][rr][q] += d1 * xL[q];
                }
            }
        }
#pragma unroll
        for (int cs = 0; cs < 2; cs++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                double val[nr], jmp[nr];
#pragma unroll
                for (int rr = 0; rr < nr; rr++) {
                    double vl = b[cs == 0 ? 0 : 1][rr][q], vr = b[cs == 0 ? 2 : 3][rr][q];
                    if (r0 + rr == 0) {
                        vl += cs == 0 ? uL[q] : uR[q];
                        vr += cs == 0 ? uR[q] : uL[q];
                    }
                    val[rr] = (vl + vr) * 0.5;
                    jmp[rr] = (vl - vr) * 0.5;
                }
                ff_group<G>(P, jmp, sums[cs][q][0]);
                ff_group<G>(P, val, sums[cs][q][1]);
            }
    };
    pass(std::integral_constant<int, 0>{});
    pass(std::integral_constant<int, 1>{});
    pass(std::integral_constant<int, 2>{});
    pass(std::integral_constant<int, 3>{});
#pragma unroll
    for (int cs = 0; cs < 2; cs++) {
        double out[4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            double a = sums[cs][q][0], b = sums[cs][q][1];
            a *= nl, b *= nl;
            out[2 * q] = a * 2.0;
            out[2 * q + 1] = b * 2.0;
        }
        // into the cell's own slot of this face, [slot][4][cell], so that the
        // indicator reads its terms as a coalesced stream
        const int32_t cell = cs ? R : L;
        const int slot = cs ? (wpos[f] >> 4) : (wpos[f] & 15);
#pragma unroll
        for (int k = 0; k < 4; k++)
            wslot[(int64_t(slot) * 4 + k) * ldc + cell] = out[k];
    }
}

// WBAP V3 smoothness indicator: sum over the cell's faces in cell2face order
// (DNDS_FV_VR.hpp:2125-2137).  Cells below WBAP_SmoothIndicatorScale (unless
// ifAll) keep their reconstruction; the others are appended to `list`
// (one atomic per wavefront; the order of the list does not affect any
// result, each cell is limited independently).
template <int NF>
__global__ __launch_bounds__(kBlock) void k_wbap_ind(int32_t nc, int64_t ldc, const int32_t *__restrict__ nbr,
                                                     const double *__restrict__ wslot, const double *__restrict__ x,
                                                     double *__restrict__ xo, double *__restrict__ ind, double scale,
                                                     int if_all, int32_t *__restrict__ count,
                                                     int32_t *__restrict__ list, uint8_t *__restrict__ limf) {
    const int32_t i = xcd_block() * blockDim.x + threadIdx.x;
    bool lim = false;
    if (i < nc) {
        double sum[4] = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < NF; s++) {
            const int32_t j = nbr[s * ldc + i];
            if (j < 0)
                continue;
#pragma unroll
            for (int k = 0; k < 4; k++)
                sum[k] += wslot[(int64_t(s) * 4 + k) * ldc + i];
        }
        const double si0 = sum[0] / (sum[1] + kVerySmallReal);
        const double si1 = sum[2] / (sum[3] + kVerySmallReal);
        const double flag = sqrt(fmax(fabs(si0), fabs(si1))) * (3 * 3);
        ind[i] = flag;
        lim = !(flag < scale && !if_all);
        if (limf)
            limf[i] = uint8_t(lim);
        else if (!lim) {
#pragma unroll
            for (int k = 0; k < 36; k++)
                xo[k * ldc + i] = x[k * ldc + i];
        }
    }
    const uint64_t mask = __ballot(lim);
    if (mask == 0)
        return;
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((unsigned long long)mask) - 1;
    int32_t base = 0;
    if (lane == first)
        base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, first);
    if (lim)
        list[base + __popcll(mask & ((1ull << lane) - 1))] = i;
}

// WBAP V3 per-order characteristic limiting of the listed cells
// (DNDS_FV_VR.hpp:2152-2317).  Reads x, writes xo.  uF is per-cell scratch
// holding the limited face values of rows 2..8 of the orders already done.
template <int NF>
__global__ __launch_bounds__(kBlock) void k_wbap_lim(int64_t ldc, int64_t ldf, const int32_t *__restrict__ count,
                                                     const int32_t *__restrict__ list,
                                                     const int32_t *__restrict__ nbr,
                                                     const int32_t *__restrict__ sface,
                                                     const int32_t *__restrict__ f2c,
                                                     const double *__restrict__ fnc,
                                                     const double *__restrict__ msec, const double *__restrict__ u,
                                                     const double *__restrict__ x, double *__restrict__ xo,
                                                     double *__restrict__ uF, double gamma, double nStd, double eps4,
                                                     int32_t *errw) {
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= count[0])
        return;
    const int32_t i = list[tid];
    double ui[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        ui[v] = u[4 * int64_t(i) + v];
    bool ok = true;
#pragma unroll
    for (int order = 3; order >= 1; order--) {
        const int ls = order == 3 ? 5 : (order == 2 ? 2 : 0);
        const int le = order == 3 ? 8 : (order == 2 ? 4 : 1);
        constexpr int kMaxR = 4;
        const int nr = le - ls + 1;
        double u0[kMaxR][4], up[kMaxR][4], down[kMaxR][4];
#pragma unroll
        for (int r = 0; r < kMaxR; r++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                u0[r][v] = r < nr ? x[((ls + r) + 9 * v) * ldc + i] : 0.0;
                up[r][v] = 0.0, down[r][v] = 0.0;
            }
#pragma unroll 1
        for (int s = 0; s < NF; s++) {
            const int32_t j = nbr[s * ldc + i];
            if (j < 0)
                continue;
            const int32_t f = sface[s * ldc + i];
            const int side = f2c[f] == i ? 0 : 1;
            double un0, un1;
            stable_normalized(fnc[f], fnc[ldf + f], un0, un1);
            const double *ms = msec + int64_t(81 * side) * ldf + f;
            const double *msO = msec + int64_t(81 * (1 - side)) * ldf + f;
            double oIn[kMaxR][4];
#pragma unroll
            for (int r = 0; r < kMaxR; r++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    oIn[r][v] = 0.0;
            // uOtherOther rows: the neighbour's coefficients up to le, above le
            // the limited face values of the higher orders mapped to its basis
#pragma unroll
            for (int c = 0; c < 9; c++) {
                double uoo[4];
                if (c <= le) {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        uoo[v] = x[(c + 9 * v) * ldc + j];
                } else {
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
#pragma unroll
                        for (int cc = le + 1; cc < 9; cc++)
                            acc += msO[(9 * c + cc) * ldf] * uF[((s * 7 + (cc - 2)) * 4 + v) * ldc + i];
                        uoo[v] = acc;
                    }
                }
#pragma unroll
                for (int r = 0; r < kMaxR; r++)
                    if (r < nr) {
                        const double m = ms[(9 * (ls + r) + c) * ldf];
#pragma unroll
                        for (int v = 0; v < 4; v++)
                            oIn[r][v] += m * uoo[v];
                    }
            }
            const double *uLp = side ? u + 4 * int64_t(j) : ui;
            const double *uRp = side ? ui : u + 4 * int64_t(j);
            double UL[4], UR[4];
#pragma unroll
            for (int v = 0; v < 4; v++)
                UL[v] = uLp[v], UR[v] = uRp[v];
            double ML[16], MR[16];
            ok &= char_matrices(gamma, UL, UR, un0, un1, ML, MR);
#pragma unroll
            for (int r = 0; r < kMaxR; r++) {
                if (r >= nr)
                    continue;
                double oC[4], tC[4], out[4];
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    double p = 0, q = 0;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        p += ML[4 * a + b] * oIn[r][b];
                        q += ML[4 * a + b] * u0[r][b];
                    }
                    oC[a] = p, tC[a] = q;
                }
                // FWBAP_L2_Biway(uThisIn, uOtherIn, 1)
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    const double u1 = tC[a], u2 = oC[a];
                    const double frac = u1 / (fabs(u2) + 1e-12) * sgn(u2);
                    const double th1 = frac * frac * frac;
                    const double th2 = (frac * frac) * (frac * frac);
                    out[a] = u1 * (1.0 + th1) / (1.0 + th2);
                }
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    double xv = 0;
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        xv += MR[4 * a + b] * out[b];
                    if (ls + r >= 2)
                        uF[((s * 7 + (ls + r - 2)) * 4 + a) * ldc + i] = xv;
                    // FWBAP_L2_Multiway accumulation, neighbours in cell2face order
                    const double ti = u0[r][a] / (fabs(xv) + eps4) * sgn(xv);
                    down[r][a] += (ti * ti) * (ti * ti);
                    up[r][a] += ti * ti * ti;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kMaxR; r++)
            if (r < nr)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    xo[((ls + r) + 9 * v) * ldc + i] = u0[r][v] * (up[r][v] + nStd) / (down[r][v] + nStd);
    }
    if (!ok)
        atomicOr(errw, ERRBIT_NONPOSITIVE);
}

// WBAP V2 limiting of the listed cells for one order (DNDS_FV_VR.hpp:1843-2040,
// rows ls..le of the order).  xo is the reference's uRecNewBuf: before this
// launch it holds x for the unlimited cells and, for the limited ones, the rows
// of the higher orders as limited by the earlier launches; rows ls..le and
// below are still x everywhere.  Each neighbour's full 9-row buffer is mapped
// to this cell's basis (rows <= le read from x, rows > le from xo), so the
// launch reads xo only above le and writes it only at ls..le: no cell reads
// what another writes.  FWBAP_L2_Biway with n = nStd per neighbour, then
// FWBAP_L2_Multiway centred on the first neighbour's result with n1 = 1.
template <int NF, int ORDER>
__global__ __launch_bounds__(kBlock) void k_wbap_lim_v2(int64_t ldc, int64_t ldf, const int32_t *__restrict__ count,
                                                        const int32_t *__restrict__ list,
                                                        const int32_t *__restrict__ nbr,
                                                        const int32_t *__restrict__ sface,
                                                        const int32_t *__restrict__ f2c,
                                                        const double *__restrict__ fnc,
                                                        const double *__restrict__ msec,
                                                        const double *__restrict__ u, const double *__restrict__ x,
                                                        double *__restrict__ xo, double gamma, double nStd,
                                                        double eps4, int32_t *errw,
                                                        const uint8_t *__restrict__ limf) {
    constexpr int ls = ORDER == 3 ? 5 : (ORDER == 2 ? 2 : 0);
    constexpr int le = ORDER == 3 ? 8 : (ORDER == 2 ? 4 : 1);
    constexpr int nr = le - ls + 1;
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= count[0])
        return;
    const int32_t i = list[tid];
    double ui[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        ui[v] = u[4 * int64_t(i) + v];
    double uC[nr][4], first[nr][4], up[nr][4], down[nr][4];
#pragma unroll
    for (int r = 0; r < nr; r++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            uC[r][v] = x[((ls + r) + 9 * v) * ldc + i];
            first[r][v] = 0.0, up[r][v] = 0.0, down[r][v] = 0.0;
        }
    bool ok = true;
    int nOther = 0;
#pragma unroll 1
    for (int s = 0; s < NF; s++) {
        const int32_t j = nbr[s * ldc + i];
        if (j < 0)
            continue;
        const int32_t f = sface[s * ldc + i];
        const int side = f2c[f] == i ? 0 : 1;
        double un0, un1;
        stable_normalized(fnc[f], fnc[ldf + f], un0, un1);
        const double *ms = msec + int64_t(81 * side) * ldf + f;
        // the neighbour's rows above le: limited (xo) or, with limf, as given (x)
        const double *above = (limf && !limf[j]) ? x : xo;
        double oIn[nr][4];
#pragma unroll
        for (int r = 0; r < nr; r++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                oIn[r][v] = 0.0;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const double *src = c <= le ? x : above;
            double uj[4];
#pragma unroll
            for (int v = 0; v < 4; v++)
                uj[v] = src[(c + 9 * v) * ldc + j];
#pragma unroll
            for (int r = 0; r < nr; r++) {
                const double m = ms[(9 * (ls + r) + c) * ldf];
#pragma unroll
                for (int v = 0; v < 4; v++)
                    oIn[r][v] += m * uj[v];
            }
        }
        const double *uLp = side ? u + 4 * int64_t(j) : ui;
        const double *uRp = side ? ui : u + 4 * int64_t(j);
        double UL[4], UR[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
            UL[v] = uLp[v], UR[v] = uRp[v];
        double ML[16], MR[16];
        ok &= char_matrices(gamma, UL, UR, un0, un1, ML, MR);
#pragma unroll
        for (int r = 0; r < nr; r++) {
            double oC[4], tC[4], out[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                double p = 0, q = 0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    p += ML[4 * a + b] * oIn[r][b];
                    q += ML[4 * a + b] * uC[r][b];
                }
                oC[a] = p, tC[a] = q;
            }
            // FWBAP_L2_Biway(uThisIn, uOtherIn, nStd)
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const double u1 = tC[a], u2 = oC[a];
                const double frac = u1 / (fabs(u2) + 1e-12) * sgn(u2);
                const double th1 = frac * frac * frac;
                const double th2 = (frac * frac) * (frac * frac);
                out[a] = u1 * (nStd + th1) / (nStd + th2);
            }
#pragma unroll
            for (int a = 0; a < 4; a++) {
                double xv = 0;
#pragma unroll
                for (int b = 0; b < 4; b++)
                    xv += MR[4 * a + b] * out[b];
                if (nOther == 0) {
                    first[r][a] = xv;
                } else {
                    const double ti = first[r][a] / (fabs(xv) + eps4) * sgn(xv);
                    down[r][a] += (ti * ti) * (ti * ti);
                    up[r][a] += ti * ti * ti;
                }
            }
        }
        nOther++;
    }
#pragma unroll
    for (int r = 0; r < nr; r++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            xo[((ls + r) + 9 * v) * ldc + i] =
                nOther ? first[r][v] * (up[r][v] + 1.0) / (down[r][v] + 1.0) : uC[r][v];
    if (!ok)
        atomicOr(errw, ERRBIT_NONPOSITIVE);
}

__global__ void k_zero_int2(int32_t *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        p[0] = 0;
}

__global__ __launch_bounds__(kBlock) void k_set(int64_t n, double *__restrict__ p, double v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_zero_int(int32_t *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        p[0] = 0;
}

// ========================================================================
// host side
// ========================================================================

namespace dndsmi_host {

int check_errw(dndsmi_ctx *c) {
    int32_t h = 0;
    DNDSMI_TRY(c->hip(hipMemcpyAsync(&h, c->errw, sizeof(h), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync"));
    DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    if (h == 0)
        return DNDSMI_OK;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->errw);
    if (h & ERRBIT_NONPOSITIVE) {
        c->err = "non-positive density or sound speed in a flux or boundary state";
        return DNDSMI_ERR_NONPOSITIVE_STATE;
    }
    if (h & ERRBIT_NAN) {
        c->err = "non-finite value in the right-hand side or time step";
        return DNDSMI_ERR_NAN;
    }
    c->err = "boundary face of an unsupported kind";
    return DNDSMI_ERR_UNSUPPORTED;
}

bool same_ctx(const dndsmi_ctx *c, const dndsmi_array *a, int32_t kind) { return a && a->ctx == c && a->kind == kind; }

} // namespace dndsmi_host

extern "C" int dndsmi_ctx_create(const dndsmi_mesh_view *mesh, const dndsmi_vr_view *vr,
                                 const dndsmi_vr_setting *vs, const dndsmi_euler_setting *es, void *stream,
                                 dndsmi_ctx **out) {
    if (!mesh || !vr || !vs || !es || !out)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    {
        std::string why;
        if (check_mesh_view(mesh, why) != DNDSMI_OK)
            return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    if (vr->n_cell != mesh->n_cell || vr->n_face != mesh->n_face || vr->n_c2f != mesh->cell_ptr[mesh->n_cell] ||
        mesh->n_cell > (int64_t(1) << 30) || mesh->n_face > (int64_t(1) << 30))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (!vr->face_gauss_xy || !vr->cell_scale || !vr->base_moment || !vr->face_d0)
        return DNDSMI_ERR_INVALID_ARGUMENT; // the device evaluates the face bases from these
    if (!vr->volume || !vr->face_area || !vr->face_center || !vr->face_norm_center || !vr->face_norm ||
        !vr->face_weight || !vr->ainv || !vr->ainv_b || !vr->ainv_bvec)
        return DNDSMI_ERR_INVALID_ARGUMENT; // matrix_secondary alone may be absent (no WBAP then)
    if (es->riemann_solver != DNDSMI_RS_ROE && es->riemann_solver != DNDSMI_RS_HLLC &&
        es->riemann_solver != DNDSMI_RS_HLLEP &&
        !(es->riemann_solver >= DNDSMI_RS_ROE_M1 && es->riemann_solver <= DNDSMI_RS_ROE_M5))
        return DNDSMI_ERR_UNSUPPORTED;
    // local coordinates: the kernels evaluate the bases' values only
    if (vr->cell_frame && (es->mu_gas > 0 || vr->matrix_secondary))
        return DNDSMI_ERR_UNSUPPORTED;
    if (!(es->mu_gas >= 0) ||
        (es->mu_gas > 0 && !(es->pr_gas > 0 && es->r_gas > 0 && es->t_ref > 0 && es->c_sutherland >= 0)))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    dndsmi_ctx *c = new (std::nothrow) dndsmi_ctx();
    if (!c)
        return DNDSMI_ERR_OUT_OF_MEMORY;
    auto fail = [&](int rc) {
        dndsmi_ctx_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess)
        return fail(DNDSMI_ERR_HIP);
    if (noalloc_scratch_init(c) != DNDSMI_OK)
        return fail(DNDSMI_ERR_HIP);
    if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
            return fail(DNDSMI_ERR_HIP);
        c->own_stream = true;
    }
    const int64_t nc = mesh->n_cell, nf = mesh->n_face;
    c->nc = nc, c->nf = nf, c->ldc = pad64(nc), c->ldf = pad64(nf);
    c->relax = vs->jacobi_relax;
    c->sor = vs->sor_instead != 0;
    c->sor_inverse = c->sor && vs->sor_inverse_scanning != 0;
    c->gas.gamma = es->gamma;
    for (int k = 0; k < 4; k++)
        c->gas.far[k] = es->far_field[k];
    c->gas.rs = es->riemann_solver;
    c->gas.mu = es->mu_gas;
    c->gas.pr = es->pr_gas;
    c->gas.cp = es->r_gas * es->gamma / (es->gamma - 1);
    c->gas.tref = es->t_ref;
    c->gas.csuth = es->c_sutherland;
    int nfmax = 0;
    for (int64_t i = 0; i < nc; i++)
        nfmax = std::max<int>(nfmax, int(mesh->cell_ptr[i + 1] - mesh->cell_ptr[i]));
    if (nfmax != 3 && nfmax != 4)
        return fail(DNDSMI_ERR_INVALID_ARGUMENT);
    c->nfmax = nfmax;
    const int64_t ldc = c->ldc, ldf = c->ldf;
    try {
        // ---- device numbering.  Every device array is indexed through c2d /
        // f2d (host arrays keep the mesh numbering).  The mesh's own order is
        // kept: on the structured-as-unstructured meshes it is row-major, and
        // measured faster for the sweeps than Morton or strip orders (whose
        // neighbour loads coalesce worse); a mesh read in a scattered order
        // would be renumbered here.  On a mesh of triangles and quads, each
        // window of 64 consecutive cells (one wavefront) puts its triangles
        // before its quads, keeping their order: the fourth slot's A^-1 B
        // vectors of a block are then contiguous over its quads, and the
        // 128-byte lines of that stream hold no unused triangle slots.
        //
        // With SOR_Instead the numbering is colour-major instead: each colour
        // is a contiguous range (its cells in mesh order), so that a colour's
        // pass streams whole lines of the A^-1 B pairs and coefficients; in
        // each colour the cells with terms on their own partial sum (far
        // field, Wall_Euler, Wall_NoSlip) come last and take their own kernel
        // variant.  The windows of 64 then start at each range's start.
        //
        // Colouring (multicolour LU-SGS and SOR): each cell takes the
        // smallest colour that no lower-numbered neighbour holds.
        std::vector<int32_t> col(size_t(nc), -1);
        int32_t ncol = 0;
        for (int64_t i = 0; i < nc; i++) { // mesh numbering
            uint32_t used = 0;
            for (int64_t q = mesh->cell_ptr[i]; q < mesh->cell_ptr[i + 1]; q++) {
                const int64_t f = mesh->cell2face[q];
                const int64_t L = mesh->face2cell[2 * f], R = mesh->face2cell[2 * f + 1];
                const int64_t j = L == i ? R : L;
                if (j != DNDSMI_NO_CELL && j < i && col[j] < 32)
                    used |= 1u << col[j];
            }
            int32_t k = 0;
            while (used & (1u << k))
                k++;
            col[i] = k;
            ncol = std::max(ncol, k + 1);
        }
        std::vector<int32_t> &c2d = c->c2d, &f2d = c->f2d;
        {
            // the numbering's segments (mesh cells in order) and their ends
            std::vector<int64_t> seq(static_cast<size_t>(nc)), seg_end;
            if (c->sor) {
                std::vector<uint8_t> self(size_t(nc), 0);
                for (int64_t i = 0; i < nc; i++)
                    for (int64_t q = mesh->cell_ptr[i]; q < mesh->cell_ptr[i + 1]; q++) {
                        const int64_t f = mesh->cell2face[q];
                        const int64_t L = mesh->face2cell[2 * f], R = mesh->face2cell[2 * f + 1];
                        if ((L == i ? R : L) == DNDSMI_NO_CELL && mesh->face_bc[f] != DNDSMI_BC_WALL)
                            self[size_t(i)] = 1;
                    }
                std::vector<int64_t> cnt(2 * size_t(ncol) + 1, 0);
                for (int64_t i = 0; i < nc; i++)
                    cnt[2 * col[i] + self[size_t(i)] + 1]++;
                for (size_t g = 0; g + 1 < cnt.size(); g++)
                    cnt[g + 1] += cnt[g];
                seg_end.assign(cnt.begin() + 1, cnt.end());
                std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
                for (int64_t i = 0; i < nc; i++)
                    seq[fill[2 * col[i] + self[size_t(i)]]++] = i;
                c->sor_start.assign(size_t(ncol) + 1, nc);
                c->sor_wall.assign(size_t(ncol), nc);
                for (int32_t k = 0; k < ncol; k++)
                    c->sor_start[k] = cnt[2 * k], c->sor_wall[k] = cnt[2 * k + 1];
            } else {
                for (int64_t i = 0; i < nc; i++)
                    seq[i] = i;
                seg_end.push_back(nc);
            }
            c2d.resize(static_cast<size_t>(nc));
            int64_t a = 0;
            for (const int64_t b : seg_end) {
                for (int64_t w0 = a; w0 < b; w0 += 64) {
                    const int64_t w1 = std::min<int64_t>(b, w0 + 64);
                    int64_t d = w0;
                    for (int pass = 0; pass < 2; pass++)
                        for (int64_t p = w0; p < w1; p++) {
                            const int64_t i = seq[p];
                            const bool quad = mesh->cell_ptr[i + 1] - mesh->cell_ptr[i] == 4;
                            if (quad == (pass == 1))
                                c2d[i] = int32_t(d++);
                        }
                }
                a = b;
            }
            // faces in the order of their first cell in the mesh (the
            // partition of a window above moves cells by less than a
            // wavefront, and the face kernels gather best in mesh order)
            std::vector<std::pair<int64_t, int64_t>> fk(static_cast<size_t>(nf));
            for (int64_t f = 0; f < nf; f++)
                fk[f] = {mesh->face2cell[2 * f], f};
            std::sort(fk.begin(), fk.end());
            f2d.assign(static_cast<size_t>(nf), 0);
            for (int64_t d = 0; d < nf; d++)
                f2d[fk[d].second] = int32_t(d);
        }
        // ---- cell streams
        std::vector<int32_t> nbr(size_t(nfmax) * ldc, -2), sbc(size_t(nfmax) * ldc, 0), sface(size_t(nfmax) * ldc, -1);
        std::vector<int32_t> gat(size_t(nfmax) * ldc, -1), widx(size_t(nfmax) * ldc, -1);
        std::vector<uint8_t> wpos(ldf, 0);
        std::vector<double> M(size_t(nfmax) * kMPairs * 2 * ldc, 0.0), bv(size_t(nfmax) * 9 * ldc, 0.0), vol(ldc, 1.0);
        std::vector<double> cmom(9 * size_t(ldc), 0.0), cscale(2 * size_t(ldc), 1.0);
        std::vector<int64_t> wall_cell, wall_face, wall_c2f;
        for (int64_t i = 0; i < nc; i++) {
            const int64_t b0 = mesh->cell_ptr[i];
            const int n = int(mesh->cell_ptr[i + 1] - b0);
            const int64_t ci = c2d[i]; // device column of mesh cell i
            vol[ci] = vr->volume[i];
            for (int k = 0; k < 9; k++)
                cmom[k * ldc + ci] = vr->base_moment[10 * i + 1 + k];
            cscale[2 * ci] = vr->cell_scale[2 * i], cscale[2 * ci + 1] = vr->cell_scale[2 * i + 1];
            std::vector<std::pair<int64_t, int>> order;
            for (int s = 0; s < n; s++) {
                const int64_t f = mesh->cell2face[b0 + s];
                const int64_t L = mesh->face2cell[2 * f], R = mesh->face2cell[2 * f + 1];
                const int side = L == i ? 0 : 1;
                const int64_t j = side == 0 ? R : L;
                nbr[s * ldc + ci] = j == DNDSMI_NO_CELL ? -1 : c2d[j];
                sbc[s * ldc + ci] = mesh->face_bc[f];
                sface[s * ldc + ci] = f2d[f];
                if (j != DNDSMI_NO_CELL)
                    wpos[f2d[f]] |= uint8_t(s << (4 * side));
                for (int k = 0; k < 81; k++)
                    M[size_t(m_vec_index(s, k / 2, ci, ldc, kMPairs)) * 2 + k % 2] = vr->ainv_b[81 * (b0 + s) + k];
                for (int k = 0; k < 9; k++)
                    bv[(size_t(s) * 9 + k) * ldc + ci] = vr->ainv_bvec[9 * (b0 + s) + k];
                order.push_back({f, side});
                if (j == DNDSMI_NO_CELL &&
                    (mesh->face_bc[f] == DNDSMI_BC_WALL_EULER || mesh->face_bc[f] == DNDSMI_BC_WALL_NOSLIP)) {
                    widx[s * ldc + ci] = int32_t(wall_cell.size());
                    wall_cell.push_back(i), wall_face.push_back(f), wall_c2f.push_back(b0 + s);
                }
            }
            // the gather visits the faces in the mesh's face order (the
            // reference's scatter order), whatever their device numbers
            std::sort(order.begin(), order.end());
            for (int s = 0; s < n; s++)
                gat[s * ldc + ci] = int32_t((int64_t(f2d[order[s].first]) << 1) | order[s].second);
        }
        DNDSMI_TRY(c->alloc(c->nbr, nfmax * ldc));
        DNDSMI_TRY(c->upload(c->nbr, nbr));
        DNDSMI_TRY(c->alloc(c->sbc, nfmax * ldc));
        DNDSMI_TRY(c->upload(c->sbc, sbc));
        DNDSMI_TRY(c->alloc(c->sface, nfmax * ldc));
        DNDSMI_TRY(c->upload(c->sface, sface));
        DNDSMI_TRY(c->alloc(c->gat, nfmax * ldc));
        DNDSMI_TRY(c->upload(c->gat, gat));
        DNDSMI_TRY(c->alloc(c->wpos, ldf));
        DNDSMI_TRY(c->upload(c->wpos, wpos));
        {
            std::vector<uint8_t> fpos(ldf, 0);
            for (int s = 0; s < nfmax; s++)
                for (int64_t ci = 0; ci < nc; ci++) {
                    const int32_t g = gat[s * ldc + ci];
                    if (g >= 0)
                        fpos[g >> 1] |= uint8_t(s << (4 * (g & 1)));
                }
            DNDSMI_TRY(c->alloc(c->fpos, ldf));
            DNDSMI_TRY(c->upload(c->fpos, fpos));
        }
        DNDSMI_TRY(c->alloc(c->wall_idx, nfmax * ldc));
        DNDSMI_TRY(c->upload(c->wall_idx, widx));
        DNDSMI_TRY(c->alloc(c->M, int64_t(nfmax) * kMPairs * 2 * ldc));
        DNDSMI_TRY(c->upload(c->M, M));
        M.clear(), M.shrink_to_fit();
        DNDSMI_TRY(c->alloc(c->bv, int64_t(nfmax) * 9 * ldc));
        DNDSMI_TRY(c->upload(c->bv, bv));
        DNDSMI_TRY(c->alloc(c->vol, ldc));
        DNDSMI_TRY(c->upload(c->vol, vol));
        DNDSMI_TRY(c->alloc(c->cmom, 9 * ldc));
        DNDSMI_TRY(c->upload(c->cmom, cmom));
        DNDSMI_TRY(c->alloc(c->cscale, 2 * ldc));
        DNDSMI_TRY(c->upload(c->cscale, cscale));
        // wall faces (Wall_Euler / Wall_NoSlip): compact arrays
        const int64_t nw = int64_t(wall_cell.size());
        c->n_wall = std::max<int64_t>(nw, 1);
        {
            std::vector<double> wa(81 * size_t(c->n_wall), 0.0), wd(27 * size_t(c->n_wall), 0.0),
                wn(6 * size_t(c->n_wall), 0.0), w0(size_t(c->n_wall), 0.0);
            for (int64_t w = 0; w < nw; w++) {
                const int64_t i = wall_cell[w], f = wall_face[w];
                for (int k = 0; k < 81; k++)
                    wa[k * c->n_wall + w] = vr->ainv[81 * i + k];
                for (int g = 0; g < 3; g++)
                    for (int k = 0; k < 9; k++)
                        wd[(g * 9 + k) * c->n_wall + w] = vr->face_d0[54 * f + 18 * g + k];
                for (int k = 0; k < 6; k++)
                    wn[k * c->n_wall + w] = vr->face_norm[6 * f + k];
                w0[w] = vr->face_weight[10 * f];
            }
            DNDSMI_TRY(c->alloc(c->wall_ainv, 81 * c->n_wall));
            DNDSMI_TRY(c->upload(c->wall_ainv, wa));
            DNDSMI_TRY(c->alloc(c->wall_d0, 27 * c->n_wall));
            DNDSMI_TRY(c->upload(c->wall_d0, wd));
            DNDSMI_TRY(c->alloc(c->wall_nrm, 6 * c->n_wall));
            DNDSMI_TRY(c->upload(c->wall_nrm, wn));
            DNDSMI_TRY(c->alloc(c->wall_w0, c->n_wall));
            DNDSMI_TRY(c->upload(c->wall_w0, w0));
            std::vector<int32_t> wc(size_t(c->n_wall), 0), wb(size_t(c->n_wall), 0);
            for (int64_t w = 0; w < nw; w++)
                wc[w] = c2d[wall_cell[w]], wb[w] = mesh->face_bc[wall_face[w]];
            DNDSMI_TRY(c->alloc(c->wall_cell, c->n_wall));
            DNDSMI_TRY(c->upload(c->wall_cell, wc));
            DNDSMI_TRY(c->alloc(c->wall_bc, c->n_wall));
            DNDSMI_TRY(c->upload(c->wall_bc, wb));
            DNDSMI_TRY(c->alloc(c->wall_corr, 36 * c->n_wall));
            c->n_wall_faces = nw;
        }
        // ---- face streams
        std::vector<int32_t> f2c(2 * size_t(ldf), -1), fbc(ldf, DNDSMI_BC_INNER);
        bool straight = true;
        std::vector<double> fxy(12 * size_t(ldf), 0.0), fn(6 * size_t(ldf), 0.0), fnc(2 * size_t(ldf), 0.0),
            farea(ldf, 0.0), fctr(2 * size_t(ldf), 0.0);
        for (int64_t fo = 0; fo < nf; fo++) {
            const int64_t f = f2d[fo]; // device column of mesh face fo
            f2c[f] = c2d[mesh->face2cell[2 * fo]];
            const int64_t R = mesh->face2cell[2 * fo + 1];
            f2c[ldf + f] = R == DNDSMI_NO_CELL ? -1 : c2d[R];
            fbc[f] = mesh->face_bc[fo];
            for (int g = 0; g < 3; g++) // rows [side][g] of (X, Y) pairs
                for (int sd = 0; sd < 2; sd++)
                    for (int k = 0; k < 2; k++)
                        fxy[(size_t(sd * 3 + g) * ldf + f) * 2 + k] = vr->face_gauss_xy[12 * fo + 4 * g + 2 * sd + k];
            for (int k = 0; k < 6; k++)
                fn[k * ldf + f] = vr->face_norm[6 * fo + k];
            for (int k = 2; k < 6; k++)
                straight &= std::memcmp(&vr->face_norm[6 * fo + k], &vr->face_norm[6 * fo + k % 2], 8) == 0;
            fnc[f] = vr->face_norm_center[2 * fo], fnc[ldf + f] = vr->face_norm_center[2 * fo + 1];
            farea[f] = vr->face_area[fo];
            fctr[f] = vr->face_center[2 * fo], fctr[ldf + f] = vr->face_center[2 * fo + 1];
        }
        DNDSMI_TRY(c->alloc(c->f2c, 2 * ldf));
        DNDSMI_TRY(c->upload(c->f2c, f2c));
        DNDSMI_TRY(c->alloc(c->fbc, ldf));
        DNDSMI_TRY(c->upload(c->fbc, fbc));
        DNDSMI_TRY(c->alloc(c->fxy, 12 * ldf));
        DNDSMI_TRY(c->upload(c->fxy, fxy));
        DNDSMI_TRY(c->alloc(c->fn, 6 * ldf));
        DNDSMI_TRY(c->upload(c->fn, fn));
        c->straight_faces = straight;
        DNDSMI_TRY(c->alloc(c->fnc, 2 * ldf));
        DNDSMI_TRY(c->upload(c->fnc, fnc));
        DNDSMI_TRY(c->alloc(c->farea, ldf));
        DNDSMI_TRY(c->upload(c->farea, farea));
        DNDSMI_TRY(c->alloc(c->fctr, 2 * ldf));
        DNDSMI_TRY(c->upload(c->fctr, fctr));
        std::vector<int32_t> bfl;
        for (int64_t f = 0; f < nf; f++)
            if (mesh->face2cell[2 * f + 1] == DNDSMI_NO_CELL)
                bfl.push_back(f2d[f]);
        std::sort(bfl.begin(), bfl.end());
        c->n_bface = int64_t(bfl.size());
        if (bfl.empty())
            bfl.push_back(0);
        DNDSMI_TRY(c->alloc(c->bfaces, int64_t(bfl.size())));
        DNDSMI_TRY(c->upload(c->bfaces, bfl));
        // ---- WBAP data
        c->tang_weight = vs->tang_weight;
        c->twd_norm = vs->tang_weight_direction == 0;
        c->wbap_scale = vs->wbap_smooth_indicator_scale;
        c->wbap_nstd = vs->wbap_n_std;
        if (vr->matrix_secondary) {
            c->has_limiter = true;
            std::vector<double> buf;
            buf.assign(162 * size_t(ldf), 0.0);
            for (int64_t f = 0; f < nf; f++)
                for (int k = 0; k < 162; k++)
                    buf[k * ldf + f2d[f]] = vr->matrix_secondary[162 * f + k];
            DNDSMI_TRY(c->alloc(c->msec, 162 * ldf));
            DNDSMI_TRY(c->upload(c->msec, buf));
            buf.assign(10 * size_t(ldf), 0.0);
            for (int64_t f = 0; f < nf; f++)
                for (int k = 0; k < 10; k++)
                    buf[k * ldf + f2d[f]] = vr->face_weight[10 * f + k];
            DNDSMI_TRY(c->alloc(c->fw, 10 * ldf));
            DNDSMI_TRY(c->upload(c->fw, buf));
            DNDSMI_TRY(c->alloc(c->wslot, int64_t(nfmax) * 4 * ldc));
            DNDSMI_TRY(c->alloc(c->uF, int64_t(nfmax) * 28 * ldc));
            DNDSMI_TRY(c->alloc(c->ind, ldc));
            DNDSMI_TRY(c->alloc(c->xlim, 36 * ldc));
            DNDSMI_TRY(c->alloc(c->limf, ldc));
            DNDSMI_TRY(c->alloc(c->lim_list, ldc));
            DNDSMI_TRY(c->alloc(c->lim_count, 1));
        }
        // ---- multicolour scan lists (col and ncol from before the numbering)
        {
            c->n_colour = ncol;
            c->col_start.assign(size_t(ncol) + 1, 0);
            for (int64_t i = 0; i < nc; i++)
                c->col_start[col[i] + 1]++;
            for (int32_t k = 0; k < ncol; k++)
                c->col_start[k + 1] += c->col_start[k];
            std::vector<int32_t> cells(static_cast<size_t>(nc)), fill(c->col_start.begin(), c->col_start.end() - 1);
            for (int64_t i = 0; i < nc; i++) // ascending mesh index within a colour
                cells[fill[col[i]]++] = c2d[i];
            std::vector<int32_t> colp(size_t(ldc), 0);
            for (int64_t i = 0; i < nc; i++)
                colp[c2d[i]] = col[i];
            c->colour_host = col;
            DNDSMI_TRY(c->alloc(c->colour, ldc));
            DNDSMI_TRY(c->upload(c->colour, colp));
            DNDSMI_TRY(c->alloc(c->col_cells, nc));
            DNDSMI_TRY(c->upload(c->col_cells, cells));
            DNDSMI_TRY(c->alloc(c->jinv, ldc));
        }
        // ---- scratch
        DNDSMI_TRY(c->alloc(c->lambda_face, ldf));
        DNDSMI_TRY(c->alloc(c->dlambda_face, ldf));
        DNDSMI_TRY(c->alloc(c->fslot, int64_t(nfmax) * 8 * ldc));
        DNDSMI_TRY(c->alloc(c->dt_part, int64_t(grid_for(nc)) + 1));
        DNDSMI_TRY(c->alloc(c->dt_min, 1));
        DNDSMI_TRY(c->alloc(c->errw, 1));
        DNDSMI_TRY(c->alloc(c->xlast, 4 * ldc));
        DNDSMI_TRY(c->alloc(c->dt_tmp, ldc));
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->errw);
        hipLaunchKernelGGL(k_set, dim3(grid_for(ldf)), dim3(kBlock), 0, c->stream, ldf, c->dlambda_face, 0.0);
        DNDSMI_TRY(c->launch_check("init kernels"));
        DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    } catch (const std::bad_alloc &) {
        return fail(DNDSMI_ERR_OUT_OF_MEMORY);
    }
    *out = c;
    return DNDSMI_OK;
}

// defined with the warm start (warmstart.hip) and the BDF integrator
// (lusgs.hip): forget the context
void dndsmi_warmstart_detach(dndsmi_warmstart *w);
void dndsmi_bdf_detach(dndsmi_bdf *b);

extern "C" void dndsmi_ctx_destroy(dndsmi_ctx *c) {
    if (!c)
        return;
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    // arrays and warm starts may outlive their context (any destruction
    // order is allowed): free the arrays' storage and detach them
    for (dndsmi_array *a : c->arrays) {
        (void)hipFree(a->d);
        a->d = nullptr;
        a->ctx = nullptr;
    }
    c->arrays.clear();
    for (dndsmi_warmstart *w : c->warmstarts)
        dndsmi_warmstart_detach(w);
    c->warmstarts.clear();
    for (dndsmi_bdf *b : c->bdfs)
        dndsmi_bdf_detach(b);
    c->bdfs.clear();
    for (void *p : c->allocs)
        (void)hipFree(p);
    for (hipEvent_t e : c->ev_pool)
        (void)hipEventDestroy(e);
    if (c->own_stream && c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int dndsmi_last_error(const dndsmi_ctx *c, char *buf, size_t n) {
    if (!c || !buf || n == 0)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    std::snprintf(buf, n, "%s", c->err.c_str());
    return DNDSMI_OK;
}

extern "C" int dndsmi_sync(dndsmi_ctx *c) {
    if (!c)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return check_errw(c);
}

extern "C" int64_t dndsmi_ctx_device_bytes(const dndsmi_ctx *c) { return c ? c->bytes : -1; }

extern "C" int dndsmi_array_create(dndsmi_ctx *c, int32_t kind, dndsmi_array **out) {
    if (!c || !out || kind < DNDSMI_ARRAY_DOF || kind > DNDSMI_ARRAY_SCALAR)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    dndsmi_array *a = new (std::nothrow) dndsmi_array();
    if (!a)
        return DNDSMI_ERR_OUT_OF_MEMORY;
    a->ctx = c;
    a->kind = kind;
    a->elems = kind == DNDSMI_ARRAY_DOF ? 4 * c->ldc : (kind == DNDSMI_ARRAY_REC ? 36 * c->ldc : c->ldc);
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * a->elems) != hipSuccess) {
        delete a;
        c->err = "hipMalloc of an array failed";
        return DNDSMI_ERR_OUT_OF_MEMORY;
    }
    a->d = static_cast<double *>(p);
    hipLaunchKernelGGL(k_set, dim3(grid_for(a->elems)), dim3(kBlock), 0, c->stream, a->elems, a->d, 0.0);
    int rc = c->launch_check("k_set");
    if (rc) {
        (void)hipFree(a->d);
        delete a;
        return rc;
    }
    c->arrays.insert(a);
    *out = a;
    return DNDSMI_OK;
}

extern "C" void dndsmi_array_destroy(dndsmi_array *a) {
    if (!a)
        return;
    if (a->ctx) { // else its context was destroyed first and freed the storage
        (void)hipStreamSynchronize(a->ctx->stream);
        (void)hipFree(a->d);
        a->ctx->arrays.erase(a);
    }
    delete a;
}

extern "C" int dndsmi_array_upload(dndsmi_array *a, const double *host) {
    if (!a || !host || !a->ctx) // no context: it was destroyed before the array
        return DNDSMI_ERR_INVALID_ARGUMENT;
    dndsmi_ctx *c = a->ctx;
    const int64_t nc = c->nc, ldc = c->ldc;
    const std::vector<int32_t> &c2d = c->c2d;
    if (a->kind == DNDSMI_ARRAY_REC) {
        std::vector<double> buf(36 * size_t(ldc), 0.0);
        for (int64_t i = 0; i < nc; i++)
            for (int k = 0; k < 36; k++)
                buf[k * ldc + c2d[i]] = host[36 * i + k];
        return c->upload(a->d, buf);
    }
    const int w = a->kind == DNDSMI_ARRAY_DOF ? 4 : 1;
    std::vector<double> buf(size_t(w) * ldc, 0.0);
    for (int64_t i = 0; i < nc; i++)
        for (int v = 0; v < w; v++)
            buf[size_t(w) * c2d[i] + v] = host[w * i + v];
    return c->upload(a->d, buf);
}

extern "C" int dndsmi_array_download(const dndsmi_array *a, double *host) {
    if (!a || !host || !a->ctx)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    dndsmi_ctx *c = a->ctx;
    const int64_t nc = c->nc, ldc = c->ldc;
    if (a->kind == DNDSMI_ARRAY_REC) {
        std::vector<double> buf(36 * size_t(ldc));
        DNDSMI_TRY(c->hip(hipMemcpyAsync(buf.data(), a->d, sizeof(double) * buf.size(), hipMemcpyDeviceToHost,
                                         c->stream),
                          "download"));
        DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
        for (int64_t i = 0; i < nc; i++)
            for (int k = 0; k < 36; k++)
                host[36 * i + k] = buf[k * ldc + c->c2d[i]];
        return DNDSMI_OK;
    }
    const int w = a->kind == DNDSMI_ARRAY_DOF ? 4 : 1;
    std::vector<double> buf(size_t(w) * ldc);
    DNDSMI_TRY(c->hip(hipMemcpyAsync(buf.data(), a->d, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, c->stream),
                      "download"));
    DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    for (int64_t i = 0; i < nc; i++)
        for (int v = 0; v < w; v++)
            host[w * i + v] = buf[size_t(w) * c->c2d[i] + v];
    return DNDSMI_OK;
}

extern "C" int dndsmi_array_copy(dndsmi_array *dst, const dndsmi_array *src) {
    if (!dst || !src || !dst->ctx || dst->ctx != src->ctx || dst->kind != src->kind)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    return dst->ctx->hip(hipMemcpyAsync(dst->d, src->d, sizeof(double) * dst->elems, hipMemcpyDeviceToDevice,
                                        dst->ctx->stream),
                         "copy");
}

extern "C" int dndsmi_array_set_constant(dndsmi_array *a, double value) {
    if (!a || !a->ctx)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(k_set, dim3(grid_for(a->elems)), dim3(kBlock), 0, a->ctx->stream, a->elems, a->d, value);
    return a->ctx->launch_check("k_set");
}

namespace dndsmi_host {

// SOR_Instead: colour by colour in place on x (forward; with
// SOR_InverseScanning then backward, DNDS_FV_VR.hpp:1493-1502)
int launch_sor(dndsmi_ctx *c, const double *u, double *x, double t) {
    const int passes = c->sor_inverse ? 2 : 1;
    for (int pass = 0; pass < passes; pass++)
        for (int32_t kk = 0; kk < c->n_colour; kk++) {
            const int32_t col = pass == 0 ? kk : c->n_colour - 1 - kk;
            for (int w = 0; w < 2; w++) {
                const int64_t b = w ? c->sor_wall[col] : c->sor_start[col];
                const int64_t n = (w ? c->sor_start[col + 1] : c->sor_wall[col]) - b;
                if (n == 0)
                    continue;
                auto kern = c->nfmax == 3 ? (w ? k_sor<3, true> : k_sor<3, false>)
                                          : (w ? k_sor<4, true> : k_sor<4, false>);
                // the boundary cells are a few thousand: one wave per
                // workgroup spreads them over as many CUs as there are waves
                const int block = w ? 64 : kBlock;
                prof_begin(c, DNDSMI_K_JACOBI);
                hipLaunchKernelGGL(kern, dim3(unsigned((n + block - 1) / block)), dim3(block), 0, c->stream,
                                   int32_t(n), int32_t(b),
                                   c->ldc, c->nbr, c->sbc, c->wall_idx, c->M, c->bv, u, x, x, c->relax,
                                   c->wall_ainv, c->wall_d0, c->wall_nrm, c->wall_w0, c->n_wall, c->gas, t, c->errw);
                prof_end(c);
                DNDSMI_TRY(c->launch_check("k_sor"));
            }
        }
    return DNDSMI_OK;
}

int step_sweep(dndsmi_ctx *c, const double *u, double *&x, double *&xn, double t) {
    if (c->sor)
        return launch_sor(c, u, x, t);
    DNDSMI_TRY(launch_jacobi(c, u, x, xn, t));
    std::swap(x, xn);
    return DNDSMI_OK;
}

int launch_jacobi(dndsmi_ctx *c, const double *u, const double *x, double *xn, double t) {
    if (c->sor) {
        // the steps expect the new iterate in xn: start it from x, then update in place
        DNDSMI_TRY(c->hip(hipMemcpyAsync(xn, x, sizeof(double) * 36 * c->ldc, hipMemcpyDeviceToDevice, c->stream),
                          "copy uRec"));
        return launch_sor(c, u, xn, t);
    }
    const int nc = int(c->nc);
    if (c->n_wall_faces > 0) {
        prof_begin(c, DNDSMI_K_WALL_CORR);
        hipLaunchKernelGGL(k_wall_corr, dim3(grid_for(c->n_wall_faces)), dim3(kBlock), 0, c->stream,
                           int32_t(c->n_wall_faces), c->ldc, c->wall_cell, c->wall_bc, c->wall_ainv, c->wall_d0,
                           c->wall_nrm, c->wall_w0, c->n_wall, u, x, c->relax, c->gas, t, c->wall_corr, c->errw);
        prof_end(c);
        DNDSMI_TRY(c->launch_check("k_wall_corr"));
    }
    prof_begin(c, DNDSMI_K_JACOBI);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_jacobi<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->nbr, c->sbc,
                           c->wall_idx, c->M, c->bv, u, x, xn, c->relax, c->wall_corr, c->n_wall, c->errw);
    else
        hipLaunchKernelGGL(k_jacobi<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->nbr, c->sbc,
                           c->wall_idx, c->M, c->bv, u, x, xn, c->relax, c->wall_corr, c->n_wall, c->errw);
    prof_end(c);
    return c->launch_check("k_jacobi");
}

int launch_dt(dndsmi_ctx *c, const double *u, double cfl, double max_dt, double *dt_out) {
    const int nc = int(c->nc), nf = int(c->nf);
    prof_begin(c, DNDSMI_K_DT_FACE);
    hipLaunchKernelGGL(k_dt_face, dim3(grid_for(nf)), dim3(kBlock), 0, c->stream, nf, c->ldf, c->f2c, c->fnc, u,
                       c->gas, c->farea, c->vol, c->lambda_face, c->dlambda_face, c->errw);
    prof_end(c);
    const unsigned g = grid_for(nc);
    prof_begin(c, DNDSMI_K_DT_CELL);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_dt_cell<3>, dim3(g), dim3(kBlock), 0, c->stream, nc, c->ldc, c->gat, c->lambda_face,
                           c->farea, c->vol, cfl, max_dt, dt_out, c->dt_part);
    else
        hipLaunchKernelGGL(k_dt_cell<4>, dim3(g), dim3(kBlock), 0, c->stream, nc, c->ldc, c->gat, c->lambda_face,
                           c->farea, c->vol, cfl, max_dt, dt_out, c->dt_part);
    prof_end(c);
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kBlock), 0, c->stream, int32_t(g), c->dt_part, c->dt_min, c->errw);
    return c->launch_check("dt kernels");
}

int launch_rhs(dndsmi_ctx *c, const double *u, const double *x, double t, double *rhs, int stage, double *xs,
               const double *dt, double coefA, double coefB, const double *xalt, const uint8_t *limf) {
    const int nc = int(c->nc), nf = int(c->nf);
    prof_begin(c, DNDSMI_K_RHS_FACE);
    if (c->gas.mu > 0) {
        hipLaunchKernelGGL((c->straight_faces ? k_rhs_face<false, false, true, true>
                                              : k_rhs_face<false, false, false, true>),
                           dim3(grid_for(nf)), dim3(kBlock), 0, c->stream, nf, c->ldf, c->ldc, c->bfaces, c->f2c,
                           c->fbc, c->fxy, c->cmom, c->fn, c->fctr, c->dlambda_face, u, x, c->gas, t, c->fpos,
                           c->fslot, c->errw, nullptr, c->cscale, c->vol, c->farea, xalt, limf);
    } else {
        hipLaunchKernelGGL((c->straight_faces ? k_rhs_face<false, false, true> : k_rhs_face<false, false, false>),
                           dim3(grid_for(nf)), dim3(kBlock), 0, c->stream, nf, c->ldf, c->ldc, c->bfaces, c->f2c,
                           c->fbc, c->fxy, c->cmom, c->fn, c->fctr, c->dlambda_face, u, x, c->gas, t, c->fpos,
                           c->fslot, c->errw, nullptr, nullptr, nullptr, nullptr, xalt, limf);
    }
    prof_end(c);
    if (c->n_bface > 0) {
        prof_begin(c, DNDSMI_K_RHS_BFACE);
        if (c->gas.mu > 0)
            hipLaunchKernelGGL((k_rhs_face<true, false, false, true>), dim3(grid_for(c->n_bface)), dim3(kBlock), 0,
                               c->stream, int32_t(c->n_bface), c->ldf, c->ldc, c->bfaces, c->f2c, c->fbc, c->fxy,
                               c->cmom, c->fn, c->fctr, c->dlambda_face, u, x, c->gas, t, c->fpos, c->fslot, c->errw,
                               nullptr, c->cscale, c->vol, c->farea, xalt, limf);
        else
            hipLaunchKernelGGL(k_rhs_face<true>, dim3(grid_for(c->n_bface)), dim3(kBlock), 0, c->stream,
                               int32_t(c->n_bface), c->ldf, c->ldc, c->bfaces, c->f2c, c->fbc, c->fxy, c->cmom, c->fn,
                               c->fctr, c->dlambda_face, u, x, c->gas, t, c->fpos, c->fslot, c->errw, nullptr,
                               nullptr, nullptr, nullptr, xalt, limf);
        prof_end(c);
    }
    prof_begin(c, DNDSMI_K_RHS_CELL);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_rhs_cell<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->gat,
                           c->fslot, c->vol, rhs, stage, xs, c->xlast, dt, c->dt_min, coefA, coefB, c->errw);
    else
        hipLaunchKernelGGL(k_rhs_cell<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->gat,
                           c->fslot, c->vol, rhs, stage, xs, c->xlast, dt, c->dt_min, coefA, coefB, c->errw);
    prof_end(c);
    return c->launch_check("rhs kernels");
}

int launch_wbap(dndsmi_ctx *c, const double *u, const double *x, double *xo, double *ind, int if_all,
                int procedure, bool full_copy) {
    c->wbap_full_copy = full_copy;
    if (!c->has_limiter) {
        c->err = "WBAP needs a setup built with build_limiter_data";
        return DNDSMI_ERR_UNSUPPORTED;
    }
    const int nc = int(c->nc), nf = int(c->nf);
    const double eps4 = std::pow(1e-200, 1.0 / 4); // verySmallReal_pDiP, DNDS_FV_Limiters.hpp:345
    prof_begin(c, DNDSMI_K_WBAP_FACE);
    hipLaunchKernelGGL(k_wbap_face, dim3(grid_for(nf)), dim3(kBlock), 0, c->stream, nf, c->ldf, c->ldc, c->f2c,
                       c->fxy, c->cmom, c->cscale, c->fw, c->fn, u, x, c->tang_weight, c->twd_norm, c->wpos,
                       c->wslot);
    prof_end(c);
    hipLaunchKernelGGL(k_zero_int2, dim3(1), dim3(64), 0, c->stream, c->lim_count);
    prof_begin(c, DNDSMI_K_WBAP_CELL);
    if (c->nfmax == 3)
        hipLaunchKernelGGL(k_wbap_ind<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->nbr,
                           c->wslot, x, xo, ind ? ind : c->ind, c->wbap_scale, if_all, c->lim_count, c->lim_list,
                           full_copy ? nullptr : c->limf);
    else
        hipLaunchKernelGGL(k_wbap_ind<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->nbr,
                           c->wslot, x, xo, ind ? ind : c->ind, c->wbap_scale, if_all, c->lim_count, c->lim_list,
                           full_copy ? nullptr : c->limf);
    prof_end(c);
    prof_begin(c, DNDSMI_K_WBAP_LIM);
    if (procedure == 1) {
        if (c->nfmax == 3)
            hipLaunchKernelGGL(k_wbap_lim<3>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, c->ldc, c->ldf,
                               c->lim_count, c->lim_list, c->nbr, c->sface, c->f2c, c->fnc, c->msec, u, x, xo, c->uF,
                               c->gas.gamma, c->wbap_nstd, eps4, c->errw);
        else
            hipLaunchKernelGGL(k_wbap_lim<4>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, c->ldc, c->ldf,
                               c->lim_count, c->lim_list, c->nbr, c->sface, c->f2c, c->fnc, c->msec, u, x, xo, c->uF,
                               c->gas.gamma, c->wbap_nstd, eps4, c->errw);
    } else {
        // orders 3, 2, 1 in stream order: each launch reads the rows the
        // previous ones wrote
        auto order = [&](auto nf_c, auto order_c) {
            constexpr int NF = decltype(nf_c)::value, O = decltype(order_c)::value;
            hipLaunchKernelGGL((k_wbap_lim_v2<NF, O>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, c->ldc,
                               c->ldf, c->lim_count, c->lim_list, c->nbr, c->sface, c->f2c, c->fnc, c->msec, u, x,
                               xo, c->gas.gamma, c->wbap_nstd, eps4, c->errw, full_copy ? nullptr : c->limf);
        };
        using I3 = std::integral_constant<int, 3>;
        using I4 = std::integral_constant<int, 4>;
        using O3 = std::integral_constant<int, 3>;
        using O2 = std::integral_constant<int, 2>;
        using O1 = std::integral_constant<int, 1>;
        if (c->nfmax == 3)
            order(I3{}, O3{}), order(I3{}, O2{}), order(I3{}, O1{});
        else
            order(I4{}, O3{}), order(I4{}, O2{}), order(I4{}, O1{});
    }
    prof_end(c);
    return c->launch_check("wbap kernels");
}

int ensure_mixed(dndsmi_ctx *c) {
    if (c->M32)
        return DNDSMI_OK;
    const int64_t nm = int64_t(c->nfmax) * kMQuads * 4 * c->ldc;
    DNDSMI_TRY(c->alloc(c->M32, nm));
    DNDSMI_TRY(c->alloc(c->r32, 36 * c->ldc));
    DNDSMI_TRY(c->alloc(c->dA32, 36 * c->ldc));
    DNDSMI_TRY(c->alloc(c->dB32, 36 * c->ldc));
    DNDSMI_TRY(c->alloc(c->fslot32, int64_t(c->nfmax) * 8 * c->ldc));
    DNDSMI_TRY(c->alloc(c->wcorr32, 36 * c->n_wall));
    hipLaunchKernelGGL(k_m_to_f32, dim3(grid_for(nm / 4)), dim3(kBlock), 0, c->stream, int32_t(c->nfmax), c->ldc,
                       c->M, c->M32);
    DNDSMI_TRY(c->launch_check("k_to_f32"));
    return c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// one FP64 step that yields r = J(x) - x, then n - 1 FP32 correction sweeps,
// the last adding the correction to x in place
int launch_jacobi_mixed(dndsmi_ctx *c, const double *u, double *x, int n_sweeps, double t) {
    const int nc = int(c->nc);
    if (c->n_wall_faces > 0) {
        hipLaunchKernelGGL(k_wall_corr, dim3(grid_for(c->n_wall_faces)), dim3(kBlock), 0, c->stream,
                           int32_t(c->n_wall_faces), c->ldc, c->wall_cell, c->wall_bc, c->wall_ainv, c->wall_d0,
                           c->wall_nrm, c->wall_w0, c->n_wall, u, x, c->relax, c->gas, t, c->wall_corr, c->errw);
        DNDSMI_TRY(c->launch_check("k_wall_corr"));
    }
    prof_begin(c, DNDSMI_K_JACOBI);
    if (c->nfmax == 3)
        hipLaunchKernelGGL((k_jacobi<3, true>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->nbr,
                           c->sbc, c->wall_idx, c->M, c->bv, u, x, nullptr, c->relax, c->wall_corr, c->n_wall,
                           c->errw, c->r32, c->dA32);
    else
        hipLaunchKernelGGL((k_jacobi<4, true>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc, c->nbr,
                           c->sbc, c->wall_idx, c->M, c->bv, u, x, nullptr, c->relax, c->wall_corr, c->n_wall,
                           c->errw, c->r32, c->dA32);
    prof_end(c);
    DNDSMI_TRY(c->launch_check("k_jacobi (residual)"));
    float *din = c->dA32, *dout = c->dB32;
    const float w = float(c->relax);
    for (int k = 1; k < n_sweeps; k++) {
        const bool last = k == n_sweeps - 1;
        if (c->n_wall_faces > 0) {
            hipLaunchKernelGGL(k_wall_corr_lin, dim3(grid_for(c->n_wall_faces)), dim3(kBlock), 0, c->stream,
                               int32_t(c->n_wall_faces), c->ldc, c->wall_cell, c->wall_bc, c->wall_ainv, c->wall_d0,
                               c->wall_nrm, c->wall_w0, c->n_wall, din, w, c->wcorr32);
            DNDSMI_TRY(c->launch_check("k_wall_corr_lin"));
        }
        prof_begin(c, DNDSMI_K_DELTA32);
        if (c->nfmax == 3) {
            if (last)
                hipLaunchKernelGGL((k_delta32<3, true>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc,
                                   c->nbr, c->sbc, c->M32, c->r32, din, nullptr, w, x, c->wall_idx, c->wcorr32,
                                   c->n_wall);
            else
                hipLaunchKernelGGL((k_delta32<3, false>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc,
                                   c->ldc, c->nbr, c->sbc, c->M32, c->r32, din, dout, w, nullptr, c->wall_idx,
                                   c->wcorr32, c->n_wall);
        } else {
            if (last)
                hipLaunchKernelGGL((k_delta32<4, true>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc,
                                   c->nbr, c->sbc, c->M32, c->r32, din, nullptr, w, x, c->wall_idx, c->wcorr32,
                                   c->n_wall);
            else
                hipLaunchKernelGGL((k_delta32<4, false>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc,
                                   c->ldc, c->nbr, c->sbc, c->M32, c->r32, din, dout, w, nullptr, c->wall_idx,
                                   c->wcorr32, c->n_wall);
        }
        prof_end(c);
        DNDSMI_TRY(c->launch_check("k_delta32"));
        std::swap(din, dout);
    }
    return DNDSMI_OK;
}

int launch_rhs_mixed(dndsmi_ctx *c, const double *u, const double *x, double t, double *rhs, int stage, double *xs,
                     const double *dt, double coefA, double coefB, const double *xalt, const uint8_t *limf) {
    const int nc = int(c->nc), nf = int(c->nf);
    prof_begin(c, DNDSMI_K_RHS_FACE32);
    hipLaunchKernelGGL((c->straight_faces ? k_rhs_face<false, true, true> : k_rhs_face<false, true, false>),
                       dim3(grid_for(nf)), dim3(kBlock), 0, c->stream, nf, c->ldf, c->ldc, c->bfaces, c->f2c, c->fbc,
                       c->fxy, c->cmom, c->fn, c->fctr, c->dlambda_face, u, x, c->gas, t, c->fpos, nullptr, c->errw,
                       c->fslot32, nullptr, nullptr, nullptr, xalt, limf);
    prof_end(c);
    if (c->n_bface > 0)
        hipLaunchKernelGGL((k_rhs_face<true, true>), dim3(grid_for(c->n_bface)), dim3(kBlock), 0, c->stream,
                           int32_t(c->n_bface), c->ldf, c->ldc, c->bfaces, c->f2c, c->fbc, c->fxy, c->cmom, c->fn,
                           c->fctr, c->dlambda_face, u, x, c->gas, t, c->fpos, nullptr, c->errw, c->fslot32,
                           nullptr, nullptr, nullptr, xalt, limf);
    prof_begin(c, DNDSMI_K_RHS_CELL32);
    if (c->nfmax == 3)
        hipLaunchKernelGGL((k_rhs_cell<3, float>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc,
                           c->gat, c->fslot32, c->vol, rhs, stage, xs, c->xlast, dt, c->dt_min, coefA, coefB,
                           c->errw);
    else
        hipLaunchKernelGGL((k_rhs_cell<4, float>), dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, nc, c->ldc,
                           c->gat, c->fslot32, c->vol, rhs, stage, xs, c->xlast, dt, c->dt_min, coefA, coefB,
                           c->errw);
    prof_end(c);
    return c->launch_check("mixed rhs kernels");
}

} // namespace dndsmi_host

namespace {

int wbap_entry(dndsmi_ctx *c, const dndsmi_array *u, const dndsmi_array *uRec, dndsmi_array *uRecNewBuf,
               dndsmi_array *uRecNewBuf1, dndsmi_array *ifUseLimiter, int32_t ifAll, int procedure) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNewBuf, DNDSMI_ARRAY_REC) || uRec == uRecNewBuf ||
        (uRecNewBuf1 && (!same_ctx(c, uRecNewBuf1, DNDSMI_ARRAY_REC) || uRecNewBuf1 == uRec ||
                         uRecNewBuf1 == uRecNewBuf)) ||
        (ifUseLimiter && !same_ctx(c, ifUseLimiter, DNDSMI_ARRAY_SCALAR)))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(launch_wbap(c, u->d, uRec->d, uRecNewBuf->d, ifUseLimiter ? ifUseLimiter->d : nullptr, ifAll,
                           procedure));
    if (uRecNewBuf1)
        DNDSMI_TRY(c->hip(hipMemcpyAsync(uRecNewBuf1->d, uRecNewBuf->d, sizeof(double) * uRecNewBuf->elems,
                                         hipMemcpyDeviceToDevice, c->stream),
                          "copy"));
    return DNDSMI_OK;
}

} // namespace

extern "C" int dndsmi_wbap_limit_facial_v3(dndsmi_ctx *c, const dndsmi_array *u, const dndsmi_array *uRec,
                                           dndsmi_array *uRecNewBuf, dndsmi_array *uRecNewBuf1,
                                           dndsmi_array *ifUseLimiter, int32_t ifAll) {
    return wbap_entry(c, u, uRec, uRecNewBuf, uRecNewBuf1, ifUseLimiter, ifAll, 1);
}

extern "C" int dndsmi_wbap_limit_facial_v2(dndsmi_ctx *c, const dndsmi_array *u, const dndsmi_array *uRec,
                                           dndsmi_array *uRecNewBuf, dndsmi_array *uRecNewBuf1,
                                           dndsmi_array *ifUseLimiter, int32_t ifAll) {
    return wbap_entry(c, u, uRec, uRecNewBuf, uRecNewBuf1, ifUseLimiter, ifAll, 0);
}

extern "C" int dndsmi_set_limiter_procedure(dndsmi_ctx *c, int32_t limiterProcedure) {
    if (!c || (limiterProcedure != 0 && limiterProcedure != 1))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    c->limiter_procedure = limiterProcedure;
    return DNDSMI_OK;
}

extern "C" int dndsmi_set_limiter_if_all(dndsmi_ctx *c, int32_t ifAll) {
    if (!c)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    c->limiter_if_all = ifAll != 0;
    return DNDSMI_OK;
}

extern "C" int dndsmi_reconstruction_jacobi_step(dndsmi_ctx *c, const dndsmi_array *u, dndsmi_array *uRec,
                                                 dndsmi_array *uRecNewBuf, double t) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNewBuf, DNDSMI_ARRAY_REC) || uRec == uRecNewBuf)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (c->sor) // in place on uRec; uRecNewBuf is not touched (DNDS_FV_VR.hpp:1507-1510, 1707)
        return launch_sor(c, u->d, uRec->d, t);
    DNDSMI_TRY(launch_jacobi(c, u->d, uRec->d, uRecNewBuf->d, t));
    // uRec = uRecNewBuf (DNDS_FV_VR.hpp:1707-1711): afterwards both arrays
    // hold the new iterate, as in the reference.
    return c->hip(hipMemcpyAsync(uRec->d, uRecNewBuf->d, sizeof(double) * uRec->elems, hipMemcpyDeviceToDevice,
                                 c->stream),
                  "copy uRecNewBuf to uRec");
}

extern "C" int dndsmi_evaluate_dt(dndsmi_ctx *c, dndsmi_array *dt, const dndsmi_array *u, double CFL,
                                  double *dtMinall, double MaxDt, int32_t UseLocaldt) {
    if (!c || !same_ctx(c, dt, DNDSMI_ARRAY_SCALAR) || !same_ctx(c, u, DNDSMI_ARRAY_DOF))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(launch_dt(c, u->d, CFL, MaxDt, dt->d));
    if (!UseLocaldt) {
        hipLaunchKernelGGL(k_fill_from, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, int32_t(c->nc), c->dt_min,
                           dt->d);
        DNDSMI_TRY(c->launch_check("k_fill_from"));
    }
    if (dtMinall) {
        DNDSMI_TRY(c->hip(hipMemcpyAsync(dtMinall, c->dt_min, sizeof(double), hipMemcpyDeviceToHost, c->stream),
                          "hipMemcpyAsync"));
        DNDSMI_TRY(check_errw(c));
    }
    return DNDSMI_OK;
}

extern "C" int dndsmi_evaluate_rhs(dndsmi_ctx *c, dndsmi_array *rhs, const dndsmi_array *u, const dndsmi_array *uRec,
                                   double t) {
    if (!c || !same_ctx(c, rhs, DNDSMI_ARRAY_DOF) || !same_ctx(c, u, DNDSMI_ARRAY_DOF) ||
        !same_ctx(c, uRec, DNDSMI_ARRAY_REC))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    return launch_rhs(c, u->d, uRec->d, t, rhs->d, 0, nullptr, nullptr, 0, 0);
}

extern "C" int dndsmi_ssprk3_step(dndsmi_ctx *c, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                                  int32_t n_sweeps, int32_t use_limiter, double CFL, double MaxDt, int32_t UseLocaldt,
                                  double t, double *dt_out) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNew, DNDSMI_ARRAY_REC) || uRec == uRecNew || n_sweeps < 0)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (use_limiter && !c->has_limiter) {
        c->err = "WBAP needs a setup built with build_limiter_data";
        return DNDSMI_ERR_UNSUPPORTED;
    }
    // fdt: EvaluateDt; the stage update reads dt_i (local) or the device minimum
    DNDSMI_TRY(launch_dt(c, u->d, CFL, MaxDt, c->dt_tmp));
    // the global minimum sets the stage times (one host read per step)
    double dtmin = 0;
    DNDSMI_TRY(c->hip(hipMemcpyAsync(&dtmin, c->dt_min, sizeof(double), hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync"));
    DNDSMI_TRY(check_errw(c));
    if (dt_out)
        *dt_out = dtmin;
    const double Coef[3] = {1.0, 0.25, 2.0 / 3.0};
    const double ct[3] = {1, 0.25, 2. / 3.};
    for (int stage = 0; stage < 3; stage++) {
        const double ts = t + ct[stage] * dtmin;
        for (int k = 0; k < n_sweeps; k++) {
            DNDSMI_TRY(step_sweep(c, u->d, uRec->d, uRecNew->d, ts));
        }
        const double cA = stage ? Coef[stage] / (1 - Coef[stage]) : 0.0;
        const double cB = stage ? 1 - Coef[stage] : 0.0;
        // the RHS sees the limited cells from xlim (flagged), the others from
        // uRec, which keeps the unlimited iterate
        if (use_limiter)
            DNDSMI_TRY(launch_wbap(c, u->d, uRec->d, c->xlim, nullptr, c->limiter_if_all, c->limiter_procedure, false));
        DNDSMI_TRY(launch_rhs(c, u->d, uRec->d, ts, nullptr, stage + 1, u->d, UseLocaldt ? c->dt_tmp : nullptr, cA,
                              cB, use_limiter ? c->xlim : nullptr, use_limiter ? c->limf : nullptr));
    }
    return DNDSMI_OK;
}

namespace {

// the three stages of a mixed-precision SSP-RK3 step on (u, x) after
// EvaluateDt (c->dt_tmp and dtmin set); `ws`: the stages run the learned
// warm start's hook, adding its correction when `correct`
int mixed_stages(dndsmi_ctx *c, double *u, double *x, int32_t n_sweeps, int32_t use_limiter, int32_t UseLocaldt,
                 double t, double dtmin, bool ws, bool correct) {
    const double Coef[3] = {1.0, 0.25, 2.0 / 3.0};
    const double ct[3] = {1, 0.25, 2. / 3.};
    for (int stage = 0; stage < 3; stage++) {
        const double ts = t + ct[stage] * dtmin;
        if (ws)
            DNDSMI_TRY(ws_stage_begin(c, u, x, correct));
        DNDSMI_TRY(launch_jacobi_mixed(c, u, x, n_sweeps, ts));
        const double cA = stage ? Coef[stage] / (1 - Coef[stage]) : 0.0;
        const double cB = stage ? 1 - Coef[stage] : 0.0;
        if (use_limiter)
            DNDSMI_TRY(launch_wbap(c, u, x, c->xlim, nullptr, c->limiter_if_all, c->limiter_procedure, false));
        DNDSMI_TRY(launch_rhs_mixed(c, u, x, ts, nullptr, stage + 1, u, UseLocaldt ? c->dt_tmp : nullptr, cA, cB,
                                    use_limiter ? c->xlim : nullptr, use_limiter ? c->limf : nullptr));
    }
    return DNDSMI_OK;
}

} // namespace

extern "C" int dndsmi_ssprk3_step_mixed(dndsmi_ctx *c, dndsmi_array *u, dndsmi_array *uRec, dndsmi_array *uRecNew,
                                        int32_t n_sweeps, int32_t use_limiter, double CFL, double MaxDt,
                                        int32_t UseLocaldt, double t, double *dt_out) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNew, DNDSMI_ARRAY_REC) || uRec == uRecNew || n_sweeps < 2)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (use_limiter && !c->has_limiter) {
        c->err = "WBAP needs a setup built with build_limiter_data";
        return DNDSMI_ERR_UNSUPPORTED;
    }
    if (c->sor) {
        c->err = "the mixed-precision step iterates the Jacobi map (SOR_Instead must be off)";
        return DNDSMI_ERR_UNSUPPORTED;
    }
    if (c->gas.mu > 0) {
        c->err = "the mixed-precision step covers the Euler model (muGas = 0) only";
        return DNDSMI_ERR_UNSUPPORTED;
    }
    DNDSMI_TRY(ensure_mixed(c));
    // the learned warm start: is this step in a check window (the plain run
    // on shadow copies beside the corrected one)?
    const bool ws = c->step_ws && !c->ws_rejected;
    bool window = false;
    if (ws) {
        if (c->ws_win_left == 0 && (c->ws_first || (c->ws_recheck > 0 && c->ws_since_check >= c->ws_recheck))) {
            DNDSMI_TRY(ws_window_begin(c, u->d, uRec->d));
            c->ws_win_left = c->ws_check_steps;
            c->ws_first = false;
        }
        window = c->ws_win_left > 0;
    }
    // the step's dt: from the plain run's state in a window, used by both runs
    DNDSMI_TRY(launch_dt(c, window ? c->ws_sh_u : u->d, CFL, MaxDt, c->dt_tmp));
    double dtmin = 0;
    DNDSMI_TRY(c->hip(hipMemcpyAsync(&dtmin, c->dt_min, sizeof(double), hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync"));
    DNDSMI_TRY(check_errw(c));
    if (dt_out)
        *dt_out = dtmin;
    if (!window) {
        if (ws)
            c->ws_since_check++;
        return mixed_stages(c, u->d, uRec->d, n_sweeps, use_limiter, UseLocaldt, t, dtmin, ws, ws);
    }
    DNDSMI_TRY(mixed_stages(c, c->ws_sh_u, c->ws_sh_x, n_sweeps, use_limiter, UseLocaldt, t, dtmin, false, false));
    DNDSMI_TRY(check_errw(c)); // the plain run's own failures are the method's
    int rc = mixed_stages(c, u->d, uRec->d, n_sweeps, use_limiter, UseLocaldt, t, dtmin, true, true);
    if (rc == DNDSMI_OK)
        rc = check_errw(c);
    const bool failed = rc == DNDSMI_ERR_NONPOSITIVE_STATE || rc == DNDSMI_ERR_NAN;
    if (rc != DNDSMI_OK && !failed)
        return rc;
    c->ws_win_left--;
    if (failed || c->ws_win_left == 0) {
        DNDSMI_TRY(ws_window_end(c, u->d, uRec->d, failed));
        c->ws_since_check = 0;
        c->ws_win_left = 0;
    }
    return DNDSMI_OK;
}

// ---- profiling ---------------------------------------------------------

extern "C" int dndsmi_read_probe(dndsmi_ctx *c, int32_t reps) {
    if (!c || reps < 1)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (!c->probe_out)
        DNDSMI_TRY(c->alloc(c->probe_out, c->ldc));
    for (int k = 0; k < reps; k++) {
        prof_begin(c, DNDSMI_K_READ_PROBE);
        if (c->nfmax == 3)
            hipLaunchKernelGGL(k_read_probe<3>, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, int32_t(c->nc),
                               c->ldc, c->nbr, c->M, c->probe_out);
        else
            hipLaunchKernelGGL(k_read_probe<4>, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, int32_t(c->nc),
                               c->ldc, c->nbr, c->M, c->probe_out);
        prof_end(c);
        DNDSMI_TRY(c->launch_check("k_read_probe"));
    }
    return DNDSMI_OK;
}

extern "C" int dndsmi_profile_enable(dndsmi_ctx *c, int32_t on) {
    if (!c)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    c->prof_on = on != 0;
    c->prof_recs.clear();
    c->ev_used = 0;
    return DNDSMI_OK;
}

extern "C" int dndsmi_profile_query(dndsmi_ctx *c, int32_t kernel, int64_t *count, double *total_ms) {
    if (!c || !count || !total_ms || kernel < 0 || kernel >= DNDSMI_K_COUNT)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    int64_t n = 0;
    double tot = 0;
    for (auto &r : c->prof_recs) {
        if (r.first != kernel)
            continue;
        float ms = 0;
        DNDSMI_TRY(c->hip(hipEventElapsedTime(&ms, c->ev_pool[r.second], c->ev_pool[r.second + 1]),
                          "hipEventElapsedTime"));
        tot += ms;
        n++;
    }
    *count = n;
    *total_ms = tot;
    return DNDSMI_OK;
}

// Algorithmic (unique) HBM bytes of one launch of a kernel: every stream the
// kernel needs read once and every output written once, from mesh counts.
extern "C" int64_t dndsmi_kernel_bytes(const dndsmi_ctx *c, int32_t kernel) {
    if (!c)
        return -1;
    const int64_t nc = c->nc, nf = c->nf, nif = nf - c->n_bface, nbf = c->n_bface;
    int64_t c2f = 0; // sum of faces per cell = 2 nif + nbf
    c2f = 2 * nif + nbf;
    switch (kernel) {
    case DNDSMI_K_JACOBI: // M + b per slot, slot neighbour index, x read + x_new write, u
        return c2f * (81 + 9) * 8 + c2f * 4 + nc * (36 * 8 * 2 + 4 * 8);
    case DNDSMI_K_DT_FACE: // f2c, centre normal, lambda + dlambda out; u once per cell
        return nf * (8 + 16 + 16) + nc * 32;
    case DNDSMI_K_DT_CELL: // gather index, lambda + area per slot, volume, dt out
        return c2f * (4 + 16) + nc * (8 + 8);
    case DNDSMI_K_RHS_FACE: // f2c, slot positions, point coordinates, normals, dlambda, 2 x 8 terms out;
        // x, u and the moments once per cell; viscous: face area, both volumes and cell scales
        return nif * (8 + 1 + 96 + (c->straight_faces ? 16 : 48) + 8 + 128 + (c->gas.mu > 0 ? 8 + 16 + 32 : 0)) +
               nc * (36 * 8 + 32 + 72);
    case DNDSMI_K_RHS_CELL: // 8 terms per slot, volume, x read+write, xLast (+ last-slot index for 4 slots)
        return c2f * 64 + nc * (8 + 32 * 3 + (c->nfmax == 4 ? 4 : 0));
    case DNDSMI_K_WBAP_FACE: // point-0 coordinates of both sides, f2c, slot positions, weights, normal, 8 terms
        // out; x (rho, E), u (rho, E), moments and scales once per cell
        return nif * (32 + 8 + 1 + 80 + 16 + 64) + nc * (18 * 8 + 16 + 72 + 16);
    case DNDSMI_K_WBAP_CELL: // terms and neighbour index per slot, indicator out; x read and copied out for the
        // unlimited cells of a full copy (an upper bound: as if no cell were limited), else a flag per cell
        return c2f * (32 + 4) + nc * (8 + (c->wbap_full_copy ? 36 * 8 * 2 : 1));
    case DNDSMI_K_WBAP_LIM: // depends on the number of limited cells; not a bandwidth kernel
        return 0;
    case DNDSMI_K_RHS_BFACE: // per boundary face: index, f2c, slot position, kind, one side's point
        // coordinates, normals, centre, dlambda, 8 terms out, the cell's x, u and moments
        return c->n_bface * (4 + 8 + 1 + 4 + 48 + 48 + 16 + 8 + 64 + 36 * 8 + 32 + 72 +
                             (c->gas.mu > 0 ? 8 + 8 + 16 : 0));
    case DNDSMI_K_LUSGS_MATVEC: // per scan position cell, D, volume, x own in, Ax out; per slot both
        // neighbour indices, face, normal, area, u and x of the neighbour, lambda
        return nc * (4 + 8 + 8 + 32 + 32) + c2f * (8 + 4 + 16 + 8 + 32 + 32 + 8);
    case DNDSMI_K_LUSGS_SWEEP: // one whole forward (or backward) sweep, all colours together: per scan position
        // cell, volume, 1/D, rhs (or the increment) in, increment out; per slot the neighbour index, and per used
        // slot (half of them) face, normal, area, u and the increment of the neighbour, lambda
        return nc * (4 + 8 + 8 + 32 + 32) + c2f * 4 + c2f / 2 * (4 + 16 + 8 + 32 + 32 + 8);
    case DNDSMI_K_DELTA32: // FP32 M per slot + slot index; d read + r read + d write (FP32)
        return c2f * (81 * 4 + 4) + nc * (36 * 4 * 3);
    case DNDSMI_K_RHS_FACE32: // as DNDSMI_K_RHS_FACE with 2 x 8 FP32 terms out
        return nif * (8 + 1 + 96 + (c->straight_faces ? 16 : 48) + 8 + 64) + nc * (36 * 8 + 32 + 72);
    case DNDSMI_K_RHS_CELL32: // 8 FP32 terms per slot, volume, x read+write, xLast (+ last-slot index)
        return c2f * 32 + nc * (8 + 32 * 3 + (c->nfmax == 4 ? 4 : 0));
    case DNDSMI_K_READ_PROBE: // the sweep's A^-1 B pairs (41 x 16 B per slot) and slot index, one double out
        return c2f * (kMPairs * 16 + 4) + nc * 8;
    default:
        return 0;
    }
}
