// gfx950 device code of the FP64 hot path.
//
// Layouts (all device arrays padded to ld = n rounded up to 64):
//   DOF arrays   u[cell][4]             (AoS, 32 B per cell)
//   REC arrays   x[k][cell], k = r + 9 v (SoA: lane-contiguous per element)
//   A^-1 B       M[slot][e/2][cell][e%2], e = r*9 + c (pairs of doubles)
//   A^-1 b       bv[(slot*9 + r)][cell]
//   face rows    d0[(g*18 + side*9 + c)][face]
// One thread owns one cell (sweep, gather) or one face (fluxes); the cell
// gathers visit a cell's faces in ascending face index, which reproduces the
// reference's face-loop scatter order without atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dndsmi_dev {

constexpr double kVerySmallReal = 1e-200;
constexpr double kLine3W0 = 5. / 9., kLine3W1 = 8. / 9., kLine3W2 = 5. / 9.;

enum : int32_t {
    BC_INNER = -1, BC_WALL = 1, BC_FARFIELD = 2, BC_WALL_EULER = 3, BC_WALL_NOSLIP = 4,
    BC_DMRFAR = -2, BC_RTFAR = -3, BC_IVFAR = -4, BC_2DRIEMANNFAR = -5
};
// bits of the device error word
enum : int32_t { ERRBIT_NONPOSITIVE = 1, ERRBIT_NAN = 2, ERRBIT_BC = 4 };

__device__ inline bool is_far(int32_t bc) {
    return bc == BC_FARFIELD || bc == BC_DMRFAR || bc == BC_RTFAR || bc == BC_IVFAR || bc == BC_2DRIEMANNFAR;
}

__device__ inline double sqr(double x) { return x * x; }

// MI355X dispatches workgroups round-robin over its 8 XCDs (workgroup w runs
// on XCD w % 8), and each XCD has its own L2.  xcd_block() renumbers the
// workgroups so that every XCD runs one contiguous range of blocks: cells and
// faces that gather the same neighbours then meet in the same L2.  The map is
// a bijection of [0, gridDim.x) for any grid size.
constexpr int32_t kNumXCD = 8;

__device__ __forceinline__ int32_t xcd_block() {
    const int32_t n = int32_t(gridDim.x), w = int32_t(blockIdx.x);
    const int32_t x = w % kNumXCD, q = w / kNumXCD;
    const int32_t base = n / kNumXCD, extra = n % kNumXCD; // XCD x < extra runs base + 1 blocks
    return x * base + (x < extra ? x : extra) + q;
}

// 16-byte vectors for streaming loads of the A^-1 B blocks: 81 entries per
// slot stored as 41 FP64 pairs (FP64 path) or 21 FP32 quads (mixed path)
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));
constexpr int kMPairs = 41, kMQuads = 21;

// Layout of the A^-1 B stream: per slot, blocks of 64 consecutive cells,
// and within a block the 16-byte vectors (pair e or quad q) of all 64 cells
// contiguous, vector after vector.  A wavefront (64 cells) then reads one
// contiguous 41 KiB (FP64) or 21 KiB (FP32) run per slot instead of 41 or 21
// runs of 1 KiB spread over the whole array.  ldc is a multiple of 64.
__host__ __device__ __forceinline__ int64_t m_vec_index(int s, int e, int64_t i, int64_t ldc, int nvec) {
    return ((int64_t(s) * (ldc >> 6) + (i >> 6)) * nvec + e) * 64 + (i & 63);
}

struct GasParams {
    double gamma;
    double far[4];
    int32_t rs; // RiemannSolverType: 1 Roe, 2 HLLC, 3 HLLEP, 11-15 Roe_M1-M5
    // laminar viscosity (Sutherland) and conduction; mu = 0: inviscid
    double mu, pr, cp, tref, csuth;
};

// ---------------------------------------------------------------- gas ----
__device__ inline void ideal_gas_thermal(double E, double rho, double vSqr, double gamma, double &p, double &asqr,
                                         double &H) {
    p = (gamma - 1) * (E - rho * 0.5 * vSqr);
    asqr = gamma * p / rho;
    H = (E + p) / rho;
}

__device__ inline void cons2prim(const double U[4], double prim[4], double gamma) {
    for (int k = 0; k < 4; k++)
        prim[k] = U[k] / U[0];
    double v0 = U[1] / U[0], v1 = U[2] / U[0];
    double vSqr = v0 * v0 + v1 * v1;
    double p = (gamma - 1) * (U[3] - U[0] * 0.5 * vSqr);
    prim[0] = U[0];
    prim[3] = p;
}

__device__ inline void prim2cons(const double prim[4], double U[4], double gamma) {
    for (int k = 0; k < 4; k++)
        U[k] = prim[k] * prim[0];
    double vSqr = prim[1] * prim[1] + prim[2] * prim[2];
    double E = prim[3] / (gamma - 1) + prim[0] * 0.5 * vSqr;
    U[0] = prim[0];
    U[3] = E;
}

__device__ inline void inviscid_flux(const double U[4], double vx, double p, double F[4]) {
    for (int k = 0; k < 4; k++)
        F[k] = U[k] * vx;
    F[1] += p;
    F[3] += vx * p;
}

// Roe flux, face frame (DNDS_Gas.hpp:405-603).  scheme 0: the Harten-Yee
// entropy fix (riemannSolverType Roe); 1..5: the eigenvalue variants
// Roe_M1..Roe_M5 (:461-518): 1 cLLF_M (per-side speeds, sound speeds capped at
// |u| / 0.2), 2 the local Lax flux, 3 / 4 the low- / increased-dissipation
// Roe_M of Fleischmann et al. (2020), 5 no fix.  std::min / std::max are
// written as the comparisons they evaluate.
__device__ inline bool roe_flux(const double UL[4], const double UR[4], double gamma, double dLambda, double F[4],
                                int scheme = 0) {
    if (!(UL[0] > 0 && UR[0] > 0))
        return false;
    const double vL0 = UL[1] / UL[0], vL1 = UL[2] / UL[0];
    const double vR0 = UR[1] / UR[0], vR1 = UR[2] / UR[0];
    double asqrL, asqrR, pL, pR, HL, HR;
    ideal_gas_thermal(UL[3], UL[0], vL0 * vL0 + vL1 * vL1, gamma, pL, asqrL, HL);
    ideal_gas_thermal(UR[3], UR[0], vR0 * vR0 + vR1 * vR1, gamma, pR, asqrR, HR);
    const double sL = sqrt(UL[0]), sR = sqrt(UR[0]);
    const double vRoe0 = (sL * vL0 + sR * vR0) / (sL + sR);
    const double vRoe1 = (sL * vL1 + sR * vR1) / (sL + sR);
    const double vsqrRoe = vRoe0 * vRoe0 + vRoe1 * vRoe1;
    const double HRoe = (sL * HL + sR * HR) / (sL + sR);
    const double asqrRoe = (gamma - 1) * (HRoe - 0.5 * vsqrRoe);
    double FL[4], FR[4];
    inviscid_flux(UL, vL0, pL, FL);
    inviscid_flux(UR, vR0, pR, FR);
    if (!(asqrRoe > 0))
        return false;
    const double aRoe = sqrt(asqrRoe);
    double lam0 = fabs(vRoe0 - aRoe);
    double lam123 = fabs(vRoe0);
    double lam4 = fabs(vRoe0 + aRoe);
    auto dmin = [](double a, double b) { return b < a ? b : a; };
    auto dmax = [](double a, double b) { return a < b ? b : a; };
    if (scheme == 0) {
        const double thr = fmax(0.05 * (sqrt(vsqrRoe) + aRoe), dLambda);
        const double thrS = thr * thr;
        if (lam0 < thr)
            lam0 = (sqr(lam0) + thrS) / (2 * thr);
        if (lam4 < thr)
            lam4 = (sqr(lam4) + thrS) / (2 * thr);
    } else if (scheme == 1) {
        const double aL = dmin(sqrt(asqrL), fabs(vL0) / 0.2), aR = dmin(sqrt(asqrR), fabs(vR0) / 0.2);
        lam0 = dmax(fabs(vL0 - aL), fabs(vR0 - aR));
        lam123 = dmax(fabs(vL0), fabs(vR0));
        lam4 = dmax(fabs(vL0 + aL), fabs(vR0 + aR));
    } else if (scheme == 2) {
        const double lam = dmax(fabs(vL0) + sqrt(asqrL), fabs(vR0) + sqrt(asqrR));
        for (int k = 0; k < 4; k++)
            F[k] = (FL[k] + FR[k]) * 0.5 - 0.5 * lam * (UR[k] - UL[k]);
        return true;
    } else if (scheme == 3) {
        const double aStar = dmin(fabs(vRoe0) / 0.2, aRoe);
        lam0 = fabs(vRoe0 - aStar);
        lam4 = fabs(vRoe0 + aStar);
    } else if (scheme == 4) {
        const double sgn = vRoe0 > 0 ? 1.0 : (vRoe0 < 0 ? -1.0 : 0.0);
        const double uStar = sgn * dmax(aRoe * 0.2, fabs(vRoe0));
        lam0 = fabs(uStar - aRoe);
        lam123 = fabs(uStar);
        lam4 = fabs(uStar + aRoe);
    }
    double inc[4];
    for (int k = 0; k < 4; k++)
        inc[k] = UR[k] - UL[k];
    double a2 = inc[2] - vRoe1 * inc[0];
    double inc4b = inc[3] - a2 * vRoe1;
    double a1 = (gamma - 1) / asqrRoe * (inc[0] * (HRoe - vRoe0 * vRoe0) + vRoe0 * inc[1] - inc4b);
    double a0 = (inc[0] * (vRoe0 + aRoe) - inc[1] - aRoe * a1) / (2 * aRoe);
    double a3 = inc[0] - (a0 + a1);
    const double la0 = lam0 * a0, la1 = lam123 * a1, la2 = lam123 * a2, la3 = lam4 * a3;
    // right eigenvectors (EulerGasRightEigenVector), rows dotted with lam*alpha in column order
    double r;
    r = 0.0;
    r += 1.0 * la0, r += 1.0 * la1, r += 0.0 * la2, r += 1.0 * la3;
    F[0] = (FL[0] + FR[0]) * 0.5 - 0.5 * r;
    r = 0.0;
    r += (vRoe0 - aRoe) * la0, r += vRoe0 * la1, r += 0.0 * la2, r += (vRoe0 + aRoe) * la3;
    F[1] = (FL[1] + FR[1]) * 0.5 - 0.5 * r;
    r = 0.0;
    r += vRoe1 * la0, r += vRoe1 * la1, r += 1.0 * la2, r += vRoe1 * la3;
    F[2] = (FL[2] + FR[2]) * 0.5 - 0.5 * r;
    r = 0.0;
    r += (HRoe - vRoe0 * aRoe) * la0, r += (0.5 * vsqrRoe) * la1, r += vRoe1 * la2, r += (HRoe + vRoe0 * aRoe) * la3;
    F[3] = (FL[3] + FR[3]) * 0.5 - 0.5 * r;
    return true;
}

// HLLEP flux in the reference's V1 form (DNDS_Gas.hpp:178-310), face frame:
// Roe's wave strengths and right eigenvectors with the eigenvalues replaced by
// the HLL-type speeds eV = ((S+ + S-) lam - 2 (1 - delta) S+ S-) / (S+ - S-),
// delta = 0 on the acoustic waves and a / (a + a + verySmallReal) on the
// contact and shear waves.  Comparisons are written as std::min / std::max
// evaluate them, so signed zeros come out as in the reference.
__device__ inline bool hllep_flux(const double UL[4], const double UR[4], double gamma, double F[4]) {
    if (!(UL[0] > 0 && UR[0] > 0))
        return false;
    const double vL0 = UL[1] / UL[0], vL1 = UL[2] / UL[0];
    const double vR0 = UR[1] / UR[0], vR1 = UR[2] / UR[0];
    double asqrL, asqrR, pL, pR, HL, HR;
    ideal_gas_thermal(UL[3], UL[0], vL0 * vL0 + vL1 * vL1, gamma, pL, asqrL, HL);
    ideal_gas_thermal(UR[3], UR[0], vR0 * vR0 + vR1 * vR1, gamma, pR, asqrR, HR);
    const double sL = sqrt(UL[0]), sR = sqrt(UR[0]);
    const double vRoe0 = (sL * vL0 + sR * vR0) / (sL + sR);
    const double vRoe1 = (sL * vL1 + sR * vR1) / (sL + sR);
    const double vsqrRoe = vRoe0 * vRoe0 + vRoe1 * vRoe1;
    const double HRoe = (sL * HL + sR * HR) / (sL + sR);
    const double asqrRoe = (gamma - 1) * (HRoe - 0.5 * vsqrRoe);
    if (!(asqrRoe > 0 && asqrL > 0 && asqrR > 0))
        return false;
    const double aRoe = sqrt(asqrRoe);
    const double lam0 = vRoe0 - aRoe, lam4 = vRoe0 + aRoe;
    double inc[4];
    for (int k = 0; k < 4; k++)
        inc[k] = UR[k] - UL[k];
    const double a2 = inc[2] - vRoe1 * inc[0];
    const double inc4b = inc[3] - a2 * vRoe1;
    const double a1 = (gamma - 1) / asqrRoe * (inc[0] * (HRoe - vRoe0 * vRoe0) + vRoe0 * inc[1] - inc4b);
    const double a0 = (inc[0] * lam4 - inc[1] - aRoe * a1) / (2 * aRoe);
    const double a3 = inc[0] - (a0 + a1);
    const double cL = vL0 - sqrt(asqrL), cR = vR0 + sqrt(asqrR);
    const double SL = cL < lam0 ? cL : lam0;
    const double SR = lam4 < cR ? cR : lam4;
    double FL[4], FR[4];
    inviscid_flux(UL, vL0, pL, FL);
    inviscid_flux(UR, vR0, pR, FR);
    const double Sp = SR < 0.0 ? 0.0 : SR, Sn = 0.0 < SL ? 0.0 : SL;
    const double delta1 = aRoe / (aRoe + aRoe + kVerySmallReal);
    const double eV1 = ((Sp + Sn) * vRoe0 - 2.0 * (1.0 - delta1) * (Sp * Sn)) / (Sp - Sn);
    const double eV2 = ((Sp + Sn) * (vRoe0 + aRoe) - 2.0 * (1.0 - 0.0) * (Sp * Sn)) / (Sp - Sn);
    const double eV3 = ((Sp + Sn) * (vRoe0 - aRoe) - 2.0 * (1.0 - 0.0) * (Sp * Sn)) / (Sp - Sn);
    const double la0 = fabs(eV3) * a0, la1 = fabs(eV1) * a1, la2 = fabs(eV1) * a2, la3 = fabs(eV2) * a3;
    double r;
    r = 0.0;
    r += 1.0 * la0, r += 1.0 * la1, r += 0.0 * la2, r += 1.0 * la3;
    F[0] = (FL[0] + FR[0]) * 0.5 - 0.5 * r;
    r = 0.0;
    r += (vRoe0 - aRoe) * la0, r += vRoe0 * la1, r += 0.0 * la2, r += (vRoe0 + aRoe) * la3;
    F[1] = (FL[1] + FR[1]) * 0.5 - 0.5 * r;
    r = 0.0;
    r += vRoe1 * la0, r += vRoe1 * la1, r += 1.0 * la2, r += vRoe1 * la3;
    F[2] = (FL[2] + FR[2]) * 0.5 - 0.5 * r;
    r = 0.0;
    r += (HRoe - vRoe0 * aRoe) * la0, r += (0.5 * vsqrRoe) * la1, r += vRoe1 * la2, r += (HRoe + vRoe0 * aRoe) * la3;
    F[3] = (FL[3] + FR[3]) * 0.5 - 0.5 * r;
    return true;
}

// HLLC flux with Einfeldt speeds, face frame (DNDS_Gas.hpp:312-403)
__device__ inline bool hllc_flux(const double UL[4], const double UR[4], double gamma, double F[4]) {
    if (!(UL[0] > 0 && UR[0] > 0))
        return false;
    const double vL0 = UL[1] / UL[0], vL1 = UL[2] / UL[0];
    const double vR0 = UR[1] / UR[0], vR1 = UR[2] / UR[0];
    double asqrL, asqrR, pL, pR, HL, HR;
    ideal_gas_thermal(UL[3], UL[0], vL0 * vL0 + vL1 * vL1, gamma, pL, asqrL, HL);
    ideal_gas_thermal(UR[3], UR[0], vR0 * vR0 + vR1 * vR1, gamma, pR, asqrR, HR);
    const double sL = sqrt(UL[0]), sR = sqrt(UR[0]);
    const double vRoe0 = (sL * vL0 + sR * vR0) / (sL + sR);
    const double eta2 = 0.5 * (sL * sR) / sqr(sL + sR);
    const double dsqr = (asqrL * sL + asqrR * sR) / (sL + sR) + eta2 * sqr(vR0 - vL0);
    if (!(dsqr > 0))
        return false;
    const double SL = vRoe0 - sqrt(dsqr);
    const double SR = vRoe0 + sqrt(dsqr);
    double FL[4], FR[4];
    inviscid_flux(UL, vL0, pL, FL);
    inviscid_flux(UR, vR0, pR, FR);
    if (0 <= SL) {
        for (int k = 0; k < 4; k++)
            F[k] = FL[k];
        return true;
    }
    if (SR <= 0) {
        for (int k = 0; k < 4; k++)
            F[k] = FR[k];
        return true;
    }
    double SS = 0;
    const double div = (UL[0] * (SL - vL0) - UR[0] * (SR - vR0));
    if (fabs(div) > kVerySmallReal)
        SS = (pR - pL + UL[1] * (SL - vL0) - UR[1] * (SR - vR0)) / div;
    const double DS[4] = {0, 1, 0, SS};
    if (SS >= 0) {
        const double dv = SL - SS;
        if (fabs(dv) < kVerySmallReal) {
            for (int k = 0; k < 4; k++)
                F[k] = FL[k];
        } else {
            const double X = (pL + UL[0] * (SL - vL0) * (SS - vL0)) * SL;
            for (int k = 0; k < 4; k++)
                F[k] = ((UL[k] * SL - FL[k]) * SS + DS[k] * X) / dv;
        }
    } else {
        const double dv = SR - SS;
        if (fabs(dv) < kVerySmallReal) {
            for (int k = 0; k < 4; k++)
                F[k] = FR[k];
        } else {
            const double X = (pR + UR[0] * (SR - vR0) * (SS - vR0)) * SR;
            for (int k = 0; k < 4; k++)
                F[k] = ((UR[k] * SR - FR[k]) * SS + DS[k] * X) / dv;
        }
    }
    return true;
}

// generateBoundaryValue (DNDS_FV_EulerEvaluator.hpp:958-1275); returns false
// on a state or position the reference asserts on.
__device__ inline bool boundary_value(const GasParams &gp, const double UL[4], double n0, double n1, double px,
                                      double py, double t, int32_t bc, double UR[4]) {
    if (!(UL[0] > 0))
        return false;
    const double gamma = gp.gamma;
    if (bc == BC_FARFIELD) {
        const double un = (UL[1] * n0 + UL[2] * n1) / UL[0];
        const double vsqr = sqr(UL[1] / UL[0]) + sqr(UL[2] / UL[0]);
        double asqr, H, p;
        ideal_gas_thermal(UL[3], UL[0], vsqr, gamma, p, asqr, H);
        if (!(asqr >= 0))
            return false;
        const double a = sqrt(asqr);
        if (un - a > 0) {
            for (int k = 0; k < 4; k++)
                UR[k] = UL[k];
        } else if (un > 0) {
            double fp[4], lp[4];
            cons2prim(gp.far, fp, gamma);
            cons2prim(UL, lp, gamma);
            lp[3] = fp[3];
            prim2cons(lp, UR, gamma);
        } else if (un + a > 0) {
            double fp[4], lp[4];
            cons2prim(gp.far, fp, gamma);
            cons2prim(UL, lp, gamma);
            fp[3] = lp[3];
            prim2cons(fp, UR, gamma);
        } else {
            for (int k = 0; k < 4; k++)
                UR[k] = gp.far[k];
        }
        return true;
    }
    if (bc == BC_DMRFAR) {
        const double pi = 3.141592653589793;
        if (((px - 10.0 / sin(pi / 3) * t - 1. / 6.) - py / tan(pi / 3)) > 0) {
            UR[0] = 1.4, UR[1] = 0, UR[2] = 0, UR[3] = 2.5;
        } else {
            UR[0] = 8, UR[1] = 57.157676649772960, UR[2] = -33, UR[3] = 5.635e2;
        }
        return true;
    }
    if (bc == BC_RTFAR) {
        const double pi = 3.141592653589793;
        const double vsqr = sqr(UL[1] / UL[0]) + sqr(UL[2] / UL[0]);
        double asqr, H, p;
        ideal_gas_thermal(UL[3], UL[0], vsqr, gamma, p, asqr, H);
        if (!(asqr >= 0))
            return false;
        const double a = sqrt(asqr);
        const double v = -0.025 * a * cos(px * 8 * pi);
        const double rho = py < 0.5 ? 2 : 1;
        const double pr = py < 0.5 ? 1 : 2.5;
        UR[0] = rho, UR[1] = 0, UR[2] = rho * v, UR[3] = 0.5 * rho * sqr(v) + pr / (gamma - 1);
        return true;
    }
    if (bc == BC_IVFAR) {
        const double pi = 3.141592653589793;
        const double chi = 5, xc = 5 + t, yc = 5 + t;
        const double r = sqrt(sqr(px - xc) + sqr(py - yc));
        const double dT = -(gamma - 1) / (8 * gamma * sqr(pi)) * sqr(chi) * exp(1 - sqr(r));
        const double dux = chi / 2 / pi * exp((1 - sqr(r)) / 2) * -(py - xc);
        const double duy = chi / 2 / pi * exp((1 - sqr(r)) / 2) * +(px - yc);
        const double T = dT + 1, ux = dux + 1, uy = duy + 1;
        const double rho = pow(T / 1.0, 1 / (gamma - 1));
        const double pr = T * rho;
        UR[0] = rho, UR[1] = rho * ux, UR[2] = rho * uy;
        UR[3] = 0.5 * (sqr(ux) + sqr(uy)) * rho + pr / (gamma - 1);
        return true;
    }
    if (bc == BC_2DRIEMANNFAR) {
        const double phi2 = -0.422115882408869;
        const double p2 = 0.8 + phi2 * t;
        double lp[4];
        cons2prim(UL, lp, gamma);
        double fp[4] = {lp[0], lp[1], lp[2], lp[3]};
        const double bTol = 1e-9;
        double rho, u, v, pre;
        if (fabs(px - 0.0) < bTol) {
            if (py <= p2)
                rho = 0.137992831541219, u = 1.206045378311055, v = 1.206045378311055, pre = 0.029032258064516;
            else
                rho = 0.532258064516129, u = 1.206045378311055, v = 0.0, pre = 0.3;
        } else if (fabs(px - 1.0) < bTol) {
            rho = lp[0], u = -lp[1], v = lp[2], pre = lp[3];
        } else if (fabs(py - 1.0) < bTol) {
            rho = lp[0], u = lp[1], v = -lp[2], pre = lp[3];
        } else if (fabs(py - 0.0) < bTol) {
            if (px <= p2)
                rho = 0.137992831541219, u = 1.206045378311055, v = 1.206045378311055, pre = 0.029032258064516;
            else
                rho = 0.532258064516129, u = 0.0, v = 1.206045378311055, pre = 0.3;
        } else {
            return false;
        }
        fp[0] = rho, fp[1] = u, fp[2] = v, fp[3] = pre;
        prim2cons(fp, UR, gamma);
        return true;
    }
    if (bc == BC_WALL_EULER) {
        for (int k = 0; k < 4; k++)
            UR[k] = UL[k];
        const double d = UR[1] * n0 + UR[2] * n1;
        UR[1] -= d * n0;
        UR[2] -= d * n1;
        return true;
    }
    if (bc == BC_WALL_NOSLIP) {
        for (int k = 0; k < 4; k++)
            UR[k] = UL[k];
        UR[1] *= -1;
        UR[2] *= -1;
        return true;
    }
    return false;
}

// CompressRecPart (DNDS_FV_EulerEvaluator.hpp:1277-1360)
__device__ inline void compress_rec_part(const double um[4], const double inc[4], double ret[4]) {
    for (int k = 0; k < 4; k++)
        ret[k] = um[k] + inc[k];
    if (ret[0] < 0)
        for (int k = 0; k < 4; k++)
            ret[k] = um[k];
    const double eK = (ret[1] * ret[1] + ret[2] * ret[2]) * 0.5 / (kVerySmallReal + ret[0]);
    const double e = ret[3] - eK;
    if (e < 0)
        for (int k = 0; k < 4; k++)
            ret[k] = um[k];
}

__device__ inline void normalized(double x, double y, double &ox, double &oy) {
    const double z = x * x + y * y;
    if (z > 0) {
        const double s = sqrt(z);
        ox = x / s, oy = y / s;
    } else
        ox = x, oy = y;
}

__device__ inline void stable_normalized(double x, double y, double &ox, double &oy) {
    const double w = fmax(fabs(x), fabs(y));
    const double zx = x / w, zy = y / w;
    const double z = zx * zx + zy * zy;
    if (z > 0) {
        const double s = sqrt(z);
        ox = (x / s) / w, oy = (y / s) / w;
    } else
        ox = x, oy = y;
}

// fluxFace (DNDS_FV_EulerEvaluator.hpp:371-591) for muGas = 0, in two parts.
// balance_fluxes: the flux-balance fluxes FLfix, FRfix of the two cell means
// in the face frame, rotated back (:487-506).
__device__ inline bool balance_fluxes(const GasParams &gp, const double ULm[4], const double URm[4], double n0,
                                      double n1, double FLfix[4], double FRfix[4]) {
    double Lm[4], Rm[4], wL[4], wR[4];
    Lm[0] = ULm[0], Lm[1] = n0 * ULm[1] + n1 * ULm[2], Lm[2] = (-n1) * ULm[1] + n0 * ULm[2], Lm[3] = ULm[3];
    Rm[0] = URm[0], Rm[1] = n0 * URm[1] + n1 * URm[2], Rm[2] = (-n1) * URm[1] + n0 * URm[2], Rm[3] = URm[3];
    if (!(Lm[0] > 0) || !(Rm[0] > 0))
        return false;
    cons2prim(Lm, wL, gp.gamma);
    cons2prim(Rm, wR, gp.gamma);
    inviscid_flux(Lm, wL[1], wL[3], FLfix);
    inviscid_flux(Rm, wR[1], wR[3], FRfix);
    double a = n0 * FLfix[1] + (-n1) * FLfix[2], b = n1 * FLfix[1] + n0 * FLfix[2];
    FLfix[1] = a, FLfix[2] = b;
    a = n0 * FRfix[1] + (-n1) * FRfix[2], b = n1 * FRfix[1] + n0 * FRfix[2];
    FRfix[1] = a, FRfix[2] = b;
    return true;
}

// riemann_face: the Riemann flux in the face frame, rotated back and negated
// (:519-591).
__device__ inline bool riemann_face(const GasParams &gp, const double ULxy[4], const double URxy[4], double n0,
                                    double n1, double dLambda, double out[4]) {
    double UL[4], UR[4];
    UL[0] = ULxy[0], UL[1] = n0 * ULxy[1] + n1 * ULxy[2], UL[2] = (-n1) * ULxy[1] + n0 * ULxy[2], UL[3] = ULxy[3];
    UR[0] = URxy[0], UR[1] = n0 * URxy[1] + n1 * URxy[2], UR[2] = (-n1) * URxy[1] + n0 * URxy[2], UR[3] = URxy[3];
    double f[4];
    const bool ok = gp.rs == 2   ? hllc_flux(UL, UR, gp.gamma, f)
                    : gp.rs == 3 ? hllep_flux(UL, UR, gp.gamma, f)
                                 : roe_flux(UL, UR, gp.gamma, dLambda, f, gp.rs > 10 ? gp.rs - 10 : 0);
    if (!ok)
        return false;
    const double a = n0 * f[1] + (-n1) * f[2];
    const double b = n1 * f[1] + n0 * f[2];
    out[0] = -f[0], out[1] = -a, out[2] = -b, out[3] = -f[3];
    return true;
}

__device__ inline bool flux_face(const GasParams &gp, const double ULxy[4], const double URxy[4],
                                 const double ULm[4], const double URm[4], double n0, double n1, double dLambda,
                                 double out[4], double FLfix[4], double FRfix[4]) {
    return balance_fluxes(gp, ULm, URm, n0, n1, FLfix, FRfix) && riemann_face(gp, ULxy, URxy, n0, n1, dLambda, out);
}

// ---------------------------------------------------------- viscous ----
// Sutherland's law at the mean state (DNDS_FV_EulerEvaluator.hpp:397-412,
// DNDS_FV_EulerEvaluator.cpp:79-85), T = p / ((gamma - 1) / gamma Cp rho);
// (T / TRef)^1.5 is formed as t sqrt(t), within rounding of the power.
__device__ inline double sutherland_mu(const GasParams &gp, double p, double rho) {
    const double T = p / ((gp.gamma - 1) / gp.gamma * gp.cp * rho);
    const double tr = T / gp.tref;
    return gp.mu * (tr * sqrt(tr)) * (gp.tref + gp.csuth) / (T + gp.csuth);
}

// Gas::ViscousFlux_IdealGas<2> (DNDS_Gas.hpp:795-826) dotted with the unit
// normal: vn[v] = sum_d Flux[d][v] n_d for the mean state U and the
// conservative gradient G[d][v]; adiabatic removes the normal part of grad T.
__device__ inline void viscous_flux_n(const double U[4], const double G[2][4], double n0, double n1, bool adiabatic,
                                      double gamma, double mu, double k, double Cp, double vn[4]) {
    const double velo0 = U[1] / U[0], velo1 = U[2] / U[0];
    const double lambda = -2. / 3.;
    double S[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
            S[i][j] = (1.0 / (U[0] * U[0])) * (U[0] * G[i][1 + j] - G[i][0] * U[1 + j]);
    double GradP[2];
#pragma unroll
    for (int i = 0; i < 2; i++)
        GradP[i] = (gamma - 1) *
                   (G[i][3] - 0.5 * ((G[i][1] * velo0 + G[i][2] * velo1) + (S[i][0] * U[1] + S[i][1] * U[2])));
    const double vSqr = velo0 * velo0 + velo1 * velo1;
    const double p = (gamma - 1) * (U[3] - U[0] * 0.5 * vSqr);
    double GradT[2];
#pragma unroll
    for (int i = 0; i < 2; i++)
        GradT[i] = (gamma / ((gamma - 1) * Cp * U[0] * U[0])) * (U[0] * GradP[i] - p * G[i][0]);
    if (adiabatic) {
        const double d = GradT[0] * n0 + GradT[1] * n1;
        GradT[0] -= d * n0, GradT[1] -= d * n1;
    }
    const double tr = S[0][0] + S[1][1];
    double F[2][4];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        F[i][0] = 0.0;
#pragma unroll
        for (int j = 0; j < 2; j++)
            F[i][1 + j] = (S[i][j] + S[j][i]) * mu + (i == j ? 1.0 : 0.0) * (lambda * mu * tr);
        F[i][3] = (F[i][1] * velo0 + F[i][2] * velo1) + k * GradT[i];
    }
#pragma unroll
    for (int v = 0; v < 4; v++)
        vn[v] = (F[0][v] * n0 + F[1][v] * n1) * 1;
}

// ------------------------------------------------------------ bases ----
// The scaled cubic monomials of FDiffBaseValue in global coordinates
// (DNDS_FV_VR.hpp:300-356), evaluated at a face point from its scaled
// coordinates (X, Y) = ((px - cx) / sx, (py - cy) / sy) instead of being
// stored per face: entry (r, ib) of the 10 x 10 DiBj block is
//   ((coef(r, ib) X^(bx - rx)) Y^(by - ry)) / (sx^rx sy^ry)   (r > 0)
//   X^bx Y^by - moment[ib]                                    (r = 0)
// with (bx, by), (rx, ry) the exponents of base ib and derivative r. Integer
// powers are formed by repeated multiplication and every operation is the
// host setup's, in its order, so each value equals the setup's bit for bit.
constexpr int kBX[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0}, kBY[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
constexpr int kDFac[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 2, 0}, {1, 3, 6, 6}};

__host__ __device__ constexpr int dibj_coef(int r, int ib) {
    return (kBX[r] <= kBX[ib] ? kDFac[kBX[ib]][kBX[r]] : 0) * (kBY[r] <= kBY[ib] ? kDFac[kBY[ib]][kBY[r]] : 0);
}

__device__ __forceinline__ void int_powers(double a, double p[4]) {
    p[0] = 1.0, p[1] = a, p[2] = a * a, p[3] = p[2] * a;
}

// ----------------------------------------------------------- WBAP ----

struct FFParams {
    double W0, n0, n1, t0, t1, w2r, w3r;
};

// parameters of FFaceFunctional with USE_NORM_FUNCTIONAL (DNDS_FV_VR.hpp:1071-1098)
__device__ inline FFParams ff_params(const double W[10], double gx, double gy, double tang_weight, int twd_norm) {
    FFParams p;
    const double epsR = 1e-50;
    double w2r, w3r, length;
    if (fabs(W[1]) >= epsR || fabs(W[2]) >= epsR) {
        w2r = fabs(W[1]) >= fabs(W[2]) ? W[3] / (W[1] * W[1]) : W[5] / (W[2] * W[2]);
        w3r = fabs(W[1]) >= fabs(W[2]) ? W[6] / (W[1] * W[1] * W[1]) : W[9] / (W[2] * W[2] * W[2]);
        length = sqrt(W[1] * W[1] + W[2] * W[2]);
    } else {
        w2r = w3r = 0.0;
        length = 0;
    }
    double n0, n1;
    if (twd_norm) {
        double ux, uy;
        stable_normalized(gx, gy, ux, uy);
        n0 = ux * length, n1 = uy * length;
    } else {
        n0 = W[1], n1 = W[2];
    }
    p.W0 = W[0];
    p.n0 = n0, p.n1 = n1;
    p.t0 = -n1 * (tang_weight * 1.0), p.t1 = n0 * (tang_weight * 1.0);
    p.w2r = w2r, p.w3r = w3r;
    return p;
}

// Conj(col, col) of FFaceFunctional for one column of a 10-row block
// FFaceFunctional's self product of one 10-row derivative vector d, split by
// the groups of rows that its terms mix -- {0}, {1, 2}, {3, 4, 5}, {6..9}:
// ff_group<G> adds group G's terms to c (d points at the group's first row),
// in the order of the full sum, so that calling it for G = 0..3 on the same
// c gives ff_self's value bit for bit.
template <int G>
__device__ inline void ff_group(const FFParams &p, const double *dg, double &c) {
    const double n0 = p.n0, n1 = p.n1, t0 = p.t0, t1 = p.t1, w2r = p.w2r, w3r = p.w3r;
    if (G == 0) {
        const double a0 = dg[0];
        c += a0 * a0 * p.W0 * p.W0;
    } else if (G == 1) {
        const double a1 = dg[0] * n0 + dg[1] * n1;
        const double a2 = dg[0] * t0 + dg[1] * t1;
        c += a1 * a1;
        c += a2 * a2 * 1.0;
    } else if (G == 2) {
        const double d3 = dg[0], d4 = dg[1], d5 = dg[2];
        const double a3 = (d3 * n0 * n0 + d4 * n0 * n1 * 2 + d5 * n1 * n1) * w2r;
        const double a4 = (d3 * n0 * t0 + d4 * (n0 * t1 + n1 * t0) + d5 * n1 * t1) * w2r;
        const double a5 = (d3 * t0 * t0 + d4 * t0 * t1 * 2 + d5 * t1 * t1) * w2r;
        c += a3 * a3;
        c += a4 * a4 * 2 * 1.0;
        c += a5 * a5 * 1.0;
    } else {
        const double d6 = dg[0], d7 = dg[1], d8 = dg[2], d9 = dg[3];
        const double a6 = (d6 * n0 * n0 * n0 + d7 * n0 * n0 * n1 * 3 + d8 * n0 * n1 * n1 * 3 + d9 * n1 * n1 * n1) * w3r;
        const double a7 = (d6 * n0 * n0 * t0 + d7 * (n0 * n0 * t1 + n0 * t0 * n1 + t0 * n0 * n1) +
                           d8 * (n0 * n1 * t1 + n0 * t1 * n1 + t0 * n1 * n1) + d9 * n1 * n1 * t1) *
                          w3r;
        const double a8 = (d6 * n0 * t0 * t0 + d7 * (t0 * n0 * t1 + n0 * t0 * t1 + t0 * t0 * n1) +
                           d8 * (t0 * n1 * t1 + n0 * t1 * t1 + t0 * t1 * n1) + d9 * n1 * t1 * t1) *
                          w3r;
        const double a9 = (d6 * t0 * t0 * t0 + d7 * t0 * t0 * t1 * 3 + d8 * t0 * t1 * t1 * 3 + d9 * t1 * t1 * t1) * w3r;
        c += a6 * a6;
        c += a7 * a7 * 3 * 1.0;
        c += a8 * a8 * 3 * 1.0;
        c += a9 * a9 * 1.0;
    }
}

__device__ inline double ff_self(const FFParams &p, const double d[10]) {
    double c = 0;
    ff_group<0>(p, d, c);
    ff_group<1>(p, d + 1, c);
    ff_group<2>(p, d + 3, c);
    ff_group<3>(p, d + 6, c);
    return c;
}

// fML / fMR (DNDS_FV_EulerSolver.hpp:749-795): left / right eigenvector
// matrices at the mean of UL and UR in the face frame, mapped back to the
// global frame.  ML, MR row-major 4x4.
__device__ inline bool char_matrices(double gamma, const double UL[4], const double UR[4], double n0, double n1,
                                     double ML[16], double MR[16]) {
    double UC[4];
    for (int k = 0; k < 4; k++)
        UC[k] = (UL[k] + UR[k]) * 0.5;
    const double r1 = n0 * UC[1] + n1 * UC[2];
    const double r2 = (-n1) * UC[1] + n0 * UC[2];
    UC[1] = r1, UC[2] = r2;
    if (!(UC[0] > 0))
        return false;
    const double v0 = UC[1] / UC[0], v1 = UC[2] / UC[0];
    const double vsqr = v0 * v0 + v1 * v1;
    double asqr, p, H;
    ideal_gas_thermal(UC[3], UC[0], vsqr, gamma, p, asqr, H);
    if (!(asqr >= 0))
        return false;
    const double a = sqrt(asqr);
    const double gb = gamma - 1;
    double L[16];
    L[0] = H + a / gb * (v0 - a), L[1] = -v0 - a / gb, L[2] = -v1, L[3] = 1;
    L[4] = -2 * H + 4 / gb * (a * a), L[5] = v0 * 2, L[6] = v1 * 2, L[7] = -2;
    L[8] = v1 * (-2 * (a * a) / gb), L[9] = 0, L[10] = 2 * (a * a) / gb, L[11] = 0;
    L[12] = H - a / gb * (v0 + a), L[13] = -v0 + a / gb, L[14] = -v1, L[15] = 1;
    const double sc = gb / (2 * a * a);
    for (int k = 0; k < 16; k++)
        L[k] *= sc;
    for (int r = 0; r < 4; r++) {
        const double c1 = L[4 * r + 1] * n0 + L[4 * r + 2] * (-n1);
        const double c2 = L[4 * r + 1] * n1 + L[4 * r + 2] * n0;
        ML[4 * r] = L[4 * r], ML[4 * r + 1] = c1, ML[4 * r + 2] = c2, ML[4 * r + 3] = L[4 * r + 3];
    }
    double R[16];
    R[0] = 1, R[1] = 1, R[2] = 0, R[3] = 1;
    R[4] = v0 - a, R[5] = v0, R[6] = 0, R[7] = v0 + a;
    R[8] = v1, R[9] = v1, R[10] = 1, R[11] = v1;
    R[12] = H - v0 * a, R[13] = 0.5 * vsqr, R[14] = v1, R[15] = H + v0 * a;
    for (int c = 0; c < 4; c++) {
        const double a1 = n0 * R[4 + c] + (-n1) * R[8 + c];
        const double a2 = n1 * R[4 + c] + n0 * R[8 + c];
        MR[c] = R[c], MR[4 + c] = a1, MR[8 + c] = a2, MR[12 + c] = R[12 + c];
    }
    return true;
}

__device__ inline double sgn(double x) { return double((x > 0) - (x < 0)); }

} // namespace dndsmi_dev
