// CPU FP64 restatement of DNDS's 2D VR finite-volume hot path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP kernels and the
// CPU baseline of the benchmark; no product code uses it.  Every function
// cites the reference lines it restates; loops run in the reference's order
// (face loop with scatter in face order for the RHS and time step, cell loop
// in cell2face order for the Jacobi sweep) and expressions keep the
// reference's operation order.  Dense products whose order is fixed by Eigen
// internals in the reference (GEMM blocking, vectorised reductions) are
// summed in ascending index order here.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kVerySmallReal = 1e-200;   // DNDS_Defines.h:70
constexpr double kVeryLargeReal = 3e200;    // DNDS_Defines.h:67
const double kPi = std::acos(-1);
const double kLine3W[3] = {5. / 9., 8. / 9., 5. / 9.};
constexpr int64_t kNoCell = INT64_MIN;

enum {
    BC_INNER = -1, BC_WALL = 1, BC_FARFIELD = 2, BC_WALL_EULER = 3, BC_WALL_NOSLIP = 4,
    BC_DMRFAR = -2, BC_RTFAR = -3, BC_IVFAR = -4, BC_2DRIEMANNFAR = -5
};
enum { ERR_OK = 0, ERR_INVALID = 1, ERR_NAN = 3, ERR_NONPOSITIVE = 4, ERR_UNSUPPORTED = 5 };

inline double sqr(double x) { return x * x; }

// Worker threads for the independent per-cell / per-face loops (the
// benchmark's CPU baseline uses the host's cores as the reference uses MPI
// ranks).  Each item is computed exactly as in the serial loop; the scatters
// stay serial, so the results do not depend on the thread count.
int g_threads = 1;

template <class F>
int parallel_items(int64_t n, F &&body) {
    const int T = int(std::min<int64_t>(g_threads, std::max<int64_t>(1, n / 2048)));
    if (T <= 1) {
        for (int64_t i = 0; i < n; i++)
            if (int rc = body(i))
                return rc;
        return 0;
    }
    std::vector<int> rcs(T, 0);
    std::vector<std::thread> pool;
    const int64_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; t++)
        pool.emplace_back([&, t] {
            for (int64_t i = t * chunk; i < std::min(n, (t + 1) * chunk); i++)
                if ((rcs[t] = body(i)))
                    return;
        });
    for (auto &th : pool)
        th.join();
    for (int rc : rcs)
        if (rc)
            return rc;
    return 0;
}

bool is_far(int32_t bc) {
    return bc == BC_FARFIELD || bc == BC_DMRFAR || bc == BC_RTFAR || bc == BC_IVFAR || bc == BC_2DRIEMANNFAR;
}

// Gas::IdealGasThermal, DNDS_Gas.hpp:73-79
inline void ideal_gas_thermal(double E, double rho, double vSqr, double gamma, double &p, double &asqr, double &H) {
    p = (gamma - 1) * (E - rho * 0.5 * vSqr);
    asqr = gamma * p / rho;
    H = (E + p) / rho;
}

// IdealGasThermalConservative2Primitive, DNDS_Gas.hpp:81-92
inline void cons2prim(const double U[4], double prim[4], double gamma) {
    for (int k = 0; k < 4; k++)
        prim[k] = U[k] / U[0];
    double v0 = U[1] / U[0], v1 = U[2] / U[0];
    double vSqr = v0 * v0 + v1 * v1;
    double rho = U[0];
    double E = U[3];
    double p = (gamma - 1) * (E - rho * 0.5 * vSqr);
    prim[0] = rho;
    prim[3] = p;
}

// IdealGasThermalPrimitive2Conservative, DNDS_Gas.hpp:94-105
inline void prim2cons(const double prim[4], double U[4], double gamma) {
    for (int k = 0; k < 4; k++)
        U[k] = prim[k] * prim[0];
    double vSqr = prim[1] * prim[1] + prim[2] * prim[2];
    double rho = prim[0];
    double p = prim[3];
    double E = p / (gamma - 1) + rho * 0.5 * vSqr;
    U[0] = rho;
    U[3] = E;
}

// GasInviscidFlux, DNDS_Gas.hpp:113-119 (x direction)
inline void inviscid_flux(const double U[4], double vx, double p, double F[4]) {
    for (int k = 0; k < 4; k++)
        F[k] = U[k] * vx;
    F[1] += p;
    F[3] += vx * p;
}

// EulerGasRightEigenVector / EulerGasLeftEigenVector, DNDS_Gas.hpp:17-71, dim = 2
void right_eigen(const double velo[2], double vsqr, double H, double a, double R[4][4]) {
    std::memset(R, 0, sizeof(double) * 16);
    R[0][0] = R[0][1] = R[0][3] = 1;
    R[2][2] = 1;
    R[1][0] = velo[0], R[1][1] = velo[0], R[1][3] = velo[0];
    R[2][0] = velo[1], R[2][1] = velo[1], R[2][3] = velo[1];
    R[1][0] -= a;
    R[1][3] += a;
    R[3][0] = H - velo[0] * a;
    R[3][3] = H + velo[0] * a;
    R[3][1] = 0.5 * vsqr;
    R[3][2] = velo[1];
}

void left_eigen(const double velo[2], double vsqr, double H, double a, double gamma, double L[4][4]) {
    (void)vsqr;
    std::memset(L, 0, sizeof(double) * 16);
    double gammaBar = gamma - 1;
    L[0][0] = H + a / gammaBar * (velo[0] - a);
    L[0][1] = -velo[0] - a / gammaBar;
    L[0][2] = -velo[1];
    L[0][3] = 1;
    L[1][0] = -2 * H + 4 / gammaBar * (a * a);
    L[1][1] = velo[0] * 2;
    L[1][2] = velo[1] * 2;
    L[1][3] = -2;
    L[2][0] = velo[1] * (-2 * (a * a) / gammaBar);
    L[2][2] = 2 * (a * a) / gammaBar;
    L[3][0] = H - a / gammaBar * (velo[0] + a);
    L[3][1] = -velo[0] + a / gammaBar;
    L[3][2] = -velo[1];
    L[3][3] = 1;
    const double sc = gammaBar / (2 * a * a);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            L[i][j] *= sc;
}

// Gas::RoeFlux_IdealGas_HartenYee<2, eigScheme>, DNDS_Gas.hpp:405-603.
// eigScheme 0 is the Roe solver with the Harten-Yee fix (riemannSolverType
// Roe); 1..5 are Roe_M1..Roe_M5 (:461-518): 1 cLLF_M, 2 the local Lax
// flux (early exit), 3 and 4 the low- and increased-dissipation Roe_M of
// Fleischmann et al. (2020), 5 no fix.
int roe_flux(const double UL[4], const double UR[4], double gamma, double dLambda, double F[4], int eigScheme) {
    const double scaleHartenYee = 0.05;
    const double scaleLD = 0.2;
    if (!(UL[0] > 0 && UR[0] > 0))
        return ERR_NONPOSITIVE;
    double veloL[2] = {UL[1] / UL[0], UL[2] / UL[0]};
    double veloR[2] = {UR[1] / UR[0], UR[2] / UR[0]};
    double asqrL, asqrR, pL, pR, HL, HR;
    double vsqrL = veloL[0] * veloL[0] + veloL[1] * veloL[1];
    double vsqrR = veloR[0] * veloR[0] + veloR[1] * veloR[1];
    ideal_gas_thermal(UL[3], UL[0], vsqrL, gamma, pL, asqrL, HL);
    ideal_gas_thermal(UR[3], UR[0], vsqrR, gamma, pR, asqrR, HR);
    double sqrtRhoL = std::sqrt(UL[0]);
    double sqrtRhoR = std::sqrt(UR[0]);
    double veloRoe[2];
    for (int k = 0; k < 2; k++)
        veloRoe[k] = (sqrtRhoL * veloL[k] + sqrtRhoR * veloR[k]) / (sqrtRhoL + sqrtRhoR);
    double vsqrRoe = veloRoe[0] * veloRoe[0] + veloRoe[1] * veloRoe[1];
    double HRoe = (sqrtRhoL * HL + sqrtRhoR * HR) / (sqrtRhoL + sqrtRhoR);
    double asqrRoe = (gamma - 1) * (HRoe - 0.5 * vsqrRoe);
    double FL[4], FR[4];
    inviscid_flux(UL, veloL[0], pL, FL);
    inviscid_flux(UR, veloR[0], pR, FR);
    if (!(asqrRoe > 0))
        return ERR_NONPOSITIVE;
    double aRoe = std::sqrt(asqrRoe);
    double lam0 = std::abs(veloRoe[0] - aRoe);
    double lam123 = std::abs(veloRoe[0]);
    double lam4 = std::abs(veloRoe[0] + aRoe);
    if (eigScheme == 0) {
        double thresholdHartenYee = std::max(scaleHartenYee * (std::sqrt(vsqrRoe) + aRoe), dLambda);
        double thresholdHartenYeeS = thresholdHartenYee * thresholdHartenYee;
        if (lam0 < thresholdHartenYee)
            lam0 = (sqr(lam0) + thresholdHartenYeeS) / (2 * thresholdHartenYee);
        if (lam4 < thresholdHartenYee)
            lam4 = (sqr(lam4) + thresholdHartenYeeS) / (2 * thresholdHartenYee);
    } else if (eigScheme == 1) {
        double aL = std::min(std::sqrt(asqrL), std::abs(veloL[0]) / scaleLD);
        double aR = std::min(std::sqrt(asqrR), std::abs(veloR[0]) / scaleLD);
        lam0 = std::max(std::abs(veloL[0] - aL), std::abs(veloR[0] - aR));
        lam123 = std::max(std::abs(veloL[0]), std::abs(veloR[0]));
        lam4 = std::max(std::abs(veloL[0] + aL), std::abs(veloR[0] + aR));
    } else if (eigScheme == 2) {
        lam0 = lam123 = lam4 = std::max(std::abs(veloL[0]) + std::sqrt(asqrL), std::abs(veloR[0]) + std::sqrt(asqrR));
        for (int k = 0; k < 4; k++)
            F[k] = (FL[k] + FR[k]) * 0.5 - 0.5 * lam0 * (UR[k] - UL[k]);
        return ERR_OK;
    } else if (eigScheme == 3) {
        double LDthreshold = std::abs(veloRoe[0]) / scaleLD;
        double aRoeStar = std::min(LDthreshold, aRoe);
        lam0 = std::abs(veloRoe[0] - aRoeStar);
        lam4 = std::abs(veloRoe[0] + aRoeStar);
    } else if (eigScheme == 4) {
        // sign(0) = 0 (DNDS_Defines.h:164-167; USE_SIGN_MINUS_AT_ROE_M4_FLUX is off)
        const double sgn = veloRoe[0] > 0 ? 1 : (veloRoe[0] < 0 ? -1 : 0);
        double uStar = sgn * std::max(aRoe * scaleLD, std::abs(veloRoe[0]));
        lam0 = std::abs(uStar - aRoe);
        lam123 = std::abs(uStar);
        lam4 = std::abs(uStar + aRoe);
    } else if (eigScheme != 5) {
        return ERR_UNSUPPORTED;
    }
    double lam[4] = {lam0, lam123, lam123, lam4};
    double R[4][4];
    right_eigen(veloRoe, vsqrRoe, HRoe, aRoe, R);
    double incU[4];
    for (int k = 0; k < 4; k++)
        incU[k] = UR[k] - UL[k];
    double alpha[4];
    alpha[2] = incU[2] - veloRoe[1] * incU[0];
    double incU4b = incU[3] - alpha[2] * veloRoe[1];
    alpha[1] = (gamma - 1) / asqrRoe * (incU[0] * (HRoe - veloRoe[0] * veloRoe[0]) + veloRoe[0] * incU[1] - incU4b);
    alpha[0] = (incU[0] * (veloRoe[0] + aRoe) - incU[1] - aRoe * alpha[1]) / (2 * aRoe);
    alpha[3] = incU[0] - (alpha[0] + alpha[1]);
    double la[4];
    for (int k = 0; k < 4; k++)
        la[k] = lam[k] * alpha[k];
    for (int r = 0; r < 4; r++) {
        double incF = 0;
        for (int c = 0; c < 4; c++)
            incF += R[r][c] * la[c];
        F[r] = (FL[r] + FR[r]) * 0.5 - 0.5 * incF;
    }
    return ERR_OK;
}

// Gas::HLLCFlux_IdealGas_HartenYee<2>, DNDS_Gas.hpp:312-403
int hllc_flux(const double UL[4], const double UR[4], double gamma, double dLambda, double F[4]) {
    (void)dLambda;
    if (!(UL[0] > 0 && UR[0] > 0))
        return ERR_NONPOSITIVE;
    double veloL[2] = {UL[1] / UL[0], UL[2] / UL[0]};
    double veloR[2] = {UR[1] / UR[0], UR[2] / UR[0]};
    double asqrL, asqrR, pL, pR, HL, HR;
    double vsqrL = veloL[0] * veloL[0] + veloL[1] * veloL[1];
    double vsqrR = veloR[0] * veloR[0] + veloR[1] * veloR[1];
    ideal_gas_thermal(UL[3], UL[0], vsqrL, gamma, pL, asqrL, HL);
    ideal_gas_thermal(UR[3], UR[0], vsqrR, gamma, pR, asqrR, HR);
    double sqrtRhoL = std::sqrt(UL[0]);
    double sqrtRhoR = std::sqrt(UR[0]);
    double veloRoe0 = (sqrtRhoL * veloL[0] + sqrtRhoR * veloR[0]) / (sqrtRhoL + sqrtRhoR);
    double eta2 = 0.5 * (sqrtRhoL * sqrtRhoR) / sqr(sqrtRhoL + sqrtRhoR);
    double dsqr = (asqrL * sqrtRhoL + asqrR * sqrtRhoR) / (sqrtRhoL + sqrtRhoR) + eta2 * sqr(veloR[0] - veloL[0]);
    if (!(dsqr > 0))
        return ERR_NONPOSITIVE;
    double SL = veloRoe0 - std::sqrt(dsqr);
    double SR = veloRoe0 + std::sqrt(dsqr);
    double FL[4], FR[4];
    inviscid_flux(UL, veloL[0], pL, FL);
    inviscid_flux(UR, veloR[0], pR, FR);
    if (0 <= SL) {
        for (int k = 0; k < 4; k++)
            F[k] = FL[k];
        return ERR_OK;
    }
    if (SR <= 0) {
        for (int k = 0; k < 4; k++)
            F[k] = FR[k];
        return ERR_OK;
    }
    double SS = 0;
    double div = (UL[0] * (SL - veloL[0]) - UR[0] * (SR - veloR[0]));
    if (std::abs(div) > kVerySmallReal)
        SS = (pR - pL + UL[1] * (SL - veloL[0]) - UR[1] * (SR - veloR[0])) / div;
    double DS[4] = {0, 1, 0, SS};
    if (SS >= 0) {
        double dv = SL - SS;
        if (std::abs(dv) < kVerySmallReal)
            for (int k = 0; k < 4; k++)
                F[k] = FL[k];
        else {
            double X = (pL + UL[0] * (SL - veloL[0]) * (SS - veloL[0])) * SL;
            for (int k = 0; k < 4; k++)
                F[k] = ((UL[k] * SL - FL[k]) * SS + DS[k] * X) / dv;
        }
    } else {
        double dv = SR - SS;
        if (std::abs(dv) < kVerySmallReal)
            for (int k = 0; k < 4; k++)
                F[k] = FR[k];
        else {
            double X = (pR + UR[0] * (SR - veloR[0]) * (SS - veloR[0])) * SR;
            for (int k = 0; k < 4; k++)
                F[k] = ((UR[k] * SR - FR[k]) * SS + DS[k] * X) / dv;
        }
    }
    return ERR_OK;
}

// Gas::HLLEPFlux_IdealGas<2> in its V1 form (DNDS_GAS_HLLEP_USE_V1 is
// defined at DNDS_Gas.hpp:178), DNDS_Gas.hpp:179-310: Roe's wave strengths
// and right eigenvectors, each eigenvalue replaced by
//   eV = ((S+ + S-) lam - 2 (1 - delta) S+ S-) / (S+ - S-),
// S+ = max(SR, 0), S- = min(SL, 0), SL = min(u - a, uL - aL),
// SR = max(u + a, uR + aR) (Roe averages u, a); delta = 0 for the acoustic
// waves and delta = a / (a + a + verySmallReal) for the contact and shear
// waves (the reference sets the |u| of that coefficient to a, :291-292).
// std::min / std::max are written out as the ternaries they are.
int hllep_flux(const double UL[4], const double UR[4], double gamma, double dLambda, double F[4]) {
    (void)dLambda; // the Harten-Yee lines are commented out in the reference (:224-231)
    if (!(UL[0] > 0 && UR[0] > 0))
        return ERR_NONPOSITIVE;
    double veloL[2] = {UL[1] / UL[0], UL[2] / UL[0]};
    double veloR[2] = {UR[1] / UR[0], UR[2] / UR[0]};
    double asqrL, asqrR, pL, pR, HL, HR;
    double vsqrL = veloL[0] * veloL[0] + veloL[1] * veloL[1];
    double vsqrR = veloR[0] * veloR[0] + veloR[1] * veloR[1];
    ideal_gas_thermal(UL[3], UL[0], vsqrL, gamma, pL, asqrL, HL);
    ideal_gas_thermal(UR[3], UR[0], vsqrR, gamma, pR, asqrR, HR);
    double sqrtRhoL = std::sqrt(UL[0]);
    double sqrtRhoR = std::sqrt(UR[0]);
    double veloRoe[2];
    for (int k = 0; k < 2; k++)
        veloRoe[k] = (sqrtRhoL * veloL[k] + sqrtRhoR * veloR[k]) / (sqrtRhoL + sqrtRhoR);
    double vsqrRoe = veloRoe[0] * veloRoe[0] + veloRoe[1] * veloRoe[1];
    double HRoe = (sqrtRhoL * HL + sqrtRhoR * HR) / (sqrtRhoL + sqrtRhoR);
    double asqrRoe = (gamma - 1) * (HRoe - 0.5 * vsqrRoe);
    if (!(asqrRoe > 0 && asqrL > 0 && asqrR > 0))
        return ERR_NONPOSITIVE;
    double aRoe = std::sqrt(asqrRoe);
    double lam0 = veloRoe[0] - aRoe;
    double lam4 = veloRoe[0] + aRoe;
    double R[4][4];
    right_eigen(veloRoe, vsqrRoe, HRoe, aRoe, R);
    double incU[4];
    for (int k = 0; k < 4; k++)
        incU[k] = UR[k] - UL[k];
    double alpha[4];
    alpha[2] = incU[2] - veloRoe[1] * incU[0];
    double incU4b = incU[3] - alpha[2] * veloRoe[1];
    alpha[1] = (gamma - 1) / asqrRoe * (incU[0] * (HRoe - veloRoe[0] * veloRoe[0]) + veloRoe[0] * incU[1] - incU4b);
    alpha[0] = (incU[0] * lam4 - incU[1] - aRoe * alpha[1]) / (2 * aRoe);
    alpha[3] = incU[0] - (alpha[0] + alpha[1]);
    const double aL = std::sqrt(asqrL), aR = std::sqrt(asqrR);
    const double cL = veloL[0] - aL, cR = veloR[0] + aR;
    double SL = cL < lam0 ? cL : lam0;     // std::min(lam0, uL - aL)
    double SR = lam4 < cR ? cR : lam4;     // std::max(lam4, uR + aR)
    double FL[4], FR[4];
    inviscid_flux(UL, veloL[0], pL, FL);
    inviscid_flux(UR, veloR[0], pR, FR);
    double Sp = SR < 0.0 ? 0.0 : SR;       // std::max(SR, 0.0)
    double Sn = 0.0 < SL ? 0.0 : SL;       // std::min(SL, 0.0)
    double delta1 = aRoe / (aRoe + aRoe + kVerySmallReal);
    double delta2 = 0.0, delta3 = 0.0;
    double un = veloRoe[0];
    double eV1 = ((Sp + Sn) * un - 2.0 * (1.0 - delta1) * (Sp * Sn)) / (Sp - Sn);
    double eV2 = ((Sp + Sn) * (un + aRoe) - 2.0 * (1.0 - delta2) * (Sp * Sn)) / (Sp - Sn);
    double eV3 = ((Sp + Sn) * (un - aRoe) - 2.0 * (1.0 - delta3) * (Sp * Sn)) / (Sp - Sn);
    double lam[4] = {std::abs(eV3), std::abs(eV1), std::abs(eV1), std::abs(eV2)};
    double la[4];
    for (int k = 0; k < 4; k++)
        la[k] = lam[k] * alpha[k];
    for (int r = 0; r < 4; r++) {
        double incF = 0;
        for (int c = 0; c < 4; c++)
            incF += R[r][c] * la[c];
        F[r] = (FL[r] + FR[r]) * 0.5 - 0.5 * incF;
    }
    return ERR_OK;
}

// generateBoundaryValue, DNDS_FV_EulerEvaluator.hpp:958-1275 (dim = 2)
int boundary_value(const oracle_setting &cfg, const double UL[4], const double n[2], const double p[2], double t,
                   int32_t bc, double UR[4]) {
    if (!(UL[0] > 0))
        return ERR_NONPOSITIVE;
    const double gamma = cfg.gamma;
    if (bc == BC_FARFIELD) {
        const double *far = cfg.far_field;
        double un = (UL[1] * n[0] + UL[2] * n[1]) / UL[0];
        double vsqr = sqr(UL[1] / UL[0]) + sqr(UL[2] / UL[0]);
        double asqr, H, pp;
        ideal_gas_thermal(UL[3], UL[0], vsqr, gamma, pp, asqr, H);
        if (!(asqr >= 0))
            return ERR_NONPOSITIVE;
        double a = std::sqrt(asqr);
        if (un - a > 0) {
            for (int k = 0; k < 4; k++)
                UR[k] = UL[k];
        } else if (un > 0) {
            double farPrim[4], ULPrim[4];
            cons2prim(far, farPrim, gamma);
            cons2prim(UL, ULPrim, gamma);
            ULPrim[3] = farPrim[3];
            prim2cons(ULPrim, UR, gamma);
        } else if (un + a > 0) {
            double farPrim[4], ULPrim[4];
            cons2prim(far, farPrim, gamma);
            cons2prim(UL, ULPrim, gamma);
            farPrim[3] = ULPrim[3];
            prim2cons(farPrim, UR, gamma);
        } else {
            for (int k = 0; k < 4; k++)
                UR[k] = far[k];
        }
        return ERR_OK;
    }
    if (bc == BC_DMRFAR) {
        const double uShock = 10;
        if (((p[0] - uShock / std::sin(kPi / 3) * t - 1. / 6.) - p[1] / std::tan(kPi / 3)) > 0) {
            UR[0] = 1.4, UR[1] = 0, UR[2] = 0, UR[3] = 2.5;
        } else {
            UR[0] = 8, UR[1] = 57.157676649772960, UR[2] = -33, UR[3] = 5.635e2;
        }
        return ERR_OK;
    }
    if (bc == BC_RTFAR) {
        double far[4] = {cfg.far_field[0], cfg.far_field[1], cfg.far_field[2], cfg.far_field[3]};
        double vsqr = sqr(UL[1] / UL[0]) + sqr(UL[2] / UL[0]);
        double asqr, H, pp;
        ideal_gas_thermal(UL[3], UL[0], vsqr, gamma, pp, asqr, H);
        if (!(asqr >= 0))
            return ERR_NONPOSITIVE;
        double a = std::sqrt(asqr);
        double v = -0.025 * a * std::cos(p[0] * 8 * kPi);
        double rho = p[1] < 0.5 ? 2 : 1;
        double pr = p[1] < 0.5 ? 1 : 2.5;
        far[0] = rho;
        far[1] = 0;
        far[2] = rho * v;
        far[3] = 0.5 * rho * sqr(v) + pr / (gamma - 1);
        for (int k = 0; k < 4; k++) // the reference overrides every branch with `far`
            UR[k] = far[k];
        return ERR_OK;
    }
    if (bc == BC_IVFAR) {
        double chi = 5;
        double xc = 5 + t;
        double yc = 5 + t;
        double r = std::sqrt(sqr(p[0] - xc) + sqr(p[1] - yc));
        double dT = -(gamma - 1) / (8 * gamma * sqr(kPi)) * sqr(chi) * std::exp(1 - sqr(r));
        double dux = chi / 2 / kPi * std::exp((1 - sqr(r)) / 2) * -(p[1] - xc);
        double duy = chi / 2 / kPi * std::exp((1 - sqr(r)) / 2) * +(p[0] - yc);
        double T = dT + 1;
        double ux = dux + 1;
        double uy = duy + 1;
        double S = 1;
        double rho = std::pow(T / S, 1 / (gamma - 1));
        double pr = T * rho;
        double E = 0.5 * (sqr(ux) + sqr(uy)) * rho + pr / (gamma - 1);
        UR[0] = rho, UR[1] = rho * ux, UR[2] = rho * uy, UR[3] = E;
        return ERR_OK;
    }
    if (bc == BC_2DRIEMANNFAR) {
        const double bdL = 0.0, bdR = 1.0, bdD = 0.0, bdU = 1.0;
        const double phi1 = -0.663324958071080;
        const double phi2 = -0.422115882408869;
        const double location = 0.8;
        double p1 = location + phi1 * t;
        double p2 = location + phi2 * t;
        double rho, u, v, pre;
        double ULPrim[4];
        cons2prim(UL, ULPrim, gamma);
        double rhoL = ULPrim[0], uL = ULPrim[1], vL = ULPrim[2], preL = ULPrim[3];
        double farPrim[4] = {ULPrim[0], ULPrim[1], ULPrim[2], ULPrim[3]};
        const double bTol = 1e-9;
        if (std::abs(p[0] - bdL) < bTol) {
            if (p[1] <= p2)
                rho = 0.137992831541219, u = 1.206045378311055, v = 1.206045378311055, pre = 0.029032258064516;
            else
                rho = 0.532258064516129, u = 1.206045378311055, v = 0.0, pre = 0.3;
        } else if (std::abs(p[0] - bdR) < bTol) {
            rho = rhoL, u = -uL, v = vL, pre = preL;
        } else if (std::abs(p[1] - bdU) < bTol) {
            rho = rhoL, u = uL, v = -vL, pre = preL;
        } else if (std::abs(p[1] - bdD) < bTol) {
            if (p[0] <= p2)
                rho = 0.137992831541219, u = 1.206045378311055, v = 1.206045378311055, pre = 0.029032258064516;
            else
                rho = 0.532258064516129, u = 0.0, v = 1.206045378311055, pre = 0.3;
        } else {
            return ERR_INVALID; // not a valid boundary position
        }
        (void)p1;
        farPrim[0] = rho;
        farPrim[1] = u, farPrim[2] = v;
        farPrim[3] = pre;
        prim2cons(farPrim, UR, gamma);
        return ERR_OK;
    }
    if (bc == BC_WALL_EULER) {
        for (int k = 0; k < 4; k++)
            UR[k] = UL[k];
        double d = UR[1] * n[0] + UR[2] * n[1];
        UR[1] -= d * n[0];
        UR[2] -= d * n[1];
        return ERR_OK;
    }
    if (bc == BC_WALL_NOSLIP) {
        for (int k = 0; k < 4; k++)
            UR[k] = UL[k];
        UR[1] *= -1;
        UR[2] *= -1;
        return ERR_OK;
    }
    return ERR_UNSUPPORTED; // Wall is not a proper flux boundary (DNDS_FV_EulerEvaluator.hpp:1265-1269)
}

// CompressRecPart, DNDS_FV_EulerEvaluator.hpp:1277-1360
inline void compress_rec_part(const double umean[4], const double inc[4], double ret[4]) {
    for (int k = 0; k < 4; k++)
        ret[k] = umean[k] + inc[k];
    if (ret[0] < 0)
        for (int k = 0; k < 4; k++)
            ret[k] = umean[k];
    double eK = (ret[1] * ret[1] + ret[2] * ret[2]) * 0.5 / (kVerySmallReal + ret[0]);
    double e = ret[3] - eK;
    if (e < 0)
        for (int k = 0; k < 4; k++)
            ret[k] = umean[k];
}

// Eigen's stableNormalized of a 2-vector
inline void stable_normalized(double x, double y, double &ox, double &oy) {
    double w = std::max(std::fabs(x), std::fabs(y));
    double zx = x / w, zy = y / w;
    double z = zx * zx + zy * zy;
    if (z > 0) {
        double s = std::sqrt(z);
        ox = (x / s) / w, oy = (y / s) / w;
    } else
        ox = x, oy = y;
}

// Eigen's normalized(): v / sqrt(squaredNorm)
inline void normalized(double x, double y, double &ox, double &oy) {
    double z = x * x + y * y;
    if (z > 0) {
        double s = std::sqrt(z);
        ox = x / s, oy = y / s;
    } else
        ox = x, oy = y;
}

// EulerEvaluator::fluxFace, DNDS_FV_EulerEvaluator.hpp:371-591, Euler model
// (the viscous flux vanishes for muGas = 0).  Returns -F and the
// flux-balance fluxes FLfix / FRfix of the cell means.
int flux_face(const oracle_setting &cfg, const double ULxy[4], const double URxy[4], const double ULMeanXy[4],
              const double URMeanXy[4], const double n[2], double dLambda, double out[4], double FLfix[4],
              double FRfix[4]) {
    const double gamma = cfg.gamma;
    // normBase = [n, (-n1, n0)] (NormBuildLocalBaseV, DNDS_Elements.hpp:1466-1488)
    auto rotT = [&](const double in[4], double o[4]) {
        o[0] = in[0];
        o[1] = n[0] * in[1] + n[1] * in[2];
        o[2] = (-n[1]) * in[1] + n[0] * in[2];
        o[3] = in[3];
    };
    auto rot = [&](double v[4]) {
        double a = n[0] * v[1] + (-n[1]) * v[2];
        double b = n[1] * v[1] + n[0] * v[2];
        v[1] = a, v[2] = b;
    };
    double UL[4], UR[4];
    rotT(ULxy, UL);
    rotT(URxy, UR);
    {
        double ULMean[4], URMean[4], wL[4], wR[4];
        rotT(ULMeanXy, ULMean);
        rotT(URMeanXy, URMean);
        cons2prim(ULMean, wL, gamma);
        cons2prim(URMean, wR, gamma);
        if (!(ULMean[0] > 0) || !(URMean[0] > 0))
            return ERR_NONPOSITIVE;
        inviscid_flux(ULMean, wL[1], wL[3], FLfix);
        inviscid_flux(URMean, wR[1], wR[3], FRfix);
        rot(FLfix);
        rot(FRfix);
    }
    double finc[4];
    int rc = cfg.riemann_solver == 2   ? hllc_flux(UL, UR, gamma, dLambda, finc)
             : cfg.riemann_solver == 1 ? roe_flux(UL, UR, gamma, dLambda, finc, 0)
             : (cfg.riemann_solver >= 11 && cfg.riemann_solver <= 15)
                 ? roe_flux(UL, UR, gamma, dLambda, finc, cfg.riemann_solver - 10)
             : cfg.riemann_solver == 3 ? hllep_flux(UL, UR, gamma, dLambda, finc)
                                       : ERR_UNSUPPORTED;
    if (rc)
        return rc;
    rot(finc);
    for (int k = 0; k < 4; k++) {
        if (!std::isfinite(finc[k]))
            return ERR_NAN;
        out[k] = -finc[k];
    }
    return ERR_OK;
}

} // namespace

extern "C" int oracle_roe_flux(const double *UL, const double *UR, double gamma, double dlambda, double *F) {
    return roe_flux(UL, UR, gamma, dlambda, F, 0);
}
extern "C" int oracle_roe_m_flux(const double *UL, const double *UR, double gamma, double dlambda, int scheme,
                                 double *F) {
    return roe_flux(UL, UR, gamma, dlambda, F, scheme);
}
extern "C" int oracle_hllc_flux(const double *UL, const double *UR, double gamma, double dlambda, double *F) {
    return hllc_flux(UL, UR, gamma, dlambda, F);
}
extern "C" int oracle_hllep_flux(const double *UL, const double *UR, double gamma, double dlambda, double *F) {
    return hllep_flux(UL, UR, gamma, dlambda, F);
}
extern "C" int oracle_eigenvectors(const double *U, double gamma, double *L, double *R) {
    if (!(U[0] > 0))
        return ERR_NONPOSITIVE;
    double velo[2] = {U[1] / U[0], U[2] / U[0]};
    double vsqr = velo[0] * velo[0] + velo[1] * velo[1];
    double asqr, p, H;
    ideal_gas_thermal(U[3], U[0], vsqr, gamma, p, asqr, H);
    if (!(asqr >= 0))
        return ERR_NONPOSITIVE;
    double a = std::sqrt(asqr);
    right_eigen(velo, vsqr, H, a, reinterpret_cast<double(*)[4]>(R));
    left_eigen(velo, vsqr, H, a, gamma, reinterpret_cast<double(*)[4]>(L));
    return ERR_OK;
}
extern "C" int oracle_boundary_value(const oracle_setting *cfg, double *UL, const double *unit_norm, const double *p,
                                     double t, int32_t bc, int32_t fix_ul, double *UR) {
    (void)fix_ul;
    return boundary_value(*cfg, UL, unit_norm, p, t, bc, UR);
}

// ReconstructionJacobiStep, DNDS_FV_VR.hpp:1474-1713, SOR_Instead = false
extern "C" int oracle_jacobi_step(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                  const double *u, double *urec, double *urec_new, double t) {
    (void)t;
    const double relax = cfg->jacobi_relax;
    const int rc0 = parallel_items(m->n_cell, [&](int64_t i) -> int {
        double *nw = urec_new + 36 * i;
        const double *xi = urec + 36 * i;
        for (int k = 0; k < 36; k++)
            nw[k] = (1 - relax) * xi[k];
        const int64_t b0 = m->cell_ptr[i];
        const int nf = int(m->cell_ptr[i + 1] - b0);
        for (int lf = 0; lf < nf; lf++) {
            const int64_t f = m->cell2face[b0 + lf];
            const int64_t j = m->face2cell[2 * f] == i ? m->face2cell[2 * f + 1] : m->face2cell[2 * f];
            const double *M = s->ainv_b + 81 * (b0 + lf);
            const double *bv = s->ainv_bvec + 9 * (b0 + lf);
            const int32_t bc = m->face_bc[f];
            if (j != kNoCell) {
                const double *xj = urec + 36 * j;
                double du[4];
                for (int v = 0; v < 4; v++)
                    du[v] = u[4 * j + v] - u[4 * i + v];
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++) {
                        double tt = 0;
                        for (int c = 0; c < 9; c++)
                            tt += M[9 * r + c] * xj[c + 9 * v];
                        nw[r + 9 * v] += relax * (tt + du[v] * bv[r]);
                    }
            } else if (bc == BC_WALL) {
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++)
                        nw[r + 9 * v] += relax * ((0.0 - u[4 * i + v]) * bv[r]);
            } else if (is_far(bc)) {
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++) {
                        double tt = 0;
                        for (int c = 0; c < 9; c++)
                            tt += M[9 * r + c] * xi[c + 9 * v];
                        nw[r + 9 * v] += relax * tt;
                    }
            } else if (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) {
                // face-Gauss boundary correction, DNDS_FV_VR.hpp:1577-1657
                double BC[36] = {0};
                const double w0 = s->face_weight[10 * f];
                for (int g = 0; g < 3; g++) {
                    const double *d = s->face_d0 + 54 * f + 18 * g; // left side (the cell itself)
                    const double nx = s->face_norm[6 * f + 2 * g], ny = s->face_norm[6 * f + 2 * g + 1];
                    double uBL[4];
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += d[c] * xi[c + 9 * v];
                        uBL[v] = acc;
                    }
                    for (int v = 0; v < 4; v++)
                        uBL[v] += u[4 * i + v];
                    double uBV[4];
                    double n0, n1;
                    stable_normalized(nx, ny, n0, n1);
                    if (bc == BC_WALL_EULER) {
                        for (int v = 0; v < 4; v++)
                            uBV[v] = uBL[v];
                        uBV[0] = uBL[0];
                        double dd = n0 * uBV[1] + n1 * uBV[2];
                        uBV[1] = uBL[1] - n0 * dd;
                        uBV[2] = uBL[2] - n1 * dd;
                        uBV[3] = uBL[3];
                    } else {
                        double nrm[2] = {n0, n1}, pp[2] = {0, 0};
                        int rc = boundary_value(*cfg, uBL, nrm, pp, t, bc, uBV);
                        if (rc)
                            return rc;
                    }
                    const double nl = std::sqrt(nx * nx + ny * ny);
                    for (int r = 0; r < 9; r++)
                        for (int v = 0; v < 4; v++) {
                            double inc = d[r] * w0 * w0 * (uBV[v] - u[4 * i + v]);
                            inc *= nl;
                            BC[r + 9 * v] += inc * kLine3W[g];
                        }
                }
                const double *Ai = s->ainv + 81 * i;
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += Ai[9 * r + c] * BC[c + 9 * v];
                        nw[r + 9 * v] += relax * acc;
                    }
            } else {
                return ERR_UNSUPPORTED;
            }
        }
        return ERR_OK;
    });
    if (rc0)
        return rc0;
    std::memcpy(urec, urec_new, sizeof(double) * 36 * m->n_cell);
    return ERR_OK;
}

// ReconstructionJacobiStep with SOR_Instead = true, DNDS_FV_VR.hpp:1495-1662:
// one in-place Gauss-Seidel sweep over the cells in the order `scan`
// (SOR_iScan2iCell; the identity, or a colour-major order).  Every term is
// added to uRec[iCell] itself, so the far-field term M uRec[iCell] and the
// Wall_Euler / Wall_NoSlip corrections see the cell's partial sum after
// (1 - relax) scaling and the faces before them in cell2face order; the
// product M uRec[iCell] is formed before it is added (Eigen evaluates a
// matrix product into a temporary).  Neighbours are read in place: those
// earlier in the scan carry this sweep's values.
extern "C" int oracle_sor_step(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                               const double *u, double *urec, const int64_t *scan, double t) {
    const double relax = cfg->jacobi_relax;
    for (int64_t is = 0; is < m->n_cell; is++) {
        const int64_t i = scan[is];
        double *xi = urec + 36 * i;
        for (int k = 0; k < 36; k++)
            xi[k] = (1 - relax) * xi[k];
        const int64_t b0 = m->cell_ptr[i];
        const int nf = int(m->cell_ptr[i + 1] - b0);
        for (int lf = 0; lf < nf; lf++) {
            const int64_t f = m->cell2face[b0 + lf];
            const int64_t j = m->face2cell[2 * f] == i ? m->face2cell[2 * f + 1] : m->face2cell[2 * f];
            const double *M = s->ainv_b + 81 * (b0 + lf);
            const double *bv = s->ainv_bvec + 9 * (b0 + lf);
            const int32_t bc = m->face_bc[f];
            if (j != kNoCell) {
                const double *xj = urec + 36 * j;
                double du[4];
                for (int v = 0; v < 4; v++)
                    du[v] = u[4 * j + v] - u[4 * i + v];
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++) {
                        double tt = 0;
                        for (int c = 0; c < 9; c++)
                            tt += M[9 * r + c] * xj[c + 9 * v];
                        xi[r + 9 * v] += relax * (tt + du[v] * bv[r]);
                    }
            } else if (bc == BC_WALL) {
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++)
                        xi[r + 9 * v] += relax * ((0.0 - u[4 * i + v]) * bv[r]);
            } else if (is_far(bc)) {
                double tt[36];
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += M[9 * r + c] * xi[c + 9 * v];
                        tt[r + 9 * v] = acc;
                    }
                for (int k = 0; k < 36; k++)
                    xi[k] += relax * tt[k];
            } else if (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) {
                double BC[36] = {0};
                const double w0 = s->face_weight[10 * f];
                for (int g = 0; g < 3; g++) {
                    const double *d = s->face_d0 + 54 * f + 18 * g;
                    const double nx = s->face_norm[6 * f + 2 * g], ny = s->face_norm[6 * f + 2 * g + 1];
                    double uBL[4];
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += d[c] * xi[c + 9 * v];
                        uBL[v] = acc;
                    }
                    for (int v = 0; v < 4; v++)
                        uBL[v] += u[4 * i + v];
                    double uBV[4];
                    double n0, n1;
                    stable_normalized(nx, ny, n0, n1);
                    if (bc == BC_WALL_EULER) {
                        for (int v = 0; v < 4; v++)
                            uBV[v] = uBL[v];
                        double dd = n0 * uBV[1] + n1 * uBV[2];
                        uBV[1] = uBL[1] - n0 * dd;
                        uBV[2] = uBL[2] - n1 * dd;
                    } else {
                        double nrm[2] = {n0, n1}, pp[2] = {0, 0};
                        int rc = boundary_value(*cfg, uBL, nrm, pp, t, bc, uBV);
                        if (rc)
                            return rc;
                    }
                    const double nl = std::sqrt(nx * nx + ny * ny);
                    for (int r = 0; r < 9; r++)
                        for (int v = 0; v < 4; v++) {
                            double inc = d[r] * w0 * w0 * (uBV[v] - u[4 * i + v]);
                            inc *= nl;
                            BC[r + 9 * v] += inc * kLine3W[g];
                        }
                }
                const double *Ai = s->ainv + 81 * i;
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += Ai[9 * r + c] * BC[c + 9 * v];
                        xi[r + 9 * v] += relax * acc;
                    }
            } else {
                return ERR_UNSUPPORTED;
            }
        }
    }
    return ERR_OK;
}

// One reconstruction sweep of the steps: Jacobi, or with a scan order the
// SOR_Instead sweep (forward, then backward with SOR_InverseScanning,
// DNDS_FV_VR.hpp:1493-1502)
static int rec_sweep(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                     double *urec, double *urec_new, double t) {
    if (!cfg->sor_scan)
        return oracle_jacobi_step(m, s, cfg, u, urec, urec_new, t);
    int rc = oracle_sor_step(m, s, cfg, u, urec, cfg->sor_scan, t);
    if (rc || !cfg->sor_inverse)
        return rc;
    std::vector<int64_t> rev(cfg->sor_scan, cfg->sor_scan + m->n_cell);
    std::reverse(rev.begin(), rev.end());
    return oracle_sor_step(m, s, cfg, u, urec, rev.data(), t);
}

// EvaluateDt, DNDS_FV_EulerEvaluator.cpp:14-183 (nTimeFilterPass = 0, muGas = 0)
extern "C" int oracle_evaluate_dt(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                  const double *u, double cfl, double max_dt, int32_t use_local_dt, double *dt,
                                  double *dt_min, double *lambda_face, double *delta_lambda_face) {
    const double gamma = cfg->gamma;
    std::vector<double> lambdaCell(m->n_cell, 0.0);
    for (int64_t f = 0; f < m->n_face; f++) {
        const int64_t L = m->face2cell[2 * f], R = m->face2cell[2 * f + 1];
        double n0, n1;
        normalized(s->face_norm_center[2 * f], s->face_norm_center[2 * f + 1], n0, n1);
        const double *UL = u + 4 * L;
        double uMean[4] = {UL[0], UL[1], UL[2], UL[3]};
        double pL, asqrL, HL, pR, asqrR, HR;
        double vL[2] = {UL[1] / UL[0], UL[2] / UL[0]};
        double vR[2] = {vL[0], vL[1]};
        ideal_gas_thermal(UL[3], UL[0], vL[0] * vL[0] + vL[1] * vL[1], gamma, pL, asqrL, HL);
        pR = pL, HR = HL, asqrR = asqrL;
        if (R != kNoCell) {
            const double *UR = u + 4 * R;
            for (int k = 0; k < 4; k++)
                uMean[k] = (uMean[k] + UR[k]) * 0.5;
            vR[0] = UR[1] / UR[0], vR[1] = UR[2] / UR[0];
            ideal_gas_thermal(UR[3], UR[0], vR[0] * vR[0] + vR[1] * vR[1], gamma, pR, asqrR, HR);
        }
        if (!(uMean[0] > 0))
            return ERR_NONPOSITIVE;
        double veloNMean = 0.5 * ((vL[0] + vR[0]) * n0 + (vL[1] + vR[1]) * n1);
        double pMean = (pL + pR) * 0.5;
        double aMean = std::sqrt(gamma * pMean / uMean[0]);
        double lambdaConvection = std::abs(veloNMean) + aMean;
        const double area = s->face_area[f];
        double lf = lambdaConvection;
        if (cfg->mu_gas > 0) { // DNDS_FV_EulerEvaluator.cpp:79-119 (NS_2D: no SA factor)
            const double cp = cfg->r_gas * gamma / (gamma - 1);
            const double T = pMean / ((gamma - 1) / gamma * cp * uMean[0]);
            const double muf = cfg->mu_gas * std::pow(T / cfg->t_ref, 1.5) * (cfg->t_ref + cfg->c_sutherland) /
                               (T + cfg->c_sutherland);
            const double lamVis = muf / uMean[0] * std::max(4. / 3., gamma / cfg->pr_gas);
            const double volR = R != kNoCell ? s->volume[R] : s->volume[L];
            lf = lambdaConvection + lamVis * area * (1. / s->volume[L] + 1. / volR);
        }
        if (lambda_face)
            lambda_face[f] = lf;
        lambdaCell[L] += lf * area;
        if (R != kNoCell)
            lambdaCell[R] += lf * area;
        double dlf = std::abs((vR[0] - vL[0]) * n0 + (vR[1] - vL[1]) * n1) + std::sqrt(std::abs(asqrR - asqrL)) * 0.7071;
        if (delta_lambda_face)
            delta_lambda_face[f] = dlf;
    }
    double dtMin = kVeryLargeReal;
    for (int64_t i = 0; i < m->n_cell; i++) {
        dt[i] = std::min(cfl * s->volume[i] / (lambdaCell[i] + 1e-100), max_dt);
        dtMin = std::min(dtMin, dt[i]);
    }
    if (!std::isfinite(dtMin))
        return ERR_NAN;
    if (dt_min)
        *dt_min = dtMin;
    if (!use_local_dt)
        for (int64_t i = 0; i < m->n_cell; i++)
            dt[i] = dtMin;
    return ERR_OK;
}

namespace {

// Rows 1 and 2 (d/dx, d/dy) of the DiBj block of `cell` at a face point from
// the point's scaled coordinates (X, Y) and the cell's scales (sx, sy)
// (FDiffBaseValue, global coordinates, DNDS_FV_VR.hpp:328-356): for base
// ib = c + 1 with exponents (bx, by), d/dx = bx X^(bx-1) Y^by / sx and
// d/dy = by X^bx Y^(by-1) / sy.
void face_grad_rows(const double XY[2], const double sc[2], double D[2][9]) {
    static const int BX[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0}, BY[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
    double pX[4], pY[4];
    pX[0] = 1.0, pX[1] = XY[0], pX[2] = XY[0] * XY[0], pX[3] = pX[2] * XY[0];
    pY[0] = 1.0, pY[1] = XY[1], pY[2] = XY[1] * XY[1], pY[3] = pY[2] * XY[1];
    for (int c = 0; c < 9; c++) {
        const int bx = BX[c + 1], by = BY[c + 1];
        D[0][c] = bx > 0 ? double(bx) * pX[bx - 1] * pY[by] * 1.0 / (sc[0] * 1.0 * 1.0) : 0.0;
        D[1][c] = by > 0 ? double(by) * pX[bx] * pY[by - 1] * 1.0 / (1.0 * sc[1] * 1.0) : 0.0;
    }
}

// Gas::ViscousFlux_IdealGas<2> (DNDS_Gas.hpp:795-826): U the mean state,
// G[d][v] the gradient of the conservative variables, Flux[d][v].
void viscous_flux(const double U[4], const double G[2][4], const double n[2], bool adiabatic, double gamma,
                  double mu, double k, double Cp, double Flux[2][4]) {
    const double velo[2] = {U[1] / U[0], U[2] / U[0]};
    const double lambda = -2. / 3.;
    double S[2][2]; // S[i][j] = d u_j / d x_i
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            S[i][j] = (1.0 / (U[0] * U[0])) * (U[0] * G[i][1 + j] - G[i][0] * U[1 + j]);
    double GradP[2];
    for (int i = 0; i < 2; i++)
        GradP[i] = (gamma - 1) * (G[i][3] - 0.5 * ((G[i][1] * velo[0] + G[i][2] * velo[1]) +
                                                   (S[i][0] * U[1] + S[i][1] * U[2])));
    const double vSqr = velo[0] * velo[0] + velo[1] * velo[1];
    const double p = (gamma - 1) * (U[3] - U[0] * 0.5 * vSqr);
    double GradT[2];
    for (int i = 0; i < 2; i++)
        GradT[i] = (gamma / ((gamma - 1) * Cp * U[0] * U[0])) * (U[0] * GradP[i] - p * G[i][0]);
    if (adiabatic) {
        const double d = GradT[0] * n[0] + GradT[1] * n[1];
        GradT[0] -= d * n[0], GradT[1] -= d * n[1];
    }
    const double tr = S[0][0] + S[1][1];
    for (int i = 0; i < 2; i++) {
        Flux[i][0] = 0.0;
        for (int j = 0; j < 2; j++)
            Flux[i][1 + j] = (S[i][j] + S[j][i]) * mu + (i == j ? 1.0 : 0.0) * (lambda * mu * tr);
        Flux[i][3] = (Flux[i][1] * velo[0] + Flux[i][2] * velo[1]) + k * GradT[i];
    }
}

} // namespace

// EvaluateRHS, DNDS_FV_EulerEvaluator.cpp:197-640 (source term zero for NS_2D
// with constMassForce = 0)
extern "C" int oracle_evaluate_rhs(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                   const double *delta_lambda_face, const double *u, const double *urec, double t,
                                   double *rhs) {
    const int64_t nc = m->n_cell;
    for (int64_t i = 0; i < nc; i++)
        for (int v = 0; v < 4; v++)
            rhs[4 * i + v] = 0.0;
    std::vector<double> fbuf(12 * size_t(m->n_face));
    const int rc0 = parallel_items(m->n_face, [&](int64_t f) -> int {
        const int64_t L = m->face2cell[2 * f], R = m->face2cell[2 * f + 1];
        double fluxEs[3][4] = {{0}};
        for (int g = 0; g < 3; g++) {
            const double nx = s->face_norm[6 * f + 2 * g], ny = s->face_norm[6 * f + 2 * g + 1];
            double un[2];
            normalized(nx, ny, un[0], un[1]);
            const double *dL = s->face_d0 + 54 * f + 18 * g;
            double inc[4];
            for (int v = 0; v < 4; v++) {
                double acc = 0;
                for (int c = 0; c < 9; c++)
                    acc += dL[c] * urec[36 * L + c + 9 * v];
                inc[v] = acc;
            }
            double ULxy[4], URxy[4], URMeanXy[4];
            compress_rec_part(u + 4 * L, inc, ULxy);
            const double *ULMeanXy = u + 4 * L;
            if (R != kNoCell) {
                const double *dR = s->face_d0 + 54 * f + 18 * g + 9;
                for (int v = 0; v < 4; v++) {
                    double acc = 0;
                    for (int c = 0; c < 9; c++)
                        acc += dR[c] * urec[36 * R + c + 9 * v];
                    inc[v] = acc;
                }
                compress_rec_part(u + 4 * R, inc, URxy);
                for (int v = 0; v < 4; v++)
                    URMeanXy[v] = u[4 * R + v];
            } else {
                const double *pc = s->face_center + 2 * f;
                int rc = boundary_value(*cfg, ULxy, un, pc, t, m->face_bc[f], URxy);
                if (rc)
                    return rc;
                rc = boundary_value(*cfg, ULMeanXy, un, pc, t, m->face_bc[f], URMeanXy);
                if (rc)
                    return rc;
            }
            double F[4], FL[4], FR[4];
            int rc = flux_face(*cfg, ULxy, URxy, ULMeanXy, URMeanXy, un, delta_lambda_face[f], F, FL, FR);
            if (rc)
                return rc;
            if (cfg->mu_gas > 0) {
                // gradients at the point (DNDS_FV_EulerEvaluator.cpp:297-346), the
                // face gradient with the jump term (:381-395), fluxFace's viscous
                // part (DNDS_FV_EulerEvaluator.hpp:397-430, 574-577)
                double GL[2][4], GR[2][4], D[2][9];
                face_grad_rows(s->face_gauss_xy + 12 * f + 4 * g, s->cell_scale + 2 * L, D);
                for (int d = 0; d < 2; d++)
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += D[d][c] * urec[36 * L + c + 9 * v];
                        GL[d][v] = acc;
                    }
                double minVol = s->volume[L];
                if (R != kNoCell) {
                    face_grad_rows(s->face_gauss_xy + 12 * f + 4 * g + 2, s->cell_scale + 2 * R, D);
                    for (int d = 0; d < 2; d++)
                        for (int v = 0; v < 4; v++) {
                            double acc = 0;
                            for (int c = 0; c < 9; c++)
                                acc += D[d][c] * urec[36 * R + c + 9 * v];
                            GR[d][v] = acc;
                        }
                    minVol = std::min(minVol, s->volume[R]);
                } else {
                    std::memcpy(GR, GL, sizeof(GL));
                }
                const int32_t bc = m->face_bc[f];
                double distGRP = minVol / s->face_area[f] * 2;
                distGRP += (bc == BC_WALL_EULER || bc == BC_WALL_NOSLIP) ? kVeryLargeReal : 0.0;
                double GM[2][4], UM[4];
                for (int d = 0; d < 2; d++)
                    for (int v = 0; v < 4; v++)
                        GM[d][v] = (GR[d][v] + GL[d][v]) * 0.5 + (1.0 / distGRP) * (un[d] * (URxy[v] - ULxy[v]));
                for (int v = 0; v < 4; v++)
                    UM[v] = 0.5 * (ULxy[v] + URxy[v]);
                const double gamma = cfg->gamma, cp = cfg->r_gas * gamma / (gamma - 1);
                const double vm0 = UM[1] / UM[0], vm1 = UM[2] / UM[0];
                double pMean, asqrMean, Hmean;
                ideal_gas_thermal(UM[3], UM[0], vm0 * vm0 + vm1 * vm1, gamma, pMean, asqrMean, Hmean);
                const double T = pMean / ((gamma - 1) / gamma * cp * UM[0]);
                const double muf = cfg->mu_gas * std::pow(T / cfg->t_ref, 1.5) * (cfg->t_ref + cfg->c_sutherland) /
                                   (T + cfg->c_sutherland);
                const double k = cp * (muf - muf) / 0.9 + cp * muf / cfg->pr_gas;
                double VF[2][4];
                viscous_flux(UM, GM, un, bc == BC_WALL_NOSLIP, gamma, muf, k, cp, VF);
                for (int v = 0; v < 4; v++)
                    F[v] += (VF[0][v] * un[0] + VF[1][v] * un[1]) * 1;
            }
            const double nl = std::sqrt(nx * nx + ny * ny);
            for (int v = 0; v < 4; v++) {
                double a = F[v] * nl, b = FL[v] * nl, c = FR[v] * nl;
                fluxEs[0][v] += a * kLine3W[g];
                fluxEs[1][v] += b * kLine3W[g];
                fluxEs[2][v] += c * kLine3W[g];
            }
        }
        std::memcpy(&fbuf[12 * f], fluxEs, sizeof(fluxEs));
        return ERR_OK;
    });
    if (rc0)
        return rc0;
    for (int64_t f = 0; f < m->n_face; f++) { // scatter in face order
        const int64_t L = m->face2cell[2 * f], R = m->face2cell[2 * f + 1];
        double fluxEs[3][4];
        std::memcpy(fluxEs, &fbuf[12 * f], sizeof(fluxEs));
        const double VL = s->volume[L];
        for (int v = 0; v < 4; v++)
            rhs[4 * L + v] += fluxEs[0][v] / VL;
        if (R != kNoCell) {
            const double VR = s->volume[R];
            for (int v = 0; v < 4; v++)
                rhs[4 * R + v] -= fluxEs[0][v] / VR;
        }
        for (int v = 0; v < 4; v++)
            rhs[4 * L + v] -= fluxEs[1][v] / VL;
        if (R != kNoCell) {
            const double VR = s->volume[R];
            for (int v = 0; v < 4; v++)
                rhs[4 * R + v] += fluxEs[2][v] / VR;
        }
    }
    for (int64_t i = 0; i < 4 * nc; i++)
        if (!std::isfinite(rhs[i]))
            return ERR_NAN;
    return ERR_OK;
}

namespace {

// ---------------------------------------------------------------- WBAP ----

// FFaceFunctional with USE_NORM_FUNCTIONAL for 10-row blocks
// (DNDS_FV_VR.hpp:1071-1210; tangWeightModMin = 1 so every tangModifier
// factor is 1): returns Conj(i, i) for the two columns of D (10 x 2).
void face_functional_diag(const oracle_setting &cfg, const double W[10], const double ng[2], const double D[10][2],
                          double out[2]) {
    const double epsR = 1e-50;
    double w2r, w3r, length;
    if (std::fabs(W[1]) >= epsR || std::fabs(W[2]) >= epsR) {
        w2r = std::fabs(W[1]) >= std::fabs(W[2]) ? W[3] / (W[1] * W[1]) : W[5] / (W[2] * W[2]);
        w3r = std::fabs(W[1]) >= std::fabs(W[2]) ? W[6] / (W[1] * W[1] * W[1]) : W[9] / (W[2] * W[2] * W[2]);
        length = std::sqrt(W[1] * W[1] + W[2] * W[2]);
    } else {
        w2r = w3r = 0.0;
        length = 0;
    }
    double n0, n1;
    if (cfg.tang_weight_direction == 0) {
        double ux, uy;
        stable_normalized(ng[0], ng[1], ux, uy);
        n0 = ux * length, n1 = uy * length;
    } else {
        n0 = W[1], n1 = W[2];
    }
    const double t0 = -n1 * (cfg.tang_weight * 1.0), t1 = n0 * (cfg.tang_weight * 1.0);
    for (int col = 0; col < 2; col++) {
        const double d0 = D[0][col], d1 = D[1][col], d2 = D[2][col], d3 = D[3][col], d4 = D[4][col];
        const double d5 = D[5][col], d6 = D[6][col], d7 = D[7][col], d8 = D[8][col], d9 = D[9][col];
        double a[10];
        a[0] = d0;
        a[1] = d1 * n0 + d2 * n1;
        a[2] = d1 * t0 + d2 * t1;
        a[3] = (d3 * n0 * n0 + d4 * n0 * n1 * 2 + d5 * n1 * n1) * w2r;
        a[4] = (d3 * n0 * t0 + d4 * (n0 * t1 + n1 * t0) + d5 * n1 * t1) * w2r;
        a[5] = (d3 * t0 * t0 + d4 * t0 * t1 * 2 + d5 * t1 * t1) * w2r;
        a[6] = (d6 * n0 * n0 * n0 + d7 * n0 * n0 * n1 * 3 + d8 * n0 * n1 * n1 * 3 + d9 * n1 * n1 * n1) * w3r;
        a[7] = (d6 * n0 * n0 * t0 + d7 * (n0 * n0 * t1 + n0 * t0 * n1 + t0 * n0 * n1) +
                d8 * (n0 * n1 * t1 + n0 * t1 * n1 + t0 * n1 * n1) + d9 * n1 * n1 * t1) *
               w3r;
        a[8] = (d6 * n0 * t0 * t0 + d7 * (t0 * n0 * t1 + n0 * t0 * t1 + t0 * t0 * n1) +
                d8 * (t0 * n1 * t1 + n0 * t1 * t1 + t0 * t1 * n1) + d9 * n1 * t1 * t1) *
               w3r;
        a[9] = (d6 * t0 * t0 * t0 + d7 * t0 * t0 * t1 * 3 + d8 * t0 * t1 * t1 * 3 + d9 * t1 * t1 * t1) * w3r;
        double c = 0;
        c += a[0] * a[0] * W[0] * W[0];
        c += a[1] * a[1];
        c += a[2] * a[2] * 1.0;
        c += a[3] * a[3];
        c += a[4] * a[4] * 2 * 1.0;
        c += a[5] * a[5] * 1.0;
        c += a[6] * a[6];
        c += a[7] * a[7] * 3 * 1.0;
        c += a[8] * a[8] * 3 * 1.0;
        c += a[9] * a[9] * 1.0;
        out[col] = c;
    }
}

// fML / fMR of the solver (DNDS_FV_EulerSolver.hpp:749-795): left and right
// eigenvector matrices at the face mean state, in the global frame.
int char_matrices(const oracle_setting &cfg, const double UL[4], const double UR[4], const double n[2], double ML[4][4],
                  double MR[4][4]) {
    double UC[4];
    for (int k = 0; k < 4; k++)
        UC[k] = (UL[k] + UR[k]) * 0.5;
    double r1 = n[0] * UC[1] + n[1] * UC[2];
    double r2 = (-n[1]) * UC[1] + n[0] * UC[2];
    UC[1] = r1, UC[2] = r2;
    if (!(UC[0] > 0))
        return ERR_NONPOSITIVE;
    double velo[2] = {UC[1] / UC[0], UC[2] / UC[0]};
    double vsqr = velo[0] * velo[0] + velo[1] * velo[1];
    double asqr, p, H;
    ideal_gas_thermal(UC[3], UC[0], vsqr, cfg.gamma, p, asqr, H);
    if (!(asqr >= 0))
        return ERR_NONPOSITIVE;
    double a = std::sqrt(asqr);
    double L[4][4], R[4][4];
    left_eigen(velo, vsqr, H, a, cfg.gamma, L);
    right_eigen(velo, vsqr, H, a, R);
    // M(:, 1:2) *= normBase^T ; normBase = [[n0, -n1], [n1, n0]]
    for (int r = 0; r < 4; r++) {
        double c1 = L[r][1] * n[0] + L[r][2] * (-n[1]);
        double c2 = L[r][1] * n[1] + L[r][2] * n[0];
        for (int k = 0; k < 4; k++)
            ML[r][k] = L[r][k];
        ML[r][1] = c1, ML[r][2] = c2;
    }
    // M(1:2, :) = normBase * M(1:2, :)
    for (int c = 0; c < 4; c++) {
        double r1c = n[0] * R[1][c] + (-n[1]) * R[2][c];
        double r2c = n[1] * R[1][c] + n[0] * R[2][c];
        MR[0][c] = R[0][c];
        MR[1][c] = r1c, MR[2][c] = r2c;
        MR[3][c] = R[3][c];
    }
    return ERR_OK;
}

inline double sgn(double x) { return double((x > 0) - (x < 0)); }

} // namespace

namespace {

// Step 0 of both WBAP procedures (V2: DNDS_FV_VR.hpp:1770-1831, V3:
// DNDS_FV_VR.hpp:2075-2138, the same code): the smoothness indicator from the
// (rho, E) jumps at the first LINE_3 point with the LINE_1 weight 2.
void wbap_indicator(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                    const double *urec, double *if_use_limiter) {
    const int64_t nc = m->n_cell;
    const int P = 3;
    parallel_items(nc, [&](int64_t i) -> int {
        const int64_t b0 = m->cell_ptr[i];
        const int nf = int(m->cell_ptr[i + 1] - b0);
        double sum[2][2] = {{0, 0}, {0, 0}};
        for (int lf = 0; lf < nf; lf++) {
            const int64_t f = m->cell2face[b0 + lf];
            const int64_t L = m->face2cell[2 * f], R = m->face2cell[2 * f + 1];
            const int64_t j = L == i ? R : L;
            double val[10][2] = {{0}}, jump[10][2] = {{0}};
            if (j != kNoCell) {
                const double *D0 = s->face_dibj_g0 + 180 * f;      // side 0 basis
                const double *D1 = s->face_dibj_g0 + 180 * f + 90; // side 1 basis
                double vL[10][2], vR[10][2];
                const int vars[2] = {0, 3};
                for (int r = 0; r < 10; r++)
                    for (int q = 0; q < 2; q++) {
                        double aL = 0, aR = 0;
                        for (int c = 0; c < 9; c++) {
                            aL += D0[9 * r + c] * urec[36 * i + c + 9 * vars[q]];
                            aR += D1[9 * r + c] * urec[36 * j + c + 9 * vars[q]];
                        }
                        vL[r][q] = aL, vR[r][q] = aR;
                    }
                for (int q = 0; q < 2; q++) {
                    vL[0][q] += u[4 * i + vars[q]];
                    vR[0][q] += u[4 * j + vars[q]];
                }
                for (int r = 0; r < 10; r++)
                    for (int q = 0; q < 2; q++) {
                        val[r][q] = (vL[r][q] + vR[r][q]) * 0.5;
                        jump[r][q] = (vL[r][q] - vR[r][q]) * 0.5;
                    }
            }
            double ISI[2], IJI[2];
            face_functional_diag(*cfg, s->face_weight + 10 * f, s->face_norm + 6 * f, val, ISI);
            face_functional_diag(*cfg, s->face_weight + 10 * f, s->face_norm + 6 * f, jump, IJI);
            const double nl = std::sqrt(s->face_norm[6 * f] * s->face_norm[6 * f] +
                                        s->face_norm[6 * f + 1] * s->face_norm[6 * f + 1]);
            double inc[2][2] = {{IJI[0], ISI[0]}, {IJI[1], ISI[1]}};
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    inc[a][b] *= nl;
                    sum[a][b] += inc[a][b] * 2.0;
                }
        }
        double si0 = sum[0][0] / (sum[0][1] + kVerySmallReal);
        double si1 = sum[1][0] / (sum[1][1] + kVerySmallReal);
        double sImax = std::max(std::abs(si0), std::abs(si1));
        if_use_limiter[i] = std::sqrt(sImax) * (P * P);
        return ERR_OK;
    });
}

} // namespace

// ReconstructionWBAPLimitFacialV3 (DNDS_FV_VR.hpp:2053-2329) with the
// characteristic matrices fML / fMR and FWBAP_L2_Biway / FWBAP_L2_Multiway
// (DNDS_FV_Limiters.hpp:341-449).  urec is read only; urec_new receives the
// limited reconstruction (urec for cells below the smoothness threshold);
// if_use_limiter receives the indicator.
extern "C" int oracle_wbap_v3(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                              const double *urec, double *urec_new, double *if_use_limiter, int32_t if_all) {
    const int64_t nc = m->n_cell;
    wbap_indicator(m, s, cfg, u, urec, if_use_limiter);
    std::memcpy(urec_new, urec, sizeof(double) * 36 * nc);
    const double eps4 = std::pow(kVerySmallReal, 1.0 / 4); // verySmallReal_pDiP
    const double nStd = cfg->wbap_n_std;
    // Step 1: per-order limiting (DNDS_FV_VR.hpp:2152-2317)
    const int rc1 = parallel_items(nc, [&](int64_t i) -> int {
        if (if_use_limiter[i] < cfg->wbap_smooth_indicator_scale && !if_all)
            return ERR_OK;
        double lim[36];
        const int64_t b0 = m->cell_ptr[i];
        const int nf = int(m->cell_ptr[i + 1] - b0);
        double uFaces[4][36];
        for (int k = 0; k < 4; k++)
            for (int q = 0; q < 36; q++)
                uFaces[k][q] = 0.0;
        for (int order = 3; order >= 1; order--) {
            const int ls = order == 3 ? 5 : (order == 2 ? 2 : 0);
            const int le = order == 3 ? 8 : (order == 2 ? 4 : 1);
            const int nr = le - ls + 1;
            double others[7][4][4]; // [neighbour + 1][row][var]
            int nOther = 0;
            for (int r = 0; r < nr; r++)
                for (int v = 0; v < 4; v++)
                    others[0][r][v] = urec[36 * i + (ls + r) + 9 * v];
            nOther = 1;
            for (int lf = 0; lf < nf; lf++) {
                const int64_t f = m->cell2face[b0 + lf];
                const int64_t L = m->face2cell[2 * f], R = m->face2cell[2 * f + 1];
                const int64_t j = L == i ? R : L;
                if (j == kNoCell)
                    continue;
                const int side = L == i ? 0 : 1;
                double un[2];
                stable_normalized(s->face_norm_center[2 * f], s->face_norm_center[2 * f + 1], un[0], un[1]);
                const double *ms = s->matrix_secondary + 162 * f + 81 * side;            // to this cell's basis
                const double *msO = s->matrix_secondary + 162 * f + 81 * (1 - side);     // to the other's basis
                double uOO[9][4];
                for (int r = 0; r < 9; r++)
                    for (int v = 0; v < 4; v++)
                        uOO[r][v] = urec[36 * j + r + 9 * v];
                if (le < 8)
                    for (int r = le + 1; r < 9; r++)
                        for (int v = 0; v < 4; v++) {
                            double acc = 0;
                            for (int c = le + 1; c < 9; c++)
                                acc += msO[9 * r + c] * uFaces[lf][c + 9 * v];
                            uOO[r][v] = acc;
                        }
                double oIn[4][4], tIn[4][4];
                for (int r = 0; r < nr; r++)
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += ms[9 * (ls + r) + c] * uOO[c][v];
                        oIn[r][v] = acc;
                        tIn[r][v] = others[0][r][v];
                    }
                const double *uR = side ? u + 4 * i : u + 4 * j;
                const double *uL = side ? u + 4 * j : u + 4 * i;
                double ML[4][4], MR[4][4];
                int rc = char_matrices(*cfg, uL, uR, un, ML, MR);
                if (rc)
                    return rc;
                double oC[4][4], tC[4][4], out[4][4];
                for (int r = 0; r < nr; r++)
                    for (int a = 0; a < 4; a++) {
                        double x = 0, y = 0;
                        for (int b = 0; b < 4; b++) {
                            x += ML[a][b] * oIn[r][b];
                            y += ML[a][b] * tIn[r][b];
                        }
                        oC[r][a] = x, tC[r][a] = y;
                    }
                // FWBAP_L2_Biway(uThisIn, uOtherIn, n = 1)
                for (int r = 0; r < nr; r++)
                    for (int a = 0; a < 4; a++) {
                        const double u1 = tC[r][a], u2 = oC[r][a];
                        const double frac = u1 / (std::abs(u2) + 1e-12) * sgn(u2);
                        const double th1 = frac * frac * frac;
                        const double th2 = (frac * frac) * (frac * frac);
                        out[r][a] = u1 * (1.0 + th1) / (1.0 + th2);
                    }
                for (int r = 0; r < nr; r++)
                    for (int a = 0; a < 4; a++) {
                        double x = 0;
                        for (int b = 0; b < 4; b++)
                            x += MR[a][b] * out[r][b];
                        others[nOther][r][a] = x;
                        uFaces[lf][(ls + r) + 9 * a] = x;
                    }
                nOther++;
            }
            // FWBAP_L2_Multiway(uOthers, n = WBAP_nStd)
            for (int r = 0; r < nr; r++)
                for (int v = 0; v < 4; v++) {
                    double up = 0, down = 0;
                    const double u0 = others[0][r][v];
                    for (int k = 1; k < nOther; k++) {
                        const double uo = others[k][r][v];
                        const double ti = u0 / (std::abs(uo) + eps4) * sgn(uo);
                        down += (ti * ti) * (ti * ti);
                        up += ti * ti * ti;
                    }
                    lim[(ls + r) + 9 * v] = u0 * (up + nStd) / (down + nStd);
                }
        }
        for (int k = 0; k < 36; k++)
            urec_new[36 * i + k] = lim[k];
        return ERR_OK;
    });
    if (rc1)
        return rc1;
    for (int64_t k = 0; k < 36 * nc; k++)
        if (!std::isfinite(urec_new[k]))
            return ERR_NAN;
    return ERR_OK;
}

// ReconstructionWBAPLimitFacialV2 (DNDS_FV_VR.hpp:1749-2047, limiterProcedure
// 0) with fML / fMR and FWBAP_L2_Biway / FWBAP_L2_Multiway
// (DNDS_FV_Limiters.hpp:341-449).  Per order 3, 2, 1 every limited cell maps
// each neighbour's current buffer (rows of the higher orders already limited)
// through the full matrixSecondary to its own basis, limits its own rows
// against it in characteristic variables (biway, n = WBAP_nStd), and combines
// the per-neighbour results by the multiway centred on the first neighbour's
// (n1 = 1); the orders' results are copied back between orders.  A cell
// without neighbours keeps its rows (the reference indexes an empty list
// there).  urec is read only; urec_new receives the result.
extern "C" int oracle_wbap_v2(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const double *u,
                              const double *urec, double *urec_new, double *if_use_limiter, int32_t if_all) {
    const int64_t nc = m->n_cell;
    wbap_indicator(m, s, cfg, u, urec, if_use_limiter);
    double *buf = urec_new; // uRecNewBuf
    std::memcpy(buf, urec, sizeof(double) * 36 * nc);
    std::vector<double> buf1(36 * nc, 0.0); // uRecNewBuf1
    const double eps4 = std::pow(kVerySmallReal, 1.0 / 4);
    const double nStd = cfg->wbap_n_std;
    auto limited = [&](int64_t i) { return !(if_use_limiter[i] < cfg->wbap_smooth_indicator_scale && !if_all); };
    for (int order = 3; order >= 1; order--) {
        const int ls = order == 3 ? 5 : (order == 2 ? 2 : 0);
        const int le = order == 3 ? 8 : (order == 2 ? 4 : 1);
        const int nr = le - ls + 1;
        const int rc = parallel_items(nc, [&](int64_t i) -> int {
            if (!limited(i))
                return ERR_OK;
            const int64_t b0 = m->cell_ptr[i];
            const int nf = int(m->cell_ptr[i + 1] - b0);
            double uC[4][4];
            for (int r = 0; r < nr; r++)
                for (int v = 0; v < 4; v++)
                    uC[r][v] = buf[36 * i + (ls + r) + 9 * v];
            double first[4][4], up[4][4], down[4][4];
            for (int r = 0; r < 4; r++)
                for (int v = 0; v < 4; v++)
                    first[r][v] = up[r][v] = down[r][v] = 0.0;
            int nOther = 0;
            for (int lf = 0; lf < nf; lf++) {
                const int64_t f = m->cell2face[b0 + lf];
                const int64_t L = m->face2cell[2 * f], R = m->face2cell[2 * f + 1];
                const int64_t j = L == i ? R : L;
                if (j == kNoCell)
                    continue;
                const int side = L == i ? 0 : 1;
                double un[2];
                stable_normalized(s->face_norm_center[2 * f], s->face_norm_center[2 * f + 1], un[0], un[1]);
                const double *ms = s->matrix_secondary + 162 * f + 81 * side;
                double oIn[4][4];
                for (int r = 0; r < nr; r++)
                    for (int v = 0; v < 4; v++) {
                        double acc = 0;
                        for (int c = 0; c < 9; c++)
                            acc += ms[9 * (ls + r) + c] * buf[36 * j + c + 9 * v];
                        oIn[r][v] = acc;
                    }
                const double *uR = side ? u + 4 * i : u + 4 * j;
                const double *uL = side ? u + 4 * j : u + 4 * i;
                double ML[4][4], MR[4][4];
                int rc2 = char_matrices(*cfg, uL, uR, un, ML, MR);
                if (rc2)
                    return rc2;
                for (int r = 0; r < nr; r++) {
                    double oC[4], tC[4], out[4];
                    for (int a = 0; a < 4; a++) {
                        double x = 0, y = 0;
                        for (int b = 0; b < 4; b++) {
                            x += ML[a][b] * oIn[r][b];
                            y += ML[a][b] * uC[r][b];
                        }
                        oC[a] = x, tC[a] = y;
                    }
                    // FWBAP_L2_Biway(uThisIn, uOtherIn, n = WBAP_nStd)
                    for (int a = 0; a < 4; a++) {
                        const double u1 = tC[a], u2 = oC[a];
                        const double frac = u1 / (std::abs(u2) + 1e-12) * sgn(u2);
                        const double th1 = frac * frac * frac;
                        const double th2 = (frac * frac) * (frac * frac);
                        out[a] = u1 * (nStd + th1) / (nStd + th2);
                    }
                    for (int a = 0; a < 4; a++) {
                        double x = 0;
                        for (int b = 0; b < 4; b++)
                            x += MR[a][b] * out[b];
                        // FWBAP_L2_Multiway(uOthers, n1 = 1): the first entry is the centre
                        if (nOther == 0) {
                            first[r][a] = x;
                        } else {
                            const double ti = first[r][a] / (std::abs(x) + eps4) * sgn(x);
                            down[r][a] += (ti * ti) * (ti * ti);
                            up[r][a] += ti * ti * ti;
                        }
                    }
                }
                nOther++;
            }
            for (int r = 0; r < nr; r++)
                for (int v = 0; v < 4; v++)
                    buf1[36 * i + (ls + r) + 9 * v] =
                        nOther ? first[r][v] * (up[r][v] + 1.0) / (down[r][v] + 1.0) : uC[r][v];
            return ERR_OK;
        });
        if (rc)
            return rc;
        for (int64_t i = 0; i < nc; i++)
            if (limited(i))
                for (int r = ls; r <= le; r++)
                    for (int v = 0; v < 4; v++)
                        buf[36 * i + r + 9 * v] = buf1[36 * i + r + 9 * v];
    }
    for (int64_t k = 0; k < 36 * nc; k++)
        if (!std::isfinite(urec_new[k]))
            return ERR_NAN;
    return ERR_OK;
}

// ExplicitSSPRK3LocalDt::Step, DNDS_ODE.hpp:131-178; frhs as the Eikonal
// solver composes it (DNDS_FV_EikonalCRSolver.hpp:948-999): sweeps then RHS.
extern "C" int oracle_ssprk3_steps(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                   int32_t n_steps, int32_t n_sweeps, int32_t use_limiter, double cfl, double max_dt,
                                   int32_t use_local_dt, double *u, double *urec, double *t) {
    const int64_t nc = m->n_cell;
    std::vector<double> dt(nc), rhs(4 * nc), xLast(4 * nc), urecNew(36 * nc), dlf(m->n_face);
    std::vector<double> urecLim(use_limiter ? 36 * nc : 0), ifUse(use_limiter ? nc : 0);
    const double Coef[3] = {1.0, 0.25, 2.0 / 3.0};
    for (int step = 0; step < n_steps; step++) {
        double dtMin;
        int rc = oracle_evaluate_dt(m, s, cfg, u, cfl, max_dt, use_local_dt, dt.data(), &dtMin, nullptr, dlf.data());
        if (rc)
            return rc;
        std::memcpy(xLast.data(), u, sizeof(double) * 4 * nc);
        const double ct[3] = {1, 0.25, 2. / 3.};
        for (int stage = 0; stage < 3; stage++) {
            const double ts = *t + ct[stage] * dtMin;
            for (int k = 0; k < n_sweeps; k++) {
                rc = rec_sweep(m, s, cfg, u, urec, urecNew.data(), ts);
                if (rc)
                    return rc;
            }
            if (use_limiter) {
                rc = cfg->limiter_procedure == 1
                         ? oracle_wbap_v3(m, s, cfg, u, urec, urecLim.data(), ifUse.data(), 0)
                         : oracle_wbap_v2(m, s, cfg, u, urec, urecLim.data(), ifUse.data(), 0);
                if (rc)
                    return rc;
            }
            rc = oracle_evaluate_rhs(m, s, cfg, dlf.data(), u, use_limiter ? urecLim.data() : urec, ts, rhs.data());
            if (rc)
                return rc;
            for (int64_t i = 0; i < nc; i++)
                for (int v = 0; v < 4; v++) {
                    double r = rhs[4 * i + v] * dt[i];
                    double x = u[4 * i + v] + r;
                    if (stage > 0) {
                        x *= Coef[stage] / (1 - Coef[stage]);
                        x += xLast[4 * i + v];
                        x *= 1 - Coef[stage];
                    }
                    u[4 * i + v] = x;
                }
        }
        *t += dtMin;
    }
    return ERR_OK;
}

// ---------------------------------------------------------------------------
// LU-SGS (DNDS_FV_EulerEvaluator.cpp:995-1357) with jacobianTypeCode 0 and the
// Euler model (the source Jacobian is zero), and the implicit-Euler dual-time
// driver around it (ImplicitEulerDualTimeStep, DNDS_ODE.hpp:201-265, with the
// solver's frhs / fdtau / fsolve / fstop, DNDS_FV_EulerSolver.hpp:713-1000).
// ---------------------------------------------------------------------------
namespace {

// fluxJacobian0_Right_Times_du, DNDS_FV_EulerEvaluator.hpp:928-956, with
// IdealGasUIncrement and GasInviscidFluxFacialIncrement (DNDS_Gas.hpp:122-146)
void flux_jacobian_times_du(double gamma, const double U[4], const double n[2], const double dU[4],
                            double lambdaMain, double dF[4]) {
    const double velo[2] = {U[1] / U[0], U[2] / U[0]};
    double p, asqr, H;
    ideal_gas_thermal(U[3], U[0], velo[0] * velo[0] + velo[1] * velo[1], gamma, p, asqr, H);
    double dVelo[2];
    for (int d = 0; d < 2; d++)
        dVelo[d] = (dU[1 + d] - U[1 + d] / U[0] * dU[0]) / U[0];
    const double dp = (gamma - 1) * (dU[3] - 0.5 * ((dU[1] * velo[0] + dU[2] * velo[1]) +
                                                    (U[1] * dVelo[0] + U[2] * dVelo[1])));
    const double vn = velo[0] * n[0] + velo[1] * n[1];
    const double dvn = dVelo[0] * n[0] + dVelo[1] * n[1];
    dF[0] = dU[1] * n[0] + dU[2] * n[1];
    for (int d = 0; d < 2; d++)
        dF[1 + d] = dU[1 + d] * vn + U[1 + d] * dvn + n[d] * dp;
    dF[3] = (dU[3] + dp) * vn + (U[3] + p) * dvn;
    for (int k = 0; k < 4; k++)
        dF[k] -= lambdaMain * dU[k];
}

// one LU-SGS half sweep: forward visits the scan order ascending and uses the
// neighbours scanned earlier, backward descending with those scanned later
int lusgs_sweep(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg, const int64_t *scan2cell,
                const double *jinv, const double *lambda_face, double alphaDiag, const double *rhs, const double *u,
                const double *uInc, double *uIncNew, bool forward) {
    const int64_t nc = m->n_cell;
    std::vector<int64_t> cell2scan(nc);
    for (int64_t k = 0; k < nc; k++)
        cell2scan[scan2cell ? scan2cell[k] : k] = k;
    for (int64_t kk = 0; kk < nc; kk++) {
        const int64_t iScan = forward ? kk : nc - 1 - kk;
        const int64_t i = scan2cell ? scan2cell[iScan] : iScan;
        double buf[4];
        for (int v = 0; v < 4; v++)
            buf[v] = forward ? rhs[4 * i + v] : 0.0;
        for (int64_t q = m->cell_ptr[i]; q < m->cell_ptr[i + 1]; q++) {
            const int64_t f = m->cell2face[q];
            const int side = m->face2cell[2 * f] == i ? 0 : 1;
            const int64_t j = m->face2cell[2 * f + 1 - side];
            if (j == kNoCell)
                continue;
            const int64_t jScan = cell2scan[j];
            if (forward ? !(jScan < iScan) : !(jScan > iScan))
                continue;
            double n[2];
            normalized(s->face_norm_center[2 * f], s->face_norm_center[2 * f + 1], n[0], n[1]);
            const double sg = side ? -1.0 : 1.0;
            n[0] *= sg, n[1] *= sg;
            double dF[4];
            flux_jacobian_times_du(cfg->gamma, u + 4 * j, n, uInc + 4 * j, lambda_face[f], dF);
            const double coef = (0.5 * alphaDiag) * s->face_area[f] / s->volume[i];
            for (int v = 0; v < 4; v++)
                buf[v] -= coef * dF[v];
        }
        for (int v = 0; v < 4; v++) {
            if (forward)
                uIncNew[4 * i + v] = jinv[i] * buf[v];
            else
                uIncNew[4 * i + v] += jinv[i] * buf[v];
        }
        for (int v = 0; v < 4; v++)
            if (!std::isfinite(uIncNew[4 * i + v]))
                return ERR_NAN;
    }
    return ERR_OK;
}

} // namespace

// LUSGSMatrixInit (DNDS_FV_EulerEvaluator.cpp:995-1047): the diagonal
// 1/dTau + 1/dt + sum_f (alpha/2) S_f lambda_f / V, inverted
extern "C" int oracle_lusgs_matrix_init(const oracle_mesh *m, const oracle_setup *s, const double *dTau, double dt,
                                        double alphaDiag, const double *lambda_face, double *jinv) {
    for (int64_t i = 0; i < m->n_cell; i++) {
        double fp = 1.0 / dTau[i] + 1.0 / dt;
        for (int64_t q = m->cell_ptr[i]; q < m->cell_ptr[i + 1]; q++) {
            const int64_t f = m->cell2face[q];
            fp += (0.5 * alphaDiag) * s->face_area[f] * lambda_face[f] / s->volume[i];
        }
        jinv[i] = 1.0 / (fp + alphaDiag * 0.0);
    }
    return ERR_OK;
}

// UpdateLUSGSForward (DNDS_FV_EulerEvaluator.cpp:1149-1276)
extern "C" int oracle_lusgs_forward(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                    const int64_t *scan2cell, const double *jinv, const double *lambda_face,
                                    double alphaDiag, const double *rhs, const double *u, const double *uInc,
                                    double *uIncNew) {
    return lusgs_sweep(m, s, cfg, scan2cell, jinv, lambda_face, alphaDiag, rhs, u, uInc, uIncNew, true);
}

// UpdateLUSGSBackward (DNDS_FV_EulerEvaluator.cpp:1285-1357)
extern "C" int oracle_lusgs_backward(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                     const int64_t *scan2cell, const double *jinv, const double *lambda_face,
                                     double alphaDiag, const double *rhs, const double *u, const double *uInc,
                                     double *uIncNew) {
    return lusgs_sweep(m, s, cfg, scan2cell, jinv, lambda_face, alphaDiag, rhs, u, uInc, uIncNew, false);
}

// LUSGSMatrixVec (DNDS_FV_EulerEvaluator.cpp:1062-1145), jacobianTypeCode 0:
// Ax_i = D_i x_i - sum_f (alpha/2) S_f / V_i dF(u_j, x_j) over every
// neighbour, faces in cell2face order
extern "C" int oracle_lusgs_matrix_vec(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                       const double *diag, const double *lambda_face, double alphaDiag,
                                       const double *u, const double *x, double *Ax) {
    for (int64_t i = 0; i < m->n_cell; i++) {
        double buf[4] = {0, 0, 0, 0};
        for (int64_t q = m->cell_ptr[i]; q < m->cell_ptr[i + 1]; q++) {
            const int64_t f = m->cell2face[q];
            const int side = m->face2cell[2 * f] == i ? 0 : 1;
            const int64_t j = m->face2cell[2 * f + 1 - side];
            if (j == kNoCell)
                continue;
            double n[2];
            normalized(s->face_norm_center[2 * f], s->face_norm_center[2 * f + 1], n[0], n[1]);
            const double sg = side ? -1.0 : 1.0;
            n[0] *= sg, n[1] *= sg;
            double dF[4];
            flux_jacobian_times_du(cfg->gamma, u + 4 * j, n, x + 4 * j, lambda_face[f], dF);
            const double coef = (0.5 * alphaDiag) * s->face_area[f] / s->volume[i];
            for (int v = 0; v < 4; v++)
                buf[v] -= coef * dF[v];
        }
        for (int v = 0; v < 4; v++)
            Ax[4 * i + v] = diag[i] * x[4 * i + v] - buf[v];
    }
    return ERR_OK;
}

namespace {

double norm2(const std::vector<double> &a) {
    double q = 0;
    for (double v : a)
        q += v * v;
    return std::sqrt(q);
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double q = 0;
    for (size_t k = 0; k < a.size(); k++)
        q += a[k] * b[k];
    return q;
}

// least squares min |H y - beta e1| for the (k+1) x k Hessenberg H (row-major),
// by Givens rotations
std::vector<double> hessenberg_lsq(std::vector<double> H, int k, double beta) {
    std::vector<double> g(size_t(k) + 1, 0.0);
    g[0] = beta;
    for (int j = 0; j < k; j++) {
        const double a = H[size_t(j) * k + j], b = H[size_t(j + 1) * k + j];
        const double r = std::hypot(a, b), c = r > 0 ? a / r : 1.0, sn = r > 0 ? b / r : 0.0;
        for (int col = j; col < k; col++) {
            const double x = H[size_t(j) * k + col], y = H[size_t(j + 1) * k + col];
            H[size_t(j) * k + col] = c * x + sn * y;
            H[size_t(j + 1) * k + col] = -sn * x + c * y;
        }
        const double x = g[size_t(j)], y = g[size_t(j) + 1];
        g[size_t(j)] = c * x + sn * y, g[size_t(j) + 1] = -sn * x + c * y;
    }
    std::vector<double> yv(size_t(k), 0.0);
    for (int j = k - 1; j >= 0; j--) {
        double acc = g[size_t(j)];
        for (int col = j + 1; col < k; col++)
            acc -= H[size_t(j) * k + col] * yv[size_t(col)];
        yv[size_t(j)] = acc / H[size_t(j) * k + j];
    }
    return yv;
}

} // namespace

// the solver's fsolve (DNDS_FV_EulerSolver.hpp:852-975), jacobianTypeCode 0,
// nSGSIterationInternal 0, CompressInc the identity for the Euler model
extern "C" int oracle_fsolve(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                             const int64_t *scan2cell, const double *dTau, double dt, double alphaDiag,
                             const double *lambda_face, const double *u, double *rhs, double *xinc) {
    const int64_t nc = m->n_cell, n = 4 * nc;
    std::vector<double> jinv(nc), diag(nc);
    std::fill(xinc, xinc + n, 0.0);
    oracle_lusgs_matrix_init(m, s, dTau, dt, alphaDiag, lambda_face, jinv.data());
    for (int64_t i = 0; i < nc; i++) { // jacobianCell = fpDivisor I + alphaDiag 0
        double fp = 1.0 / dTau[i] + 1.0 / dt;
        for (int64_t q = m->cell_ptr[i]; q < m->cell_ptr[i + 1]; q++) {
            const int64_t f = m->cell2face[q];
            fp += (0.5 * alphaDiag) * s->face_area[f] * lambda_face[f] / s->volume[i];
        }
        diag[i] = fp + alphaDiag * 0.0;
    }
    for (int64_t i = 0; i < nc; i++)
        for (int v = 0; v < 4; v++)
            rhs[4 * i + v] = (rhs[4 * i + v] * dTau[i]) / dTau[i];
    int rc;
    auto FML = [&](const double *b, double *out) -> int { // LU-SGS forward + backward, in place
        int r = oracle_lusgs_forward(m, s, cfg, scan2cell, jinv.data(), lambda_face, alphaDiag, b, u, out, out);
        return r ? r : oracle_lusgs_backward(m, s, cfg, scan2cell, jinv.data(), lambda_face, alphaDiag, b, u, out, out);
    };
    if (cfg->gmres_code == 0 || cfg->gmres_code == 2)
        if ((rc = FML(rhs, xinc)))
            return rc;
    if (cfg->gmres_code != 0) {
        // GMRES_LeftPreconditioned::solve (DNDS_Linear.hpp:48-104), FStop never true
        const int k = cfg->n_gmres_space;
        const uint32_t nRestart = uint32_t(cfg->n_gmres_iter);
        std::vector<std::vector<double>> Vs(static_cast<size_t>(k) + 1, std::vector<double>(static_cast<size_t>(n)));
        std::vector<double> Vt(static_cast<size_t>(n)), MLb(static_cast<size_t>(n), 0.0),
            H(static_cast<size_t>(k + 1) * static_cast<size_t>(k), 0.0);
        if ((rc = FML(rhs, MLb.data())))
            return rc;
        for (uint32_t iRestart = 0; iRestart <= nRestart; iRestart++) {
            oracle_lusgs_matrix_vec(m, s, cfg, diag.data(), lambda_face, alphaDiag, u, xinc, Vt.data());
            std::fill(Vs[0].begin(), Vs[0].end(), 0.0);
            if ((rc = FML(Vt.data(), Vs[0].data())))
                return rc;
            for (int64_t q = 0; q < n; q++)
                Vs[0][size_t(q)] += MLb[size_t(q)] * -1.0;
            const double beta = norm2(Vs[0]);
            if (std::fabs(beta) < 1e-200 || iRestart == nRestart)
                break;
            for (double &v : Vs[0])
                v *= -1.0 / beta;
            for (int j = 0; j < k; j++) {
                oracle_lusgs_matrix_vec(m, s, cfg, diag.data(), lambda_face, alphaDiag, u, Vs[size_t(j)].data(),
                                        Vt.data());
                std::fill(Vs[size_t(j) + 1].begin(), Vs[size_t(j) + 1].end(), 0.0);
                if ((rc = FML(Vt.data(), Vs[size_t(j) + 1].data())))
                    return rc;
                for (int i = 0; i <= j; i++)
                    H[size_t(i) * k + j] = dot(Vs[size_t(j) + 1], Vs[size_t(i)]);
                for (int i = 0; i <= j; i++)
                    for (int64_t q = 0; q < n; q++)
                        Vs[size_t(j) + 1][size_t(q)] += Vs[size_t(i)][size_t(q)] * -H[size_t(i) * k + j];
                H[size_t(j + 1) * k + j] = norm2(Vs[size_t(j) + 1]);
                for (double &v : Vs[size_t(j) + 1])
                    v *= 1.0 / H[size_t(j + 1) * k + j];
            }
            const std::vector<double> y = hessenberg_lsq(H, k, beta);
            for (int j = 0; j < k; j++)
                for (int64_t q = 0; q < n; q++)
                    xinc[q] += Vs[size_t(j)][size_t(q)] * y[size_t(j)];
        }
    }
    for (int64_t q = 0; q < n; q++)
        if (!std::isfinite(xinc[q]))
            return ERR_NAN;
    return ERR_OK;
}

// Implicit-Euler dual-time iterations (DNDS_ODE.hpp:248-263) with
// frhs = n_sweeps Jacobi sweeps + EvaluateRHS, fdtau = EvaluateDt / alpha,
// fsolve = LU-SGS forward + backward in place (nSGSIterationInternal = 0,
// CompressInc the identity for the Euler model) and fstop on the L1 norm of
// the density increment relative to its running maximum
// (DNDS_FV_EulerSolver.hpp:984-999).  res_hist receives the L1 increment
// norms (4 per iteration); returns the iterations done in *n_done.
extern "C" int oracle_implicit_euler(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                     const int64_t *scan2cell, int32_t max_iter, int32_t n_sweeps, double cfl,
                                     double dt, int32_t use_local_dt, double threshold, double t, double *u,
                                     double *urec, double *res_hist, int32_t *n_done) {
    const int64_t nc = m->n_cell;
    std::vector<double> dTau(nc), rhs(4 * nc), xLast(u, u + 4 * nc), xinc(4 * nc), urecNew(36 * nc);
    std::vector<double> lf(m->n_face), dlf(m->n_face), jinv(nc);
    const double alphaDiag = 1.0;
    double resBase[4] = {0, 0, 0, 0};
    *n_done = 0;
    for (int32_t iter = 1; iter <= max_iter; iter++) {
        double dtMin;
        int rc = oracle_evaluate_dt(m, s, cfg, u, cfl, 1e100, use_local_dt, dTau.data(), &dtMin, lf.data(),
                                    dlf.data());
        if (rc)
            return rc;
        for (int64_t i = 0; i < nc; i++)
            dTau[i] /= alphaDiag;
        for (int k = 0; k < n_sweeps; k++)
            if ((rc = rec_sweep(m, s, cfg, u, urec, urecNew.data(), t + dt)))
                return rc;
        if ((rc = oracle_evaluate_rhs(m, s, cfg, dlf.data(), u, urec, t + dt, rhs.data())))
            return rc;
        for (int64_t k = 0; k < 4 * nc; k++) {
            double r = xLast[k];
            r -= u[k];
            r *= 1.0 / dt;
            rhs[k] = r + rhs[k];
        }
        if ((rc = oracle_fsolve(m, s, cfg, scan2cell, dTau.data(), dt, alphaDiag, lf.data(), u, rhs.data(),
                                    xinc.data())))
            return rc;
        for (int64_t k = 0; k < 4 * nc; k++)
            u[k] += xinc[k];
        double res[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i < nc; i++)
            for (int v = 0; v < 4; v++)
                res[v] += std::fabs(xinc[4 * i + v]);
        for (int v = 0; v < 4; v++) {
            if (res_hist)
                res_hist[4 * (iter - 1) + v] = res[v];
            resBase[v] = iter == 1 ? res[v] : std::max(resBase[v], res[v]);
        }
        *n_done = iter;
        if (res[0] / resBase[0] < threshold)
            break;
    }
    return ERR_OK;
}

// ImplicitBDFDualTimeStep::Step (DNDS_ODE.hpp:441-500), n_steps steps of
// size dt from a fresh integrator of order k (step i uses order min(i, k)),
// with the same frhs / fdtau / fsolve / fstop as oracle_implicit_euler and
// alphaDiag = BDFCoefs(kCurrent - 1, 0) (DNDS_ODE.hpp:506-511).  n_iter
// receives the iterations done per step.
extern "C" int oracle_bdf_steps(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                const int64_t *scan2cell, int32_t k, int32_t n_steps, int32_t max_iter,
                                int32_t n_sweeps, double cfl, double dt, int32_t use_local_dt, double threshold,
                                double t, double *u, double *urec, int32_t *n_iter) {
    static const double coefs[4][5] = {{1. / 1., 1. / 1., 0. / 0., 0. / 0., 0. / 0.},
                                       {2. / 3., 4. / 3., -1. / 3., 0. / 0., 0. / 0.},
                                       {6. / 11., 18. / 11., -9. / 11., 2. / 11., 0. / 0.},
                                       {12. / 25., 48. / 25., -36. / 25., 16. / 25., -3. / 25.}};
    if (k < 1 || k > 4)
        return ERR_INVALID;
    const int64_t nc = m->n_cell;
    const int32_t prevSiz = k - 1;
    std::vector<std::vector<double>> xPrevs(size_t(prevSiz), std::vector<double>(4 * nc));
    int32_t cnPrev = 0, prevStart = k - 2;
    auto mod = [](int32_t a, int32_t b) { return ((a % b) + b) % b; };
    std::vector<double> dTau(nc), rhs(4 * nc), xLast(4 * nc), xinc(4 * nc), urecNew(36 * nc);
    std::vector<double> lf(m->n_face), dlf(m->n_face), jinv(nc);
    for (int32_t step = 0; step < n_steps; step++) {
        const int32_t kCurrent = cnPrev + 1;
        const double alphaDiag = coefs[kCurrent - 1][0];
        std::copy(u, u + 4 * nc, xLast.begin());
        double resBase[4] = {0, 0, 0, 0};
        n_iter[step] = 0;
        for (int32_t iter = 1; iter <= max_iter; iter++) {
            double dtMin;
            int rc = oracle_evaluate_dt(m, s, cfg, u, cfl, 1e100, use_local_dt, dTau.data(), &dtMin, lf.data(),
                                        dlf.data());
            if (rc)
                return rc;
            for (int64_t i = 0; i < nc; i++)
                dTau[i] /= alphaDiag;
            for (int q = 0; q < n_sweeps; q++)
                if ((rc = rec_sweep(m, s, cfg, u, urec, urecNew.data(), t + dt)))
                    return rc;
            if ((rc = oracle_evaluate_rhs(m, s, cfg, dlf.data(), u, urec, t + dt, rhs.data())))
                return rc;
            for (int64_t q = 0; q < 4 * nc; q++) {
                double r = rhs[q];
                r *= alphaDiag;
                r += u[q] * (-1. / dt);
                r += xLast[q] * (coefs[kCurrent - 1][1] / dt);
                for (int32_t ip = 0; ip < cnPrev; ip++)
                    r += xPrevs[size_t(mod(ip + prevStart, prevSiz))][q] * (coefs[kCurrent - 1][2 + ip] / dt);
                rhs[q] = r;
            }
            if ((rc = oracle_fsolve(m, s, cfg, scan2cell, dTau.data(), dt, alphaDiag, lf.data(), u, rhs.data(),
                                    xinc.data())))
            return rc;
            for (int64_t q = 0; q < 4 * nc; q++)
                u[q] += xinc[q] * 1.0;
            double res[4] = {0, 0, 0, 0};
            for (int64_t i = 0; i < nc; i++)
                for (int v = 0; v < 4; v++)
                    res[v] += std::fabs(xinc[4 * i + v]);
            for (int v = 0; v < 4; v++)
                resBase[v] = iter == 1 ? res[v] : std::max(resBase[v], res[v]);
            n_iter[step] = iter;
            if (res[0] / resBase[0] < threshold)
                break;
        }
        if (prevSiz) {
            prevStart = mod(prevStart - 1, prevSiz);
            xPrevs[size_t(prevStart)] = xLast;
            cnPrev = std::min(cnPrev + 1, prevSiz);
        }
        t += dt;
    }
    return ERR_OK;
}

// ImplicitSDIRK4DualTimeStep::Step (DNDS_ODE.hpp:316-371) with the same
// frhs / fdtau / fsolve / fstop as oracle_implicit_euler; frhs evaluates the
// boundary states at t + c_i dt.  n_iter receives the iterations of the 3
// stages.
extern "C" int oracle_sdirk4_step(const oracle_mesh *m, const oracle_setup *s, const oracle_setting *cfg,
                                  const int64_t *scan2cell, int32_t max_iter, int32_t n_sweeps, double cfl, double dt,
                                  int32_t use_local_dt, double threshold, double t, double *u, double *urec,
                                  int32_t *n_iter) {
    const double zeta = 0.128886400515;
    const double A[3][3] = {{zeta, 0, 0}, {0.5 - zeta, zeta, 0}, {2 * zeta, 1 - 4 * zeta, zeta}};
    const double Cc[3] = {zeta, 0.5, 1 - zeta};
    auto sqr = [](double x) { return x * x; };
    const double B[3] = {1. / (6 * sqr(2 * zeta - 1)), (4 * sqr(zeta) - 4 * zeta + 2. / 3.) / sqr(2 * zeta - 1),
                         1. / (6 * sqr(2 * zeta - 1))};
    const int64_t nc = m->n_cell;
    std::vector<double> dTau(nc), rhs(4 * nc), xLast(u, u + 4 * nc), xinc(4 * nc), urecNew(36 * nc);
    std::vector<std::vector<double>> rhsbuf(3, std::vector<double>(4 * nc));
    std::vector<double> lf(m->n_face), dlf(m->n_face), jinv(nc);
    for (int iB = 0; iB < 3; iB++) {
        std::copy(xLast.begin(), xLast.end(), u);
        double resBase[4] = {0, 0, 0, 0};
        const double alphaDiag = A[iB][iB];
        n_iter[iB] = 0;
        for (int32_t iter = 1; iter <= max_iter; iter++) {
            double dtMin;
            int rc = oracle_evaluate_dt(m, s, cfg, u, cfl, 1e100, use_local_dt, dTau.data(), &dtMin, lf.data(),
                                        dlf.data());
            if (rc)
                return rc;
            for (int64_t i = 0; i < nc; i++)
                dTau[i] /= alphaDiag;
            const double ts = t + Cc[iB] * dt;
            for (int q = 0; q < n_sweeps; q++)
                if ((rc = rec_sweep(m, s, cfg, u, urec, urecNew.data(), ts)))
                    return rc;
            if ((rc = oracle_evaluate_rhs(m, s, cfg, dlf.data(), u, urec, ts, rhsbuf[size_t(iB)].data())))
                return rc;
            for (int64_t q = 0; q < 4 * nc; q++) {
                double r = xLast[q];
                r -= u[q];
                r *= 1.0 / dt;
                for (int jB = 0; jB <= iB; jB++)
                    r += rhsbuf[size_t(jB)][q] * A[iB][jB];
                rhs[q] = r;
            }
            if ((rc = oracle_fsolve(m, s, cfg, scan2cell, dTau.data(), dt, alphaDiag, lf.data(), u, rhs.data(),
                                    xinc.data())))
            return rc;
            for (int64_t q = 0; q < 4 * nc; q++)
                u[q] += xinc[q] * 1.0;
            double res[4] = {0, 0, 0, 0};
            for (int64_t i = 0; i < nc; i++)
                for (int v = 0; v < 4; v++)
                    res[v] += std::fabs(xinc[4 * i + v]);
            for (int v = 0; v < 4; v++)
                resBase[v] = iter == 1 ? res[v] : std::max(resBase[v], res[v]);
            n_iter[iB] = iter;
            if (res[0] / resBase[0] < threshold)
                break;
        }
    }
    std::copy(xLast.begin(), xLast.end(), u);
    for (int jB = 0; jB < 3; jB++)
        for (int64_t q = 0; q < 4 * nc; q++)
            u[q] += rhsbuf[size_t(jB)][q] * (B[jB] * dt);
    return ERR_OK;
}

extern "C" void oracle_set_threads(int32_t n) { g_threads = n > 1 ? n : 1; }

extern "C" int oracle_flux_jacobian_times_du(const double *U, const double *unit_norm, const double *dU,
                                             double lambda_main, double gamma, double *dF) {
    flux_jacobian_times_du(gamma, U, unit_norm, dU, lambda_main, dF);
    return ERR_OK;
}
