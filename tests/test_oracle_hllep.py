"""Pins of the oracle's HLLEP flux (Gas::HLLEPFlux_IdealGas in its V1 form,
DNDS_Gas.hpp:178-310), the Riemann solver of the reference's own 2D
Riemann configuration (data/euler_config_2DRiemann.json, "HLLEP").  The
expected values are built from textbook pieces, not from the oracle's
expressions: Einfeldt's HLLE flux (Toro, Riemann Solvers and Numerical
Methods for Fluid Dynamics, 3rd ed., sec. 10.3) with the speed estimates
S_L = min(u~ - a~, u_L - a_L), S_R = max(u~ + a~, u_R + a_R), and the Roe
wave strengths written with pressure and velocity jumps (Toro eqs.
11.68-11.70).  CPU only."""
import numpy as np
import pytest

import oracle
from test_oracle_pins import G, cons2prim, prim, xflux


def roe_waves(UL, UR, gamma=G):
    """Roe averages and the four waves (alpha_k, R_k) in x: acoustic
    alpha = (dp -/+ rho~ a~ du) / (2 a~^2), entropy alpha = drho - dp / a~^2
    with R = (1, u, v, |V|^2 / 2), shear alpha = rho~ dv with R = (0, 0, 1, v)."""
    rl, ul, vl, pl = cons2prim(UL, gamma)
    rr, ur, vr, pr = cons2prim(UR, gamma)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    Hl, Hr = (UL[3] + pl) / rl, (UR[3] + pr) / rr
    u = (sl * ul + sr * ur) / (sl + sr)
    v = (sl * vl + sr * vr) / (sl + sr)
    H = (sl * Hl + sr * Hr) / (sl + sr)
    a = np.sqrt((gamma - 1) * (H - 0.5 * (u * u + v * v)))
    rho = sl * sr
    dp, du, dv, drho = pr - pl, ur - ul, vr - vl, rr - rl
    waves = {
        "minus": ((dp - rho * a * du) / (2 * a * a), np.array([1.0, u - a, v, H - u * a])),
        "entropy": (drho - dp / (a * a), np.array([1.0, u, v, 0.5 * (u * u + v * v)])),
        "shear": (rho * dv, np.array([0.0, 0.0, 1.0, v])),
        "plus": ((dp + rho * a * du) / (2 * a * a), np.array([1.0, u + a, v, H + u * a])),
    }
    return u, a, waves


def hlle_speeds(UL, UR, gamma=G):
    u, a, _ = roe_waves(UL, UR, gamma)
    rl, ul, _, pl = cons2prim(UL, gamma)
    rr, ur, _, pr = cons2prim(UR, gamma)
    SL = min(u - a, ul - np.sqrt(gamma * pl / rl))
    SR = max(u + a, ur + np.sqrt(gamma * pr / rr))
    return max(SR, 0.0), min(SL, 0.0)


def hlle_textbook(UL, UR, gamma=G):
    """F = (S+ F_L - S- F_R + S+ S- (U_R - U_L)) / (S+ - S-)."""
    Sp, Sn = hlle_speeds(UL, UR, gamma)
    return (Sp * xflux(UL, gamma) - Sn * xflux(UR, gamma) + Sp * Sn * (UR - UL)) / (Sp - Sn)


def hllep_textbook(UL, UR, gamma=G):
    """HLLE plus the anti-diffusion of the two linearly degenerate waves
    (the HLLEM form of Einfeldt, Munz, Roe and Sjogreen, J. Comput. Phys. 92,
    1991): their HLLE speed |((S+ + S-) u - 2 S+ S-) / (S+ - S-)| is replaced
    by |((S+ + S-) u - 2 (1 - delta) S+ S-) / (S+ - S-)|, delta = a / (2 a)
    (the reference's coefficient a / (|u| + a) with |u| taken as a)."""
    Sp, Sn = hlle_speeds(UL, UR, gamma)
    u, a, w = roe_waves(UL, UR, gamma)
    delta = a / (a + a + 1e-200)
    e_hlle = ((Sp + Sn) * u - 2 * Sp * Sn) / (Sp - Sn)
    e_ep = abs(((Sp + Sn) * u - 2 * (1 - delta) * Sp * Sn) / (Sp - Sn))
    lin = w["entropy"][0] * w["entropy"][1] + w["shear"][0] * w["shear"][1]
    return hlle_textbook(UL, UR, gamma) + 0.5 * (e_hlle - e_ep) * lin


def _pairs(seed, n, mach=0.6):
    rng = np.random.default_rng(seed)
    return [(prim(rng.uniform(0.3, 3), rng.uniform(-mach, mach), rng.uniform(-1, 1), rng.uniform(0.2, 4)),
             prim(rng.uniform(0.3, 3), rng.uniform(-mach, mach), rng.uniform(-1, 1), rng.uniform(0.2, 4)))
            for _ in range(n)]


def test_hllep_of_equal_states_is_the_physical_flux():
    for UL, _ in _pairs(0, 20, mach=3.0):
        assert np.abs(oracle.hllep_flux(UL, UL) - xflux(UL)).max() <= 1e-14 * np.abs(xflux(UL)).max()


@pytest.mark.parametrize("seed", range(3))
def test_hllep_supersonic_is_the_upwind_flux(seed):
    """S- = 0 (every wave moving right) turns each speed into the Roe speed
    itself, so F = (F_L + F_R)/2 - R Lambda alpha / 2 = F_L by the Roe
    property; mirrored, F = F_R.  Tangential velocities differ, so all four
    wave strengths and eigenvectors enter."""
    rng = np.random.default_rng(20 + seed)
    for _ in range(10):
        rl, rr = rng.uniform(0.5, 2, 2)
        pl, pr = rng.uniform(0.5, 2, 2)
        u = rng.uniform(3.0, 5.0) * np.sqrt(G * max(pl / rl, pr / rr))
        vl, vr = rng.uniform(-2, 2, 2)
        UL, UR = prim(rl, u, vl, pl), prim(rr, 1.1 * u, vr, pr)
        F = oracle.hllep_flux(UL, UR)
        assert np.abs(F - xflux(UL)).max() <= 1e-13 * np.abs(xflux(UL)).max(), (F, xflux(UL))
        ULm, URm = prim(rl, -1.1 * u, vl, pl), prim(rr, -u, vr, pr)
        Fn = oracle.hllep_flux(ULm, URm)
        assert np.abs(Fn - xflux(URm)).max() <= 1e-13 * np.abs(xflux(URm)).max()


@pytest.mark.parametrize("seed", range(4))
def test_hllep_subsonic_matches_hlle_with_linear_wave_antidiffusion(seed):
    """Subsonic pairs with jumps in every variable: the oracle equals the
    textbook HLLE flux plus the anti-diffusion of the entropy and shear
    waves, and differs from plain HLLE (the anti-diffusion is active)."""
    for UL, UR in _pairs(seed, 30):
        F = oracle.hllep_flux(UL, UR)
        ref = hllep_textbook(UL, UR)
        scale = max(np.abs(ref).max(), 1.0)
        assert np.abs(F - ref).max() <= 1e-13 * scale, (UL, UR, F, ref)
    UL, UR = prim(1.0, 0.2, 0.5, 1.0), prim(0.4, 0.1, -0.5, 1.1)
    assert np.abs(oracle.hllep_flux(UL, UR) - hlle_textbook(UL, UR)).max() > 1e-2


def test_hllep_textbook_form_separates_plausible_slips():
    """The comparison is sharp for the reference's quirks: taking delta from
    |u| (a / (|u| + a)) instead of a / (2 a), or applying delta to the
    acoustic waves too, moves the flux far beyond the tolerance."""
    UL, UR = prim(1.0, 0.5, 0.5, 1.0), prim(0.4, 0.3, -0.5, 1.6)
    F = oracle.hllep_flux(UL, UR)
    assert np.abs(F - hllep_textbook(UL, UR)).max() <= 1e-13
    Sp, Sn = hlle_speeds(UL, UR)
    u, a, w = roe_waves(UL, UR)
    lin = w["entropy"][0] * w["entropy"][1] + w["shear"][0] * w["shear"][1]
    e_hlle = ((Sp + Sn) * u - 2 * Sp * Sn) / (Sp - Sn)
    d_bad = a / (abs(u) + a)
    bad = hlle_textbook(UL, UR) + 0.5 * (e_hlle - abs(((Sp + Sn) * u - 2 * (1 - d_bad) * Sp * Sn) / (Sp - Sn))) * lin
    assert np.abs(bad - F).max() > 1e-3
    ac = w["minus"][0] * w["minus"][1] + w["plus"][0] * w["plus"][1]
    assert np.abs(0.5 * 2 * 0.5 * Sp * Sn / (Sp - Sn) * ac).max() > 1e-3


def test_hllep_rejects_nonpositive_states():
    with pytest.raises(RuntimeError):
        oracle.hllep_flux(prim(1.0, 0, 0, 1.0), np.array([-1.0, 0.0, 0.0, 1.0]))
    with pytest.raises(RuntimeError):
        oracle.hllep_flux(prim(1.0, 0, 0, 1.0), np.array([1.0, 0.0, 0.0, -1.0]))
