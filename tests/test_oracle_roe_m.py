"""Pins of the oracle's Roe_M1..Roe_M5 fluxes (riemannSolverType 11-15:
Gas::RoeFlux_IdealGas_HartenYee<2, 1..5>, DNDS_Gas.hpp:405-603, eigenvalue
variants at :461-518; Roe_M4 is the solver of the reference's SA
configuration data/eulerSA_config.json).  Each expected flux is the Roe form
F = (F_L + F_R)/2 - sum_k |lam_k| alpha_k R_k / 2 with the wave strengths and
eigenvectors written independently (Toro eqs. 11.68-11.70, via
test_oracle_hllep.roe_waves) and the variant's eigenvalues restated from
their definitions; Roe_M2 is the local Lax (Rusanov) flux.  CPU only."""
import numpy as np
import pytest

import oracle
from test_oracle_hllep import roe_waves
from test_oracle_pins import G, cons2prim, prim, xflux

LD = 0.2  # scaleLD, DNDS_Gas.hpp:411


def sign(a):
    return 1.0 if a > 0 else (-1.0 if a < 0 else 0.0)


def variant_eigenvalues(UL, UR, scheme):
    """(lam_minus, lam_linear, lam_plus) of Roe_M<scheme>."""
    u, a, _ = roe_waves(UL, UR)
    rl, ul, _, pl = cons2prim(UL)
    rr, ur, _, pr = cons2prim(UR)
    al, ar = np.sqrt(G * pl / rl), np.sqrt(G * pr / rr)
    if scheme == 1:  # cLLF_M: per-side speeds, sound speeds capped at |u| / 0.2
        aL, aR = min(al, abs(ul) / LD), min(ar, abs(ur) / LD)
        return (max(abs(ul - aL), abs(ur - aR)), max(abs(ul), abs(ur)), max(abs(ul + aL), abs(ur + aR)))
    if scheme == 3:  # low dissipation: acoustic speeds from a* = min(|u| / 0.2, a)
        s = min(abs(u) / LD, a)
        return abs(u - s), abs(u), abs(u + s)
    if scheme == 4:  # increased dissipation: u* = sign(u) max(0.2 a, |u|)
        us = sign(u) * max(a * LD, abs(u))
        return abs(us - a), abs(us), abs(us + a)
    if scheme == 5:  # no entropy fix
        return abs(u - a), abs(u), abs(u + a)
    raise ValueError(scheme)


def roe_variant_textbook(UL, UR, scheme):
    if scheme == 2:
        rl, ul, _, pl = cons2prim(UL)
        rr, ur, _, pr = cons2prim(UR)
        lam = max(abs(ul) + np.sqrt(G * pl / rl), abs(ur) + np.sqrt(G * pr / rr))
        return 0.5 * (xflux(UL) + xflux(UR)) - 0.5 * lam * (UR - UL)
    _, _, w = roe_waves(UL, UR)
    lm, ll, lp = variant_eigenvalues(UL, UR, scheme)
    diss = (lm * w["minus"][0] * w["minus"][1] + ll * (w["entropy"][0] * w["entropy"][1] + w["shear"][0] * w["shear"][1])
            + lp * w["plus"][0] * w["plus"][1])
    return 0.5 * (xflux(UL) + xflux(UR)) - 0.5 * diss


def _pairs(seed, n, mach):
    rng = np.random.default_rng(seed)
    return [(prim(rng.uniform(0.3, 3), rng.uniform(-mach, mach), rng.uniform(-1, 1), rng.uniform(0.2, 4)),
             prim(rng.uniform(0.3, 3), rng.uniform(-mach, mach), rng.uniform(-1, 1), rng.uniform(0.2, 4)))
            for _ in range(n)]


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("mach", [0.05, 0.6, 3.0])
def test_roe_variant_matches_the_textbook_form(scheme, mach):
    """Random pairs from low-Mach (where Roe_M3/M4 change the acoustic
    speeds) to supersonic, with jumps in every variable."""
    for UL, UR in _pairs(100 * scheme + int(10 * mach), 25, mach):
        F = oracle.roe_m_flux(UL, UR, scheme)
        ref = roe_variant_textbook(UL, UR, scheme)
        assert np.abs(F - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0), (scheme, UL, UR, F, ref)


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_roe_variant_of_equal_states_is_the_physical_flux(scheme):
    for UL, _ in _pairs(7, 10, 2.0):
        assert np.abs(oracle.roe_m_flux(UL, UL, scheme) - xflux(UL)).max() <= 1e-14 * np.abs(xflux(UL)).max()


def test_roe_m4_at_zero_normal_velocity_uses_sign_zero():
    """At u~ = 0 exactly (mirror-symmetric states), sign(0) = 0 gives u* = 0,
    so the linear waves carry no dissipation and the acoustic speeds are a~.
    The signM(0) = -1 of the switch the reference leaves off
    (USE_SIGN_MINUS_AT_ROE_M4_FLUX) would give u* = -0.2 a~ instead."""
    UL, UR = prim(1.0, 0.3, 0.5, 1.0), prim(1.0, -0.3, -0.2, 1.0)
    u, a, w = roe_waves(UL, UR)
    assert u == 0.0
    F = oracle.roe_m_flux(UL, UR, 4)
    ref = roe_variant_textbook(UL, UR, 4)
    assert np.abs(F - ref).max() <= 1e-14
    diss = a * (w["minus"][0] * w["minus"][1] + w["plus"][0] * w["plus"][1])
    assert np.abs(F - (0.5 * (xflux(UL) + xflux(UR)) - 0.5 * diss)).max() <= 1e-14
    us = -LD * a
    bad = 0.5 * (xflux(UL) + xflux(UR)) - 0.5 * (abs(us - a) * w["minus"][0] * w["minus"][1] + abs(us) * (
        w["entropy"][0] * w["entropy"][1] + w["shear"][0] * w["shear"][1]) + abs(us + a) * w["plus"][0] * w["plus"][1])
    assert np.abs(F - bad).max() > 1e-3


def test_roe_variants_differ_where_they_should():
    """At low Mach, Roe_M3 lowers and Roe_M4 raises the dissipation against
    the plain Roe flux without a fix (Roe_M5), and the Harten-Yee fix of the
    Roe solver changes nothing there; near a sonic point it is the fix that
    differs.  The comparisons above are sharp enough to tell them apart."""
    UL, UR = prim(1.0, 0.02, 0.1, 1.0), prim(0.9, 0.03, -0.1, 1.1)
    f3, f4, f5 = (oracle.roe_m_flux(UL, UR, s) for s in (3, 4, 5))
    assert np.abs(f3 - f5).max() > 1e-3 and np.abs(f4 - f5).max() > 1e-4
    assert np.abs(oracle.roe_m_flux(UL, UR, 0) - f5).max() == 0.0
    a = np.sqrt(G)  # rho = p = 1
    UL, UR = prim(1.0, 1.01 * a, 0.1, 1.0), prim(0.95, 0.99 * a, -0.1, 1.05)
    assert np.abs(oracle.roe_m_flux(UL, UR, 0) - oracle.roe_m_flux(UL, UR, 5)).max() > 1e-4


def test_unknown_eigen_scheme_is_refused():
    with pytest.raises(RuntimeError):
        oracle.roe_m_flux(prim(1.0, 0, 0, 1.0), prim(1.0, 0, 0, 1.0), 6)
