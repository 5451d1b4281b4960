"""Pins of the oracle's branches that the consistency tests leave open: the
subsonic HLLC star state, Roe's shear and contact waves, the positivity
fallback of the reconstructed face values, and the special far-field ghost
states.  Each expected value comes from a closed form stated in the test
(textbook formula, Rankine-Hugoniot state, exact solution) rather than from
the oracle's own expressions.  CPU only."""
import numpy as np
import pytest

import oracle
from conftest import mesh_and_setup, rel_err

G = 1.4


def prim(rho, u, v, p, gamma=G):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def cons2prim(U, gamma=G):
    rho, mu, mv, E = U
    u, v = mu / rho, mv / rho
    return np.array([rho, u, v, (gamma - 1) * (E - 0.5 * rho * (u * u + v * v))])


def xflux(U, gamma=G):
    rho, u, v, p = cons2prim(U, gamma)
    return np.array([rho * u, rho * u * u + p, rho * u * v, (U[3] + p) * u])


def einfeldt_speeds(UL, UR, gamma=G):
    """Einfeldt's HLLE wave-speed estimates (Toro, Riemann Solvers, 3rd ed.,
    eqs. 10.52-10.54): Roe-averaged normal velocity -/+ d with
    d^2 = (sqrt(rhoL) aL^2 + sqrt(rhoR) aR^2) / (sqrt(rhoL) + sqrt(rhoR))
          + eta2 (uR - uL)^2,  eta2 = 0.5 sqrt(rhoL rhoR) / (sqrt(rhoL) + sqrt(rhoR))^2."""
    rl, ul, _, pl = cons2prim(UL, gamma)
    rr, ur, _, pr = cons2prim(UR, gamma)
    al2, ar2 = gamma * pl / rl, gamma * pr / rr
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    ubar = (sl * ul + sr * ur) / (sl + sr)
    eta2 = 0.5 * sl * sr / (sl + sr) ** 2
    d = np.sqrt((sl * al2 + sr * ar2) / (sl + sr) + eta2 * (ur - ul) ** 2)
    return ubar - d, ubar + d


def hllc_textbook(UL, UR, gamma=G):
    """Toro's HLLC (Riemann Solvers, sec. 10.4): S* from equal star
    pressures (eq. 10.37), the star states U*_K of eq. 10.39 and
    F*_K = F_K + S_K (U*_K - U_K) (eq. 10.38)."""
    SL, SR = einfeldt_speeds(UL, UR, gamma)
    FL, FR = xflux(UL, gamma), xflux(UR, gamma)
    if SL >= 0:
        return FL
    if SR <= 0:
        return FR
    rl, ul, vl, pl = cons2prim(UL, gamma)
    rr, ur, vr, pr = cons2prim(UR, gamma)
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))

    def star(U, rho, u, v, p, S):
        c = rho * (S - u) / (S - Ss)
        return c * np.array([1.0, Ss, v, U[3] / rho + (Ss - u) * (Ss + p / (rho * (S - u)))])

    if Ss >= 0:
        return FL + SL * (star(UL, rl, ul, vl, pl, SL) - UL)
    return FR + SR * (star(UR, rr, ur, vr, pr, SR) - UR)


def _subsonic_pairs(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        UL = prim(rng.uniform(0.3, 3), rng.uniform(-0.6, 0.6), rng.uniform(-1, 1), rng.uniform(0.2, 4))
        UR = prim(rng.uniform(0.3, 3), rng.uniform(-0.6, 0.6), rng.uniform(-1, 1), rng.uniform(0.2, 4))
        SL, SR = einfeldt_speeds(UL, UR)
        if SL < 0 < SR and abs(cons2prim(UL)[3] - cons2prim(UR)[3]) > 0.1:
            out.append((UL, UR))
    return out


@pytest.mark.parametrize("seed", range(4))
def test_hllc_subsonic_matches_the_textbook_star_state(seed):
    """DNDS_Gas.hpp:312-403 on subsonic pairs with pL != pR (both signs of
    S* occur): the reference's star flux form (S*(S_K U_K - F_K) +
    S_K p*_K D*) / (S_K - S*) equals Toro's F_K + S_K (U*_K - U_K); the
    pressure difference in S* and the eta2 term of d both enter."""
    signs = set()
    for UL, UR in _subsonic_pairs(seed, 25):
        F = oracle.hllc_flux(UL, UR)
        ref = hllc_textbook(UL, UR)
        assert np.abs(F - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0), (UL, UR, F, ref)
        rl, ul, _, pl = cons2prim(UL)
        rr, ur, _, pr = cons2prim(UR)
        SL, SR = einfeldt_speeds(UL, UR)
        signs.add((pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur)) >= 0)
    assert signs == {True, False}


def test_hllc_textbook_form_separates_plausible_slips():
    """The comparison above is sharp: dropping the pressure jump from S* or
    the eta2 term from d changes the flux by far more than the tolerance."""
    UL, UR = prim(1.0, 0.3, 0.2, 2.0), prim(0.5, -0.4, -0.1, 0.6)
    F = oracle.hllc_flux(UL, UR)
    assert np.abs(F - hllc_textbook(UL, UR)).max() < 1e-13
    # a flux built with the pressure jump dropped from S*
    SL, SR = einfeldt_speeds(UL, UR)
    rl, ul, vl, pl = cons2prim(UL)
    rr, ur, vr, pr = cons2prim(UR)
    Ss_bad = (rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))
    assert abs(Ss_bad - Ss) > 0.1


@pytest.mark.parametrize("seed", range(4))
def test_roe_supersonic_with_shear_is_the_upwind_flux(seed):
    """DNDS_Gas.hpp:405-603: when every Roe wave speed exceeds the
    Harten-Yee threshold max(0.05 (|u~| + a~), dLambda) and has one sign,
    F = (FL + FR)/2 - R |Lambda| alpha / 2 = FL (or FR), because
    R Lambda alpha = FR - FL (the Roe property).  Both states carry
    different tangential velocities, so the shear wave's strength and
    eigenvector column enter."""
    rng = np.random.default_rng(seed)
    for _ in range(10):
        rl, rr = rng.uniform(0.5, 2, 2)
        pl, pr = rng.uniform(0.5, 2, 2)
        u = rng.uniform(3.0, 5.0) * np.sqrt(G * max(pl / rl, pr / rr))
        vl, vr = rng.uniform(-2, 2, 2)
        UL, UR = prim(rl, u, vl, pl), prim(rr, 1.1 * u, vr, pr)
        F = oracle.roe_flux(UL, UR, G, 0.0)
        assert rel_err(F, xflux(UL)) <= 1e-13, (F, xflux(UL))
        Fn = oracle.roe_flux(prim(rl, -1.1 * u, vl, pl), prim(rr, -u, vr, pr), G, 0.0)
        assert rel_err(Fn, xflux(prim(rr, -u, vr, pr))) <= 1e-13


@pytest.mark.parametrize("seed", range(4))
def test_roe_contact_and_shear_jump_is_central_minus_upwinded_jump(seed):
    """Equal pressure and normal velocity, different density and tangential
    velocity: only the linearly degenerate waves carry the jump (acoustic
    strengths are 0), so F = (FL + FR)/2 - |u| (UR - UL)/2 exactly, for any
    Harten-Yee threshold (the fix only touches the acoustic speeds)."""
    rng = np.random.default_rng(10 + seed)
    for _ in range(10):
        u, p = rng.uniform(-0.8, 0.8), rng.uniform(0.5, 2)
        UL = prim(rng.uniform(0.3, 3), u, rng.uniform(-2, 2), p)
        UR = prim(rng.uniform(0.3, 3), u, rng.uniform(-2, 2), p)
        for dl in (0.0, 0.5):
            F = oracle.roe_flux(UL, UR, G, dl)
            ref = 0.5 * (xflux(UL) + xflux(UR)) - 0.5 * abs(u) * (UR - UL)
            assert np.abs(F - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0)


def test_dmr_far_is_the_mach10_oblique_shock():
    """Special_DMRFar (DNDS_FV_EulerEvaluator.hpp:1017-1038): the
    double-Mach-reflection boundary of Woodward and Colella (J. Comput. Phys.
    54, 1984).  Ahead of the shock x_s(y, t) = 1/6 + y / tan 60 + 10 t / sin 60
    the gas is at rest with rho = 1.4, p = 1; behind it the Mach-10 post-shock
    state rho = 8, velocity 8.25 (cos 30, -sin 30), p = 116.5."""
    from dnds_amd import _capi as A
    UL = prim(1.0, 0.1, 0.2, 1.0)
    pre = prim(1.4, 0.0, 0.0, 1.0)
    post = prim(8.0, 8.25 * np.cos(np.pi / 6), -8.25 * np.sin(np.pi / 6), 116.5)
    rng = np.random.default_rng(0)
    for _ in range(40):
        x, y, t = rng.uniform(0, 4), rng.uniform(0, 1), rng.uniform(0, 0.2)
        xs = 1 / 6 + y / np.tan(np.pi / 3) + 10 * t / np.sin(np.pi / 3)
        if abs(x - xs) < 1e-9:
            continue
        ub = oracle.boundary_value(UL, [0.0, -1.0], [x, y], t, A.BC_SPECIAL_DMRFAR)
        ref = pre if x > xs else post
        assert rel_err(ub, ref) <= 1e-14, (x, y, t, ub, ref)


def test_rt_far_is_the_rayleigh_taylor_layer():
    """Special_RTFar (DNDS_FV_EulerEvaluator.hpp:1039-1098, whose last line
    overrides the characteristic switch): the ghost is the far state of the
    Rayleigh-Taylor setup, heavy gas (rho 2, p 1) below y = 0.5, light gas
    (rho 1, p 2.5) above, at rest but for the seeded normal velocity
    v = -0.025 a cos(8 pi x), with a the sound speed of the interior state."""
    from dnds_amd import _capi as A
    rng = np.random.default_rng(1)
    for _ in range(30):
        UL = prim(rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 3))
        x, y = rng.uniform(0, 0.25), rng.uniform(0, 1)
        n = rng.standard_normal(2)
        n /= np.linalg.norm(n)
        ub = oracle.boundary_value(UL, n, [x, y], 0.3, A.BC_SPECIAL_RTFAR)
        rho, _, _, p = cons2prim(UL)
        a = np.sqrt(G * p / rho)
        v = -0.025 * a * np.cos(8 * np.pi * x)
        ref = prim(2.0, 0.0, v, 1.0) if y < 0.5 else prim(1.0, 0.0, v, 2.5)
        assert rel_err(ub, ref) <= 1e-14


def test_iv_far_is_the_exact_vortex_anywhere():
    """Special_IVFar (DNDS_FV_EulerEvaluator.hpp:1100-1125) at generic points
    and times: Shu's isentropic vortex (strength 5) centred at (5 + t, 5 + t),
    T = 1 - (gamma - 1) chi^2 / (8 gamma pi^2) e^(1 - r^2), velocity
    (1, 1) + chi / (2 pi) e^((1 - r^2)/2) (-(y - yc), x - xc), rho = T^(1/(gamma-1)), p = rho T."""
    from dnds_amd import _capi as A
    rng = np.random.default_rng(2)
    chi = 5.0
    for _ in range(30):
        x, y, t = rng.uniform(2, 9), rng.uniform(2, 9), rng.uniform(0, 1)
        xc = yc = 5 + t
        r2 = (x - xc) ** 2 + (y - yc) ** 2
        T = 1 - (G - 1) * chi ** 2 / (8 * G * np.pi ** 2) * np.exp(1 - r2)
        f = chi / (2 * np.pi) * np.exp((1 - r2) / 2)
        rho = T ** (1 / (G - 1))
        ref = prim(rho, 1 - f * (y - yc), 1 + f * (x - xc), rho * T)
        ub = oracle.boundary_value(prim(1, 0, 0, 1), [1.0, 0.0], [x, y], t, A.BC_SPECIAL_IVFAR)
        assert rel_err(ub, ref) <= 1e-13


def fallback_state(mesh, vfv, cell, seed=0, scale=1e6):
    """A smooth random state whose `cell` carries huge first-order momentum
    coefficients: at every face Gauss point of that cell the momentum
    (m_x + c phi_x, m_y + c phi_y) has kinetic energy ~ c^2 |x_g - x_c|^2 /
    (2 rho s^2), far above E, since no Gauss point sits at the barycentre;
    the density and energy coefficients stay zero."""
    rng = np.random.default_rng(seed)
    n = mesh.n_cell
    u = np.stack([prim(rng.uniform(0.9, 1.1), rng.uniform(0.2, 0.4), rng.uniform(-0.2, 0.2), rng.uniform(0.9, 1.1))
                  for _ in range(n)])
    urec = 0.01 * rng.standard_normal((n, 36))
    bad = urec.copy()
    bad[cell] = 0.0
    bad[cell, 0 + 9 * 1] = scale   # d/dx of rho u
    bad[cell, 1 + 9 * 2] = scale   # d/dy of rho v
    bad[cell, 0 + 9 * 2] = scale   # d/dx of rho v
    bad[cell, 1 + 9 * 1] = -scale  # d/dy of rho u
    zeroed = urec.copy()
    zeroed[cell] = 0.0
    return u, bad, zeroed


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("riemann", [1, 2])
def test_compress_rec_part_falls_back_to_the_cell_mean(kind, riemann):
    """CompressRecPart (DNDS_FV_EulerEvaluator.hpp:1277-1360): a face value
    with e < 0 (or rho < 0) is replaced by the cell mean.  When that happens
    at every Gauss point of a cell, the cell's reconstruction drops out of
    EvaluateRHS entirely: the right-hand side equals the one with that
    cell's coefficients set to zero (whose face values are the mean,
    umean + 0), on the cell and on all its neighbours, bit for bit."""
    mesh, vfv = mesh_and_setup(kind=kind, nx=7, ny=6, periodic=True, jitter=0.0, seed=4)
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=riemann, far_field=(1.0, 0.3, 0.1, 2.6))
    cell = mesh.n_cell // 2
    u, bad, zeroed = fallback_state(mesh, vfv, cell)
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    r_bad = orc.evaluate_rhs(u, bad, dlf)
    r_zero = orc.evaluate_rhs(u, zeroed, dlf)
    assert np.array_equal(r_bad, r_zero)
    # and the fallback is what made them equal: with moderate coefficients
    # (no fallback) the same cell changes the right-hand side
    _, mild, _ = fallback_state(mesh, vfv, cell, scale=1e-2)
    assert not np.array_equal(orc.evaluate_rhs(u, mild, dlf), r_zero)
