"""The CPU oracle against the reference's printed results and against what
the mathematics fixes (consistency, conservation, freestream, limiter
identities, convergence of the full solver).  CPU only."""
import numpy as np
import pytest

import oracle
from conftest import mesh_and_setup, rel_err


def prim(rho, u, v, p, gamma=1.4):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def xflux(U, gamma=1.4):
    rho, mu, mv, E = U
    u, v = mu / rho, mv / rho
    p = (gamma - 1) * (E - 0.5 * rho * (u * u + v * v))
    return np.array([mu, mu * u + p, mu * v, (E + p) * u])


def test_roe_flux_reproduces_reference_printout():
    """test/testGas.cpp:23-27 calls the Roe flux on UL = (1, 0.5916, 0, 0,
    2.675), UR = (1, 0, 0, 0, 2.5), gamma 1.4, and records the result
    0.369291683630373 1.54891136111009 0 1.45489148963839 (testGas.cpp:68)."""
    F = oracle.roe_flux([1.0, 0.5916, 0.0, 2.675], [1.0, 0.0, 0.0, 2.5], 1.4, 0.0)
    ref = np.array([0.369291683630373, 1.54891136111009, 0.0, 1.45489148963839])
    assert np.abs(F - ref).max() <= 5e-15


@pytest.mark.parametrize("seed", range(5))
def test_fluxes_are_consistent(seed):
    rng = np.random.default_rng(seed)
    U = prim(rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.1, 5))
    for fn in (oracle.roe_flux, oracle.hllc_flux):
        assert rel_err(fn(U, U), xflux(U)) <= 1e-14


def test_supersonic_hllc_is_upwind():
    UL = prim(1.0, 3.0, 0.2, 1.0)
    UR = prim(0.5, 2.5, -0.1, 0.8)
    assert rel_err(oracle.hllc_flux(UL, UR), xflux(UL)) == 0.0
    assert rel_err(oracle.hllc_flux(prim(1.0, -3.0, 0, 1.0), prim(0.5, -3.5, 0, 0.8)),
                   xflux(prim(0.5, -3.5, 0, 0.8))) == 0.0


@pytest.mark.parametrize("seed", range(4))
def test_eigenvectors_invert(seed):
    rng = np.random.default_rng(seed)
    U = prim(rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.1, 5))
    L, R = oracle.eigenvectors(U)
    assert np.abs(L @ R - np.eye(4)).max() < 1e-13


def build(kind=0, periodic=True, bc=None, riemann=1, nx=8, ny=6):
    mesh, vfv = mesh_and_setup(kind=kind, nx=nx, ny=ny, periodic=periodic, jitter=0.0 if periodic else 0.2,
                               seed=2, bc=bc)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=riemann, far_field=(1.0, 0.3, 0.1, 2.6),
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale)
    return mesh, vfv, orc


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("riemann", [1, 2])
def test_freestream_is_preserved(kind, riemann):
    mesh, vfv, orc = build(kind, riemann=riemann)
    u = np.tile(prim(1.2, 0.7, -0.3, 0.9), (mesh.n_cell, 1))
    urec = orc.jacobi_step(u, np.zeros((mesh.n_cell, 36)))
    assert np.abs(urec).max() == 0.0
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    rhs = orc.evaluate_rhs(u, urec, dlf)
    assert np.abs(rhs).max() < 1e-13


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("riemann", [1, 2])
def test_rhs_conserves_on_periodic_meshes(kind, riemann):
    mesh, vfv, orc = build(kind, riemann=riemann)
    rng = np.random.default_rng(1)
    n = mesh.n_cell
    u = np.stack([prim(rng.uniform(0.8, 1.2), rng.uniform(0, 0.5), rng.uniform(-0.3, 0.3), rng.uniform(0.8, 1.2))
                  for _ in range(n)])
    urec = 0.02 * rng.standard_normal((n, 36))
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    rhs = orc.evaluate_rhs(u, urec, dlf)
    vol = vfv.arrays(["volume"])["volume"]
    total = (rhs * vol[:, None]).sum(0)
    assert np.abs(total).max() < 1e-13 * np.abs(rhs * vol[:, None]).sum()


def test_wbap_leaves_smooth_fields_alone_and_bounds_rough_ones():
    mesh, vfv, orc = build(0, periodic=True, nx=24, ny=24)
    st = vfv.arrays()
    bary = st["cell_bary"].reshape(-1, 2)
    u = np.stack([prim(1 + 0.01 * np.sin(2 * np.pi * x), 0.5, 0.2, 1.0) for x, _ in bary])
    urec = np.zeros((mesh.n_cell, 36))
    for _ in range(30):
        urec = orc.jacobi_step(u, urec)
    lim, ind = orc.wbap_v3(u, urec)
    assert np.all(ind < 0.1) and np.array_equal(lim, urec)
    # a jump: limited coefficients do not exceed the unlimited ones in size
    u2 = u.copy()
    u2[bary[:, 0] > 0.5] = prim(0.3, 0.5, 0.2, 0.2)
    urec2 = np.zeros_like(urec)
    for _ in range(30):
        urec2 = orc.jacobi_step(u2, urec2)
    lim2, ind2 = orc.wbap_v3(u2, urec2)
    assert (ind2 >= 0.1).sum() > 0
    assert np.abs(lim2).max() <= np.abs(urec2).max() * (1 + 1e-12)


def test_vortex_solver_converges():
    """Periodic isentropic vortex (the small CPU-runnable configuration's
    case) to t = 0.3: the L2 density error against the exact solution falls
    by at least 2^3 per mesh halving."""
    from dnds_amd import cases
    errs = []
    for n in (10, 20):
        c = cases.isentropic_vortex(n, seed=1, nthreads=4)
        orc = oracle.Oracle(c.mesh.arrays(), c.vfv.arrays(), riemann_solver=1, far_field=c.far)
        u, urec, t = c.u0.copy(), np.zeros((c.n_cell, 36)), 0.0
        dt_fixed = 0.3 / (3 * n // 2)
        # advance with the solver's steps until t >= 0.3 (CFL 0.5 global steps)
        while t < 0.3 - 1e-12:
            u, urec, t = orc.ssprk3_steps(u, urec, 1, 5, 0.5, t=t, max_dt=min(dt_fixed, 0.3 - t))
        vol = c.vfv.arrays(["volume"])["volume"]
        ex = cases.cell_means(c.mesh, vol, lambda x, y: cases.vortex_exact(x, y, t))
        errs.append(np.sqrt(np.sum((u[:, 0] - ex[:, 0]) ** 2 * vol) / vol.sum()))
    assert errs[0] / errs[1] > 8, errs


def _cons2prim(U, gamma=1.4):
    rho, mu, mv, E = U
    u, v = mu / rho, mv / rho
    return np.array([rho, u, v, (gamma - 1) * (E - 0.5 * rho * (u * u + v * v))])


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("periodic", [False, True])
def test_dt_of_a_uniform_state_is_the_closed_form(kind, periodic):
    """Uniform state: lambda_f = |v . n_f| + a on every face, so
    dt_i = CFL V_i / sum_f (|v . n_f| + a) |f|, with V_i (shoelace), |f| and
    n_f taken from the node coordinates, not from the setup arrays."""
    mesh, vfv, orc = build(kind, periodic=periodic, nx=7, ny=6)
    m = mesh.arrays()
    U = prim(1.3, 0.6, -0.4, 0.8)
    u = np.tile(U, (mesh.n_cell, 1))
    cfl = 0.7
    dt, dtmin, lf, dlf = orc.evaluate_dt(u, cfl, use_local_dt=True)
    a = np.sqrt(1.4 * 0.8 / 1.3)
    xy, f2n = m["node_xy"], m["face2node"]
    e = xy[f2n[:, 1]] - xy[f2n[:, 0]]
    length = np.linalg.norm(e, axis=1)
    nrm = np.stack([e[:, 1], -e[:, 0]], 1) / length[:, None]
    lam = np.abs(nrm @ np.array([0.6, -0.4])) + a
    assert rel_err(lf, lam) <= 1e-14
    assert np.all(dlf == 0.0)
    cp, c2n, c2f = m["cell_ptr"], m["cell2node"], m["cell2face"]
    for i in range(mesh.n_cell):
        X = xy[c2n[cp[i]:cp[i + 1]]]
        vol = 0.5 * abs(np.sum(X[:, 0] * np.roll(X[:, 1], -1) - np.roll(X[:, 0], -1) * X[:, 1]))
        fs = c2f[cp[i]:cp[i + 1]]
        ref = cfl * vol / np.sum(lam[fs] * length[fs])
        assert abs(dt[i] - ref) <= 1e-13 * ref
    assert dtmin == dt.min()
    dtg, dtmin_g, _, _ = orc.evaluate_dt(u, cfl, use_local_dt=False)
    assert np.all(dtg == dtmin)
    # MaxDt caps every cell
    dtc, _, _, _ = orc.evaluate_dt(u, cfl, max_dt=0.5 * dtmin, use_local_dt=True)
    assert np.all(dtc == 0.5 * dtmin)


def test_farfield_boundary_follows_the_characteristics():
    """Farfield (DNDS_FV_EulerEvaluator.hpp:977-1016): supersonic outflow takes
    the interior state, supersonic inflow the far state, subsonic outflow the
    interior primitives with the far pressure, subsonic inflow the far
    primitives with the interior pressure."""
    from dnds_amd import _capi as A
    far = prim(1.0, 0.2, 0.1, 1.0)
    n = np.array([0.6, 0.8])
    cases_ = {
        "sup_out": prim(1.2, 3.0 * 0.6, 3.0 * 0.8, 1.0),
        "sup_in": prim(1.2, -3.0 * 0.6, -3.0 * 0.8, 1.0),
        "sub_out": prim(1.2, 0.3 * 0.6, 0.3 * 0.8, 0.9),
        "sub_in": prim(1.2, -0.3 * 0.6, -0.3 * 0.8, 0.9),
    }
    ub = {k: oracle.boundary_value(v, n, [0.3, 0.4], 0.0, A.BC_FARFIELD, far_field=far) for k, v in cases_.items()}
    assert np.array_equal(ub["sup_out"], cases_["sup_out"])
    assert np.array_equal(ub["sup_in"], far)
    pl, pf = _cons2prim(cases_["sub_out"]), _cons2prim(far)
    assert rel_err(_cons2prim(ub["sub_out"]), np.r_[pl[:3], pf[3]]) <= 1e-14
    pl = _cons2prim(cases_["sub_in"])
    assert rel_err(_cons2prim(ub["sub_in"]), np.r_[pf[:3], pl[3]]) <= 1e-14


def test_wall_boundaries():
    """Wall_Euler removes the normal momentum and keeps the rest
    (DNDS_FV_EulerEvaluator.hpp:1247-1251); NoSlip reverses the momentum
    (:1252-1264); Wall is not a flux boundary."""
    from dnds_amd import _capi as A
    UL = prim(1.1, 0.4, -0.7, 0.9)
    n = np.array([-0.28, 0.96])
    ue = oracle.boundary_value(UL, n, [0, 0], 0.0, A.BC_WALL_EULER)
    assert abs(ue[1:3] @ n) < 1e-15
    t = np.array([n[1], -n[0]])
    assert abs(ue[1:3] @ t - UL[1:3] @ t) < 1e-15
    assert ue[0] == UL[0] and ue[3] == UL[3]
    un = oracle.boundary_value(UL, n, [0, 0], 0.0, A.BC_WALL_NOSLIP)
    assert np.array_equal(un, UL * np.array([1, -1, -1, 1]))
    with pytest.raises(RuntimeError):
        oracle.boundary_value(UL, n, [0, 0], 0.0, A.BC_WALL)


def test_riemann_and_vortex_far_states():
    """2DRiemannFar (DNDS_FV_EulerEvaluator.hpp:1126-1243): the left and
    bottom edges carry the quadrant states of the four-quadrant problem; the
    right and top edges mirror the interior.  IVFar (:1100-1125) is the exact
    isentropic vortex: freestream (1, 1, 1, 1) far from the centre and the
    closed-form temperature dip at the centre."""
    from dnds_amd import _capi as A
    UL = prim(0.9, 0.3, 0.2, 0.7)
    n = np.array([-1.0, 0.0])
    q3 = _cons2prim(oracle.boundary_value(UL, n, [0.0, 0.1], 0.0, A.BC_SPECIAL_2DRIEMANNFAR))
    assert rel_err(q3, [0.137992831541219, 1.206045378311055, 1.206045378311055, 0.029032258064516]) <= 1e-14
    q2 = _cons2prim(oracle.boundary_value(UL, n, [0.0, 0.95], 0.0, A.BC_SPECIAL_2DRIEMANNFAR))
    assert rel_err(q2, [0.532258064516129, 1.206045378311055, 0.0, 0.3]) <= 1e-14
    q4 = _cons2prim(oracle.boundary_value(UL, [0, -1.0], [0.95, 0.0], 0.0, A.BC_SPECIAL_2DRIEMANNFAR))
    assert rel_err(q4, [0.532258064516129, 0.0, 1.206045378311055, 0.3]) <= 1e-14
    pr = _cons2prim(oracle.boundary_value(UL, [1.0, 0], [1.0, 0.5], 0.0, A.BC_SPECIAL_2DRIEMANNFAR))
    assert rel_err(pr, [0.9, -0.3, 0.2, 0.7]) <= 1e-14
    far = oracle.boundary_value(UL, n, [60.0, 70.0], 0.0, A.BC_SPECIAL_IVFAR)
    assert rel_err(far, prim(1.0, 1.0, 1.0, 1.0)) <= 1e-14
    g, chi = 1.4, 5.0
    T = 1 - (g - 1) / (8 * g * np.pi ** 2) * chi ** 2 * np.e
    rho = T ** (1 / (g - 1))
    ctr = oracle.boundary_value(UL, n, [5.5, 5.5], 0.5, A.BC_SPECIAL_IVFAR)
    assert rel_err(_cons2prim(ctr), [rho, 1.0, 1.0, rho * T]) <= 1e-13


def test_threaded_oracle_is_bitwise_serial():
    """The worker threads only split independent per-cell / per-face work;
    the scatters stay serial, so results are identical."""
    from dnds_amd import cases
    case = cases.riemann_2d(64, nthreads=4)  # 8192 cells: 4 worker threads
    st = case.vfv.setting
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=2, far_field=case.far,
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale)
    z = np.zeros((case.n_cell, 36))
    out = []
    for threads in (1, 4):
        oracle.set_threads(threads)
        out.append(orc.ssprk3_steps(case.u0, z, 2, 3, 0.5, use_limiter=True))
    oracle.set_threads(1)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
