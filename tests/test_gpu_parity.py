"""GPU (HIP) vs CPU oracle parity of the FP64 hot path, through the C ABI.

Bar: per-call Jacobi sweeps, time steps and right-hand sides agree with the
oracle to <= 1e-12 relative (max abs difference over the array / max abs of
the oracle's array); SSP-RK3 trajectories to the same bar after several
steps.  Both sides consume the same host setup arrays.
"""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12

BC_SETS = {
    "periodic": None,
    "farfield": (2, 2, 2, 2),
    "walls": (3, 3, 3, 3),          # Wall_Euler everywhere
    "mixed_bc": (2, 3, 4, 3),       # Farfield, Wall_Euler, Wall_NoSlip, Wall_Euler
    "with_wall": (2, 3, 4, 1),      # ... and a Wall face (valid for the sweep only)
    "riemann": (-5, -5, -5, -5),    # Special_2DRiemannFar
    "ivfar": (-4, -4, -4, -4),      # Special_IVFar (exact vortex), on [0,10]^2
    "dmrfar": (-2, -2, -2, -2),     # Special_DMRFar (Mach-10 shock ghost), on [0,4]x[0,1]
    "rtfar": (-3, -3, -3, -3),      # Special_RTFar (Rayleigh-Taylor far states)
}

# domain (x1, y1) of the special far-field sets: the vortex centre (5, 5)
# inside the domain, the double-Mach shock line crossing the bottom edge
DOMAIN = {"ivfar": (10.0, 10.0), "dmrfar": (4.0, 1.0)}


RS_CODE = {"Roe": 1, "HLLC": 2, "HLLEP": 3, "Roe_M1": 11, "Roe_M2": 12, "Roe_M3": 13, "Roe_M4": 14, "Roe_M5": 15}


def build(kind, bcname, riemann="Roe", nx=9, ny=7, jitter=0.2, seed=5):
    import oracle
    from dnds_amd.host import EulerEvaluator
    periodic = bcname == "periodic"
    x1, y1 = DOMAIN.get(bcname, (1.0, 1.0))
    mesh, vfv = mesh_and_setup(kind=kind, nx=nx, ny=ny, periodic=periodic, jitter=0.0 if periodic else jitter,
                               seed=seed, bc=BC_SETS[bcname], x1=x1, y1=y1)
    far = (1.0, 0.3, -0.2, 2.6)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType=riemann, gamma=1.4, farFieldStaticValue=far)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=RS_CODE[riemann], gamma=1.4,
                        far_field=far, jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std)
    return mesh, vfv, ev, orc


def random_state(n, seed, amp=0.2, rec_amp=0.05):
    rng = np.random.default_rng(seed)
    rho = 1.0 + amp * rng.uniform(-1, 1, n)
    vx = 0.5 + amp * rng.uniform(-1, 1, n)
    vy = -0.3 + amp * rng.uniform(-1, 1, n)
    p = 1.0 + amp * rng.uniform(-1, 1, n)
    E = p / 0.4 + 0.5 * rho * (vx ** 2 + vy ** 2)
    u = np.stack([rho, rho * vx, rho * vy, E], axis=1)
    urec = rec_amp * rng.standard_normal((n, 36))
    return u, urec


CASES = ([(k, b) for k in (0, 1, 2) for b in ("periodic", "farfield", "walls", "mixed_bc")] + [(0, "riemann")]
         + [(k, b) for k, b in ((0, "ivfar"), (2, "ivfar"), (0, "dmrfar"), (1, "dmrfar"), (2, "rtfar"), (0, "rtfar"))])


@pytest.mark.parametrize("kind,bcname", CASES + [(2, "with_wall")])
def test_jacobi_sweeps_match_oracle(kind, bcname):
    mesh, vfv, ev, orc = build(kind, bcname)
    u, urec = random_state(mesh.n_cell, 11)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ref = urec
    for sweep in range(4):
        vfv.ReconstructionJacobiStep(du, dr, dn, t=0.0)
        ref = orc.jacobi_step(u, ref)
        got = dr.download()
        assert rel_err(got, ref) <= TOL, (sweep, rel_err(got, ref))
    ev.device.sync()


@pytest.mark.parametrize("kind,bcname", CASES)
@pytest.mark.parametrize("local", [False, True])
def test_evaluate_dt_matches_oracle(kind, bcname, local):
    mesh, vfv, ev, orc = build(kind, bcname)
    u, _ = random_state(mesh.n_cell, 12)
    du, dt = ev.BuildMean().upload(u), ev.BuildScalar()
    dtmin = ev.EvaluateDt(dt, du, CFL=0.7, MaxDt=1e100, UseLocaldt=local)
    rdt, rmin, _, _ = orc.evaluate_dt(u, 0.7, 1e100, local)
    assert abs(dtmin - rmin) <= TOL * abs(rmin)
    assert rel_err(dt.download(), rdt) <= TOL


@pytest.mark.parametrize("kind,bcname", CASES)
@pytest.mark.parametrize("riemann", ["Roe", "HLLC", "HLLEP", "Roe_M1", "Roe_M2", "Roe_M3", "Roe_M4", "Roe_M5"])
def test_evaluate_rhs_matches_oracle(kind, bcname, riemann):
    mesh, vfv, ev, orc = build(kind, bcname, riemann=riemann)
    u, urec = random_state(mesh.n_cell, 13)
    du, dr, rhs, dt = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), ev.BuildMean(), ev.BuildScalar()
    ev.EvaluateDt(dt, du, CFL=0.5)            # refreshes deltaLambdaFace
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    ev.EvaluateRHS(rhs, du, dr, t=0.05)
    ref = orc.evaluate_rhs(u, urec, dlf, t=0.05)
    assert rel_err(rhs.download(), ref) <= TOL


@pytest.mark.parametrize("kind,bcname,riemann", [(0, "periodic", "Roe"), (2, "farfield", "HLLC"),
                                                 (1, "walls", "Roe"), (0, "riemann", "HLLC"),
                                                 (2, "ivfar", "Roe"), (0, "rtfar", "HLLC"),
                                                 (0, "riemann", "HLLEP"), (1, "walls", "HLLEP"),
                                                 (2, "farfield", "Roe_M4"), (0, "periodic", "Roe_M1")])
def test_ssprk3_steps_match_oracle(kind, bcname, riemann):
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc = build(kind, bcname, riemann=riemann)
    u, urec = random_state(mesh.n_cell, 14, amp=0.1, rec_amp=0.0)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=2, CFL=0.4)
    t = 0.0
    for _ in range(5):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(u, urec, 5, 2, 0.4)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("riemann", ["Roe", "HLLC", "HLLEP"])
def test_positivity_fallback_matches_oracle(kind, riemann):
    """CompressRecPart's fallback to the cell mean (DNDS_FV_EulerEvaluator.hpp:
    1322-1351) forced at every face Gauss point of one cell (huge momentum
    coefficients, e < 0 there): the device right-hand side equals the
    oracle's, and equals the one with that cell's coefficients zeroed."""
    from test_oracle_pins import fallback_state
    mesh, vfv, ev, orc = build(kind, "farfield", riemann=riemann)
    cell = mesh.n_cell // 2
    u, bad, zeroed = fallback_state(mesh, vfv, cell)
    du, dt, rhs = ev.BuildMean().upload(u), ev.BuildScalar(), ev.BuildMean()
    ev.EvaluateDt(dt, du, CFL=0.5)
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    out = []
    for rec in (bad, zeroed):
        ev.EvaluateRHS(rhs, du, vfv.BuildRec().upload(rec), t=0.0)
        got = rhs.download()
        assert rel_err(got, orc.evaluate_rhs(u, rec, dlf)) <= TOL
        out.append(got)
    assert np.array_equal(out[0], out[1])


def test_vortex_config_full_run_matches_oracle():
    """The small CPU-runnable configuration end to end: periodic 40x40 split
    quads (3200 cells), p = 3, Roe, CFL 0.5, 5 sweeps per stage, 100 steps."""
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ExplicitSSPRK3LocalDt
    case = cases.isentropic_vortex(40, seed=1)
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType="Roe", farFieldStaticValue=case.far)
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=1, far_field=case.far)
    urec0 = np.zeros((case.n_cell, 36))
    du, dr, dn = ev.BuildMean().upload(case.u0), case.vfv.BuildRec().upload(urec0), case.vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=case.sweeps, CFL=case.cfl)
    t = 0.0
    for _ in range(case.steps):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(case.u0, urec0, case.steps, case.sweeps, case.cfl)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL


def test_wall_face_in_rhs_is_an_error_on_both_sides():
    """A `Wall` face is not a flux boundary (the reference asserts,
    DNDS_FV_EulerEvaluator.hpp:1265-1269): both implementations report it."""
    from dnds_amd.host import DNDSError
    mesh, vfv, ev, orc = build(0, "with_wall")
    u, urec = random_state(mesh.n_cell, 15)
    du, dr, rhs, dt = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), ev.BuildMean(), ev.BuildScalar()
    ev.EvaluateDt(dt, du, CFL=0.5)
    ev.EvaluateRHS(rhs, du, dr, t=0.0)
    with pytest.raises(DNDSError):
        ev.device.sync()
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    with pytest.raises(RuntimeError):
        orc.evaluate_rhs(u, urec, dlf)


@pytest.mark.parametrize("kind,bcname", [(0, "periodic"), (2, "farfield"), (1, "walls"), (0, "riemann"),
                                         (2, "mixed_bc")])
@pytest.mark.parametrize("if_all", [False, True])
@pytest.mark.parametrize("version", ["V3", "V2"])
def test_wbap_matches_oracle(kind, bcname, if_all, version):
    """ReconstructionWBAPLimitFacialV3 / V2: smoothness indicator and the
    limited reconstruction (limited and unlimited cells both present)."""
    mesh, vfv, ev, orc = build(kind, bcname)
    # smooth means everywhere, rough reconstructions only where x > 0.5, so
    # both limited and unlimited cells occur
    bary = vfv.arrays(["cell_bary"])["cell_bary"].reshape(-1, 2)
    _, urec = random_state(mesh.n_cell, 16, rec_amp=0.3)
    rho = 1.0 + 0.1 * np.sin(3 * bary[:, 0]) * np.cos(2 * bary[:, 1])
    vx, vy, p = 0.4 + 0.1 * bary[:, 1], -0.2 + 0.05 * bary[:, 0], 1.0 + 0.1 * bary[:, 0] * bary[:, 1]
    u = np.stack([rho, rho * vx, rho * vy, p / 0.4 + 0.5 * rho * (vx ** 2 + vy ** 2)], axis=1)
    urec[bary[:, 0] < 0.5] = 0.0
    du, dr, dn, ind = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec(), ev.BuildScalar()
    getattr(vfv, "ReconstructionWBAPLimitFacial" + version)(du, dr, dn, None, ind, ifAll=if_all)
    ref, rind = getattr(orc, "wbap_" + version.lower())(u, urec, if_all)
    assert rel_err(ind.download(), rind) <= TOL
    got = dn.download()
    assert rel_err(got, ref) <= TOL
    assert rel_err(dr.download(), urec) == 0.0  # uRec is read only
    limited = rind >= vfv.setting.wbap_smooth_indicator_scale
    if not if_all:
        assert 0 < limited.sum() < mesh.n_cell
        assert np.array_equal(got[~limited], urec[~limited])


@pytest.mark.parametrize("procedure", [0, 1])
def test_ssprk3_with_limiter_matches_oracle_on_riemann(procedure):
    """The four-quadrant Riemann case at 24 x 24 quads: SSP-RK3 with 5 sweeps,
    WBAP (limiterProcedure 0: V2, 1: V3) and HLLC for 10 steps against the
    oracle."""
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ExplicitSSPRK3LocalDt
    case = cases.riemann_2d(24, nthreads=4)
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType="HLLC", farFieldStaticValue=case.far)
    st = case.vfv.setting
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=2, far_field=case.far,
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, limiter_procedure=procedure)
    urec0 = np.zeros((case.n_cell, 36))
    du, dr, dn = ev.BuildMean().upload(case.u0), case.vfv.BuildRec().upload(urec0), case.vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=5, CFL=0.5, useLimiter=True, limiterProcedure=procedure)
    t = 0.0
    for _ in range(10):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(case.u0, urec0, 10, 5, 0.5, use_limiter=True)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL


def test_limited_trajectory_stays_bitwise_on_the_oracle():
    """The four-quadrant Riemann problem at 64 x 64 quads, 100 SSP-RK3 steps
    with 5 sweeps and WBAP V2 per stage, compared with the oracle after every
    20 steps: the limiter's threshold decisions could amplify any rounding
    difference, and there is none (both compute in the same order)."""
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ExplicitSSPRK3LocalDt
    case = cases.riemann_2d(64, nthreads=8)
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType="HLLC", farFieldStaticValue=case.far)
    st = case.vfv.setting
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=2, far_field=case.far,
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale)
    oracle.set_threads(8)
    z = np.zeros((case.n_cell, 36))
    du, dr, dn = ev.BuildMean().upload(case.u0), case.vfv.BuildRec().upload(z), case.vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=5, CFL=0.5, useLimiter=True)
    ou, orr, t, ot = case.u0.copy(), z.copy(), 0.0, 0.0
    for k in range(5):
        for _ in range(20):
            t += ode.Step(du, dr, dn, t)
        ou, orr, ot = orc.ssprk3_steps(ou, orr, 20, 5, 0.5, use_limiter=True, t=ot)
        assert t == ot
        assert np.array_equal(du.download(), ou) and np.array_equal(dr.download(), orr)
