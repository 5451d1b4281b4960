"""The laminar viscous terms (NS_2D, muGas > 0) on the GPU against the
oracle, through the C ABI: EvaluateDt and EvaluateRHS per call, SSP-RK3 with
WBAP, and implicit-Euler iterations whose LU-SGS diagonal carries the
viscous spectral radius.  Bar: 1e-12 relative, as for the inviscid path.
"""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err
from test_gpu_parity import BC_SETS, random_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
# muGas, Pr, Rgas, TRef, CSutherland: T near 1 on Sutherland's curve, and the
# reference's TRef / CSutherland with T near 300
GASES = {"unit": (0.05, 0.72, 1.0, 1.0, 0.5), "air": (0.05, 0.72, 1.0 / 300.0, 273.15, 110.4)}
BC_SETS = dict(BC_SETS, noslip=(4, 4, 4, 4))


def build(kind, bcname, gas, riemann="Roe", nx=9, ny=7, seed=5):
    import oracle
    from dnds_amd.host import EulerEvaluator
    periodic = bcname == "periodic"
    mesh, vfv = mesh_and_setup(kind=kind, nx=nx, ny=ny, periodic=periodic, jitter=0.0 if periodic else 0.2,
                               seed=seed, bc=BC_SETS[bcname])
    far = (1.0, 0.3, -0.2, 2.6)
    mu, pr, R, tref, cs = GASES[gas]
    ev = EulerEvaluator(mesh, vfv, riemannSolverType=riemann, gamma=1.4, farFieldStaticValue=far, muGas=mu,
                        prGas=pr, Rgas=R, TRef=tref, CSutherland=cs)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=1 if riemann == "Roe" else 2, gamma=1.4,
                        far_field=far, jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std,
                        mu_gas=mu, pr_gas=pr, r_gas=R, t_ref=tref, c_sutherland=cs)
    return mesh, vfv, ev, orc


@pytest.mark.parametrize("kind,bcname", [(k, b) for k in (0, 1, 2) for b in ("periodic", "farfield", "noslip",
                                                                              "mixed_bc")])
@pytest.mark.parametrize("gas", ["unit", "air"])
@pytest.mark.parametrize("riemann", ["Roe", "HLLC"])
def test_viscous_dt_and_rhs_match_oracle(kind, bcname, gas, riemann):
    mesh, vfv, ev, orc = build(kind, bcname, gas, riemann)
    u, urec = random_state(mesh.n_cell, 21)
    du, dr, rhs, dt = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), ev.BuildMean(), ev.BuildScalar()
    dtmin = ev.EvaluateDt(dt, u=du, CFL=0.5, MaxDt=1e100, UseLocaldt=True)
    ev.EvaluateRHS(rhs, du, dr, 0.2)
    rdt, rdtmin, _, dlf = orc.evaluate_dt(u, 0.5, use_local_dt=True)
    assert rel_err(dt.download(), rdt) <= TOL and abs(dtmin - rdtmin) <= TOL * rdtmin
    ref = orc.evaluate_rhs(u, urec, dlf, t=0.2)
    assert rel_err(rhs.download(), ref) <= TOL
    # the viscous terms are present: the inviscid right-hand side differs
    orc.cfg.mu_gas = 0.0
    assert rel_err(orc.evaluate_rhs(u, urec, dlf, t=0.2), ref) > 1e-6


def test_viscous_ssprk3_with_limiter_matches_oracle():
    """Four-quadrant Riemann problem at 16 x 16 quads with muGas = 1e-3
    (TRef 1, CSutherland 0.5): 5 SSP-RK3 steps, 3 sweeps per stage, WBAP V2."""
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ExplicitSSPRK3LocalDt
    case = cases.riemann_2d(16, nthreads=4)
    gas = dict(muGas=1e-3, prGas=0.72, Rgas=1.0, TRef=1.0, CSutherland=0.5)
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType="HLLC", farFieldStaticValue=case.far, **gas)
    st = case.vfv.setting
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=2, far_field=case.far,
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, mu_gas=1e-3, pr_gas=0.72,
                        r_gas=1.0, t_ref=1.0, c_sutherland=0.5)
    urec0 = np.zeros((case.n_cell, 36))
    du, dr, dn = ev.BuildMean().upload(case.u0), case.vfv.BuildRec().upload(urec0), case.vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=3, CFL=0.5, useLimiter=True)
    t = 0.0
    for _ in range(5):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(case.u0, urec0, 5, 3, 0.5, use_limiter=True)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL


def test_viscous_implicit_iterations_match_oracle_in_colour_order():
    from dnds_amd.host import ImplicitEulerDualTimeStep
    mesh, vfv, ev, orc = build(2, "noslip", "unit", "Roe", nx=10, ny=8)
    u0, _ = random_state(mesh.n_cell, 33, amp=0.1)
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=2, CFL=5.0, nTimeStepInternal=4,
                                    rhsThresholdInternal=0.0)
    u, r, rn = ev.BuildMean().upload(u0), vfv.BuildRec().setConstant(0.0), vfv.BuildRec()
    ode.Step(u, r, rn, dt=0.1)
    ou, orr, _ = orc.implicit_euler(u0, np.zeros((mesh.n_cell, 36)), 4, 2, 5.0, dt=0.1, threshold=0.0,
                                    scan2cell=ev.LUSGSScanOrder())
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL


def test_mixed_step_refuses_the_viscous_model():
    from dnds_amd import _capi as A
    from dnds_amd.host import DNDSError, ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc = build(0, "farfield", "unit")
    u, urec = random_state(mesh.n_cell, 3)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    with pytest.raises(DNDSError) as e:
        ExplicitSSPRK3LocalDt(ev, nInternalRecStep=3, CFL=0.5, mixedPrecision=True).Step(du, dr, dn, 0.0)
    assert e.value.code == A.ERR_UNSUPPORTED and "muGas" in str(e.value)
