"""Multicolour LU-SGS on the GPU against the oracle, through the C ABI.

* One forward and one backward sweep: the device sweeps colour by colour;
  the oracle runs the sequential sweeps of DNDS_FV_EulerEvaluator.cpp:1149-1357
  in the device's colour-major scan order.  Bar: 1e-12 relative.
* A fixed number of implicit-Euler dual-time iterations, same order: 1e-12.
* The converged implicit step against the oracle in the reference's own
  (natural) scan order: the iteration paths differ, the fixed point does not.
  Bar: 1e-8 in the relative L2 norm of the state.
"""
import numpy as np
import pytest

from conftest import rel_err
from test_gpu_parity import CASES, build, random_state

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _colouring_is_proper(mesh, col):
    m = mesh.arrays()
    f2c = m["face2cell"]
    inner = f2c[:, 1] >= 0
    return np.all(col[f2c[inner, 0]] != col[f2c[inner, 1]])


@pytest.mark.parametrize("kind,bcname", CASES)
def test_lusgs_sweeps_match_oracle_in_colour_order(kind, bcname):
    mesh, vfv, ev, orc = build(kind, bcname)
    n = mesh.n_cell
    u, _ = random_state(n, seed=11)
    col, ncol = ev.LUSGSColouring()
    assert _colouring_is_proper(mesh, col) and ncol <= 5
    scan = ev.LUSGSScanOrder()
    rng = np.random.default_rng(12)
    rhs = rng.standard_normal((n, 4))
    alpha, dt = 1.0, 0.7
    du, dtau = ev.BuildMean().upload(u), ev.BuildScalar()
    ev.EvaluateDt(dtau, du, 5.0, 1e100, True)
    ev.LUSGSMatrixInit(dtau, dt, alpha)
    drhs, dinc = ev.BuildMean().upload(rhs), ev.BuildMean().setConstant(0.0)
    ev.UpdateLUSGSForward(alpha, drhs, du, dinc, dinc)
    y = dinc.download()
    ev.UpdateLUSGSBackward(alpha, drhs, du, dinc, dinc)
    x = dinc.download()
    dto, _, lam, _ = orc.evaluate_dt(u, 5.0, use_local_dt=True)
    jinv = orc.lusgs_matrix_init(dto, dt, alpha, lam)
    yo = orc.lusgs_sweep(True, jinv, lam, alpha, rhs, u, np.zeros((n, 4)), scan2cell=scan)
    xo = orc.lusgs_sweep(False, jinv, lam, alpha, rhs, u, yo, scan2cell=scan)
    assert rel_err(y, yo) <= TOL
    assert rel_err(x, xo) <= TOL


def test_lusgs_distinct_buffers_match_oracle():
    mesh, vfv, ev, orc = build(2, "farfield")
    n = mesh.n_cell
    u, _ = random_state(n, seed=13)
    rng = np.random.default_rng(14)
    rhs, old = rng.standard_normal((n, 4)), rng.standard_normal((n, 4))
    du, dtau = ev.BuildMean().upload(u), ev.BuildScalar()
    ev.EvaluateDt(dtau, du, 5.0, 1e100, True)
    ev.LUSGSMatrixInit(dtau, 1e100, 1.0)
    drhs, dold, dnew = ev.BuildMean().upload(rhs), ev.BuildMean().upload(old), ev.BuildMean().setConstant(0.0)
    ev.UpdateLUSGSForward(1.0, drhs, du, dold, dnew)
    dto, _, lam, _ = orc.evaluate_dt(u, 5.0, use_local_dt=True)
    jinv = orc.lusgs_matrix_init(dto, 1e100, 1.0, lam)
    ref = orc.lusgs_sweep(True, jinv, lam, 1.0, rhs, u, old, scan2cell=ev.LUSGSScanOrder(), in_place=False)
    assert rel_err(dnew.download(), ref) <= TOL


def _vortex(n):
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator
    c = cases.isentropic_vortex(n)
    st = c.vfv.setting
    ev = EulerEvaluator(c.mesh, c.vfv, riemannSolverType="Roe", farFieldStaticValue=c.far)
    orc = oracle.Oracle(c.mesh.arrays(), c.vfv.arrays(), riemann_solver=1, far_field=c.far,
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction)
    return c, ev, orc


def test_implicit_iterations_match_oracle_in_colour_order():
    from dnds_amd.host import ImplicitEulerDualTimeStep
    c, ev, orc = _vortex(12)
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=6,
                                    rhsThresholdInternal=0.0)
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    hist = ode.Step(u, r, rn, dt=0.3)
    assert hist.shape == (6, 4)
    ou, orr, oh = orc.implicit_euler(c.u0, np.zeros((c.n_cell, 36)), 6, 2, 10.0, dt=0.3, threshold=0.0,
                                     scan2cell=ev.LUSGSScanOrder())
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL
    assert rel_err(hist, oh) <= 1e-10  # norms summed in a different order


def test_converged_implicit_step_matches_reference_order():
    """Both converged to a 1e-13 drop of the density increment."""
    from dnds_amd.host import ImplicitEulerDualTimeStep
    c, ev, orc = _vortex(10)
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=1, CFL=10.0, nTimeStepInternal=800,
                                    rhsThresholdInternal=1e-13)
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    hist = ode.Step(u, r, rn, dt=0.3)
    assert hist[-1, 0] < 1e-13 * hist[:, 0].max()
    ou, _, oh = orc.implicit_euler(c.u0, np.zeros((c.n_cell, 36)), 800, 1, 10.0, dt=0.3, threshold=1e-13)
    assert oh[-1, 0] < 1e-13 * oh[:, 0].max()
    gu = u.download()
    assert np.linalg.norm(gu - ou) / np.linalg.norm(ou) <= 1e-8
    assert np.abs(gu - c.u0).max() > 1e-3


@pytest.mark.parametrize("k", [2, 3])
def test_bdf_steps_match_oracle_in_colour_order(k):
    """ImplicitBDFDualTimeStep: four steps of a fresh order-k integrator (the
    first steps run at orders 1, 2, ... as in the reference), a fixed number
    of inner iterations each, against the oracle in the same scan order."""
    from dnds_amd.host import ImplicitBDFDualTimeStep
    c, ev, orc = _vortex(12)
    bdf = ImplicitBDFDualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=5, rhsThresholdInternal=0.0,
                                  k=k)
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    dt, t = 0.2, 0.0
    for step in range(4):
        hist = bdf.Step(u, r, rn, dt=dt, t=t)
        assert hist.shape == (5, 4)
        t += dt
    ou, orr, its = orc.bdf_steps(c.u0, np.zeros((c.n_cell, 36)), k, 4, 5, 2, 10.0, dt, threshold=0.0,
                                 scan2cell=ev.LUSGSScanOrder())
    assert list(its) == [5, 5, 5, 5]
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL


def test_bdf_rejects_a_changed_dt_and_outlives_nothing():
    from dnds_amd.host import ImplicitBDFDualTimeStep
    c, ev, orc = _vortex(8)
    bdf = ImplicitBDFDualTimeStep(ev, nInternalRecStep=1, CFL=10.0, nTimeStepInternal=2, rhsThresholdInternal=0.0)
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    bdf.Step(u, r, rn, dt=0.2)
    with pytest.raises(Exception):
        bdf.Step(u, r, rn, dt=0.3)  # the reference asserts a constant dt
    bdf.Step(u, r, rn, dt=0.2)


def test_sdirk4_step_matches_oracle_in_colour_order():
    """Two SDIRK4 steps, a fixed number of inner iterations per stage."""
    from dnds_amd.host import ImplicitSDIRK4DualTimeStep
    c, ev, orc = _vortex(12)
    ode = ImplicitSDIRK4DualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=4, rhsThresholdInternal=0.0)
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    ou, orr = c.u0.copy(), np.zeros((c.n_cell, 36))
    dt = 0.2
    for step in range(2):
        its, hist = ode.Step(u, r, rn, dt=dt, t=step * dt)
        assert list(its) == [4, 4, 4] and np.all(hist[:, :, 0] > 0)
        ou, orr, oits = orc.sdirk4_step(ou, orr, 4, 2, 10.0, dt, threshold=0.0, t=step * dt,
                                        scan2cell=ev.LUSGSScanOrder())
        assert list(oits) == [4, 4, 4]
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL
    assert np.abs(ou - c.u0).max() > 1e-4


def test_lusgs_matrix_vec_matches_oracle():
    c, ev, orc = _vortex(12)
    n = c.n_cell
    u, _ = random_state(n, seed=21)
    alpha, dt = 0.9, 0.4
    du, dtau = ev.BuildMean().upload(u), ev.BuildScalar()
    ev.EvaluateDt(dtau, du, 5.0, 1e100, True)
    ev.LUSGSMatrixInit(dtau, dt, alpha)
    x = np.random.default_rng(22).standard_normal((n, 4))
    dx, dax = ev.BuildMean().upload(x), ev.BuildMean()
    ev.LUSGSMatrixVec(alpha, du, dx, dax)
    dto, _, lam, _ = orc.evaluate_dt(u, 5.0, use_local_dt=True)
    jinv = orc.lusgs_matrix_init(dto, dt, alpha, lam)
    ref = orc.lusgs_matrix_vec(1.0 / jinv, lam, alpha, u, x)
    assert rel_err(dax.download(), ref) <= 1e-13


@pytest.mark.parametrize("code", [1, 2])
def test_gmres_implicit_iterations_match_oracle(code):
    """Implicit-Euler iterations whose fsolve is GMRES (gmresCode 1: from zero,
    2: from the LU-SGS increment), LU-SGS-preconditioned, 2 restarts of a
    6-vector subspace; the device's dot products and least squares round
    differently from the oracle's, hence 1e-9."""
    from dnds_amd.host import ImplicitEulerDualTimeStep
    c, ev, orc = _vortex(12)
    ev.SetLinearSolver(code, 6, 2)
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=4, rhsThresholdInternal=0.0)
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    ode.Step(u, r, rn, dt=0.3)
    import oracle
    st = c.vfv.setting
    orc_g = oracle.Oracle(c.mesh.arrays(), c.vfv.arrays(), riemann_solver=1, far_field=c.far,
                          tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                          gmres_code=code, n_gmres_space=6, n_gmres_iter=2)
    ou, orr, _ = orc_g.implicit_euler(c.u0, np.zeros((c.n_cell, 36)), 4, 2, 10.0, dt=0.3, threshold=0.0,
                                      scan2cell=ev.LUSGSScanOrder())
    assert rel_err(u.download(), ou) <= 1e-9
    # GMRES changes the iterate: not the LU-SGS-only result
    ol, _, _ = orc.implicit_euler(c.u0, np.zeros((c.n_cell, 36)), 4, 2, 10.0, dt=0.3, threshold=0.0,
                                  scan2cell=ev.LUSGSScanOrder())
    assert rel_err(ou, ol) > 1e-8
    ev.SetLinearSolver(0)
