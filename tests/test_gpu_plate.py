"""The laminar flat plate (NS_2D over a Wall_NoSlip plate, quads stretched
geometrically towards the wall, the reference's boundary-layer VR settings)
on the GPU against the oracle, at P_ORDER 3 and 1: steady implicit-Euler
iterations with local time steps at CFL 20 and an infinite physical time
step.
* The first iterations, LU-SGS in colour order on both sides: 1e-12
  relative, as for the other FP64 paths.
* Converged: the multicolour LU-SGS driven to a 1e-11 drop of the density
  increment against the oracle's steady state in the reference's natural
  order (tests/golden/make_plate.py): the steady state does not depend on the
  scan order, and the two agree to 1e-8 of the state's departure from the
  free stream."""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("p", [3, 1])
def test_steady_plate_iterations_match_oracle_in_colour_order(p):
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ImplicitEulerDualTimeStep
    case = cases.flat_plate(24, 12, reynolds=5e3, rec_order=p, nthreads=2)
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType="Roe", farFieldStaticValue=case.far, **case.gas)
    st = case.vfv.setting
    g = case.gas
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=1, far_field=case.far,
                        jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction, mu_gas=g["muGas"], pr_gas=g["prGas"],
                        r_gas=g["Rgas"], t_ref=g["TRef"], c_sutherland=g["CSutherland"])
    urec0 = np.zeros((case.n_cell, 36))
    u, r, rn = ev.BuildMean().upload(case.u0), case.vfv.BuildRec().upload(urec0), case.vfv.BuildRec()
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=1, CFL=case.cfl, nTimeStepInternal=8,
                                    rhsThresholdInternal=0.0)
    hist = ode.Step(u, r, rn, dt=1e100)
    ou, orr, ohist = orc.implicit_euler(case.u0, urec0, 8, 1, case.cfl, dt=1e100, threshold=0.0,
                                        scan2cell=ev.LUSGSScanOrder())
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL
    assert len(hist) == 8 and np.all(np.isfinite(hist))
    # the plate slows the flow: the state has left the free stream
    assert rel_err(u.download(), case.u0) > 1e-6


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("p", [1, 3])
def test_plate_converges_to_the_oracle_steady_state(p):
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ImplicitEulerDualTimeStep
    path = os.path.join(GOLDEN, f"plate_32x16_p{p}.npz")
    if not os.path.exists(path):
        pytest.skip(f"no converged oracle state for p = {p} (tests/golden/make_plate.py writes it)")
    gold = np.load(path)
    nx, ny, re, cfl, drop = gold["settings"]
    case = cases.flat_plate(int(nx), int(ny), reynolds=re, rec_order=p, cfl=cfl, nthreads=2)
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType="Roe", farFieldStaticValue=case.far, **case.gas)
    u, r, rn = ev.BuildMean().upload(case.u0), case.vfv.BuildRec().setConstant(0.0), case.vfv.BuildRec()
    chunk = 2000
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=1, CFL=cfl, nTimeStepInternal=chunk,
                                    rhsThresholdInternal=0.0)
    first, its, last = None, 0, None
    print(f"plate {int(nx)}x{int(ny)} p={p}: multicolour LU-SGS to a {drop:.0e} drop", flush=True)
    while its < 3 * int(gold["iterations"]):
        h = ode.Step(u, r, rn, dt=1e100)
        first = h[0, 0] if first is None else first
        its += len(h)
        last = h[-1, 0]
        if last <= drop * first:
            break
    print(f"plate p={p}: {its} iterations (oracle, natural order: {int(gold['iterations'])}), "
          f"drop {last / first:.2e}", flush=True)
    assert last <= drop * first
    got = u.download()
    dep = np.linalg.norm(gold["u"] - case.u0)
    assert np.linalg.norm(got - gold["u"]) <= 1e-8 * dep, np.linalg.norm(got - gold["u"]) / dep
