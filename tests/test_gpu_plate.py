"""The laminar flat plate (NS_2D over a Wall_NoSlip plate, quads stretched
geometrically towards the wall) on the GPU against the oracle: steady
implicit-Euler iterations with local time steps at CFL 20 and an infinite
physical time step, LU-SGS in colour order on both sides.  Bar: 1e-12
relative, as for the other FP64 paths."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def test_steady_plate_iterations_match_oracle_in_colour_order():
    import oracle
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, ImplicitEulerDualTimeStep
    case = cases.flat_plate(24, 12, reynolds=5e3, nthreads=2)
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
