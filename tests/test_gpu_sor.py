"""SOR_Instead (and SOR_InverseScanning) on the GPU against the oracle,
through the C ABI.  The device sweeps in place colour by colour; no two cells
of a colour are neighbours, so this is the reference's sequential in-place
scan (DNDS_FV_VR.hpp:1495-1662) in the colour-major order
EulerEvaluator.LUSGSScanOrder, which the oracle follows.  Every boundary
variant is covered: far field and the special far fields (terms on the cell's
own partial sum), Wall_Euler / Wall_NoSlip corrections, Wall.  Bar: 1e-12
relative per sweep, and on SSP-RK3 and implicit-Euler trajectories whose
reconstruction sweeps are SOR sweeps."""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err
from test_gpu_parity import BC_SETS, DOMAIN, RS_CODE, random_state

pytestmark = pytest.mark.gpu

TOL = 1e-12

SOR_CASES = [(0, "periodic"), (1, "farfield"), (2, "farfield"), (0, "walls"), (2, "mixed_bc"), (1, "with_wall"),
             (0, "riemann"), (2, "ivfar"), (1, "dmrfar")]


def build_sor(kind, bcname, inverse=False, riemann="Roe", relax=1.0, nx=9, ny=7):
    import oracle
    from dnds_amd.host import EulerEvaluator
    periodic = bcname == "periodic"
    x1, y1 = DOMAIN.get(bcname, (1.0, 1.0))
    # far weight 1 so that the far-field faces carry the M x_i term
    mesh, vfv = mesh_and_setup(kind=kind, nx=nx, ny=ny, periodic=periodic, jitter=0.0 if periodic else 0.2, seed=5,
                               bc=BC_SETS[bcname], x1=x1, y1=y1, far_weight=1.0, jacobi_relax=relax,
                               sor_instead=1, sor_inverse_scanning=int(inverse))
    far = (1.0, 0.3, -0.2, 2.6)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType=riemann, gamma=1.4, farFieldStaticValue=far)
    st = vfv.setting
    scan = ev.LUSGSScanOrder()
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=RS_CODE[riemann], gamma=1.4, far_field=far,
                        jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std,
                        sor_scan=scan, sor_inverse=inverse)
    return mesh, vfv, ev, orc, scan


def oracle_sweep(orc, u, x, scan, inverse, t=0.0):
    x = orc.sor_step(u, x, scan, t)
    return orc.sor_step(u, x, scan[::-1], t) if inverse else x


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("kind,bcname", SOR_CASES)
def test_sor_sweeps_match_oracle_in_colour_order(kind, bcname, inverse):
    mesh, vfv, ev, orc, scan = build_sor(kind, bcname, inverse, relax=0.8 if inverse else 1.0)
    u, urec = random_state(mesh.n_cell, 31)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec().setConstant(7.0)
    ref = urec
    for sweep in range(3):
        vfv.ReconstructionJacobiStep(du, dr, dn, t=0.1)
        ref = oracle_sweep(orc, u, ref, scan, inverse, t=0.1)
        assert rel_err(dr.download(), ref) <= TOL, (sweep, rel_err(dr.download(), ref))
    # in place: uRecNewBuf is not touched (DNDS_FV_VR.hpp:1707)
    assert np.all(dn.download() == 7.0)
    # and the in-place sweep differs from the Jacobi sweep
    assert rel_err(ref, orc.jacobi_step(u, urec)) > 1e-6


# p = 1 (the coefficients padded into the p = 3 slots) and the local inertia
# coordinates of the reference's isentropic-vortex vfvSetting
# (data/euler_config_IV.json:60-80), with far-field and no-slip faces
VARIANTS = {
    "p1": dict(rec_order=1, far_weight=1.0),
    "local": dict(anisotropic_lengths=1, tang_weight=0.0, tang_weight_direction=0, wall_weight=0.5, far_weight=0.0,
                  weight_len_scale=1.0, weight_scheme_geom=3, weight_scheme_dir=1, use_local_coord=1),
}


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("kind", [0, 2])
def test_sor_sweeps_on_p1_and_local_coordinate_setups(variant, kind):
    import oracle
    from dnds_amd.host import EulerEvaluator
    mesh, vfv = mesh_and_setup(kind=kind, nx=8, ny=6, jitter=0.2, seed=8, bc=(2, 4, 2, 3), limiter=False,
                               sor_instead=1, **VARIANTS[variant])
    far = (1.0, 0.3, -0.2, 2.6)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType="Roe", gamma=1.4, farFieldStaticValue=far)
    st = vfv.setting
    scan = ev.LUSGSScanOrder()
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=1, gamma=1.4, far_field=far,
                        jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction)
    u, urec = random_state(mesh.n_cell, 36)
    if variant == "p1":
        urec.reshape(-1, 4, 9)[:, :, 2:] = 0.0  # p = 1 holds the two first-order rows
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ref = urec
    for sweep in range(3):
        vfv.ReconstructionJacobiStep(du, dr, dn)
        ref = orc.sor_step(u, ref, scan)
        assert rel_err(dr.download(), ref) <= TOL, (sweep, rel_err(dr.download(), ref))


def test_sor_converges_to_the_jacobi_fixed_point_on_the_device():
    """Same equations, other iteration: the SOR and Jacobi devices meet."""
    from dnds_amd.host import EulerEvaluator
    mesh, vfv_j = mesh_and_setup(kind=2, nx=10, ny=8, periodic=True, jitter=0.0, seed=5)
    _, vfv_s = mesh_and_setup(kind=2, nx=10, ny=8, periodic=True, jitter=0.0, seed=5, sor_instead=1)
    u, _ = random_state(mesh.n_cell, 32)
    out = []
    for vfv in (vfv_j, vfv_s):
        ev = EulerEvaluator(mesh, vfv, riemannSolverType="Roe", gamma=1.4)
        du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().setConstant(0.0), vfv.BuildRec()
        for _ in range(300):
            vfv.ReconstructionJacobiStep(du, dr, dn)
        out.append(dr.download())
    assert rel_err(out[1], out[0]) <= 1e-9


@pytest.mark.parametrize("kind,bcname,inverse", [(2, "mixed_bc", False), (0, "riemann", True)])
def test_ssprk3_steps_with_sor_match_oracle(kind, bcname, inverse):
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc, scan = build_sor(kind, bcname, inverse, riemann="HLLC")
    u, urec = random_state(mesh.n_cell, 33, amp=0.1, rec_amp=0.0)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=2, CFL=0.4)
    t = 0.0
    for _ in range(4):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(u, urec, 4, 2, 0.4)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL


def test_implicit_iterations_with_sor_match_oracle():
    from dnds_amd.host import ImplicitEulerDualTimeStep
    mesh, vfv, ev, orc, scan = build_sor(1, "farfield", inverse=False)
    u, _ = random_state(mesh.n_cell, 34, amp=0.1)
    ode = ImplicitEulerDualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=5,
                                    rhsThresholdInternal=0.0)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().setConstant(0.0), vfv.BuildRec()
    hist = ode.Step(du, dr, dn, dt=0.05)
    ou, orr, oh = orc.implicit_euler(u, np.zeros((mesh.n_cell, 36)), 5, 2, 10.0, dt=0.05, threshold=0.0,
                                     scan2cell=scan)
    assert rel_err(du.download(), ou) <= TOL
    assert rel_err(dr.download(), orr) <= TOL
    assert rel_err(hist, oh) <= 1e-10


def test_mixed_precision_step_refuses_sor():
    """The FP32 correction sweeps iterate the Jacobi map; with SOR_Instead the
    step says so instead of running another method."""
    from dnds_amd.host import DNDSError, ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc, scan = build_sor(0, "farfield")
    u, urec = random_state(mesh.n_cell, 35)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=3, CFL=0.4, mixedPrecision=True)
    with pytest.raises(DNDSError, match="SOR_Instead"):
        ode.Step(du, dr, dn, 0.0)
    assert rel_err(du.download(), u) == 0.0


def test_bdf_and_sdirk4_steps_with_sor_match_oracle():
    """The dual-time BDF2 and SDIRK4 steps with SOR sweeps inside, a fixed
    number of inner iterations, against the oracle in the same scan order."""
    from dnds_amd.host import ImplicitBDFDualTimeStep, ImplicitSDIRK4DualTimeStep
    mesh, vfv, ev, orc, scan = build_sor(2, "farfield", inverse=False)
    u0, _ = random_state(mesh.n_cell, 37, amp=0.1)
    bdf = ImplicitBDFDualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=3, rhsThresholdInternal=0.0,
                                  k=2)
    u, r, rn = ev.BuildMean().upload(u0), vfv.BuildRec().setConstant(0.0), vfv.BuildRec()
    dt = 0.02
    for step in range(3):
        bdf.Step(u, r, rn, dt=dt, t=step * dt)
    ou, orr, its = orc.bdf_steps(u0, np.zeros((mesh.n_cell, 36)), 2, 3, 3, 2, 10.0, dt, threshold=0.0,
                                 scan2cell=scan)
    assert list(its) == [3, 3, 3]
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL
    ode = ImplicitSDIRK4DualTimeStep(ev, nInternalRecStep=2, CFL=10.0, nTimeStepInternal=3, rhsThresholdInternal=0.0)
    u, r = ev.BuildMean().upload(u0), vfv.BuildRec().setConstant(0.0)
    its, _ = ode.Step(u, r, rn, dt=dt, t=0.0)
    ou, orr, oits = orc.sdirk4_step(u0, np.zeros((mesh.n_cell, 36)), 3, 2, 10.0, dt, threshold=0.0, t=0.0,
                                    scan2cell=scan)
    assert list(its) == list(oits) == [3, 3, 3]
    assert rel_err(u.download(), ou) <= TOL
    assert rel_err(r.download(), orr) <= TOL
