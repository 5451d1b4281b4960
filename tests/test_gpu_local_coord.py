"""GPU parity with the local-inertia-coordinate basis (useLocalCoord) through
the C ABI: the device evaluates the face bases from the points' local
coordinates (face_gauss_xy holds xi = A (p - c)), so the sweeps, right-hand
sides and SSP-RK3 steps match the oracle, which reads the setup's stored
bases, to the FP64 bar (1e-12 relative).  Contexts on such a setup refuse
the viscous model (it needs derivative rows)."""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12

# the reference's isentropic-vortex vfvSetting (data/euler_config_IV.json:60-80)
IV = dict(anisotropic_lengths=1, tang_weight=0.0, tang_weight_direction=0, wall_weight=0.5, far_weight=0.0,
          weight_len_scale=1.0, weight_scheme_geom=3, weight_scheme_dir=1)


def build(kind, bc, riemann="Roe", periodic=False):
    import oracle
    from dnds_amd.host import EulerEvaluator
    mesh, vfv = mesh_and_setup(kind=kind, nx=9, ny=7, periodic=periodic, jitter=0.0 if periodic else 0.2, seed=6,
                               bc=bc, limiter=False, use_local_coord=1, **IV)
    far = (1.0, 0.3, -0.2, 2.6)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType=riemann, farFieldStaticValue=far)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=1 if riemann == "Roe" else 2, far_field=far,
                        tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction)
    return mesh, vfv, ev, orc


def state(n, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.1 * rng.uniform(-1, 1, n)
    vx, vy = 0.5 + 0.1 * rng.uniform(-1, 1, n), -0.3 + 0.1 * rng.uniform(-1, 1, n)
    p = 1.0 + 0.1 * rng.uniform(-1, 1, n)
    u = np.stack([rho, rho * vx, rho * vy, p / 0.4 + 0.5 * rho * (vx ** 2 + vy ** 2)], axis=1)
    return u, 0.02 * rng.standard_normal((n, 36))


CASES = [(0, (2, 2, 2, 2), False), (1, (2, 3, 4, 3), False), (2, None, True), (2, (3, 3, 3, 3), False)]


@pytest.mark.parametrize("kind,bc,periodic", CASES)
@pytest.mark.parametrize("riemann", ["Roe", "HLLC"])
def test_local_coordinates_match_oracle(kind, bc, periodic, riemann):
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc = build(kind, bc, riemann, periodic)
    print(f"local coords: kind {kind}, bc {bc}, {riemann}, {mesh.n_cell} cells", flush=True)
    u, urec = state(mesh.n_cell, 3)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ref = urec
    for _ in range(3):
        vfv.ReconstructionJacobiStep(du, dr, dn)
        ref = orc.jacobi_step(u, ref)
    assert rel_err(dr.download(), ref) <= TOL
    rhs, dt = ev.BuildMean(), ev.BuildScalar()
    ev.EvaluateDt(dt, du, CFL=0.5)
    _, _, _, dlf = orc.evaluate_dt(u, 0.5)
    ev.EvaluateRHS(rhs, du, dr, t=0.0)
    assert rel_err(rhs.download(), orc.evaluate_rhs(u, ref, dlf, t=0.0)) <= TOL
    du.upload(u), dr.upload(np.zeros_like(urec))
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=2, CFL=0.4)
    t = 0.0
    for _ in range(3):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(u, np.zeros_like(urec), 3, 2, 0.4)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL
    ev.device.sync()


def test_local_coordinates_refuse_the_viscous_model():
    from dnds_amd.host import DNDSError, EulerEvaluator
    mesh, vfv = mesh_and_setup(kind=0, nx=5, ny=4, bc=(2, 2, 4, 2), limiter=False, use_local_coord=1, **IV)
    with pytest.raises(DNDSError):
        EulerEvaluator(mesh, vfv, riemannSolverType="Roe", farFieldStaticValue=(1.0, 0.2, 0.0, 2.5), muGas=1e-3)
