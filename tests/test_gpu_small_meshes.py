"""Degenerate and ragged meshes on the GPU against the oracle: a single quad
(every face a boundary face, no neighbour at all), two triangles, a few mixed
cells, exactly one wavefront of cells, one cell more than a wavefront, and a
mixed mesh whose 64-cell windows hold both cell kinds.  Each goes through
the Jacobi sweeps, EvaluateDt, EvaluateRHS and limited SSP-RK3 steps, to the
same 1e-12 bar as the other parity tests."""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err
from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (mesh kind, nx, ny): 0 triangles, 1 quads, 2 mixed
SHAPES = [(1, 1, 1), (0, 1, 1), (2, 3, 1), (0, 4, 8), (1, 65, 1), (2, 16, 9)]


def build(kind, nx, ny, riemann):
    import oracle
    from dnds_amd.host import EulerEvaluator
    # far weight 1: with the Riemann settings' 0 a cell without neighbours
    # would have an empty functional
    mesh, vfv = mesh_and_setup(kind=kind, nx=nx, ny=ny, jitter=0.2, seed=7, far_weight=1.0)
    far = (1.0, 0.3, -0.2, 2.6)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType=riemann, gamma=1.4, farFieldStaticValue=far)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver={"Roe": 1, "HLLC": 2}[riemann], gamma=1.4,
                        far_field=far, jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std)
    return mesh, vfv, ev, orc


@pytest.mark.parametrize("kind,nx,ny", SHAPES)
def test_small_and_ragged_meshes_match_oracle(kind, nx, ny):
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc = build(kind, nx, ny, "HLLC")
    n = mesh.n_cell
    if kind != 2:
        assert n == (2 if kind == 0 else 1) * nx * ny
    u, urec = random_state(n, 21)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ref = urec
    for _ in range(3):
        vfv.ReconstructionJacobiStep(du, dr, dn, t=0.0)
        ref = orc.jacobi_step(u, ref)
    assert rel_err(dr.download(), ref) <= TOL
    dt = ev.BuildScalar()
    dtmin = ev.EvaluateDt(dt, du, CFL=0.7, MaxDt=1e100, UseLocaldt=True)
    rdt, rmin, _, dlf = orc.evaluate_dt(u, 0.7, 1e100, True)
    assert abs(dtmin - rmin) <= TOL * abs(rmin)
    assert rel_err(dt.download(), rdt) <= TOL
    rhs = ev.BuildMean()
    ev.EvaluateRHS(rhs, du, dr, t=0.0)
    assert rel_err(rhs.download(), orc.evaluate_rhs(u, ref, dlf, t=0.0)) <= TOL
    u2, r2 = random_state(n, 22, amp=0.1, rec_amp=0.0)
    du.upload(u2)
    dr.upload(r2)
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=2, CFL=0.4, useLimiter=True)
    t = 0.0
    for _ in range(3):
        t += ode.Step(du, dr, dn, t)
    ru, rr, rt = orc.ssprk3_steps(u2, r2, 3, 2, 0.4, use_limiter=True)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL
