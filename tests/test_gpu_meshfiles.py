"""Mesh files read by the Gmsh reader, through the device path against the
oracle (bar 1e-12 relative, as for generated meshes).  The files are written
by tests/golden/make_meshes.py in an external exporter's layout.

* tests/golden/tri_delaunay_a.msh (58 cells) and tri_delaunay_b.msh (238
  cells), Delaunay triangulations bounded by `Wall` faces: Jacobi sweeps (a
  `Wall` face is valid for the sweep only, not as a flux boundary).
* tests/golden/quad_6x4.msh (24 quadrangles), bounded by `Farfield` faces:
  the time step, the right-hand side with Roe and HLLC, and SSP-RK3 steps
  with the WBAP limiter.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err
from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = os.path.join(ROOT, "tests", "golden")
FAR = (1.0, 0.5, -0.3, 2.67)


def build(name, riemann="Roe"):
    import oracle
    from dnds_amd.host import (RIEMANN_VR_SETTING, CompactFacedMeshSerialRW, EulerEvaluator, VRFiniteVolume2D,
                               vr_setting)
    mesh = CompactFacedMeshSerialRW.read_gmsh(os.path.join(GOLD, name))
    vfv = VRFiniteVolume2D(mesh, vr_setting(**RIEMANN_VR_SETTING, build_limiter_data=1))
    vfv.Initialization(4)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType=riemann, gamma=1.4, farFieldStaticValue=FAR)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=1 if riemann == "Roe" else 2, gamma=1.4,
                        far_field=FAR, jacobi_relax=st.jacobi_relax, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std)
    return mesh, vfv, ev, orc


@pytest.mark.parametrize("name", ["tri_delaunay_a.msh", "tri_delaunay_b.msh", "quad_6x4.msh"])
def test_jacobi_sweeps_on_the_mesh_files_match_oracle(name):
    mesh, vfv, ev, orc = build(name)
    u, urec = random_state(mesh.n_cell, 41)
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ref = urec
    for sweep in range(6):
        vfv.ReconstructionJacobiStep(du, dr, dn, t=0.0)
        ref = orc.jacobi_step(u, ref)
        assert rel_err(dr.download(), ref) <= TOL, sweep


@pytest.mark.parametrize("riemann", ["Roe", "HLLC"])
def test_dt_and_rhs_on_the_farfield_mesh_file_match_oracle(riemann):
    mesh, vfv, ev, orc = build("quad_6x4.msh", riemann)
    u, urec = random_state(mesh.n_cell, 42)
    du, dr, rhs, dt = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), ev.BuildMean(), ev.BuildScalar()
    dtmin = ev.EvaluateDt(dt, du, CFL=0.5, MaxDt=1e100, UseLocaldt=True)
    rdt, rmin, _, dlf = orc.evaluate_dt(u, 0.5, 1e100, True)
    assert abs(dtmin - rmin) <= TOL * rmin and rel_err(dt.download(), rdt) <= TOL
    ev.EvaluateRHS(rhs, du, dr, t=0.1)
    assert rel_err(rhs.download(), orc.evaluate_rhs(u, urec, dlf, t=0.1)) <= TOL


@pytest.mark.parametrize("procedure", [0, 1])
def test_limited_ssprk3_on_the_farfield_mesh_file_matches_oracle(procedure):
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    mesh, vfv, ev, orc = build("quad_6x4.msh", "HLLC")
    u, _ = random_state(mesh.n_cell, 43, amp=0.3)
    urec = np.zeros((mesh.n_cell, 36))
    du, dr, dn = ev.BuildMean().upload(u), vfv.BuildRec().upload(urec), vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=3, CFL=0.5, useLimiter=True, limiterProcedure=procedure)
    t = 0.0
    for _ in range(5):
        t += ode.Step(du, dr, dn, t)
    orc.cfg.limiter_procedure = procedure
    ru, rr, rt = orc.ssprk3_steps(u, urec, 5, 3, 0.5, use_limiter=True)
    assert abs(t - rt) <= TOL * rt
    assert rel_err(du.download(), ru) <= TOL
    assert rel_err(dr.download(), rr) <= TOL
