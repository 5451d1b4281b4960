"""The C++ host classes (include/dndsmi.hpp): the header compiles on its own,
the example program links against the in-tree library, and on the GPU the
program's results equal the same sequence of calls made through the Python
host classes (both drive the same C ABI, so the results are identical)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_run")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "dndsmi.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_example_program_is_built_and_links():
    assert os.path.exists(BIN), "run build() first"
    r = subprocess.run(["ldd", BIN], capture_output=True, text=True)
    line = [ln for ln in r.stdout.splitlines() if "libdndsmi" in ln]
    assert line and "not found" not in line[0], r.stdout


def _python_run():
    from dnds_amd import _capi as A
    from dnds_amd.host import (CompactFacedMeshSerialRW, DNDSError, EulerEvaluator, ExplicitSSPRK3LocalDt,
                               ImplicitBDFDualTimeStep, ImplicitEulerDualTimeStep, ImplicitSDIRK4DualTimeStep,
                               VRFiniteVolume2D, vr_setting)
    mesh = CompactFacedMeshSerialRW.generate(A.MESH_TRI_SPLIT, 10, 8, 0.0, 1.0, 0.0, 1.0, jitter=0.2, seed=5,
                                             bc=(A.BC_FARFIELD,) * 4)
    vfv = VRFiniteVolume2D(mesh, vr_setting(build_limiter_data=1, far_weight=1.0))
    vfv.Initialization(2)
    far = (1.0, 0.3, -0.2, 2.6)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType="HLLC", farFieldStaticValue=far)
    bary = vfv.arrays(["cell_bary"])["cell_bary"].reshape(-1, 2)
    s = 1 + 0.1 * np.sin(2 * np.pi * bary[:, 0]) * np.cos(2 * np.pi * bary[:, 1])
    u0 = np.tile(np.array(far), (mesh.n_cell, 1))
    u0[:, 0] *= s
    u0[:, 3] *= s
    u, r, rn = ev.BuildMean().upload(u0), vfv.BuildRec().setConstant(0.0), vfv.BuildRec()
    rk = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=2, CFL=0.5, MaxDt=1e100, UseLocaldt=False, useLimiter=True)
    t = 0.0
    for _ in range(3):
        t += rk.Step(u, r, rn, t)
    imp = ImplicitEulerDualTimeStep(ev, nInternalRecStep=1, CFL=5.0, nTimeStepInternal=4, rhsThresholdInternal=0.0,
                                    UseLocaldt=True)
    imp.Step(u, r, rn, dt=0.05, t=t)
    t += 0.05
    bdf = ImplicitBDFDualTimeStep(ev, nInternalRecStep=1, CFL=5.0, nTimeStepInternal=3, rhsThresholdInternal=0.0,
                                  UseLocaldt=True, k=2)
    for _ in range(2):
        bdf.Step(u, r, rn, dt=0.05, t=t)
        t += 0.05
    ImplicitSDIRK4DualTimeStep(ev, nInternalRecStep=1, CFL=5.0, nTimeStepInternal=2, rhsThresholdInternal=0.0,
                               UseLocaldt=True).Step(u, r, rn, dt=0.05, t=t)
    t += 0.05
    rhs = ev.BuildMean()
    ev.EvaluateRHS(rhs, u, r, t)
    res = ev.EvaluateResidual(rhs, 1, False)
    dtau = []  # the reference's std::vector<real> form
    ev.EvaluateDt(dtau, u, 5.0, 1e100, True)
    inc = ev.BuildMean().setConstant(0.0)
    ev.LUSGSMatrixInit(dtau, 1e100, 1.0, u, r, 0, t)
    ev.UpdateLUSGSForward(1.0, rhs, u, inc, inc)
    ev.UpdateLUSGSBackward(1.0, rhs, u, inc, inc)
    res_inc = ev.EvaluateResidual(inc, 1, False)
    with pytest.raises(DNDSError):
        ev.LUSGSMatrixInit(dtau, 1e100, 1.0, u, r, 1, t)
    return mesh.n_cell, u.download(), r.download(), np.r_[res, res_inc], np.array(dtau)


@pytest.mark.gpu
def test_cpp_host_classes_match_python_host(tmp_path):
    out = tmp_path / "facade.bin"
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n, u, urec, res, dtau = _python_run()
    data = np.fromfile(out, dtype=np.float64)
    assert data.size == n * 4 + n * 36 + 8 + n
    assert np.array_equal(data[:4 * n].reshape(n, 4), u)
    assert np.array_equal(data[4 * n:40 * n].reshape(n, 36), urec)
    assert np.array_equal(data[40 * n:40 * n + 8], res)
    assert np.array_equal(data[40 * n + 8:], dtau)
