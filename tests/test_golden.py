"""Committed golden fixture (tests/golden/vortex8.npz, written by
tests/golden/make_golden.py): the oracle must reproduce it exactly (CPU),
and the device, fed the stored mesh and setup arrays through the C ABI, to
<= 1e-12 relative (GPU)."""
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT, rel_err

FIX = os.path.join(ROOT, "tests", "golden", "vortex8.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    mesh = {k[5:]: d[k] for k in d if k.startswith("mesh_")}
    for k in ("n_node", "n_cell", "n_face"):
        mesh[k] = int(mesh[k])
    setup = {k[6:]: d[k] for k in d if k.startswith("setup_")}
    return d, mesh, setup


def make_oracle(d, mesh, setup, rs, limiter_procedure=1):
    tw, twd, scale, nstd, relax, gamma = d["setting"]
    return oracle.Oracle(mesh, setup, riemann_solver=rs, gamma=gamma, far_field=tuple(d["far"]), jacobi_relax=relax,
                         tang_weight=tw, tang_weight_direction=int(twd), wbap_smooth_indicator_scale=scale,
                         wbap_n_std=nstd, limiter_procedure=limiter_procedure)


def test_oracle_reproduces_golden_fixture(fx):
    d, mesh, setup = fx
    orc = make_oracle(d, mesh, setup, 1)
    u, urec = d["u"], d["urec"]
    assert np.array_equal(orc.jacobi_step(u, urec), d["jacobi1"])
    dt, dtmin, _, dlf = orc.evaluate_dt(u, 0.5)
    assert np.array_equal(dt, d["dt"]) and dtmin == d["dtmin"][0] and np.array_equal(dlf, d["dlf"])
    assert np.array_equal(orc.evaluate_rhs(u, urec, dlf, t=0.1), d["rhs_roe"])
    assert np.array_equal(make_oracle(d, mesh, setup, 2).evaluate_rhs(u, urec, dlf, t=0.1), d["rhs_hllc"])
    lim, ind = orc.wbap_v3(u, urec)
    assert np.array_equal(lim, d["wbap_lim"]) and np.array_equal(ind, d["wbap_ind"])
    ru, rr, rt = orc.ssprk3_steps(d["u0"], np.zeros_like(urec), 2, 3, 0.5)
    assert np.array_equal(ru, d["rk3_u"]) and np.array_equal(rr, d["rk3_urec"]) and rt == d["rk3_t"][0]
    ru, rr, rt = orc.ssprk3_steps(d["u0"], np.zeros_like(urec), 2, 3, 0.5, use_limiter=True)
    assert np.array_equal(ru, d["rk3lim_u"]) and np.array_equal(rr, d["rk3lim_urec"]) and rt == d["rk3lim_t"][0]
    lim, ind = orc.wbap_v2(u, urec)
    assert np.array_equal(lim, d["wbap2_lim"]) and np.array_equal(ind, d["wbap_ind"])
    orc2 = make_oracle(d, mesh, setup, 1, limiter_procedure=0)
    ru, rr, rt = orc2.ssprk3_steps(d["u0"], np.zeros_like(urec), 2, 3, 0.5, use_limiter=True)
    assert np.array_equal(ru, d["rk3lim2_u"]) and np.array_equal(rr, d["rk3lim2_urec"]) and rt == d["rk3lim2_t"][0]


@pytest.mark.gpu
def test_device_reproduces_golden_fixture_from_stored_arrays(fx):
    from dnds_amd.host import EulerEvaluator, ExplicitSSPRK3LocalDt, ExternalMesh, ExternalVR, vr_setting
    d, mesh, setup = fx
    tw, twd, scale, nstd, relax, gamma = d["setting"]
    st = vr_setting(tang_weight=tw, tang_weight_direction=int(twd), wbap_smooth_indicator_scale=scale,
                    wbap_n_std=nstd, jacobi_relax=relax, build_limiter_data=1)
    em = ExternalMesh(mesh)
    evr = ExternalVR(em, setup, st)
    results = {}
    for rs in ("Roe", "HLLC"):
        ev = EulerEvaluator(em, evr, riemannSolverType=rs, gamma=gamma, farFieldStaticValue=tuple(d["far"]))
        u, r, rn, rhs, dt = (ev.BuildMean().upload(d["u"]), evr.BuildRec().upload(d["urec"]), evr.BuildRec(),
                             ev.BuildMean(), ev.BuildScalar())
        dtmin = ev.EvaluateDt(dt, u, 0.5)
        ev.EvaluateRHS(rhs, u, r, 0.1)
        results[rs] = rhs.download()
        if rs == "Roe":
            assert abs(dtmin - d["dtmin"][0]) <= 1e-12 * d["dtmin"][0]
            assert rel_err(dt.download(), d["dt"]) <= 1e-12
            ind = ev.BuildScalar()
            evr.ReconstructionWBAPLimitFacialV3(u, r, rn, None, ind)
            assert rel_err(rn.download(), d["wbap_lim"]) <= 1e-12
            assert rel_err(ind.download(), d["wbap_ind"]) <= 1e-12
            evr.ReconstructionWBAPLimitFacialV2(u, r, rn, None, ind)
            assert rel_err(rn.download(), d["wbap2_lim"]) <= 1e-12
            assert rel_err(ind.download(), d["wbap_ind"]) <= 1e-12
            evr.ReconstructionJacobiStep(u, r, rn)
            assert rel_err(r.download(), d["jacobi1"]) <= 1e-12
            for lim, proc, key in ((False, 0, "rk3"), (True, 1, "rk3lim"), (True, 0, "rk3lim2")):
                u2, r2, rn2 = ev.BuildMean().upload(d["u0"]), evr.BuildRec(), evr.BuildRec()
                ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=3, CFL=0.5, useLimiter=lim, limiterProcedure=proc)
                t = 0.0
                for _ in range(2):
                    t += ode.Step(u2, r2, rn2, t)
                assert abs(t - d[f"{key}_t"][0]) <= 1e-12 * t
                assert rel_err(u2.download(), d[f"{key}_u"]) <= 1e-12
                assert rel_err(r2.download(), d[f"{key}_urec"]) <= 1e-12
    assert rel_err(results["Roe"], d["rhs_roe"]) <= 1e-12
    assert rel_err(results["HLLC"], d["rhs_hllc"]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("rs", ["Roe", "HLLC"])
def test_device_rhs_with_distinct_gauss_point_normals(fx, rs):
    """Face normals that differ between the three LINE_3 points (as on curved
    faces; here the stored normals of points 1 and 2 are rescaled) select the
    face kernel that recomputes the unit normal, |n| and the balance fluxes
    at every point; it must match the oracle fed the same arrays."""
    from dnds_amd.host import EulerEvaluator, ExternalMesh, ExternalVR, vr_setting
    d, mesh, setup = fx
    tw, twd, scale, nstd, relax, gamma = d["setting"]
    setup = dict(setup)
    fn = np.array(setup["face_norm"], dtype=np.float64).reshape(-1, 3, 2)
    fn[:, 1] *= 1.0 + 2e-3
    fn[:, 2] *= 1.0 - 1e-3
    setup["face_norm"] = fn.ravel()
    st = vr_setting(tang_weight=tw, tang_weight_direction=int(twd), wbap_smooth_indicator_scale=scale,
                    wbap_n_std=nstd, jacobi_relax=relax, build_limiter_data=1)
    em = ExternalMesh(mesh)
    evr = ExternalVR(em, setup, st)
    ev = EulerEvaluator(em, evr, riemannSolverType=rs, gamma=gamma, farFieldStaticValue=tuple(d["far"]))
    u, r, rhs, dt = ev.BuildMean().upload(d["u"]), evr.BuildRec().upload(d["urec"]), ev.BuildMean(), ev.BuildScalar()
    ev.EvaluateDt(dt, u, 0.5)
    ev.EvaluateRHS(rhs, u, r, 0.1)
    orc = make_oracle(d, mesh, setup, 1 if rs == "Roe" else 2)
    _, _, _, dlf = orc.evaluate_dt(d["u"], 0.5)
    ref = orc.evaluate_rhs(d["u"], d["urec"], dlf, t=0.1)
    assert rel_err(rhs.download(), ref) <= 1e-12
    # the rescaled normals change the result: the per-point normals were used
    assert rel_err(ref, d["rhs_roe" if rs == "Roe" else "rhs_hllc"]) > 1e-6
