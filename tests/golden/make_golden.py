"""Writes tests/golden/vortex8.npz: a seeded 8 x 8 periodic split-quad mesh
(128 cells) of the vortex case, its VR setup, seeded states, and the
oracle's outputs for each hot-path call (Jacobi sweep, EvaluateDt,
EvaluateRHS with Roe and HLLC, WBAP V3 and V2, two SSP-RK3 steps without the
limiter and with each).  Inputs and outputs are stored, so the fixture pins the oracle
(regression) and the device (parity) independently of the host setup code.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from dnds_amd import cases  # noqa: E402


def main():
    case = cases.isentropic_vortex(8, seed=11, nthreads=1, build_limiter_data=True)
    m = case.mesh.arrays()
    s = case.vfv.arrays()
    st = case.vfv.setting
    rng = np.random.default_rng(2024)
    n = case.n_cell
    u = case.u0 * (1 + 0.05 * rng.uniform(-1, 1, (n, 1)))
    urec = 0.05 * rng.standard_normal((n, 36))
    out = {}
    for rs, name in ((1, "roe"), (2, "hllc")):
        orc = oracle.Oracle(m, s, riemann_solver=rs, far_field=case.far, tang_weight=st.tang_weight,
                            tang_weight_direction=st.tang_weight_direction,
                            wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, limiter_procedure=1)
        dt, dtmin, lf, dlf = orc.evaluate_dt(u, 0.5)
        out[f"rhs_{name}"] = orc.evaluate_rhs(u, urec, dlf, t=0.1)
        if rs == 1:
            out["jacobi1"] = orc.jacobi_step(u, urec)
            out["dt"], out["dtmin"], out["dlf"] = dt, np.array([dtmin]), dlf
            out["wbap_lim"], out["wbap_ind"] = orc.wbap_v3(u, urec)
            ru, rr, rt = orc.ssprk3_steps(case.u0, np.zeros((n, 36)), 2, 3, 0.5)
            out["rk3_u"], out["rk3_urec"], out["rk3_t"] = ru, rr, np.array([rt])
            ru, rr, rt = orc.ssprk3_steps(case.u0, np.zeros((n, 36)), 2, 3, 0.5, use_limiter=True)
            out["rk3lim_u"], out["rk3lim_urec"], out["rk3lim_t"] = ru, rr, np.array([rt])
            out["wbap2_lim"], _ = orc.wbap_v2(u, urec)
            orc2 = oracle.Oracle(m, s, riemann_solver=rs, far_field=case.far, tang_weight=st.tang_weight,
                                 tang_weight_direction=st.tang_weight_direction,
                                 wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, limiter_procedure=0)
            ru, rr, rt = orc2.ssprk3_steps(case.u0, np.zeros((n, 36)), 2, 3, 0.5, use_limiter=True)
            out["rk3lim2_u"], out["rk3lim2_urec"], out["rk3lim2_t"] = ru, rr, np.array([rt])
    setting = np.array([st.tang_weight, st.tang_weight_direction, st.wbap_smooth_indicator_scale, st.wbap_n_std,
                        st.jacobi_relax, 1.4])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "vortex8.npz"),
                        **{f"mesh_{k}": np.asarray(v) for k, v in m.items()},
                        **{f"setup_{k}": v for k, v in s.items() if v is not None},
                        far=np.array(case.far), u0=case.u0, u=u, urec=urec, setting=setting, **out)


if __name__ == "__main__":
    main()
