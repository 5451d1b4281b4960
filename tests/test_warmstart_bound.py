"""What any warm start of the stage's reconstruction can gain at a fixed sweep
count, measured on the oracle (CPU).

In the SSP-RK3 step the reconstruction uRec carries over from stage to stage
(DNDS_ODE.hpp:131-178, DNDS_FV_EulerSolver.hpp:713-837), and each stage runs
a fixed number k of Jacobi sweeps.  A learned warm start adds a correction to
the carried uRec before the sweeps.  For the periodic vortex the converged
reconstruction is linear in the cell means, so the best correction a
predictor of "the change of the converged reconstruction since the last
stage" can make is that change itself, L(u - u_prev), computed here exactly
by converging the sweep on u - u_prev.  The test runs the vortex with that
exact correction at k = 4 and shows that its density L2 error against the
exact solution still misses the plain k = 5 run's by more than 1 %: no
predictor of the stage change reaches equal L2 error with fewer sweeps.
"""
import numpy as np

import oracle

COEF = (1.0, 0.25, 2.0 / 3.0)
CT = (1.0, 0.25, 2.0 / 3.0)


def _vortex_l2(case, orc, vol, k, steps, exact_correction):
    """SSP-RK3 at CFL 0.5 in the oracle's stage order (oracle_ssprk3_steps),
    optionally adding the exact change of the converged reconstruction
    before each stage's k sweeps."""
    from dnds_amd import cases
    u, x, t, u_prev = case.u0.copy(), np.zeros((case.n_cell, 36)), 0.0, None
    for _ in range(steps):
        dt, dtmin, _, dlf = orc.evaluate_dt(u, 0.5)
        u_last = u.copy()
        for st in range(3):
            ts = t + CT[st] * dtmin
            if exact_correction and u_prev is not None:
                d = np.zeros_like(x)
                for _ in range(200):
                    d = orc.jacobi_step(u - u_prev, d)
                x = x + d
            u_prev = u.copy()
            for _ in range(k):
                x = orc.jacobi_step(u, x)
            un = u + orc.evaluate_rhs(u, x, dlf, ts) * dt[:, None]
            if st > 0:
                un = (un * (COEF[st] / (1 - COEF[st])) + u_last) * (1 - COEF[st])
            u = un
        t += dtmin
    ex = cases.cell_means(case.mesh, vol, lambda px, py: cases.vortex_exact(px, py, t))
    return float(np.sqrt(np.sum((u[:, 0] - ex[:, 0]) ** 2 * vol) / vol.sum())), u


def test_exact_stage_correction_cannot_replace_a_sweep():
    from dnds_amd import cases
    oracle.set_threads(4)
    case = cases.isentropic_vortex(20, seed=1, nthreads=4)
    orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=1, far_field=case.far)
    vol = case.vfv.arrays(["volume"])["volume"]
    steps = 30
    e5, u5 = _vortex_l2(case, orc, vol, 5, steps, False)
    # the stage loop above is the oracle's own step
    ref, _, _ = orc.ssprk3_steps(case.u0, np.zeros((case.n_cell, 36)), steps, 5, 0.5)
    assert np.array_equal(u5, ref)
    e4, _ = _vortex_l2(case, orc, vol, 4, steps, False)
    e4x, _ = _vortex_l2(case, orc, vol, 4, steps, True)
    oracle.set_threads(1)
    assert e4x < e4          # the exact correction helps ...
    assert e4x > 1.01 * e5   # ... but not to the 5-sweep run's error
