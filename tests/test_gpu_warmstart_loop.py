"""The learned warm start inside the mixed-precision SSP-RK3 step
(dndsmi_ssprk3_set_warm_start), through the C ABI on the GPU.

* A network whose parameters are all zero predicts a zero correction: the
  check finds no deviation at all (tolerance 0 passes), and the steps equal
  the plain mixed-precision steps bit for bit, re-checks included.
* A network with large random weights predicts a harmful correction: the
  check fails, and the run continues on the plain run's state, so the result
  equals the plain mixed-precision steps bit for bit.
* Attaching checks its arguments."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(n=16, seed=3):
    from dnds_amd import cases
    from dnds_amd.host import EulerEvaluator, WarmStart
    c = cases.isentropic_vortex(n, seed=seed, nthreads=2)
    ev = EulerEvaluator(c.mesh, c.vfv, riemannSolverType="Roe", farFieldStaticValue=c.far)
    ws = WarmStart(ev, seed=1, capacity=0)
    return c, ev, ws


def _run(c, ev, ode, steps):
    u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    t = 0.0
    for _ in range(steps):
        t += ode.Step(u, r, rn, t)
    return u.download(), r.download(), t


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_zero_network_passes_every_check_and_changes_nothing():
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    c, ev, ws = _setup()
    ws.set_params(np.zeros(ws.n_param, dtype=np.float32))
    print(f"warm start in the loop: zero network, {c.n_cell} cells", flush=True)
    plain = _run(c, ev, ExplicitSSPRK3LocalDt(ev, 5, 0.5, mixedPrecision=True), 9)
    ode = ExplicitSSPRK3LocalDt(ev, 5, 0.5, mixedPrecision=True, warmStart=ws, checkSteps=2, recheckInterval=2,
                                tolerance=0.0)
    got = _run(c, ev, ode, 9)
    assert _same(got, plain)
    st = ode.WarmStartStats()
    # checks at steps 1-2, 5-6 and 9 (unfinished); the first stage of the run
    # only records the means the next stage's change is taken from
    assert st["checks_passed"] == 2 and st["fallbacks"] == 0, st
    assert st["corrected_stages"] == 3 * 9 - 1 and st["last_check_deviation"] == 0.0, st


def test_harmful_network_falls_back_to_the_plain_run():
    from dnds_amd.host import ExplicitSSPRK3LocalDt
    c, ev, ws = _setup()
    rng = np.random.default_rng(7)
    ws.set_params((rng.standard_normal(ws.n_param) * 3.0).astype(np.float32))
    print(f"warm start in the loop: random network, {c.n_cell} cells", flush=True)
    plain = _run(c, ev, ExplicitSSPRK3LocalDt(ev, 5, 0.5, mixedPrecision=True), 8)
    ode = ExplicitSSPRK3LocalDt(ev, 5, 0.5, mixedPrecision=True, warmStart=ws, checkSteps=3, recheckInterval=2,
                                tolerance=1e-6)
    got = _run(c, ev, ode, 8)
    st = ode.WarmStartStats()
    assert st["fallbacks"] == 1 and st["checks_passed"] == 0, st
    assert st["last_check_deviation"] > 1e-6 or st["last_check_deviation"] == -1.0, st
    # the run went on from the plain run's state with the correction off
    assert _same(got, plain)


def test_attach_checks_its_arguments():
    from dnds_amd import _capi as A
    from dnds_amd.host import DNDSError, ExplicitSSPRK3LocalDt
    c, ev, ws = _setup(n=8)
    c2, ev2, ws2 = _setup(n=8, seed=4)
    with pytest.raises(DNDSError):
        ExplicitSSPRK3LocalDt(ev, 5, 0.5, warmStart=ws)  # FP64 step: no hook
    lib = A.load()
    assert lib.dndsmi_ssprk3_set_warm_start(ev.device.ctx, ws2.h, 3, 0, 1e-6) == A.ERR_INVALID_ARGUMENT  # other ctx
    assert lib.dndsmi_ssprk3_set_warm_start(ev.device.ctx, ws.h, 0, 0, 1e-6) == A.ERR_INVALID_ARGUMENT
    assert lib.dndsmi_ssprk3_set_warm_start(ev.device.ctx, ws.h, 2, -1, 1e-6) == A.ERR_INVALID_ARGUMENT
    assert lib.dndsmi_ssprk3_set_warm_start(ev.device.ctx, ws.h, 2, 0, -1.0) == A.ERR_INVALID_ARGUMENT
    assert lib.dndsmi_ssprk3_set_warm_start(ev.device.ctx, None, 1, 0, 0.0) == A.OK  # detach
