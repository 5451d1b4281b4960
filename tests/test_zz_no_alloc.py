"""dndsmi_evaluate_residual_noalloc (the allocation-free form of
dndsmi_evaluate_residual, whose block partials live in a static scratch array
of the library) against dndsmi_evaluate_residual through the C ABI: equal
bit for bit on two calls in a row, the second on a smaller mesh, for every
norm; the allocation-free call runs first on each context, before the other
entry point has allocated anything of its own."""
import ctypes as C

import numpy as np
import pytest

from conftest import mesh_and_setup

pytestmark = pytest.mark.gpu


def _ctx(nx, ny, seed):
    from dnds_amd.host import EulerEvaluator
    mesh, vfv = mesh_and_setup(kind=2, nx=nx, ny=ny, periodic=False, jitter=0.2, seed=seed, bc=(2, 2, 2, 2),
                               limiter=False)
    ev = EulerEvaluator(mesh, vfv, riemannSolverType="Roe", farFieldStaticValue=(1.0, 0.3, -0.2, 2.6))
    rng = np.random.default_rng(seed)
    rhs = ev.BuildMean().upload(rng.standard_normal((mesh.n_cell, 4)) * rng.uniform(0.1, 10, (mesh.n_cell, 1)))
    return mesh, vfv, ev, rhs


def _call(fn, ev, rhs, P, vol_wise):
    from dnds_amd import _capi as A
    res = np.full(4, np.nan)
    rc = fn(ev.device.ctx, rhs.h, P, vol_wise, res.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == A.OK, rc
    return res


def test_noalloc_residual_matches_the_allocating_entry_point():
    from dnds_amd import _capi as A
    lib = A.load()
    sizes = [(150, 90, 1), (37, 11, 2)]  # about 20,000 cells, then about 600
    for nx, ny, seed in sizes:
        mesh, vfv, ev, rhs = _ctx(nx, ny, seed)
        vol = vfv.arrays(["volume"])["volume"]
        print(f"no-alloc residual: mesh {nx}x{ny}, {mesh.n_cell} cells", flush=True)
        for P in (1, 2, 3):
            for vol_wise in (0, 1):
                got = _call(lib.dndsmi_evaluate_residual_noalloc, ev, rhs, P, vol_wise)
                again = _call(lib.dndsmi_evaluate_residual_noalloc, ev, rhs, P, vol_wise)
                ref = _call(lib.dndsmi_evaluate_residual, ev, rhs, P, vol_wise)
                assert np.array_equal(got, ref), (nx, ny, P, vol_wise, got, ref)
                assert np.array_equal(again, ref)
                # and against the definition on the host
                r = rhs.download()
                w = vol if (vol_wise and P < 3) else np.ones(mesh.n_cell)
                host = (np.abs(r).max(0) if P >= 3 else
                        ((np.abs(r) ** P) * w[:, None]).sum(0) ** (1.0 / P))
                assert np.allclose(got, host, rtol=1e-12, atol=0)
        ev.device.sync()

