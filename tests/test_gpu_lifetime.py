"""Objects bound to a device context may be destroyed in any order: arrays
and warm starts destroyed after their context must not touch it (the
context frees the arrays' storage and detaches them)."""
import ctypes as C

import numpy as np
import pytest

from conftest import mesh_and_setup

pytestmark = pytest.mark.gpu


def test_arrays_and_warm_start_outlive_their_context():
    from dnds_amd import _capi as A
    mesh, vfv = mesh_and_setup(kind=0, nx=6, ny=5)
    lib = A.load()
    es = A.EulerSetting(riemann_solver=A.RS_ROE, gamma=1.4, far_field=(C.c_double * 4)(1.0, 0.0, 0.0, 2.5))
    ctx = C.c_void_p()
    assert lib.dndsmi_ctx_create(C.byref(mesh.view), C.byref(vfv.view), C.byref(vfv.setting), C.byref(es), None,
                                 C.byref(ctx)) == 0
    arrays = []
    for kind in (A.ARRAY_DOF, A.ARRAY_REC, A.ARRAY_SCALAR):
        h = C.c_void_p()
        assert lib.dndsmi_array_create(ctx, kind, C.byref(h)) == 0
        arrays.append(h)
    ws = C.c_void_p()
    assert lib.dndsmi_warmstart_create(ctx, C.byref(mesh.view), C.byref(vfv.view), 1, 0, C.byref(ws)) == 0
    lib.dndsmi_array_destroy(arrays.pop())  # one before the context, the rest after
    lib.dndsmi_ctx_destroy(ctx)
    # an array whose context is gone refuses every call but destroy
    host = np.zeros(4 * mesh.n_cell)
    assert lib.dndsmi_array_upload(arrays[0], A.dptr(host)) == A.ERR_INVALID_ARGUMENT
    assert lib.dndsmi_array_download(arrays[0], A.dptr(host)) == A.ERR_INVALID_ARGUMENT
    assert lib.dndsmi_array_set_constant(arrays[0], 1.0) == A.ERR_INVALID_ARGUMENT
    assert lib.dndsmi_array_copy(arrays[1], arrays[1]) == A.ERR_INVALID_ARGUMENT
    for h in arrays:
        lib.dndsmi_array_destroy(h)
    lib.dndsmi_warmstart_destroy(ws)
