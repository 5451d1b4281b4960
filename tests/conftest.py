"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU.

Small meshes and their setup are built once per session through the
product's C ABI; the oracle consumes the same arrays.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    from dnds_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        from dnds_amd import build
        build.build()
    import oracle
    if not os.path.exists(oracle.LIB):
        oracle.build()


_ensure_built()


def mesh_and_setup(kind=0, nx=8, ny=6, periodic=False, jitter=0.2, seed=3, bc=None, x1=1.0, y1=1.0,
                   limiter=True, **vr_overrides):
    from dnds_amd import _capi as A
    from dnds_amd.host import RIEMANN_VR_SETTING, CompactFacedMeshSerialRW, VRFiniteVolume2D, vr_setting
    bc = bc if bc is not None else (A.BC_FARFIELD,) * 4
    mesh = CompactFacedMeshSerialRW.generate(kind, nx, ny, 0.0, x1, 0.0, y1, periodic=(periodic, periodic),
                                             jitter=jitter, seed=seed, bc=bc)
    s = dict(RIEMANN_VR_SETTING)
    s.update(vr_overrides)
    vfv = VRFiniteVolume2D(mesh, vr_setting(**s, build_limiter_data=int(limiter)))
    vfv.Initialization(4)
    return mesh, vfv


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.abs(b).max()
    return float(np.abs(a - b).max() / (scale if scale > 0 else 1.0))
