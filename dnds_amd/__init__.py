"""MI355X (gfx950) implementation of DNDS's 2D VR finite-volume hot path.

The native library `libdndsmi.so` (C ABI in include/dndsmi.h) holds the mesh
builder, the FP64 VR setup and the HIP kernels.  `dnds_amd.host` mirrors the
reference's VRFiniteVolume2D / EulerEvaluator / ExplicitSSPRK3LocalDt
interfaces on top of that ABI.
"""
from . import _capi  # noqa: F401

__all__ = ["_capi"]
