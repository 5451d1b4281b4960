"""CPU FP64 oracle of the hot path -- TEST INFRASTRUCTURE ONLY (the parity
checker and the benchmark's CPU baseline; never imported by the product)."""
from .oracle import (LIB, Oracle, boundary_value, build, eigenvectors, flux_jacobian_times_du,  # noqa: F401
                     hllc_flux, lib, roe_flux, set_threads)
