"""CPU FP64 oracle of the hot path -- TEST INFRASTRUCTURE ONLY (the parity
checker and the benchmark's CPU baseline; never imported by the product)."""
from .oracle import (BASELINE_FLAGS, BASELINE_LIB, LIB, Oracle, boundary_value, build,  # noqa: F401
                     build_baseline, eigenvectors, flux_jacobian_times_du, hllc_flux, hllep_flux, lib, roe_flux,
                     roe_m_flux, set_threads)
