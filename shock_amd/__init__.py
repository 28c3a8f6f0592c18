"""shock_amd: MI355X-native RHS and time-step path of the SHOCK compressible Navier-Stokes solver.

The compute path is libshockx.so (hand-written HIP kernels for gfx950) behind the C ABI in
include/shockx.h; `shock_amd.shockx` is its Python mirror.
"""
from .shockx import Case, lib, QUIRKS_REFERENCE, QUIRK_ZETA_VISC_J, QUIRK_METRIC_DIAG3D  # noqa: F401
