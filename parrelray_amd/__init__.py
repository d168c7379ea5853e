"""MI355X-native ParRelRay hot path: null-geodesic integrator and AA passes.

The compute runs in hand-written HIP kernels for gfx950 behind the C ABI of
include/prr.h (library parrelray_amd/lib/libprr.so).  This package is a thin
host-side wrapper: it loads that library and fails loudly when it is missing.
"""
from . import _abi  # noqa: F401

__all__ = ["_abi"]
