"""MI355X-native CATO hot path: gfx950 HIP kernels behind the reference's
flux-solver / reconstruction / boundary-condition / time-integrator interfaces.

The compute path is ``libcato_hip.so`` (C ABI in include/cato_hip.h); this
package is its Python host side.
"""
from ._native import CatoError, load  # noqa: F401
from .solver import SchemeConfig, Solver  # noqa: F401

__all__ = ["CatoError", "SchemeConfig", "Solver", "load"]
