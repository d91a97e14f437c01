"""natray_amd: the natray APIC water substep on MI355X (gfx950).

The hot path (binning, P2G, grid update, pressure CG, projection, velocity extension, G2P) runs
as hand-written HIP kernels in libnatray.so behind a C ABI (include/natray.h). This package is
the Python view of that ABI (`abi`, `Sim`) plus the reference's scene-construction API
(`params`, `scenes`).
"""
from . import abi, params, scenes  # noqa: F401
from .sim import Sim  # noqa: F401

__all__ = ["abi", "params", "scenes", "Sim"]
