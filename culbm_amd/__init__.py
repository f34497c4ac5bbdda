"""MI355X-native D3Q27 lattice-Boltzmann engine (gfx950 HIP kernels behind a C ABI).

Public surface:
    Lattice, LatticeConfig   one simulation context (culbm_amd.lattice)
    cases                    initial-state generators and .dat file I/O
    blocking                 halo-blocking axis math of the reference
"""
from .lattice import Lattice, LatticeConfig, validate  # noqa: F401
from . import cases  # noqa: F401
from ._lib import CulbmError  # noqa: F401

__all__ = ["Lattice", "LatticeConfig", "validate", "cases", "CulbmError"]
