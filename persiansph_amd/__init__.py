"""persiansph_amd: PersianSPH's particle-interaction hot path on MI355X (gfx950).

The product is the HIP library persiansph_amd/lib/libpersiansph_amd.so behind the C ABI of
include/persiansph_amd.h; `Solver` is its Python binding, `cases` builds the dam-break and
granular-collapse inputs, `abi` mirrors the C data types.
"""
from . import abi, cases  # noqa: F401
from .abi import FP32_MIXED, FP64, PARTICLE_DTYPE, default_config, new_particles  # noqa: F401
from .solver import LIB_PATH, Solver, SolverError, lib  # noqa: F401

__all__ = ["Solver", "SolverError", "abi", "cases", "default_config", "new_particles",
           "PARTICLE_DTYPE", "FP64", "FP32_MIXED", "LIB_PATH", "lib"]
