"""EMSolvers for MI355X: explicit time-domain Maxwell solvers on gfx950.

The compute lives in libems.so (hand-written HIP kernels behind the C ABI of
include/ems/ems.h); this package is its Python binding.
"""
from . import _lib  # noqa: F401

__version__ = "0.1.0"
