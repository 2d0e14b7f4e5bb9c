"""MI355X (gfx950) SIMPLE conjugate-heat-transfer step of bolyme/hsolver, one GPU.

The compute path is libhsolver_amd.so (hand-written HIP kernels behind a C ABI,
include/hsolver.h); this package is the thin Python host around it.
"""
from . import abi, configs  # noqa: F401

__all__ = ["abi", "configs"]
