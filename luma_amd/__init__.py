"""luma_amd -- MI355X (gfx950) lattice-Boltzmann time step with LUMA's GridObj semantics.

The compute path is libluma_amd.so (hand-written HIP kernels behind the C ABI declared in
include/lumagpu.h).  This package holds the ctypes binding (_abi), a numpy-facing grid
handle (grid.GpuGrid) and the host-side case set-up (setups).
"""
from ._abi import LIB_PATH, LumaGpuError, load  # noqa: F401

__all__ = ["LIB_PATH", "LumaGpuError", "load"]
