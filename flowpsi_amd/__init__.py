"""flowpsi_amd: flowPsi's implicit finite-volume Newton step on MI355X (gfx950).

The numerics live in libfpsi_gpu.so (HIP kernels behind the C ABI of
include/fpsi.h); this package holds the seeded grid generators, the
workload definitions and the ctypes binding.
"""
from . import grids, workloads  # noqa: F401
from .capi import FpsiError, Solver, load  # noqa: F401

__all__ = ["grids", "workloads", "Solver", "FpsiError", "load"]
