"""tecs_amd - MI355X (gfx950) engine for the per-step hot path of TECS cell simulations.

The engine is the C-ABI library tecs_amd/_lib/libtecs_amd.so (HIP kernels + C++ host code,
declared in include/tecs_amd.h). This package holds its build script, a ctypes mirror of the
ABI, numpy-facing handles and the seeded benchmark workloads.
"""
from . import abi, io, workloads
from ._lib import TecsError, device_count, load
from .engine import GpuBatch, GpuSimulation

__all__ = ["abi", "io", "workloads", "TecsError", "device_count", "load", "GpuBatch", "GpuSimulation"]
