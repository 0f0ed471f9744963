"""pawan_amd: the vortex-particle interaction and RK4 path of pawan
(punsingh/pawan) on one AMD MI355X (gfx950), behind pawan's interaction /
integration API.

The compute lives in ``libpawan_hip.so`` (hand-written HIP kernels and a C
ABI, include/pawan_hip.h); this package is its host-side mirror of the
reference's classes plus the seeded workload generators.
"""
from . import generators  # noqa: F401
from ._lib import (FP32C, FP64, KERNEL_GAUSSIAN, KERNEL_HIGHORDER, SCHEME_DEFAULT,  # noqa: F401
                   SCHEME_ONESIDED, SCHEME_SYMMETRIC, PawanError)
from .interaction import RK4, Batch, Integration, Interaction, Server, Wake, WakeFile, autotune  # noqa: F401

__all__ = ["Interaction", "Integration", "RK4", "Batch", "Server", "Wake", "WakeFile", "autotune", "generators",
           "FP64", "FP32C", "KERNEL_HIGHORDER", "KERNEL_GAUSSIAN", "SCHEME_DEFAULT",
           "SCHEME_ONESIDED", "SCHEME_SYMMETRIC", "PawanError"]
