"""MI355X (gfx950) backend for the RealTimeParticles physics hot path.

The product is librtp.so (HIP kernels + the C ABI of include/rtp.h) and the C++
Physics::Model classes in csrc/host.  This package adds the Python binding used by the
tests and the benchmark, and the seeded scene generators.
"""
from . import scenes  # noqa: F401
from .capi import (  # noqa: F401
    FIELD_ACC,
    FIELD_DENSITY,
    FIELD_LAMBDA,
    FIELD_PRED_POS,
    FIELD_VORTICITY,
    Batch,
    Boids,
    BoidsParams,
    FluidParams,
    Fluids,
    Mlp,
    RtpError,
    device_count,
    grid_build,
    lib,
    sort_pairs,
)

__all__ = [
    "Batch", "Boids", "BoidsParams", "FluidParams", "Fluids", "Mlp", "RtpError", "device_count", "grid_build",
    "lib", "sort_pairs", "scenes",
]
