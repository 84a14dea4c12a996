"""MI355X-native Grains3D DEM time step.

``pacific_amd.gdem`` binds libgdem.so (HIP kernels for gfx950 behind the C ABI
of include/gdem.h); ``pacific_amd.workloads`` builds seeded initial states.
Importing the package loads libgdem.so and fails if it has not been built.
"""
from . import gdem, workloads  # noqa: F401
from .gdem import GdemSystem, build_system  # noqa: F401

gdem.load_library()

__all__ = ["gdem", "workloads", "GdemSystem", "build_system"]
