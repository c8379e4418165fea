"""picsar_amd -- PICSAR's per-step PIC hot path on AMD MI355X (gfx950).

The kernels live in libpxrhip.so (hand-written HIP, C ABI in
include/pxrhip.h).  ``picsar_amd.picsarpy.picsar`` mirrors the reference's
Python module for the leaf routines; ``picsar_amd.sim.Simulation`` holds
device-resident fields and tiled particles and runs the reference's step
(gather+push, particle boundaries, sorting, deposition, current guards,
FDTD) on the GPU.
"""
from ._lib import PxrHipError, load  # noqa: F401

__version__ = "0.1.0"
