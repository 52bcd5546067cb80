"""phi_eng_amd — MI355X-native PhiEng 3D SPH fluid step.

The product is ``libphieng_hip.so``: the reference's World3 fluid C ABI
(``include/phieng_fluid3.h``) over hand-written HIP kernels for gfx950.
This package holds its sources (``csrc/``), the build script and a thin
ctypes host binding (``ffi``, ``world``) that mirrors the reference's
FluidState3/World3 surface for tests and benchmarks.
"""
from .ffi import FluidConfig3Ffi, lib, lib_path  # noqa: F401
from .world import World3  # noqa: F401
