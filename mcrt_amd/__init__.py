"""mcrt_amd — MI355X-native (gfx950) render hot path of the mcrt path tracer.

The compute lives in ``libmcrt_hip.so`` (hand-written HIP kernels behind the
C-ABI of ``include/mcrt_abi.h``); this package is the host side above it:
scene YAML loading, BVH building through the library's C++ builder, the
headless driver and ctypes bindings.
"""
from ._lib import Context, McrtError, build_mesh_bvh, build_sphere_bvh  # noqa: F401
from . import types  # noqa: F401

__all__ = ["Context", "McrtError", "build_mesh_bvh", "build_sphere_bvh", "types"]
