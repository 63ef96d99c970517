"""NumPy record types for the reference's GPU buffer contract.

Byte-identical to the repr(C) structs of the reference (and to
``include/mcrt_abi.h``):

* ``GpuSphere`` 32 B — src/scene.rs:22-29
* ``GpuBvhNode`` 48 B — src/bvh.rs:54-68
* ``GpuVertex`` 48 B — src/mesh.rs:24-34
* ``GpuTriangle`` 16 B — src/mesh.rs:39-46
* ``GpuMaterial`` 48 B / ``GpuMaterialData`` 3088 B — src/material.rs:35-63
* ``CameraUniform`` 112 B — src/camera.rs:35-53
* ``DenoiseParams`` 16 B — src/gpu.rs:164-183
"""
from __future__ import annotations

import numpy as np

MAX_MATERIALS = 64  # src/material.rs:17
MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_EMISSIVE = 0, 1, 2, 3  # src/material.rs:70-75
BACKFACE_CULL_FLAG = 1 << 31  # src/scene.rs:713
T_MIN = np.float32(1e-4)  # path_trace.wgsl:52
T_MAX = np.float32(1e9)  # path_trace.wgsl:55
MAX_BOUNCES = 8  # path_trace.wgsl:61
ENV_MAP_WIDTH, ENV_MAP_HEIGHT = 1024, 512  # src/texture.rs:33-36
ID_MISS = 0xFFFFFFFF
ID_SPHERE_BIT = 0x80000000

GpuSphere = np.dtype([("center_r", "<f4", 4), ("mat_and_pad", "<u4", 4)])
GpuBvhNode = np.dtype(
    [
        ("aabb_min", "<f4", 4),
        ("aabb_max", "<f4", 4),
        ("right_or_offset", "<u4"),
        ("prim_count", "<u4"),
        ("_pad", "<u4", 2),
    ]
)
GpuVertex = np.dtype([("position", "<f4", 4), ("normal", "<f4", 4), ("uv", "<f4", 4)])
GpuTriangle = np.dtype([("v", "<u4", 3), ("mat_idx", "<u4")])
GpuMaterial = np.dtype([("type_pad", "<u4", 4), ("albedo_fuzz", "<f4", 4), ("ior_pad", "<f4", 4)])
GpuMaterialData = np.dtype(
    [
        ("mat_count", "<u4"),
        ("n_emissive", "<u4"),
        ("_pad", "<u4", 2),
        ("materials", GpuMaterial, MAX_MATERIALS),
    ]
)
CameraUniform = np.dtype(
    [
        ("origin", "<f4", 4),
        ("lower_left", "<f4", 4),
        ("horizontal", "<f4", 4),
        ("vertical", "<f4", 4),
        ("defocus_u", "<f4", 4),
        ("defocus_v", "<f4", 4),
        ("frame_count", "<u4"),
        ("_pad", "<u4", 3),
    ]
)
DenoiseParams = np.dtype(
    [("radius", "<i4"), ("sigma_s", "<f4"), ("sigma_d", "<f4"), ("normal_pow", "<f4")]
)

# STUB_SPHERE, src/scene.rs:36-39.
STUB_SPHERE = np.zeros((), GpuSphere)
STUB_SPHERE["center_r"] = (1.0e9, 1.0e9, 1.0e9, 0.0)

for _dt, _size in (
    (GpuSphere, 32),
    (GpuBvhNode, 48),
    (GpuVertex, 48),
    (GpuTriangle, 16),
    (GpuMaterial, 48),
    (GpuMaterialData, 3088),
    (CameraUniform, 112),
    (DenoiseParams, 16),
):
    assert _dt.itemsize == _size, (_dt, _dt.itemsize, _size)


def default_denoise_params() -> np.ndarray:
    """DenoiseParams::default(), src/gpu.rs:177-183."""
    p = np.zeros((), DenoiseParams)
    p["radius"] = 3
    p["sigma_s"] = 2.0
    p["sigma_d"] = 0.5
    p["normal_pow"] = 8.0
    return p
