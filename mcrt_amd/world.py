"""Seeded voxel terrain for the voxel-world workload.

A 512 x 256 x 512 voxel heightmap terrain (32 x 32 chunks of 16 x 256 x 16,
the chunk layout of src/world/chunk.rs:15-57) from 7-octave value-noise fBm,
with four opaque block types: a 3-voxel surface layer chosen by height and
slope (sand near the shore line, grass, stone on steep or high ground, snow
on peaks) over stone / sand strata.  It is greedy-meshed by the library's
C++ mesher (mcrt_mesh_voxel_chunks, the semantics of src/world/mesher.rs)
into about 2.16 million triangles.

The reference's own generator (noise-rs fBm with caves and trees,
src/world/generator.rs) depends on the noise crate and is not reproduced;
this heightmap stands in for it with the same chunk and mesh contract.
"""
from __future__ import annotations

import numpy as np

from . import scenes
from ._lib import mesh_voxel_chunks
from . import types as T

CHUNK_XZ, CHUNK_Y = 16, 256
VOXEL_SIZE = 0.125  # src/world/chunk.rs:21

AIR, SAND, GRASS, STONE, SNOW = 0, 1, 2, 3, 4
BLOCK_MATERIALS = {
    SAND: ("sand", (0.76, 0.70, 0.50)),
    GRASS: ("grass", (0.30, 0.52, 0.18)),
    STONE: ("stone", (0.48, 0.47, 0.45)),
    SNOW: ("snow", (0.90, 0.91, 0.93)),
}


def heightmap(nx: int, nz: int, seed: int, base: float = 110.0, amp: float = 150.0,
              freq: float = 1.0 / 20.0, octaves: int = 7) -> np.ndarray:
    """Column heights (in voxels) from value-noise fBm."""
    xs, zs = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(nz, dtype=np.float64), indexing="ij")
    p = np.stack([xs.ravel() * freq, zs.ravel() * freq, np.full(xs.size, 0.37)], axis=1)
    n = scenes.fbm3(p, seed, octaves=octaves, base_freq=1.0).reshape(nx, nz)
    h = base + amp * n
    return np.clip(np.round(h), 1, CHUNK_Y - 1).astype(np.int32)


def terrain_blocks(chunks_x: int = 32, chunks_z: int = 32, seed: int = 11) -> np.ndarray:
    """(chunks_x, chunks_z, 16, 256, 16) uint16 block ids."""
    nx, nz = chunks_x * CHUNK_XZ, chunks_z * CHUNK_XZ
    h = heightmap(nx, nz, seed)
    gx = np.abs(np.diff(h, axis=0, append=h[-1:, :]))
    gz = np.abs(np.diff(h, axis=1, append=h[:, -1:]))
    slope = np.maximum(gx, gz)
    top = np.full((nx, nz), GRASS, np.uint16)
    top[h < 70] = SAND
    top[(slope >= 3) | (h > 130)] = STONE
    top[h > 150] = SNOW
    y = np.arange(CHUNK_Y)[None, :, None]
    hh = h[:, None, :]
    # Below the 3-voxel surface layer: stone / sandstone strata, 3 voxels each.
    strata = np.where((y // 3) % 2 == 0, STONE, SAND).astype(np.uint16)
    world = np.where(y < hh, strata, AIR).astype(np.uint16)
    surface = (y >= hh - 3) & (y < hh)
    world = np.where(surface, top[:, None, :], world).astype(np.uint16)
    # (x, y, z) -> (cx, cz, x, y, z) chunk-major, each chunk XZY-major.
    w = world.reshape(chunks_x, CHUNK_XZ, CHUNK_Y, chunks_z, CHUNK_XZ)
    return np.ascontiguousarray(w.transpose(0, 3, 1, 2, 4))


def mesh_terrain(blocks: np.ndarray, cx0: int, cz0: int, name_to_idx: dict):
    opaque = np.zeros(256, np.uint8)
    mat = np.zeros(256, np.uint32)
    for bid, (name, _) in BLOCK_MATERIALS.items():
        opaque[bid] = 1
        mat[bid] = name_to_idx[name]
    return mesh_voxel_chunks(blocks, cx0, cz0, VOXEL_SIZE, opaque, mat)


def world_materials_yaml() -> str:
    out = ["world_materials:\n"]
    for _, (name, alb) in BLOCK_MATERIALS.items():
        out.append(f"  {name}: {{type: lambertian, albedo: [{alb[0]}, {alb[1]}, {alb[2]}]}}\n")
    return "".join(out)
