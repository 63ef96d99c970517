"""The benchmark workloads as ready-to-render scenes.

* ``c1`` — Cornell box, 36 triangles, 256 x 256, 16 spp, depth 8.
* ``c2`` — Cornell walls + light around a displaced icosphere of 69,620
  triangles (8 seeded materials), 1024 x 1024, 64 spp, depth 8; its mesh is
  also the 16M-ray closest-hit microbenchmark scene.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import camera as cam_mod
from . import scene as scene_mod
from . import scenes


@dataclass
class Workload:
    name: str
    loaded: scene_mod.LoadedScene
    buffers: scene_mod.SceneBuffers
    camera: np.ndarray
    width: int
    height: int
    spp: int
    depth: int
    denoise: Optional[str] = None


def c1(width: int = 256, height: int = 256, spp: int = 16, depth: int = 8) -> Workload:
    loaded = scene_mod.load_scene_from_str(scenes.cornell_yaml(), "cornell-c1")
    sb = scene_mod.build_scene_buffers(loaded)
    return Workload("cornell_c1", loaded, sb, cam_mod.camera_for(loaded, width, height), width,
                    height, spp, depth)


def c2(width: int = 1024, height: int = 1024, spp: int = 64, depth: int = 8, n: int = 59,
       seed: int = 7) -> Workload:
    mats = scenes.random_materials_yaml(8, seed)
    text = scenes.cornell_yaml(with_boxes=False, extra_materials=mats)
    base = scene_mod.load_scene_from_str(text, "cornell-c2")
    mat0 = base.name_to_idx["m0"]
    v, t, _ = scenes.displaced_icosphere(n=n, seed=seed, mat_base=mat0, n_mats=8)
    loaded = scene_mod.load_scene_from_str(text, "cornell-c2", extra_mesh=(v, t))
    sb = scene_mod.build_scene_buffers(loaded)
    return Workload("icosphere_c2", loaded, sb, cam_mod.camera_for(loaded, width, height), width,
                    height, spp, depth)


C3_CAMERA = "camera: {look_from: [0.0, 44.0, 46.0], look_at: [0.0, 8.0, 0.0], vfov: 55.0}\n"


def c3(width: int = 1920, height: int = 1080, spp: int = 16, depth: int = 6, chunks: int = 32,
       seed: int = 11) -> Workload:
    """Voxel terrain: 512 x 256 x 512 voxels greedy-meshed to ~2M triangles,
    gradient sky + Gaussian sun at 35 degrees elevation."""
    from . import texture, world

    text = C3_CAMERA + world.world_materials_yaml()
    base = scene_mod.load_scene_from_str(text, "voxel-c3")
    blocks = world.terrain_blocks(chunks, chunks, seed)
    cx0 = -chunks // 2
    v, t = world.mesh_terrain(blocks, cx0, cx0, base.name_to_idx)
    del blocks
    loaded = scene_mod.load_scene_from_str(text, "voxel-c3", extra_mesh=(v, t))
    env = texture.rgb_to_env(texture.generate_skymap(1024, 512, sun_elevation_deg=35.0))
    sb = scene_mod.build_scene_buffers(loaded, env=env)
    return Workload("voxel_c3", loaded, sb, cam_mod.camera_for(loaded, width, height), width,
                    height, spp, depth, "bilateral")


def scene_aabb(sb: scene_mod.SceneBuffers):
    nodes = sb.mesh_bvh
    return nodes[0]["aabb_min"][:3].copy(), nodes[0]["aabb_max"][:3].copy()


def by_name(name: str, **kw) -> Workload:
    return {"c1": c1, "c2": c2, "c3": c3}[name](**kw)
