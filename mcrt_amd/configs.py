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


def scene_aabb(sb: scene_mod.SceneBuffers):
    nodes = sb.mesh_bvh
    return nodes[0]["aabb_min"][:3].copy(), nodes[0]["aabb_max"][:3].copy()


def by_name(name: str, **kw) -> Workload:
    return {"c1": c1, "c2": c2}[name](**kw)
