"""Scene YAML -> GPU buffers, the reference's scene schema unchanged.

Restates ``load_scene_from_str`` (src/scene.rs:640-978) and
``gpu_layout::build_scene_buffers`` (src/gpu_layout.rs:259-390):

* palette order: ``world_materials`` then ``materials``, first occurrence wins
  (src/scene.rs:654-663); inline materials are appended per use
  (src/scene.rs:667-680);
* material defaults albedo 1, fuzz 0, ior 1.5, strength 1
  (src/scene.rs:354-381);
* ``plane`` = one quad (2 triangles) with Gram-Schmidt tangents
  (src/scene.rs:523-590), back-face flag in bit 31 of ``mat_idx``
  (src/scene.rs:709-718); ``box`` = 6 quads (src/scene.rs:736-827);
* Euler rotation R = Rz * Ry * Rx in degrees (src/scene.rs:396-405);
* stub vertex + degenerate triangle when there is no mesh geometry
  (src/scene.rs:887-899); more than 64 materials is an error
  (src/scene.rs:902-907).

All geometry arithmetic is FP32 in the reference's operation order.  The
``mesh`` object type (OBJ files) is not supported here and raises
:class:`SceneError`.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import yaml

from . import texture
from . import types as T

F = np.float32


class SceneError(Exception):
    """src/scene.rs:597-633: Io, Parse, UnknownMaterial, MeshLoad, TooManyMaterials."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind


class _UniqueKeyLoader(yaml.SafeLoader):
    """SafeLoader that rejects duplicate mapping keys (serde/IndexMap do)."""


def _construct_mapping(loader, node, deep=False):
    keys = set()
    for k_node, _ in node.value:
        k = loader.construct_object(k_node, deep=deep)
        if k in keys:
            raise SceneError("Parse", f"duplicate key {k!r}")
        keys.add(k)
    return yaml.SafeLoader.construct_mapping(loader, node, deep)


_UniqueKeyLoader.add_constructor(yaml.resolver.BaseResolver.DEFAULT_MAPPING_TAG, _construct_mapping)


# ---------------------------------------------------------------------------
# Field validation helpers (serde's shape checks)
# ---------------------------------------------------------------------------
def _num(v, what: str) -> np.float32:
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise SceneError("Parse", f"{what}: expected a number, got {v!r}")
    return F(v)


def _vec(v, n: int, what: str) -> List[np.float32]:
    if not isinstance(v, list) or len(v) != n:
        raise SceneError("Parse", f"{what}: expected a sequence of {n} numbers, got {v!r}")
    return [_num(x, what) for x in v]


def _opt_vec(d: dict, key: str, n: int, what: str):
    return None if d.get(key) is None else _vec(d[key], n, f"{what}.{key}")


def _opt_num(d: dict, key: str, what: str):
    return None if d.get(key) is None else _num(d[key], f"{what}.{key}")


# ---------------------------------------------------------------------------
# Materials, src/scene.rs:105-127 and :354-381
# ---------------------------------------------------------------------------
_KINDS = {"lambertian": T.MAT_LAMBERTIAN, "metal": T.MAT_METAL, "dielectric": T.MAT_DIELECTRIC,
          "emissive": T.MAT_EMISSIVE}


def material_desc_to_gpu(desc, what: str = "material") -> np.ndarray:
    if not isinstance(desc, dict) or "type" not in desc:
        raise SceneError("Parse", f"{what}: a material needs a 'type'")
    kind = desc["type"]
    if kind not in _KINDS:
        raise SceneError("Parse", f"{what}: unknown material type {kind!r}")
    albedo = _opt_vec(desc, "albedo", 3, what) or [F(1.0), F(1.0), F(1.0)]
    fuzz = _opt_num(desc, "fuzz", what)
    ior = _opt_num(desc, "ior", what)
    strength = _opt_num(desc, "emission_strength", what)
    layer = desc.get("texture_layer", 0) or 0
    if isinstance(layer, bool) or not isinstance(layer, int) or layer < 0:
        raise SceneError("Parse", f"{what}.texture_layer: expected an unsigned integer")
    m = np.zeros((), T.GpuMaterial)
    t = _KINDS[kind]
    if t == T.MAT_LAMBERTIAN:
        m["type_pad"] = (t, layer, 0, 0)
        m["albedo_fuzz"] = (albedo[0], albedo[1], albedo[2], 0.0)
        m["ior_pad"] = (1.0, 0.0, 0.0, 0.0)
    elif t == T.MAT_METAL:
        m["type_pad"] = (t, layer, 0, 0)
        m["albedo_fuzz"] = (albedo[0], albedo[1], albedo[2], F(0.0) if fuzz is None else fuzz)
        m["ior_pad"] = (1.0, 0.0, 0.0, 0.0)
    elif t == T.MAT_DIELECTRIC:
        m["type_pad"] = (t, 0, 0, 0)
        m["albedo_fuzz"] = (1.0, 1.0, 1.0, 0.0)
        m["ior_pad"] = (F(1.5) if ior is None else ior, 0.0, 0.0, 0.0)
    else:
        m["type_pad"] = (t, 0, 0, 0)
        m["albedo_fuzz"] = (albedo[0], albedo[1], albedo[2], 0.0)
        m["ior_pad"] = (F(1.0) if strength is None else strength, 0.0, 0.0, 0.0)
    return m


# ---------------------------------------------------------------------------
# Rotation and transform, src/scene.rs:387-495 (FP32, reference order)
# ---------------------------------------------------------------------------
def _sinf(x: np.float32) -> np.float32:
    return F(math.sin(float(x)))


def _cosf(x: np.float32) -> np.float32:
    return F(math.cos(float(x)))


def to_rad(deg) -> np.float32:
    return F(F(deg) * F(math.pi)) / F(180.0)


def euler_to_mat3(rx, ry, rz):
    ax, ay, az = to_rad(rx), to_rad(ry), to_rad(rz)
    sx, cx = _sinf(ax), _cosf(ax)
    sy, cy = _sinf(ay), _cosf(ay)
    sz, cz = _sinf(az), _cosf(az)
    return [
        [cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
        [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
        [-sy, sx * cy, cx * cy],
    ]


def rot3(m, v):
    return [
        m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2],
        m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
        m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2],
    ]


@dataclass
class Transform:
    translate: Optional[List[np.float32]] = None
    rotate: Optional[List[np.float32]] = None
    scale: Optional[np.float32] = None

    @staticmethod
    def parse(d, what: str) -> "Transform":
        if d is None:
            return Transform()
        if not isinstance(d, dict):
            raise SceneError("Parse", f"{what}.transform: expected a mapping")
        return Transform(_opt_vec(d, "translate", 3, what + ".transform"),
                         _opt_vec(d, "rotate", 3, what + ".transform"),
                         _opt_num(d, "scale", what + ".transform"))

    def translation(self):
        return self.translate or [F(0.0), F(0.0), F(0.0)]

    def rotation_matrix(self):
        return None if self.rotate is None else euler_to_mat3(*self.rotate)

    def apply_to_normal(self, n):
        m = self.rotation_matrix()
        return rot3(m, n) if m is not None else list(n)


def _normalize3(v):
    # plane_normalize3, src/scene.rs:510-517: f32 sum, f32 sqrt, f32 divides.
    ln = F(np.sqrt(F(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
    if ln > F(1e-9):
        return [v[0] / ln, v[1] / ln, v[2] / ln]
    return [F(0.0), F(1.0), F(0.0)]


def _cross3(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def push_quad_face(center, normal, hw, hh, mat_idx, verts: list, tris: list):
    """src/scene.rs:523-590."""
    nx, ny, nz = normal
    a = [abs(nx), abs(ny), abs(nz)]
    if a[0] <= a[1] and a[0] <= a[2]:
        helper = [F(1.0), F(0.0), F(0.0)]
    elif a[1] <= a[2]:
        helper = [F(0.0), F(1.0), F(0.0)]
    else:
        helper = [F(0.0), F(0.0), F(1.0)]
    t1 = _normalize3(_cross3(helper, [nx, ny, nz]))
    t2 = _cross3([nx, ny, nz], t1)
    cx, cy, cz = center

    def corner(s1, s2):
        s1, s2 = F(s1), F(s2)
        return (cx + t1[0] * s1 * hw + t2[0] * s2 * hh,
                cy + t1[1] * s1 * hw + t2[1] * s2 * hh,
                cz + t1[2] * s1 * hw + t2[2] * s2 * hh, F(1.0))

    base = len(verts)
    n4 = (nx, ny, nz, F(0.0))
    for p, uv in ((corner(-1, -1), (0, 0)), (corner(1, -1), (1, 0)), (corner(1, 1), (1, 1)),
                  (corner(-1, 1), (0, 1))):
        verts.append((p, n4, (uv[0], uv[1], 0.0, 0.0)))
    tris.append(((base, base + 1, base + 2), mat_idx))
    tris.append(((base, base + 2, base + 3), mat_idx))


def _box_faces(hx, hy, hz):
    """The six (outward normal, 4 CCW corners) of src/scene.rs:756-802."""
    return [
        ([1.0, 0.0, 0.0], [[hx, -hy, -hz], [hx, -hy, hz], [hx, hy, hz], [hx, hy, -hz]]),
        ([-1.0, 0.0, 0.0], [[-hx, -hy, hz], [-hx, -hy, -hz], [-hx, hy, -hz], [-hx, hy, hz]]),
        ([0.0, 1.0, 0.0], [[-hx, hy, -hz], [hx, hy, -hz], [hx, hy, hz], [-hx, hy, hz]]),
        ([0.0, -1.0, 0.0], [[hx, -hy, -hz], [-hx, -hy, -hz], [-hx, -hy, hz], [hx, -hy, hz]]),
        ([0.0, 0.0, 1.0], [[hx, -hy, hz], [-hx, -hy, hz], [-hx, hy, hz], [hx, hy, hz]]),
        ([0.0, 0.0, -1.0], [[-hx, -hy, -hz], [hx, -hy, -hz], [hx, hy, -hz], [-hx, hy, -hz]]),
    ]


# ---------------------------------------------------------------------------
# Loaded scene (src/scene.rs:311-351)
# ---------------------------------------------------------------------------
@dataclass
class CameraSettings:
    look_from: List[np.float32]
    look_at: List[np.float32]
    vfov: np.float32
    aperture: np.float32
    focus_dist: np.float32


@dataclass
class LoadedScene:
    spheres: np.ndarray
    materials: np.ndarray  # () GpuMaterialData
    mesh_vertices: np.ndarray
    mesh_triangles: np.ndarray
    camera: Optional[CameraSettings]
    emissive_spheres: np.ndarray
    world_config: Optional[dict]
    env_map_path: Optional[str]
    material_names: List[str] = field(default_factory=list)
    name_to_idx: Dict[str, int] = field(default_factory=dict)


def _vertices_array(verts: list) -> np.ndarray:
    a = np.zeros(len(verts), T.GpuVertex)
    for i, (p, n, uv) in enumerate(verts):
        a[i]["position"] = p
        a[i]["normal"] = n
        a[i]["uv"] = uv
    return a


def _triangles_array(tris: list) -> np.ndarray:
    a = np.zeros(len(tris), T.GpuTriangle)
    for i, (v, m) in enumerate(tris):
        a[i]["v"] = v
        a[i]["mat_idx"] = m
    return a


def load_scene_from_str(text: str, source: str = "<inline>",
                        extra_mesh: Optional[tuple] = None) -> LoadedScene:
    """Parse a scene YAML string (src/scene.rs:640-978).

    ``extra_mesh = (vertices, triangles)`` appends generated geometry (used by
    the synthetic benchmark scenes) after the YAML objects, exactly as a
    ``mesh`` object's merge would (src/scene.rs:865-878).
    """
    try:
        doc = yaml.load(text, Loader=_UniqueKeyLoader)
    except SceneError as e:
        raise SceneError("Parse", f"failed to parse '{source}': {e}") from None
    except yaml.YAMLError as e:
        raise SceneError("Parse", f"failed to parse '{source}': {e}") from None
    if doc is None:
        doc = {}
    if not isinstance(doc, dict):
        raise SceneError("Parse", f"failed to parse '{source}': top level must be a mapping")

    name_to_idx: Dict[str, int] = {}
    gpu_mats: List[np.ndarray] = []
    for sect in ("world_materials", "materials"):
        pal = doc.get(sect) or {}
        if not isinstance(pal, dict):
            raise SceneError("Parse", f"failed to parse '{source}': {sect} must be a mapping")
        for name, desc in pal.items():
            if name in name_to_idx:
                continue
            gpu = material_desc_to_gpu(desc, f"{sect}.{name}")
            name_to_idx[str(name)] = len(gpu_mats)
            gpu_mats.append(gpu)

    def resolve(mref, what):
        if isinstance(mref, str):
            if mref not in name_to_idx:
                raise SceneError("UnknownMaterial", f"unknown material '{mref}'")
            return name_to_idx[mref]
        if isinstance(mref, dict):
            idx = len(gpu_mats)
            gpu_mats.append(material_desc_to_gpu(mref, what))
            return idx
        raise SceneError("Parse", f"{what}: material must be a name or a mapping")

    spheres = []
    verts: list = []
    tris: list = []
    objects = doc.get("objects") or []
    if not isinstance(objects, list):
        raise SceneError("Parse", f"failed to parse '{source}': objects must be a sequence")
    for oi, obj in enumerate(objects):
        what = f"objects[{oi}]"
        if not isinstance(obj, dict) or "type" not in obj:
            raise SceneError("Parse", f"{what}: missing field 'type'")
        kind = obj["type"]
        for req in {"sphere": ("center", "radius", "material"), "plane": ("normal", "half_extents", "material"),
                    "box": ("half_extents", "material"), "mesh": ("path", "material")}.get(kind, ()):
            if req not in obj:
                raise SceneError("Parse", f"{what}: missing field '{req}'")
        if kind == "sphere":
            c = _vec(obj["center"], 3, what + ".center")
            r = _num(obj["radius"], what + ".radius")
            mi = resolve(obj["material"], what)
            s = np.zeros((), T.GpuSphere)
            s["center_r"] = (c[0], c[1], c[2], r)
            s["mat_and_pad"] = (mi, 0, 0, 0)
            spheres.append(s)
        elif kind == "plane":
            normal = _vec(obj["normal"], 3, what + ".normal")
            hw, hh = _vec(obj["half_extents"], 2, what + ".half_extents")
            tr = Transform.parse(obj.get("transform"), what)
            cull = obj.get("backface_culling", False)
            if not isinstance(cull, bool):
                raise SceneError("Parse", f"{what}.backface_culling: expected a boolean")
            mi = resolve(obj["material"], what)
            mat_idx = (mi | T.BACKFACE_CULL_FLAG) if cull else mi
            push_quad_face(tr.translation(), tr.apply_to_normal(normal), hw, hh, mat_idx, verts, tris)
        elif kind == "box":
            hx, hy, hz = _vec(obj["half_extents"], 3, what + ".half_extents")
            tr = Transform.parse(obj.get("transform"), what)
            mi = resolve(obj["material"], what)
            cx, cy, cz = tr.translation()
            rot = tr.rotation_matrix() or [[F(1), F(0), F(0)], [F(0), F(1), F(0)], [F(0), F(0), F(1)]]
            uvs = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
            for ln, lc in _box_faces(hx, hy, hz):
                rn = rot3(rot, [F(x) for x in ln])
                n4 = (rn[0], rn[1], rn[2], F(0.0))
                base = len(verts)
                for i, cc in enumerate(lc):
                    rc = rot3(rot, [F(x) for x in cc])
                    verts.append(((cx + rc[0], cy + rc[1], cz + rc[2], F(1.0)), n4,
                                  (uvs[i][0], uvs[i][1], 0.0, 0.0)))
                tris.append(((base, base + 1, base + 2), mi))
                tris.append(((base, base + 2, base + 3), mi))
        elif kind == "mesh":
            resolve(obj["material"], what)
            raise SceneError("MeshLoad", f"failed to load mesh '{obj['path']}': OBJ meshes are "
                                         "not supported by this loader")
        else:
            raise SceneError("Parse", f"{what}: unknown object type {kind!r}")

    mesh_vertices = _vertices_array(verts)
    mesh_triangles = _triangles_array(tris)
    if extra_mesh is not None:
        ev, et = extra_mesh
        et = et.copy()
        et["v"] += len(mesh_vertices)
        mesh_vertices = np.concatenate([mesh_vertices, ev.astype(T.GpuVertex)])
        mesh_triangles = np.concatenate([mesh_triangles, et])
    if len(mesh_vertices) == 0:  # stub, src/scene.rs:887-899
        mesh_vertices = np.zeros(1, T.GpuVertex)
        mesh_vertices[0]["position"] = (1.0e9, 1.0e9, 1.0e9, 1.0)
        mesh_vertices[0]["normal"] = (0.0, 1.0, 0.0, 0.0)
        mesh_triangles = np.zeros(1, T.GpuTriangle)

    if len(gpu_mats) > T.MAX_MATERIALS:
        raise SceneError("TooManyMaterials",
                         f"scene uses {len(gpu_mats)} materials but limit is {T.MAX_MATERIALS}")
    mat_data = np.zeros((), T.GpuMaterialData)
    mat_data["mat_count"] = len(gpu_mats)
    for i, m in enumerate(gpu_mats):
        mat_data["materials"][i] = m

    sph = np.array(spheres, T.GpuSphere) if spheres else np.zeros(0, T.GpuSphere)
    emissive = np.array(
        [s for s in spheres if s["mat_and_pad"][0] < len(gpu_mats)
         and gpu_mats[s["mat_and_pad"][0]]["type_pad"][0] == T.MAT_EMISSIVE], T.GpuSphere
    ) if spheres else np.zeros(0, T.GpuSphere)
    mat_data["n_emissive"] = len(emissive)

    cam = None
    if doc.get("camera") is not None:
        cd = doc["camera"]
        if not isinstance(cd, dict) or "look_from" not in cd or "look_at" not in cd:
            raise SceneError("Parse", "camera: look_from and look_at are required")
        lf = _vec(cd["look_from"], 3, "camera.look_from")
        la = _vec(cd["look_at"], 3, "camera.look_at")
        d = [lf[0] - la[0], lf[1] - la[1], lf[2] - la[2]]
        dist = F(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
        vfov = _opt_num(cd, "vfov", "camera")
        ap = _opt_num(cd, "aperture", "camera")
        fd = _opt_num(cd, "focus_dist", "camera")
        cam = CameraSettings(lf, la, F(60.0) if vfov is None else vfov,
                             F(0.0) if ap is None else ap, dist if fd is None else fd)

    world = None
    if doc.get("world") is not None:
        w = doc["world"]
        world = {"seed": int(w.get("seed", 0)), "view_distance": int(w.get("view_distance", 4)),
                 "voxels_per_block": int(w.get("voxels_per_block", 8))}

    names = [f"inline-{i}" for i in range(len(gpu_mats))]
    for n, i in name_to_idx.items():
        names[i] = n
    env_map = doc.get("env_map")
    return LoadedScene(sph, mat_data, mesh_vertices, mesh_triangles, cam, emissive, world,
                       None if env_map is None else str(env_map), names, dict(name_to_idx))


def load_scene_from_yaml(path: str, extra_mesh: Optional[tuple] = None) -> LoadedScene:
    """src/scene.rs:993-999."""
    try:
        with open(path, "rb") as f:
            raw = f.read()
    except OSError as e:
        raise SceneError("Io", f"I/O error: {e}") from None
    try:
        text = raw.decode("utf-8")
    except UnicodeDecodeError as e:
        raise SceneError("Io", f"I/O error: {e}") from None
    return load_scene_from_str(text, path, extra_mesh)


# ---------------------------------------------------------------------------
# build_scene_buffers, src/gpu_layout.rs:259-390
# ---------------------------------------------------------------------------
@dataclass
class SceneBuffers:
    spheres: np.ndarray      # BVH order, or [STUB_SPHERE]
    sphere_orig: np.ndarray
    sphere_bvh: np.ndarray   # or [zeroed node]
    vertices: np.ndarray
    triangles: np.ndarray    # BVH leaf order
    tri_orig: np.ndarray
    mesh_bvh: np.ndarray
    materials: np.ndarray    # () GpuMaterialData
    env: np.ndarray          # (H, W, 4) float32
    emissive_spheres: np.ndarray
    n_real_spheres: int = 0
    n_real_tris: int = 0


def build_scene_buffers(loaded: LoadedScene, env: Optional[np.ndarray] = None,
                        cwd: Optional[str] = None) -> SceneBuffers:
    from ._lib import build_mesh_bvh, build_sphere_bvh

    if len(loaded.spheres):
        snodes, sord, sorig = build_sphere_bvh(loaded.spheres)
    else:
        snodes = np.zeros(1, T.GpuBvhNode)
        sord = np.array([T.STUB_SPHERE], T.GpuSphere)
        sorig = np.zeros(1, np.uint32)
    mnodes, mord, morig = build_mesh_bvh(loaded.mesh_vertices, loaded.mesh_triangles)
    if len(mnodes) == 0:
        mnodes = np.zeros(1, T.GpuBvhNode)
        mord = np.zeros(1, T.GpuTriangle)
        morig = np.zeros(1, np.uint32)
    if env is None:
        env = texture.load_env_map_data(loaded.env_map_path, cwd=cwd)
    emissive = loaded.emissive_spheres if len(loaded.emissive_spheres) else np.array(
        [T.STUB_SPHERE], T.GpuSphere)
    return SceneBuffers(sord, sorig, snodes, loaded.mesh_vertices, mord, morig, mnodes,
                        loaded.materials, np.ascontiguousarray(env, np.float32), emissive,
                        len(loaded.spheres), len(loaded.mesh_triangles))
