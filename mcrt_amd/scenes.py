"""Synthetic benchmark scenes, all seeded and deterministic.

* ``cornell_yaml`` — the Cornell box in the reference's YAML schema (layout of
  assets/cornell-box.yaml:1-151): 5 back-face-culled walls (10 triangles),
  an axis-aligned box and a box rotated 15 degrees (24 triangles), one
  emissive quad of strength 15 (2 triangles): 36 triangles.
* ``displaced_icosphere`` — an icosphere of 20*n^2 faces (n = 59 gives
  69,620), vertices pushed radially by 3-octave value noise of amplitude 0.08,
  area-weighted vertex normals, one material slot per base face.
* ``microbench_rays`` — origins uniform in a box, directions uniform on the
  sphere (z = 1 - 2u, phi = 2 pi v), from a counter-based hash.
"""
from __future__ import annotations

import numpy as np

from . import types as T

CORNELL_TEMPLATE = """\
camera:
  look_from: [0.0, 4.75, 8.0]
  look_at:   [0.0, 4.0, 0.0]
  vfov:      60.0
  aperture:  0.0

materials:
  white:
    type: lambertian
    albedo: [0.73, 0.73, 0.73]
  red:
    type: lambertian
    albedo: [0.65, 0.05, 0.05]
  green:
    type: lambertian
    albedo: [0.12, 0.45, 0.15]
  light:
    type: emissive
    albedo: [1.0, 1.0, 1.0]
    emission_strength: 15.0
{extra_materials}
objects:
  - type: plane
    normal: [0.0, 1.0, 0.0]
    half_extents: [5, 5]
    material: white
    backface_culling: true
    transform:
      translate: [0.0, 0.0, 0.0]
  - type: plane
    normal: [0.0, -1.0, 0.0]
    half_extents: [5, 5]
    material: white
    backface_culling: true
    transform:
      translate: [0.0, 10.0, 0.0]
  - type: plane
    normal: [0.0, 0.0, 1.0]
    half_extents: [5, 5]
    material: white
    backface_culling: true
    transform:
      translate: [0.0, 5.0, -5.0]
  - type: plane
    normal: [1.0, 0.0, 0.0]
    half_extents: [5, 5]
    material: red
    backface_culling: true
    transform:
      translate: [-5.0, 5.0, 0.0]
  - type: plane
    normal: [-1.0, 0.0, 0.0]
    half_extents: [5, 5]
    material: green
    backface_culling: true
    transform:
      translate: [5.0, 5.0, 0.0]
  - type: plane
    normal: [0.0, -1.0, 0.0]
    half_extents: [1.5, 1.5]
    material: light
    transform:
      translate: [0.0, 9.9, -2.5]
{boxes}"""

CORNELL_BOXES = """\
  - type: box
    half_extents: [1.2, 2.0, 1.2]
    material: white
    transform:
      translate: [-2.2, 2.0, -3.2]
  - type: box
    half_extents: [1.0, 1.4, 1.0]
    material: white
    transform:
      translate: [2.2, 1.4, -0.5]
      rotate: [0.0, 15.0, 0.0]
"""


def cornell_yaml(with_boxes: bool = True, extra_materials: str = "") -> str:
    return CORNELL_TEMPLATE.format(boxes=CORNELL_BOXES if with_boxes else "",
                                   extra_materials=extra_materials)


# ---------------------------------------------------------------------------
# Counter-based hashing (seeded, platform independent)
# ---------------------------------------------------------------------------
def _mix32(x: np.ndarray) -> np.ndarray:
    """lowbias32 integer hash on uint32 arrays (multiplications wrap mod 2^32)."""
    x = np.atleast_1d(np.asarray(x)).astype(np.uint32)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint32(16)
        x = (x * np.uint32(0x7FEB352D)).astype(np.uint32)
        x ^= x >> np.uint32(15)
        x = (x * np.uint32(0x846CA68B)).astype(np.uint32)
        x ^= x >> np.uint32(16)
    return x


def _u01(counter: np.ndarray, seed: int, stream: int) -> np.ndarray:
    h = _mix32(counter.astype(np.uint32) ^ _mix32(np.uint32((seed * 0x9E3779B1 + stream * 0x85EBCA6B) & 0xFFFFFFFF)))
    h = _mix32(h + np.uint32(stream))
    return (h >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))


def _lattice_value(ix, iy, iz, seed: int) -> np.ndarray:
    h = _mix32(ix.astype(np.uint32) * np.uint32(0x8DA6B343)
               ^ iy.astype(np.uint32) * np.uint32(0xD8163841)
               ^ iz.astype(np.uint32) * np.uint32(0xCB1AB31F)
               ^ np.uint32(seed & 0xFFFFFFFF))
    return (h >> np.uint32(8)).astype(np.float64) * (2.0 / (1 << 24)) - 1.0


def value_noise3(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise with smoothstep fade, values in [-1, 1]."""
    f = np.floor(p)
    t = p - f
    t = t * t * (3.0 - 2.0 * t)
    i = f.astype(np.int64)
    acc = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1]) * \
                    (t[:, 2] if dz else 1 - t[:, 2])
                acc = acc + w * _lattice_value(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return acc


def fbm3(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 2.0) -> np.ndarray:
    acc = np.zeros(len(p))
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        acc += amp * value_noise3(p * freq, seed + 1013 * o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return acc / norm


# ---------------------------------------------------------------------------
# Displaced icosphere
# ---------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def displaced_icosphere(n: int = 59, center=(0.0, 3.0, -1.5), radius: float = 2.0,
                        amplitude: float = 0.08, seed: int = 7, mat_base: int = 0, n_mats: int = 8):
    """Returns (GpuVertex array, GpuTriangle array, base-face index per tri)."""
    V, Fc = _icosahedron()
    # Lattice points of each base face: (i, j) with i + j <= n.
    ij = [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]
    ij = np.array(ij)
    idx_of = {tuple(x): k for k, x in enumerate(ij)}
    tri_local = []
    for i in range(n):
        for j in range(n - i):
            a, b, c = idx_of[(i, j)], idx_of[(i + 1, j)], idx_of[(i, j + 1)]
            tri_local.append((a, b, c))
            if j < n - i - 1:
                d = idx_of[(i + 1, j + 1)]
                tri_local.append((b, d, c))
    tri_local = np.array(tri_local)
    pts, tris, face_of = [], [], []
    for fi, (a, b, c) in enumerate(Fc):
        wa = 1.0 - (ij[:, 0] + ij[:, 1]) / n
        wb = ij[:, 0] / n
        wc = ij[:, 1] / n
        p = wa[:, None] * V[a] + wb[:, None] * V[b] + wc[:, None] * V[c]
        base = len(ij) * fi
        pts.append(p)
        tris.append(tri_local + base)
        face_of.append(np.full(len(tri_local), fi))
    P = np.concatenate(pts)
    Tr = np.concatenate(tris)
    face_of = np.concatenate(face_of)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    # Merge duplicated edge/corner points.
    key = np.round(P * 1e6).astype(np.int64)
    _, uniq, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    P = P[uniq]
    Tr = inv.reshape(-1)[Tr]
    disp = 1.0 + amplitude * fbm3(P, seed)
    pos = P * (radius * disp)[:, None] + np.asarray(center)[None, :]
    # Orient every triangle outward, then area-weighted vertex normals.
    e1 = pos[Tr[:, 1]] - pos[Tr[:, 0]]
    e2 = pos[Tr[:, 2]] - pos[Tr[:, 0]]
    fn = np.cross(e1, e2)
    outward = np.einsum("ij,ij->i", fn, pos[Tr[:, 0]] - np.asarray(center)) > 0
    Tr[~outward] = Tr[~outward][:, [0, 2, 1]]
    fn[~outward] *= -1
    vn = np.zeros_like(pos)
    for k in range(3):
        np.add.at(vn, Tr[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    verts = np.zeros(len(pos), T.GpuVertex)
    verts["position"][:, :3] = pos.astype(np.float32)
    verts["position"][:, 3] = 1.0
    verts["normal"][:, :3] = vn.astype(np.float32)
    tris = np.zeros(len(Tr), T.GpuTriangle)
    tris["v"] = Tr.astype(np.uint32)
    tris["mat_idx"] = (mat_base + face_of % n_mats).astype(np.uint32)
    return verts, tris, face_of


def random_materials_yaml(n: int, seed: int, prefix: str = "m") -> str:
    """n palette entries drawn by seed: lambertian (albedo U[0.1,0.9]), metal
    (fuzz U[0,0.5]) or dielectric (ior 1.5)."""
    c = np.arange(n * 8, dtype=np.uint32)
    u = _u01(c, seed, 11).reshape(n, 8)
    out = []
    for i in range(n):
        kind = int(u[i, 0] * 3)
        alb = [0.1 + 0.8 * u[i, 1], 0.1 + 0.8 * u[i, 2], 0.1 + 0.8 * u[i, 3]]
        name = f"{prefix}{i}"
        if kind == 0:
            out.append(f"  {name}:\n    type: lambertian\n    albedo: [{alb[0]:.6f}, {alb[1]:.6f}, {alb[2]:.6f}]\n")
        elif kind == 1:
            out.append(f"  {name}:\n    type: metal\n    albedo: [{alb[0]:.6f}, {alb[1]:.6f}, {alb[2]:.6f}]\n"
                       f"    fuzz: {0.5 * u[i, 4]:.6f}\n")
        else:
            out.append(f"  {name}:\n    type: dielectric\n    ior: 1.5\n")
    return "".join(out)


def microbench_rays(n: int, lo, hi, seed: int = 1234, start: int = 0):
    """Origins U(box), directions uniform on S^2; float32 (n, 3) each."""
    c = np.arange(start, start + n, dtype=np.uint64).astype(np.uint32)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    o = np.empty((n, 3), np.float32)
    for k in range(3):
        o[:, k] = (lo[k] + (hi[k] - lo[k]) * _u01(c, seed, k)).astype(np.float32)
    z = 1.0 - 2.0 * _u01(c, seed, 3)
    phi = 2.0 * np.pi * _u01(c, seed, 4)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)
    return o, d
