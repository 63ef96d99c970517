"""CPU oracle of the render hot path — TEST INFRASTRUCTURE ONLY.

A plain-C FP32 restatement of src/shaders/path_trace.wgsl and
src/shaders/denoise.wgsl (``mcrt_oracle.c``) plus an FP64 converged-image
oracle (``oracle_fp64.c``), with ctypes bindings.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import this package.  The product (``mcrt_amd``) never does.

The reference itself (Rust + wgpu + WGSL) cannot be built here (no cargo,
no Vulkan ICD), so there is no ``oracle/_ref``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
LIB = os.path.join(BUILD, "liboracle_mcrt.so")
SOURCES = ["mcrt_oracle.c", "oracle_fp64.c"]
CFLAGS = ["-O3", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-pthread", "-std=c11"]


def build() -> str:
    """Compile the oracle with gcc (FP32 steps round as written)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", *CFLAGS, "-o", tmp, *srcs, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


class _Scene(C.Structure):
    _fields_ = [
        ("spheres", C.c_void_p), ("n_spheres", C.c_uint32), ("sphere_orig", C.c_void_p),
        ("sphere_bvh", C.c_void_p), ("n_sphere_nodes", C.c_uint32),
        ("verts", C.c_void_p), ("n_verts", C.c_uint32),
        ("tris", C.c_void_p), ("n_tris", C.c_uint32), ("tri_orig", C.c_void_p),
        ("mesh_bvh", C.c_void_p), ("n_mesh_nodes", C.c_uint32),
        ("mats", C.c_void_p), ("env", C.c_void_p), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
    ]


class Counts(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("tris", C.c_uint64), ("spheres", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        L.oracle_pcg_hash.restype = u32
        L.oracle_pcg_hash.argtypes = [u32]
        L.oracle_rand_f32.restype = f32
        L.oracle_rand_f32.argtypes = [C.POINTER(u32)]
        L.oracle_pixel_seed.restype = u32
        L.oracle_pixel_seed.argtypes = [u32, u32, u32, u32]
        L.oracle_ray_triangle.argtypes = [vp] * 5 + [f32, f32, vp]
        L.oracle_ray_aabb.restype = C.c_int
        L.oracle_ray_aabb.argtypes = [vp] * 4 + [f32, f32]
        L.oracle_trace_batch.argtypes = [vp, vp, vp, C.c_uint64, f32, f32, C.c_int, vp, vp, vp, C.c_int]
        L.oracle_render.argtypes = [vp, vp, u32, u32, u32, u32, u32, vp, vp, vp, vp, vp, C.c_int]
        L.oracle_sample.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32, vp, vp]
        L.oracle_denoise.argtypes = [vp, u32, u32, vp, vp, vp, C.c_int]
        L.oracle_tonemap_u8.restype = C.c_uint8
        L.oracle_tonemap_u8.argtypes = [f32]
        L.oracle_gradient_env.argtypes = [u32, u32, vp]
        if hasattr(L, "oracle64_render"):
            L.oracle64_render.argtypes = [vp, vp, u32, u32, u32, u32, u32, vp, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def default_threads() -> int:
    return max(1, min(os.cpu_count() or 1, 16))


class Scene:
    """Holds numpy buffers of a SceneBuffers-like object and the C view."""

    def __init__(self, sb):
        self.sb = sb
        self._keep = [np.ascontiguousarray(x) for x in (
            sb.spheres, sb.sphere_orig, sb.sphere_bvh, sb.vertices, sb.triangles, sb.tri_orig,
            sb.mesh_bvh, sb.materials, sb.env)]
        sp, so, sn, v, t, to, mn, m, e = self._keep
        self.c = _Scene(_p(sp), len(sp), _p(so), _p(sn), len(sn), _p(v), len(v), _p(t), len(t),
                        _p(to), _p(mn), len(mn), _p(m), _p(e), e.shape[1], e.shape[0])

    def ref(self):
        return C.byref(self.c)


def trace(scene: Scene, origins, dirs, tmin=1e-4, tmax=1e9, mode="exact", threads=None,
          counts=False):
    L = lib()
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = len(o)
    ids = np.empty(n, np.uint32)
    ts = np.empty(n, np.float32)
    cnt = Counts()
    m = {"ref": 0, "exact": 1, "brute": 2}[mode]
    L.oracle_trace_batch(scene.ref(), _p(o), _p(d), n, tmin, tmax, m, _p(ids), _p(ts),
                         C.byref(cnt) if counts else None, threads or default_threads())
    if counts:
        return ids, ts, (cnt.nodes, cnt.tris, cnt.spheres)
    return ids, ts


def render(scene: Scene, camera, width, height, spp, first_frame=0, max_depth=8, accum_in=None,
           threads=None, counts=False):
    L = lib()
    cam = np.ascontiguousarray(camera)
    out = np.empty((height, width, 4), np.float32)
    gb = np.empty((height, width, 4), np.float32)
    ain = None if accum_in is None else np.ascontiguousarray(accum_in, np.float32)
    cnt = Counts()
    seg = C.c_uint64(0)
    L.oracle_render(scene.ref(), _p(cam), width, height, spp, first_frame, max_depth, _p(ain),
                    _p(out), _p(gb), C.byref(cnt) if counts else None,
                    C.byref(seg) if counts else None, threads or default_threads())
    if counts:
        return out, gb, (cnt.nodes, cnt.tris, cnt.spheres, seg.value)
    return out, gb


def render_nee(scene: Scene, camera, width, height, spp, first_frame=0, max_depth=8, threads=None):
    """oracle_render with next-event estimation + MIS; (accum, gbuf)."""
    L = lib()
    L.oracle_render_nee.argtypes = [C.c_void_p, C.c_void_p] + [C.c_uint32] * 5 + [C.c_void_p] * 3 + [C.c_int]
    cam = np.ascontiguousarray(camera)
    out = np.empty((height, width, 4), np.float32)
    gb = np.empty((height, width, 4), np.float32)
    L.oracle_render_nee(scene.ref(), _p(cam), width, height, spp, first_frame, max_depth, None, _p(out),
                        _p(gb), threads or default_threads())
    return out, gb


def sample(scene: Scene, camera, width, height, x, y, frame, max_depth=8):
    L = lib()
    cam = np.ascontiguousarray(camera)
    rgb = np.empty(3, np.float32)
    gb = np.empty(4, np.float32)
    L.oracle_sample(scene.ref(), _p(cam), width, height, x, y, frame, max_depth, _p(rgb), _p(gb))
    return rgb, gb


def denoise(accum, gbuf, params, threads=None):
    L = lib()
    a = np.ascontiguousarray(accum, np.float32)
    g = np.ascontiguousarray(gbuf, np.float32)
    p = np.ascontiguousarray(params)
    out = np.empty_like(a)
    h, w = a.shape[:2]
    L.oracle_denoise(_p(p), w, h, _p(a), _p(g), _p(out), threads or default_threads())
    return out


def tonemap_u8(v: float) -> int:
    return int(lib().oracle_tonemap_u8(float(v)))


def gradient_env(w=1024, h=512):
    out = np.empty((h, w, 4), np.float32)
    lib().oracle_gradient_env(w, h, _p(out))
    return out


def sample_env(scene: Scene, direction):
    """Environment radiance the shader looks up for a direction (rgb)."""
    L = lib()
    L.oracle_sample_env.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    d = np.ascontiguousarray(direction, np.float32)
    out = np.empty(3, np.float32)
    L.oracle_sample_env(scene.ref(), _p(d), _p(out))
    return out


def pcg_hash(v: int) -> int:
    return int(lib().oracle_pcg_hash(v & 0xFFFFFFFF))


def pixel_seed(x, y, width, frame) -> int:
    return int(lib().oracle_pixel_seed(x, y, width, frame))


def rand_f32(state: int):
    s = C.c_uint32(state)
    f = lib().oracle_rand_f32(C.byref(s))
    return float(np.float32(f)), int(s.value)


def ray_triangle(o, d, v0, v1, v2, tmin=1e-4, tmax=1e9):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    out = np.empty(4, np.float32)
    lib().oracle_ray_triangle(*[_p(a) for a in arrs], tmin, tmax, _p(out))
    return out


def diff_render64(scene: Scene, camera, width, height, spp, first_frame, max_depth, params, pmap,
                  threads=None):
    """FP64 image (H, W, 3) of the extended estimator for params (n, 4)."""
    L = lib()
    L.oracle_diff_render64.argtypes = [C.c_void_p, C.c_void_p] + [C.c_uint32] * 5 + [
        C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int]
    cam = np.ascontiguousarray(camera)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 4)
    m = np.ascontiguousarray(pmap, np.int32)
    out = np.empty((height, width, 3), np.float64)
    L.oracle_diff_render64(scene.ref(), _p(cam), width, height, spp, first_frame, max_depth, _p(p),
                           len(p), _p(m), _p(out), threads or default_threads())
    return out


def diff_grad64(scene: Scene, camera, width, height, spp, first_frame, max_depth, params, pmap,
                d_image, threads=None):
    """Gradient (n, 4) of sum(d_image * image) w.r.t. params, FP64."""
    L = lib()
    L.oracle_diff_grad64.argtypes = [C.c_void_p, C.c_void_p] + [C.c_uint32] * 5 + [
        C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    cam = np.ascontiguousarray(camera)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 4)
    m = np.ascontiguousarray(pmap, np.int32)
    dimg = np.ascontiguousarray(np.asarray(d_image)[..., :3], np.float64)
    g = np.empty((len(p), 4), np.float64)
    L.oracle_diff_grad64(scene.ref(), _p(cam), width, height, spp, first_frame, max_depth, _p(p),
                         len(p), _p(m), _p(dimg), _p(g), threads or default_threads())
    return g


def reference64_window(scene: Scene, camera, width, height, x0, y0, cw, ch, spp, first_frame,
                       max_depth, threads=None):
    """Converged FP64-accumulated reference of a (ch, cw) window, (ch, cw, 3)."""
    L = lib()
    L.oracle_reference64_window.argtypes = [C.c_void_p, C.c_void_p] + [C.c_uint32] * 9 + [
        C.c_void_p, C.c_int]
    cam = np.ascontiguousarray(camera)
    out = np.empty((ch, cw, 3), np.float64)
    L.oracle_reference64_window(scene.ref(), _p(cam), width, height, x0, y0, cw, ch, spp, first_frame,
                                max_depth, _p(out), threads or default_threads())
    return out


def scatter(material, rd, n, front: bool, rng: int):
    """One scatter event; returns (dir, att, absorbed, rng_out)."""
    L = lib()
    L.oracle_scatter.argtypes = [C.c_void_p] * 3 + [C.c_int, C.POINTER(C.c_uint32)] + [C.c_void_p] * 3
    m = np.ascontiguousarray(material)
    a = np.ascontiguousarray(rd, np.float32)
    b = np.ascontiguousarray(n, np.float32)
    d = np.empty(3, np.float32)
    att = np.empty(3, np.float32)
    ab = C.c_int(0)
    s = C.c_uint32(rng)
    L.oracle_scatter(_p(m), _p(a), _p(b), 1 if front else 0, C.byref(s), _p(d), _p(att), C.byref(ab))
    return d, att, bool(ab.value), int(s.value)


def ray_aabb(o, d, bmin, bmax, tmin=1e-4, tmax=1e9) -> bool:
    arrs = [np.ascontiguousarray(x, np.float32) for x in (o, d, bmin, bmax)]
    return bool(lib().oracle_ray_aabb(*[_p(a) for a in arrs], tmin, tmax))
