"""ctypes binding of ``libmcrt_hip.so`` (the C-ABI of ``include/mcrt_abi.h``).

The library is built in-tree by :func:`mcrt_amd.build.build_library`.  There
is no fallback: if the library is missing or cannot be loaded every entry
point raises :class:`McrtError`.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import types as T

_HERE = os.path.dirname(os.path.abspath(__file__))
# MCRT_HIP_LIB selects another build of the library (e.g. to compare two).
LIB_PATH = os.environ.get("MCRT_HIP_LIB") or os.path.join(_HERE, "libmcrt_hip.so")

MCRT_OK = 0
_ERR_NAMES = {
    -1: "NO_DEVICE",
    -2: "BAD_ARG",
    -3: "OOM",
    -4: "HIP_ERROR",
    -5: "LIMIT",
    -6: "NO_SCENE",
}


class McrtError(RuntimeError):
    """Error returned by the C-ABI (status code + per-context message)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {message}")
        self.code = code


_lib: Optional[C.CDLL] = None
_vp = C.c_void_p
_u32 = C.c_uint32
_f32 = C.c_float


def load() -> C.CDLL:
    """Load the library (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise McrtError(-4, f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    # One HIP runtime per process: torch bundles its own libamdhip64 (same
    # SONAME as ROCm's).  Loading torch first makes the library's dependency
    # resolve to that already-loaded runtime, so device pointers and streams
    # from torch tensors are shared with the library's kernels.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    sig = {
        "mcrt_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
        "mcrt_destroy": (None, [_vp]),
        "mcrt_last_error": (C.c_char_p, [_vp]),
        "mcrt_upload_scene": (
            C.c_int,
            [_vp, _vp, _u32, _vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32,
             _u32, _vp, _u32, _u32],
        ),
        "mcrt_render": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp, _vp]),
        "mcrt_render_dev": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp]),
        "mcrt_denoise_bilateral": (C.c_int, [_vp, _vp, _u32, _u32, _vp, _vp, _vp]),
        "mcrt_denoise_bilateral_dev": (C.c_int, [_vp, _vp, _u32, _u32, _vp, _vp, _vp, _vp]),
        "mcrt_trace_closest": (C.c_int, [_vp, _vp, _vp, _u32, _f32, _f32, _vp, _vp]),
        "mcrt_trace_closest_dev": (C.c_int, [_vp, _vp, _vp, _u32, _f32, _f32, _vp, _vp, _vp]),
        "mcrt_trace_closest_noalloc_dev": (C.c_int, [_vp, _vp, _vp, _u32, _f32, _f32, _vp, _vp, _vp]),
        "mcrt_build_mesh_bvh": (C.c_int, [_vp, _u32, _vp, _u32, _vp, C.POINTER(_u32), _vp, _vp]),
        "mcrt_build_sphere_bvh": (C.c_int, [_vp, _u32, _vp, C.POINTER(_u32), _vp, _vp]),
        "mcrt_validate_bvh": (C.c_int, [_vp, _u32, _u32]),
        "mcrt_set_profiling": (C.c_int, [_vp, C.c_int]),
        "mcrt_set_nee": (C.c_int, [_vp, C.c_int]),
        "mcrt_kernel_stats": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
        "mcrt_reset_stats": (C.c_int, [_vp]),
        "mcrt_render_counters": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "mcrt_measure_peaks": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "mcrt_diff_render": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _u32, _vp, _vp]),
        "mcrt_diff_render_dev": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _u32, _vp, _vp, _vp]),
        "mcrt_diff_adjoint": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _u32, _vp, _vp, _vp]),
        "mcrt_diff_adjoint_dev": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _u32, _vp, _vp, _vp,
                                            _vp]),
        "mcrt_diff_set_mode": (C.c_int, [_vp, C.c_int]),
        "mcrt_render_aux_dev": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
        "mcrt_denoiser_create": (C.c_int, [_vp, _u32, C.POINTER(_vp)]),
        "mcrt_denoiser_destroy": (None, [_vp]),
        "mcrt_denoiser_param_count": (_u32, []),
        "mcrt_denoiser_get_weights": (C.c_int, [_vp, _vp]),
        "mcrt_denoiser_set_weights": (C.c_int, [_vp, _vp]),
        "mcrt_denoiser_set_loss": (C.c_int, [_vp, C.c_int]),
        "mcrt_denoiser_infer_dev": (C.c_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
        "mcrt_denoiser_train_step_dev": (C.c_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _f32, _vp, _vp]),
        "mcrt_denoiser_gradient": (C.c_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
        "mcrt_mesh_voxel_chunks": (
            C.c_int,
            [_vp, _u32, _u32, C.c_int32, C.c_int32, _f32, _vp, _vp, _vp, C.c_uint64, _vp, C.c_uint64,
             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        ),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list:
    return [
        "mcrt_create", "mcrt_destroy", "mcrt_last_error", "mcrt_upload_scene", "mcrt_render",
        "mcrt_render_dev", "mcrt_denoise_bilateral", "mcrt_denoise_bilateral_dev",
        "mcrt_trace_closest", "mcrt_trace_closest_dev", "mcrt_trace_closest_noalloc_dev", "mcrt_build_mesh_bvh",
        "mcrt_build_sphere_bvh", "mcrt_validate_bvh", "mcrt_set_profiling", "mcrt_set_nee", "mcrt_kernel_stats", "mcrt_reset_stats",
        "mcrt_render_counters", "mcrt_measure_peaks", "mcrt_mesh_voxel_chunks", "mcrt_diff_render", "mcrt_diff_render_dev",
        "mcrt_diff_adjoint", "mcrt_diff_adjoint_dev", "mcrt_diff_set_mode", "mcrt_render_aux_dev", "mcrt_denoiser_create",
        "mcrt_denoiser_destroy", "mcrt_denoiser_param_count", "mcrt_denoiser_get_weights",
        "mcrt_denoiser_set_weights", "mcrt_denoiser_infer_dev", "mcrt_denoiser_train_step_dev",
        "mcrt_denoiser_gradient", "mcrt_denoiser_set_loss",
    ]


def measure_peaks(device: int = 0, copy_bytes: int = 1 << 30):
    """(copy GB/s, BF16 MFMA TFLOP/s with 32x32x16, with 16x16x32, non-packed
    FP32 VALU TFLOP/s) measured on `device`."""
    lib = load()
    g, t, t16, v = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
    rc = lib.mcrt_measure_peaks(device, copy_bytes, C.byref(g), C.byref(t), C.byref(t16), C.byref(v))
    if rc != MCRT_OK:
        raise McrtError(rc, "mcrt_measure_peaks failed")
    return g.value, t.value, t16.value, v.value


def diff_tape_bytes(width: int, height: int, spp: int, max_depth: int) -> int:
    """Device memory of the tape adjoint (MCRT_DIFF_TAPE) for a render size:
    48 B per vertex for every thread of its grid (frame groups until about
    2^22 threads, 256-thread blocks)."""
    wh = width * height
    groups = max(1, min(spp, (1 << 22) // wh))
    grid = (wh * groups + 255) // 256
    return grid * 256 * max_depth * 48


def validate_bvh(nodes: np.ndarray, n_prims: int) -> None:
    """mcrt_validate_bvh: raises McrtError with the library's reason for a
    malformed reference pre-order tree."""
    lib = load()
    nodes = np.ascontiguousarray(nodes, T.GpuBvhNode)
    rc = lib.mcrt_validate_bvh(_ptr(nodes) if len(nodes) else None, len(nodes), n_prims)
    if rc != MCRT_OK:
        raise McrtError(rc, lib.mcrt_last_error(None).decode())


def mesh_voxel_chunks(blocks: np.ndarray, cx0: int, cz0: int, voxel_size: float,
                      opaque: np.ndarray, block_material: np.ndarray):
    """Greedy-mesh a (chunks_x, chunks_z, 16, 256, 16) uint16 block grid.

    Returns (GpuVertex array, GpuTriangle array)."""
    lib = load()
    blocks = np.ascontiguousarray(blocks, np.uint16)
    if blocks.ndim != 5 or blocks.shape[2:] != (16, 256, 16):
        raise ValueError("blocks must have shape (chunks_x, chunks_z, 16, 256, 16)")
    op = np.ascontiguousarray(opaque, np.uint8)
    bm = np.ascontiguousarray(block_material, np.uint32)
    if op.shape != (256,) or bm.shape != (256,):
        raise ValueError("opaque and block_material need 256 entries")
    nv, nt = C.c_uint64(0), C.c_uint64(0)
    args = (_ptr(blocks), blocks.shape[0], blocks.shape[1], cx0, cz0, voxel_size, _ptr(op), _ptr(bm))
    rc = lib.mcrt_mesh_voxel_chunks(*args, None, 0, None, 0, C.byref(nv), C.byref(nt))
    if rc != MCRT_OK:
        raise McrtError(rc, "mcrt_mesh_voxel_chunks rejected its input")
    verts = np.zeros(nv.value, T.GpuVertex)
    tris = np.zeros(nt.value, T.GpuTriangle)
    rc = lib.mcrt_mesh_voxel_chunks(*args, _ptr(verts), len(verts), _ptr(tris), len(tris),
                                    C.byref(nv), C.byref(nt))
    if rc != MCRT_OK:
        raise McrtError(rc, "mcrt_mesh_voxel_chunks failed")
    return verts, tris


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def build_mesh_bvh(verts: np.ndarray, tris: np.ndarray):
    """C++ binned-SAH mesh BVH (src/mesh.rs:82).  Returns (nodes, ordered, orig)."""
    lib = load()
    verts = np.ascontiguousarray(verts, T.GpuVertex)
    tris = np.ascontiguousarray(tris, T.GpuTriangle)
    n = len(tris)
    nodes = np.zeros(max(2 * n - 1, 1), T.GpuBvhNode)
    ordered = np.zeros(max(n, 1), T.GpuTriangle)
    orig = np.zeros(max(n, 1), np.uint32)
    nn = _u32(0)
    rc = lib.mcrt_build_mesh_bvh(_ptr(verts), len(verts), _ptr(tris), n, _ptr(nodes), C.byref(nn),
                                 _ptr(ordered), _ptr(orig))
    if rc != MCRT_OK:
        raise McrtError(rc, "mcrt_build_mesh_bvh rejected its input")
    return nodes[: nn.value].copy(), ordered[:n].copy(), orig[:n].copy()


def build_sphere_bvh(spheres: np.ndarray):
    """C++ binned-SAH sphere BVH (src/bvh.rs:239).  Returns (nodes, ordered, orig)."""
    lib = load()
    spheres = np.ascontiguousarray(spheres, T.GpuSphere)
    n = len(spheres)
    nodes = np.zeros(max(2 * n - 1, 1), T.GpuBvhNode)
    ordered = np.zeros(max(n, 1), T.GpuSphere)
    orig = np.zeros(max(n, 1), np.uint32)
    nn = _u32(0)
    rc = lib.mcrt_build_sphere_bvh(_ptr(spheres), n, _ptr(nodes), C.byref(nn), _ptr(ordered),
                                   _ptr(orig))
    if rc != MCRT_OK:
        raise McrtError(rc, "mcrt_build_sphere_bvh rejected its input")
    return nodes[: nn.value].copy(), ordered[:n].copy(), orig[:n].copy()


class Context:
    """One mcrt_ctx bound to a HIP device."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = _vp()
        rc = self._lib.mcrt_create(C.byref(h), device)
        if rc != MCRT_OK:
            raise McrtError(rc, self._lib.mcrt_last_error(None).decode())
        self._h = h
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mcrt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != MCRT_OK:
            raise McrtError(rc, self._lib.mcrt_last_error(self._h).decode())

    def upload_scene(self, scene) -> None:
        """Upload a :class:`mcrt_amd.scene.SceneBuffers`."""
        s = scene
        self._check(
            self._lib.mcrt_upload_scene(
                self._h,
                _ptr(s.spheres), len(s.spheres), _ptr(s.sphere_orig),
                _ptr(s.sphere_bvh), len(s.sphere_bvh),
                _ptr(s.vertices), len(s.vertices),
                _ptr(s.triangles), len(s.triangles), _ptr(s.tri_orig),
                _ptr(s.mesh_bvh), len(s.mesh_bvh),
                _ptr(s.materials),
                _ptr(s.env), s.env.shape[1], s.env.shape[0],
                None, 0, 0,
            )
        )

    def render(self, camera: np.ndarray, width: int, height: int, spp: int, first_frame: int = 0,
               max_depth: int = T.MAX_BOUNCES, accum_in: Optional[np.ndarray] = None,
               want_gbuf: bool = True):
        cam = np.ascontiguousarray(camera, T.CameraUniform)
        out = np.empty((height, width, 4), np.float32)
        gb = np.empty((height, width, 4), np.float32) if want_gbuf else None
        ain = None if accum_in is None else np.ascontiguousarray(accum_in, np.float32)
        self._check(self._lib.mcrt_render(self._h, _ptr(cam), width, height, spp, first_frame,
                                          max_depth, _ptr(ain), _ptr(out), _ptr(gb)))
        return out, gb

    def render_dev(self, camera, width, height, spp, first_frame, max_depth, accum_in_ptr,
                   out_accum_ptr, out_gbuf_ptr, stream_ptr) -> None:
        cam = np.ascontiguousarray(camera, T.CameraUniform)
        self._check(self._lib.mcrt_render_dev(self._h, _ptr(cam), width, height, spp, first_frame,
                                              max_depth, accum_in_ptr, out_accum_ptr, out_gbuf_ptr,
                                              stream_ptr))

    def denoise_bilateral(self, accum: np.ndarray, gbuf: np.ndarray, params=None) -> np.ndarray:
        accum = np.ascontiguousarray(accum, np.float32)
        gbuf = np.ascontiguousarray(gbuf, np.float32)
        h, w = accum.shape[:2]
        out = np.empty_like(accum)
        p = None if params is None else np.ascontiguousarray(params, T.DenoiseParams)
        self._check(self._lib.mcrt_denoise_bilateral(self._h, _ptr(p), w, h, _ptr(accum),
                                                     _ptr(gbuf), _ptr(out)))
        return out

    def denoise_bilateral_dev(self, params, width, height, accum_ptr, gbuf_ptr, out_ptr,
                              stream_ptr) -> None:
        p = None if params is None else np.ascontiguousarray(params, T.DenoiseParams)
        self._check(self._lib.mcrt_denoise_bilateral_dev(self._h, _ptr(p), width, height, accum_ptr,
                                                         gbuf_ptr, out_ptr, stream_ptr))

    def trace_closest(self, origins: np.ndarray, dirs: np.ndarray, tmin: float = 1e-4,
                      tmax: float = 1e9):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(o)
        ids = np.empty(n, np.uint32)
        ts = np.empty(n, np.float32)
        self._check(self._lib.mcrt_trace_closest(self._h, _ptr(o), _ptr(d), n, tmin, tmax,
                                                 _ptr(ids), _ptr(ts)))
        return ids, ts

    def trace_closest_dev(self, o_ptr, d_ptr, n, tmin, tmax, id_ptr, t_ptr, stream_ptr) -> None:
        self._check(self._lib.mcrt_trace_closest_dev(self._h, o_ptr, d_ptr, n, tmin, tmax, id_ptr,
                                                     t_ptr, stream_ptr))

    def trace_closest_noalloc_dev(self, o_ptr, d_ptr, n, tmin, tmax, id_ptr, t_ptr, stream_ptr) -> None:
        """mcrt_trace_closest_dev without device allocation (static scratch)."""
        self._check(self._lib.mcrt_trace_closest_noalloc_dev(self._h, o_ptr, d_ptr, n, tmin, tmax, id_ptr,
                                                             t_ptr, stream_ptr))

    # --- differentiable render (extended estimator, see include/mcrt_abi.h) ---
    @staticmethod
    def _diff_inputs(camera, params, pmap):
        cam = np.ascontiguousarray(camera, T.CameraUniform)
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 4)
        m = np.ascontiguousarray(pmap, np.int32)
        if m.shape != (64,):
            raise ValueError("param_of_material needs 64 entries")
        return cam, p, m

    def diff_render(self, camera, width, height, spp, first_frame, max_depth, params, pmap):
        cam, p, m = self._diff_inputs(camera, params, pmap)
        out = np.empty((height, width, 4), np.float32)
        self._check(self._lib.mcrt_diff_render(self._h, _ptr(cam), width, height, spp, first_frame,
                                               max_depth, _ptr(p), len(p), _ptr(m), _ptr(out)))
        return out

    def diff_adjoint(self, camera, width, height, spp, first_frame, max_depth, params, pmap, d_image):
        cam, p, m = self._diff_inputs(camera, params, pmap)
        di = np.ascontiguousarray(d_image, np.float32)
        g = np.empty((len(p), 4), np.float32)
        self._check(self._lib.mcrt_diff_adjoint(self._h, _ptr(cam), width, height, spp, first_frame,
                                                max_depth, _ptr(p), len(p), _ptr(m), _ptr(di), _ptr(g)))
        return g

    def diff_render_dev(self, camera, width, height, spp, first_frame, max_depth, params, pmap,
                        out_ptr, stream_ptr) -> None:
        cam, p, m = self._diff_inputs(camera, params, pmap)
        self._check(self._lib.mcrt_diff_render_dev(self._h, _ptr(cam), width, height, spp, first_frame,
                                                   max_depth, _ptr(p), len(p), _ptr(m), out_ptr,
                                                   stream_ptr))

    def diff_adjoint_dev(self, camera, width, height, spp, first_frame, max_depth, params, pmap,
                         d_image_ptr, grad_ptr, stream_ptr) -> None:
        cam, p, m = self._diff_inputs(camera, params, pmap)
        self._check(self._lib.mcrt_diff_adjoint_dev(self._h, _ptr(cam), width, height, spp,
                                                    first_frame, max_depth, _ptr(p), len(p), _ptr(m),
                                                    d_image_ptr, grad_ptr, stream_ptr))

    def render_aux_dev(self, camera, width, height, spp, first_frame, max_depth, accum_in_ptr,
                       out_accum_ptr, out_gbuf_ptr, out_albedo_ptr, stream_ptr) -> None:
        cam = np.ascontiguousarray(camera, T.CameraUniform)
        self._check(self._lib.mcrt_render_aux_dev(self._h, _ptr(cam), width, height, spp, first_frame,
                                                  max_depth, accum_in_ptr, out_accum_ptr, out_gbuf_ptr,
                                                  out_albedo_ptr, stream_ptr))

    def set_profiling(self, enabled: bool) -> None:
        self._check(self._lib.mcrt_set_profiling(self._h, 1 if enabled else 0))

    def set_diff_mode(self, mode: str) -> None:
        """Adjoint form: "replay" (default, no tape) or "tape"."""
        self._check(self._lib.mcrt_diff_set_mode(self._h, {"replay": 0, "tape": 1}[mode]))

    def set_nee(self, enabled: bool) -> None:
        """Next-event estimation with MIS for later renders (default off)."""
        self._check(self._lib.mcrt_set_nee(self._h, 1 if enabled else 0))

    def kernel_stats(self, name: str):
        ms = C.c_double(0.0)
        n = C.c_uint64(0)
        self._check(self._lib.mcrt_kernel_stats(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_stats(self) -> None:
        self._check(self._lib.mcrt_reset_stats(self._h))

    def render_counters(self):
        seg = C.c_uint64(0)
        smp = C.c_uint64(0)
        self._check(self._lib.mcrt_render_counters(self._h, C.byref(seg), C.byref(smp)))
        return seg.value, smp.value
