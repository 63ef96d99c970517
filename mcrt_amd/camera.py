"""Thin-lens camera, src/camera.rs:76-127, in FP32 with glam's operation order.

glam 0.32 (scalar Vec3): ``normalize(v) = v * (1 / sqrt(dot(v, v)))``,
``cross(a, b) = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)``.
"""
from __future__ import annotations

import math

import numpy as np

from . import types as T

F = np.float32
DEFAULT_VFOV = F(60.0)  # src/camera.rs:15
INIT_LOOK_FROM = (F(0.0), F(1.0), F(3.5))  # src/camera.rs:18
INIT_LOOK_AT = (F(0.0), F(0.0), F(0.0))  # src/camera.rs:21


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return (a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1])


def _normalize(a):
    r = F(1.0) / F(np.sqrt(_dot(a, a)))
    return (a[0] * r, a[1] * r, a[2] * r)


def _scale(s, a):
    return (s * a[0], s * a[1], s * a[2])


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def compute_camera(width: int, height: int, look_from, look_at, vfov_deg, aperture,
                   focus_dist) -> np.ndarray:
    """Returns a () CameraUniform with frame_count 0."""
    lf = tuple(F(x) for x in look_from)
    la = tuple(F(x) for x in look_at)
    aspect = F(width) / F(height)
    vup = (F(0.0), F(1.0), F(0.0))
    vfov_rad = F(vfov_deg) * (F(math.pi) / F(180.0))  # f32::to_radians
    h = F(math.tan(float(vfov_rad * F(0.5))))
    vh = F(2.0) * h
    vw = aspect * vh
    w = _normalize(_sub(lf, la))
    c = _cross(vup, w)
    vup_safe = vup if _dot(c, c) > F(1e-10) else (F(0.0), F(0.0), F(1.0))
    cu = _normalize(_cross(vup_safe, w))
    cv = _cross(w, cu)
    fd = F(focus_dist)
    horizontal = _scale(fd * vw, cu)
    vertical = _scale(fd * vh, cv)
    # look_from - horizontal*0.5 - vertical*0.5 - focus_dist*w
    hh = _scale(F(0.5), horizontal)
    vv = _scale(F(0.5), vertical)
    fw = _scale(fd, w)
    lower_left = _sub(_sub(_sub(lf, hh), vv), fw)
    lens_radius = F(aperture) * F(0.5)
    du = _scale(lens_radius, cu)
    dv = _scale(lens_radius, cv)
    cam = np.zeros((), T.CameraUniform)
    cam["origin"] = (*lf, 0.0)
    cam["lower_left"] = (*lower_left, 0.0)
    cam["horizontal"] = (*horizontal, 0.0)
    cam["vertical"] = (*vertical, 0.0)
    cam["defocus_u"] = (*du, 0.0)
    cam["defocus_v"] = (*dv, 0.0)
    return cam


def camera_for(loaded, width: int, height: int) -> np.ndarray:
    """Camera of a loaded scene or the built-in default (src/headless.rs:220-232)."""
    if loaded.camera is not None:
        c = loaded.camera
        return compute_camera(width, height, c.look_from, c.look_at, c.vfov, c.aperture, c.focus_dist)
    d = _sub(INIT_LOOK_FROM, INIT_LOOK_AT)
    dist = F(np.sqrt(_dot(d, d)))
    return compute_camera(width, height, INIT_LOOK_FROM, INIT_LOOK_AT, DEFAULT_VFOV, F(0.0), dist)
