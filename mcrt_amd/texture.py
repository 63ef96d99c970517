"""Environment maps: gradient sky, synthetic sun sky, Radiance HDR files.

* ``build_gradient_env_map`` restates src/texture.rs:158-199 (FP32, row 0 =
  zenith, pixel-centre rows).
* ``generate_skymap`` restates the closed form of
  tests/assets/gen_test_skymap.py:122-241 (FP64, pixel-corner angles, cast to
  FP32); it is re-implemented here, not imported.
* ``read_hdr`` / ``write_hdr`` handle Radiance RGBE files; the reader decodes
  flat and new-style RLE scanlines.  ``load_env_map_data`` follows the fallback
  chain of src/texture.rs:209-220 (YAML path -> ``textures/env.hdr`` ->
  gradient) with nearest-neighbour resampling to 1024 x 512
  (src/texture.rs:124-148).
"""
from __future__ import annotations

import math
import os
from typing import Optional

import numpy as np

from .types import ENV_MAP_HEIGHT, ENV_MAP_WIDTH

F = np.float32


def gradient_sky_rgb(elevation: np.ndarray):
    """src/texture.rs:183-199, vectorised in FP32."""
    e = elevation.astype(np.float32)
    sky = e >= F(0.0)
    t_sky = e
    t_gnd = np.minimum(-e, F(1.0))
    r = np.where(sky, F(1.0) - F(0.5) * t_sky, (F(1.0) - t_gnd) * F(0.8))
    g = np.where(sky, F(1.0) - F(0.3) * t_sky, (F(1.0) - t_gnd) * F(0.4))
    b = np.where(sky, F(1.0), (F(1.0) - t_gnd) * F(0.1))
    return r.astype(np.float32), g.astype(np.float32), b.astype(np.float32)


def build_gradient_env_map(width: int = ENV_MAP_WIDTH, height: int = ENV_MAP_HEIGHT) -> np.ndarray:
    """Procedural sky, (H, W, 4) float32 with alpha 1 (src/texture.rs:158-178)."""
    y = np.arange(height, dtype=np.float32)
    v = (y + F(0.5)) / F(height)
    e = F(1.0) - F(2.0) * v
    r, g, b = gradient_sky_rgb(e)
    out = np.empty((height, width, 4), np.float32)
    out[:, :, 0] = r[:, None]
    out[:, :, 1] = g[:, None]
    out[:, :, 2] = b[:, None]
    out[:, :, 3] = 1.0
    return out


def _hex_to_linear(hex_color: str) -> np.ndarray:
    h = hex_color.lstrip("#")
    r, g, b = (int(h[i:i + 2], 16) / 255.0 for i in (0, 2, 4))
    return np.array([r ** 2.2, g ** 2.2, b ** 2.2], dtype=np.float32)


def generate_skymap(width: int, height: int, *, horizon_hex="#0033AA", zenith_hex="#88BBFF",
                    ground_hex="#333333", sun_color_hex="#FFEEDD", sun_elevation_deg=45.0,
                    sun_azimuth_deg=30.0, sun_peak_radiance=80_000.0, sun_sigma_deg=2.0,
                    sky_horizon_nit=1.0, sky_zenith_nit=3.0, ground_nit=0.1) -> np.ndarray:
    """Gradient sky + Gaussian sun + ground fill, (H, W, 3) float32.

    Same closed form and parameters as tests/assets/gen_test_skymap.py:122-241:
    pixel-corner angles phi = (u/W - 0.5) * 2pi, theta = (0.5 - v/H) * pi.
    """
    hc, zc, gc, sc = (_hex_to_linear(x) for x in (horizon_hex, zenith_hex, ground_hex, sun_color_hex))
    el, az = math.radians(sun_elevation_deg), math.radians(sun_azimuth_deg)
    sun_dir = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    sig2 = 2.0 * math.radians(sun_sigma_deg) ** 2
    uu, vv = np.meshgrid(np.arange(width, dtype=np.float64), np.arange(height, dtype=np.float64))
    phi = (uu / width - 0.5) * (2.0 * math.pi)
    theta = (0.5 - vv / height) * math.pi
    t = np.maximum(np.sin(theta), 0.0) ** 0.6
    sky = (hc + (zc - hc) * t[:, :, None]) * (sky_horizon_nit + (sky_zenith_nit - sky_horizon_nit) * t)[:, :, None]
    dirs = np.stack([np.cos(theta) * np.cos(phi), np.sin(theta), np.cos(theta) * np.sin(phi)], axis=-1)
    ang = np.arccos(np.clip(dirs @ sun_dir, -1.0, 1.0))
    sun = np.where(theta >= 0, sun_peak_radiance * np.exp(-(ang ** 2) / sig2), 0.0)
    sky = np.where((theta >= 0)[:, :, None], sky, gc * ground_nit)
    return (sky + sun[:, :, None] * sc).astype(np.float32)


def rgb_to_env(rgb: np.ndarray) -> np.ndarray:
    """(H, W, 3) -> (H, W, 4) with alpha 1, values clamped to [0, 1e10]
    (src/texture.rs:139-146)."""
    h, w, _ = rgb.shape
    out = np.empty((h, w, 4), np.float32)
    out[:, :, :3] = np.clip(np.nan_to_num(rgb, nan=0.0, posinf=1e10), 0.0, 1e10)
    out[:, :, 3] = 1.0
    return out


# ---------------------------------------------------------------------------
# Radiance RGBE
# ---------------------------------------------------------------------------
def write_hdr(path: str, rgb: np.ndarray) -> None:
    """Flat (non-RLE) RGBE writer, the encoding of gen_test_skymap.py:44-86."""
    h, w, _ = rgb.shape
    out = np.zeros((h, w, 4), np.uint8)
    m = rgb.max(axis=-1)
    nz = m > 1e-32
    if np.any(nz):
        mant, ex = np.frexp(m[nz])
        scale = mant * 256.0 / m[nz]
        for c in range(3):
            out[nz, c] = np.clip(rgb[nz, c] * scale, 0, 255).astype(np.uint8)
        out[nz, 3] = np.clip(ex + 128, 0, 255).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode("ascii"))
        f.write(out.tobytes())


def read_hdr(path: str) -> np.ndarray:
    """Read a Radiance .hdr (flat or RLE scanlines) -> (H, W, 3) float32.

    RGBE decode: 0 when E == 0, else c * 2^(E - 136).
    """
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    lines = []
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line == b"":
            break
        lines.append(line)
    if not lines or not lines[0].startswith(b"#?"):
        raise ValueError("not a Radiance HDR file")
    end = data.index(b"\n", pos)
    parts = data[pos:end].split()
    pos = end + 1
    if len(parts) != 4 or parts[0] != b"-Y" or parts[2] != b"+X":
        raise ValueError("unsupported HDR orientation")
    h, w = int(parts[1]), int(parts[3])
    out = np.zeros((h, w, 4), np.uint8)
    buf = memoryview(data)
    for y in range(h):
        if 8 <= w < 32768 and pos + 4 <= len(data) and data[pos] == 2 and data[pos + 1] == 2 \
                and (data[pos + 2] << 8 | data[pos + 3]) == w:
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    cnt = data[pos]
                    pos += 1
                    if cnt > 128:
                        cnt -= 128
                        out[y, x:x + cnt, c] = data[pos]
                        pos += 1
                    else:
                        out[y, x:x + cnt, c] = np.frombuffer(buf[pos:pos + cnt], np.uint8)
                        pos += cnt
                    x += cnt
        else:
            out[y] = np.frombuffer(buf[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    e = out[:, :, 3].astype(np.int32)
    scale = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136)).astype(np.float32)
    return out[:, :, :3].astype(np.float32) * scale[:, :, None]


def resize_nearest(img: np.ndarray, width: int, height: int) -> np.ndarray:
    """Nearest-neighbour resample (src/texture.rs:127-134): source index
    floor((i + 0.5) * src / dst)."""
    h, w = img.shape[:2]
    if (h, w) == (height, width):
        return img
    ys = np.minimum(((np.arange(height) + 0.5) * h / height).astype(np.int64), h - 1)
    xs = np.minimum(((np.arange(width) + 0.5) * w / width).astype(np.int64), w - 1)
    return img[ys][:, xs]


def try_load_hdr(path: str) -> Optional[np.ndarray]:
    try:
        rgb = read_hdr(path)
    except (OSError, ValueError, IndexError):
        return None
    return rgb_to_env(resize_nearest(rgb, ENV_MAP_WIDTH, ENV_MAP_HEIGHT))


def load_env_map_data(scene_path: Optional[str], cwd: Optional[str] = None) -> np.ndarray:
    """Fallback chain of src/texture.rs:209-220."""
    base = cwd or os.getcwd()
    if scene_path:
        p = scene_path if os.path.isabs(scene_path) else os.path.join(base, scene_path)
        data = try_load_hdr(p)
        if data is not None:
            return data
    data = try_load_hdr(os.path.join(base, "textures", "env.hdr"))
    return data if data is not None else build_gradient_env_map()
