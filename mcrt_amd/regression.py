"""Image comparison of the reference's regression harness (src/regression.rs).

``compare_images`` (src/regression.rs:120-181): MSE, PSNR and max delta over
the R, G, B channels of two 8-bit images normalised to [0, 1]; alpha is
ignored; a size mismatch is an error; PSNR is +inf when MSE == 0.
``load_thresholds`` (src/regression.rs:296-347): ``[regression]`` of the
``.test.toml`` sidecar with defaults MSE 0.002 and PSNR 38 dB.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

DEFAULT_THRESHOLD_MSE = 0.002
DEFAULT_THRESHOLD_PSNR = 38.0


@dataclass
class DiffResult:
    mse: float
    psnr: float
    max_delta: float


class DimensionMismatch(ValueError):
    pass


def compare_images(a: np.ndarray, b: np.ndarray) -> DiffResult:
    if a.shape[:2] != b.shape[:2]:
        raise DimensionMismatch(f"image sizes differ: {a.shape[:2]} vs {b.shape[:2]}")
    if a.shape[0] == 0 or a.shape[1] == 0:
        return DiffResult(0.0, math.inf, 0.0)
    fa = a[..., :3].astype(np.float64) / 255.0
    fb = b[..., :3].astype(np.float64) / 255.0
    d = np.abs(fa - fb)
    mse = float(np.mean(d * d))
    psnr = math.inf if mse == 0.0 else 10.0 * math.log10(1.0 / mse)
    return DiffResult(mse, psnr, float(d.max()))


@dataclass
class Thresholds:
    mse: float = DEFAULT_THRESHOLD_MSE
    psnr: float = DEFAULT_THRESHOLD_PSNR

    def passes(self, r: DiffResult) -> bool:
        return r.mse <= self.mse and r.psnr >= self.psnr


def load_thresholds(scene_path: str) -> Thresholds:
    base, _ = os.path.splitext(scene_path)
    try:
        import tomli

        with open(base + ".test.toml", "rb") as f:
            doc = tomli.load(f)
    except Exception:
        return Thresholds()
    r = doc.get("regression") or {}
    return Thresholds(float(r.get("threshold_mse", DEFAULT_THRESHOLD_MSE)),
                      float(r.get("threshold_psnr", DEFAULT_THRESHOLD_PSNR)))
