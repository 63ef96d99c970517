"""Learned denoiser on the HIP BF16-MFMA kernels (no reference counterpart).

:class:`Denoiser` wraps ``mcrt_denoiser_*``: a 6-layer 3x3 conv net
(10 -> 32 x5 -> 3, ReLU, global skip in log space), trained with Adam on
(low-spp, high-spp) pairs that the renderer produces itself.  The bilateral
filter (``mcrt_denoise_bilateral_dev``) remains the fallback.

Training data: seeded variants of the Cornell box and of the displaced
icosphere scene with a random camera, random materials and a random light
position, rendered at 256 x 256 (noisy at `spp_in`, reference at `spp_ref`,
both with the first-hit albedo and G-buffer of the noisy render).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List

import numpy as np

from . import camera as cam_mod
from . import configs, scene as scene_mod, scenes
from ._lib import Context, McrtError, MCRT_OK, load

N_PARAMS = 40771


class Denoiser:
    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        self._lib = load()
        h = C.c_void_p()
        rc = self._lib.mcrt_denoiser_create(ctx._h, seed, C.byref(h))
        ctx._check(rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mcrt_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self) -> np.ndarray:
        out = np.empty(self._lib.mcrt_denoiser_param_count(), np.float32)
        self.ctx._check(self._lib.mcrt_denoiser_get_weights(self._h, out.ctypes.data))
        return out

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        if w.shape != (N_PARAMS,):
            raise ValueError(f"expected {N_PARAMS} weights")
        self.ctx._check(self._lib.mcrt_denoiser_set_weights(self._h, w.ctypes.data))

    def infer(self, noisy, albedo, gbuf, out, stream_ptr=None):
        """Tensors (B, H, W, 4) float32 on the device."""
        B, H, W = noisy.shape[:3]
        self.ctx._check(self._lib.mcrt_denoiser_infer_dev(
            self._h, B, W, H, noisy.data_ptr(), albedo.data_ptr(), gbuf.data_ptr(), out.data_ptr(),
            stream_ptr))

    def train_step(self, noisy, albedo, gbuf, ref, lr: float, want_loss: bool = True, stream_ptr=None):
        B, H, W = noisy.shape[:3]
        loss = C.c_float(0.0)
        self.ctx._check(self._lib.mcrt_denoiser_train_step_dev(
            self._h, B, W, H, noisy.data_ptr(), albedo.data_ptr(), gbuf.data_ptr(), ref.data_ptr(),
            float(lr), C.byref(loss) if want_loss else None, stream_ptr))
        return loss.value if want_loss else None

    def gradient(self, noisy, albedo, gbuf, ref):
        B, H, W = noisy.shape[:3]
        g = np.empty(N_PARAMS, np.float32)
        loss = C.c_float(0.0)
        self.ctx._check(self._lib.mcrt_denoiser_gradient(
            self._h, B, W, H, noisy.data_ptr(), albedo.data_ptr(), gbuf.data_ptr(), ref.data_ptr(),
            g.ctypes.data, C.byref(loss)))
        return g, loss.value


def layer_shapes():
    cin = [10, 32, 32, 32, 32, 32]
    cout = [32, 32, 32, 32, 32, 3]
    return list(zip(cout, cin))


def split_weights(w: np.ndarray):
    """Public weight vector -> [(W (co, ci, 3, 3), b (co,))]."""
    out, k = [], 0
    for co, ci in layer_shapes():
        n = co * ci * 9
        out.append((w[k:k + n].reshape(co, ci, 3, 3), w[k + n:k + n + co]))
        k += n + co
    return out


# ---------------------------------------------------------------------------
# Training data from the renderer
# ---------------------------------------------------------------------------
def variant_scene(seed: int, kind: str):
    """A seeded variant: random camera, material palette and light position."""
    rng = np.random.default_rng(seed)
    lx, lz = rng.uniform(-3.0, 3.0), rng.uniform(-4.0, 1.0)
    yaw = rng.uniform(-25, 25)
    dist = rng.uniform(6.5, 9.0)
    lf = (dist * np.sin(np.radians(yaw)), rng.uniform(3.0, 6.5), dist * np.cos(np.radians(yaw)))
    la = (rng.uniform(-1, 1), rng.uniform(2.5, 5.0), rng.uniform(-2.5, -0.5))
    mats = scenes.random_materials_yaml(8, seed)
    text = scenes.cornell_yaml(with_boxes=(kind == "c1"), extra_materials=mats)
    text = text.replace("translate: [0.0, 9.9, -2.5]", f"translate: [{lx:.4f}, 9.9, {lz:.4f}]")
    text = text.replace("look_from: [0.0, 4.75, 8.0]", f"look_from: [{lf[0]:.4f}, {lf[1]:.4f}, {lf[2]:.4f}]")
    text = text.replace("look_at:   [0.0, 4.0, 0.0]", f"look_at:   [{la[0]:.4f}, {la[1]:.4f}, {la[2]:.4f}]")
    base = scene_mod.load_scene_from_str(text, f"variant-{seed}")
    extra = None
    if kind == "c2":
        v, t, _ = scenes.displaced_icosphere(n=24, seed=seed, mat_base=base.name_to_idx["m0"], n_mats=8)
        extra = (v, t)
    loaded = scene_mod.load_scene_from_str(text, f"variant-{seed}", extra_mesh=extra)
    return loaded, scene_mod.build_scene_buffers(loaded)


@dataclass
class Pairs:
    noisy: object   # torch (N, H, W, 4)
    albedo: object
    gbuf: object
    ref: object


def render_pairs(ctx: Context, n_variants: int, size: int = 256, spp_in: int = 4, spp_ref: int = 1024,
                 depth: int = 8, seed0: int = 1000, log=None) -> Pairs:
    import torch

    dev = torch.device("cuda:0")
    shape = (n_variants, size, size, 4)
    noisy = torch.empty(shape, dtype=torch.float32, device=dev)
    albedo = torch.empty_like(noisy)
    gbuf = torch.empty_like(noisy)
    ref = torch.empty_like(noisy)
    s = torch.cuda.current_stream(dev).cuda_stream
    for i in range(n_variants):
        kind = "c1" if i % 2 == 0 else "c2"
        loaded, sb = variant_scene(seed0 + i, kind)
        cam = cam_mod.camera_for(loaded, size, size)
        ctx.upload_scene(sb)
        ctx.render_aux_dev(cam, size, size, spp_in, 0, depth, None, noisy[i].data_ptr(), gbuf[i].data_ptr(),
                           albedo[i].data_ptr(), s)
        # Reference from an independent frame range.
        tmp_g = torch.empty((size, size, 4), dtype=torch.float32, device=dev)
        ctx.render_dev(cam, size, size, spp_ref, 1 << 20, depth, None, ref[i].data_ptr(), tmp_g.data_ptr(), s)
        if log and (i % 16 == 0 or i == n_variants - 1):
            torch.cuda.synchronize()
            log(f"pairs {i + 1}/{n_variants}")
    torch.cuda.synchronize()
    return Pairs(noisy, albedo, gbuf, ref)


def train(dn: Denoiser, pairs: Pairs, steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 0,
          log=None) -> List[float]:
    import torch

    n = pairs.noisy.shape[0]
    rng = np.random.default_rng(seed)
    losses = []
    for it in range(steps):
        idx = torch.as_tensor(rng.choice(n, size=min(batch, n), replace=False), device=pairs.noisy.device)
        lr_t = lr * (0.1 ** (it / max(steps, 1)))
        loss = dn.train_step(pairs.noisy[idx].contiguous(), pairs.albedo[idx].contiguous(),
                             pairs.gbuf[idx].contiguous(), pairs.ref[idx].contiguous(), lr_t,
                             want_loss=(it % 20 == 0 or it == steps - 1))
        if loss is not None:
            losses.append(loss)
            if log:
                log(f"train step {it}: loss {loss:.4e}")
    return losses


def rel_mse(x: np.ndarray, y: np.ndarray) -> float:
    """mean over pixels and RGB of (x - y)^2 / (y^2 + 1e-2)."""
    x = x[..., :3].astype(np.float64)
    y = y[..., :3].astype(np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))
