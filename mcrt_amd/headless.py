"""Headless renderer: the drop-in for ``mcrt --headless`` (src/main.rs:56-80,
src/headless.rs:112-386).

``run`` resolves width / height / spp with the reference's precedence
CLI > ``[render]`` of the ``<scene>.test.toml`` sidecar > defaults 512 x 512 x
64 (src/headless.rs:36-40, :120-127), loads the scene YAML, builds the BVHs,
renders frames ``0..spp-1`` on the GPU through libmcrt_hip.so, sanitises
non-finite values, tonemaps (Reinhard + gamma 2.2, src/headless.rs:93-98) and
writes an RGBA8 PNG.  Extensions (the reference's headless path has none):
``--denoise bilateral|learned`` applies the bilateral filter or the learned
denoiser before tonemapping; ``--nee`` renders with next-event estimation.
"""
from __future__ import annotations

import os
import sys
from typing import Optional

import numpy as np

from . import camera as cam_mod
from . import scene as scene_mod
from . import types as T
from ._lib import Context, McrtError

DEFAULT_WIDTH = 512
DEFAULT_HEIGHT = 512
DEFAULT_SPP = 64


def load_sidecar(scene_path: str) -> dict:
    """src/headless.rs:69-77: ``<stem>.test.toml``; any error -> defaults."""
    base, _ = os.path.splitext(scene_path)
    path = base + ".test.toml"
    try:
        import tomli

        with open(path, "rb") as f:
            doc = tomli.load(f)
    except Exception:
        return {}
    r = doc.get("render")
    return r if isinstance(r, dict) else {}


def tonemap_to_u8(hdr: np.ndarray) -> np.ndarray:
    """Per channel: non-finite -> 0, max(v, 0), v/(v+1), pow(1/2.2), clamp,
    *255 + 0.5 truncated (src/headless.rs:93-98, :359-366)."""
    v = np.where(np.isfinite(hdr), hdr, np.float32(0.0)).astype(np.float32)
    v = np.maximum(v, np.float32(0.0))
    mapped = v / (v + np.float32(1.0))
    g = np.clip(np.power(mapped, np.float32(1.0) / np.float32(2.2)), 0.0, 1.0).astype(np.float32)
    return (g * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)


def to_rgba8(accum: np.ndarray) -> np.ndarray:
    h, w = accum.shape[:2]
    out = np.empty((h, w, 4), np.uint8)
    out[:, :, :3] = tonemap_to_u8(accum[:, :, :3])
    out[:, :, 3] = 255
    return out


def write_png(path: str, rgba: np.ndarray) -> None:
    from PIL import Image

    parent = os.path.dirname(path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    Image.fromarray(rgba, "RGBA").save(path)


def run(scene_path: str, output_path: str, width: Optional[int] = None, height: Optional[int] = None,
        spp: Optional[int] = None, denoise: Optional[str] = None, device: int = 0,
        max_depth: int = T.MAX_BOUNCES, nee: bool = False) -> dict:
    """``denoise``: None / "none", "bilateral" or "learned" (committed
    weights; the set trained on next-event-estimation renders when ``nee``).
    ``nee``: render with next-event estimation (an extension)."""
    if denoise not in (None, "none", "bilateral", "learned"):
        raise ValueError(f"unknown denoiser {denoise!r}")
    side = load_sidecar(scene_path)
    width = max(1, int(width if width is not None else side.get("width", DEFAULT_WIDTH)))
    height = max(1, int(height if height is not None else side.get("height", DEFAULT_HEIGHT)))
    spp = max(1, int(spp if spp is not None else side.get("spp", DEFAULT_SPP)))
    loaded = scene_mod.load_scene_from_yaml(scene_path)
    sb = scene_mod.build_scene_buffers(loaded)
    cam = cam_mod.camera_for(loaded, width, height)
    with Context(device) as ctx:
        ctx.upload_scene(sb)
        ctx.set_nee(nee)
        if denoise == "learned":
            accum = _render_learned(ctx, cam, width, height, spp, max_depth, device, nee)
        else:
            accum, gbuf = ctx.render(cam, width, height, spp, 0, max_depth)
            if denoise == "bilateral":
                accum = ctx.denoise_bilateral(accum, gbuf)
    write_png(output_path, to_rgba8(accum))
    return {"width": width, "height": height, "spp": spp}


def _render_learned(ctx, cam, width, height, spp, max_depth, device, nee):
    import torch

    from . import learned

    dev = torch.device(f"cuda:{device}")
    acc = torch.empty((height, width, 4), dtype=torch.float32, device=dev)
    gb, al, out = (torch.empty_like(acc) for _ in range(3))
    s = torch.cuda.current_stream(dev).cuda_stream
    ctx.render_aux_dev(cam, width, height, spp, 0, max_depth, None, acc.data_ptr(), gb.data_ptr(),
                       al.data_ptr(), s)
    dn = learned.load_trained(ctx, learned.WEIGHTS_NEE_PATH if nee else learned.WEIGHTS_PATH)
    try:
        dn.infer(acc[None], al[None], gb[None], out[None], s)
        torch.cuda.synchronize(dev)
        return out.cpu().numpy()
    finally:
        dn.close()


def get_next_after(args, flag):
    """src/main.rs:95-100."""
    try:
        i = args.index(flag)
    except ValueError:
        return None
    return args[i + 1] if i + 1 < len(args) else None


def _opt_u32(args, flag):
    v = get_next_after(args, flag)
    try:
        return int(v) if v is not None else None
    except ValueError:
        return None


def main(argv=None) -> int:
    args = list(sys.argv[1:] if argv is None else argv)
    if "--headless" not in args:
        sys.stderr.write("error: only --headless mode is provided (the interactive window is not)\n")
        return 1
    scene = get_next_after(args, "--headless")
    if scene is None:
        sys.stderr.write("error: --headless requires a scene YAML path\n")
        return 1
    output = get_next_after(args, "--output")
    if output is None:
        sys.stderr.write("error: --headless mode requires --output <path>\n")
        return 1
    try:
        run(scene, output, _opt_u32(args, "--width"), _opt_u32(args, "--height"),
            _opt_u32(args, "--spp"), get_next_after(args, "--denoise"), nee="--nee" in args)
    except (McrtError, scene_mod.SceneError, OSError, ValueError) as e:
        sys.stderr.write(f"headless render failed: {e}\n")
        return 1
    return 0
