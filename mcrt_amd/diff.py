"""Inverse rendering with the path-replay adjoint (no reference counterpart).

* :class:`DiffRender` — a thin ``torch.autograd.Function``: forward calls
  ``mcrt_diff_render_dev``, backward calls ``mcrt_diff_adjoint_dev`` with the
  incoming image gradient.  Both run on the tensors' device memory and the
  current stream; torch records nothing but the Function's inputs, and the
  adjoint re-traces its paths (path replay) instead of reading a tape.
* :func:`inverse_workload` — the displaced-icosphere scene whose 8 icosphere
  materials become parameter materials (RGB albedo + scalar emission).
* :func:`optimize` — Adam on the L2 image loss against a target rendered with
  hidden albedos.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import configs

N_PARAMS = 8


@dataclass
class InverseWorkload:
    work: configs.Workload
    pmap: np.ndarray          # (64,) int32
    hidden: np.ndarray        # (8, 4) float32 target params
    init: np.ndarray          # (8, 4) float32 start params


def inverse_workload(width: int = 512, height: int = 512, spp: int = 32, depth: int = 6,
                     seed: int = 5) -> InverseWorkload:
    w = configs.c2(width=width, height=height, spp=spp, depth=depth)
    pmap = np.full(64, -1, np.int32)
    for i in range(N_PARAMS):
        pmap[w.loaded.name_to_idx[f"m{i}"]] = i
    rng = np.random.default_rng(seed)
    hidden = np.zeros((N_PARAMS, 4), np.float32)
    hidden[:, :3] = rng.uniform(0.1, 0.9, (N_PARAMS, 3))
    init = np.zeros((N_PARAMS, 4), np.float32)
    init[:, :3] = 0.5
    init[:, 3] = 0.05
    return InverseWorkload(w, pmap, hidden, init)


def _autograd_function():
    import torch

    class DiffRender(torch.autograd.Function):
        """image = render(params); params (8, 4) float32 on the GPU."""

        @staticmethod
        def forward(fctx, params, ctx, work, pmap, first_frame, grad_frame):
            p = params.detach().float().cpu().numpy()
            img = torch.empty((work.height, work.width, 4), dtype=torch.float32, device=params.device)
            s = torch.cuda.current_stream(params.device).cuda_stream
            ctx.diff_render_dev(work.camera, work.width, work.height, work.spp, first_frame,
                                work.depth, p, pmap, img.data_ptr(), s)
            fctx.save_for_backward(params)
            fctx.meta = (ctx, work, pmap, first_frame if grad_frame is None else grad_frame, p)
            return img

        @staticmethod
        def backward(fctx, grad_img):
            (params,) = fctx.saved_tensors
            ctx, work, pmap, grad_frame, p = fctx.meta
            g = torch.empty((len(p), 4), dtype=torch.float32, device=params.device)
            gi = grad_img.contiguous().float()
            s = torch.cuda.current_stream(params.device).cuda_stream
            ctx.diff_adjoint_dev(work.camera, work.width, work.height, work.spp, grad_frame,
                                 work.depth, p, pmap, gi.data_ptr(), g.data_ptr(), s)
            return g.to(params.dtype), None, None, None, None, None

    return DiffRender


_FN = None


def render(params, ctx, work, pmap, first_frame: int = 0, grad_frame: int = None):
    """Differentiable render of `work` with parameter tensor `params` (8, 4).

    The image uses frames first_frame..first_frame+spp-1.  The backward pass
    replays frames from `grad_frame` (default: the same frames, giving the
    exact derivative of this image).  A different range decorrelates the
    image from its derivative: the gradient of a nonlinear loss such as
    mean((I - T)^2) is then unbiased, 2 (E[I] - T) E[dI]."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(params, ctx, work, pmap, first_frame, grad_frame)


def optimize(ctx, inv: InverseWorkload, steps: int = 200, lr: float = 0.02, target_frame: int = 100_000,
             log_every: int = 0, decorrelate: bool = True, target_spp: int = None):
    """Adam on mean((I - I*)^2) over RGB, the target I* rendered with the
    hidden params from PCG frames at `target_frame` with `target_spp` samples
    (default: the workload's spp).  Adam converges to E[I] = I*, so the
    target's own noise sets the floor of the albedo error.  With
    `decorrelate`, step k renders frames 2k*spp.. and replays frames
    (2k+1)*spp.. for the gradient (fresh, independent samples each step);
    otherwise every step renders and replays frames 0..spp-1.  Returns
    (params, history of (loss, max albedo error, mean albedo error))."""
    import dataclasses

    import torch

    w = inv.work
    dev = torch.device("cuda:0")
    hidden = torch.tensor(inv.hidden, device=dev)
    tw = dataclasses.replace(w, spp=target_spp or w.spp)
    with torch.no_grad():
        target = render(hidden, ctx, tw, inv.pmap, target_frame)[..., :3].clone()
    params = torch.tensor(inv.init, device=dev, requires_grad=True)
    opt = torch.optim.Adam([params], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        if decorrelate:
            img = render(params, ctx, w, inv.pmap, 2 * it * w.spp, (2 * it + 1) * w.spp)
        else:
            img = render(params, ctx, w, inv.pmap, 0)
        loss = torch.mean((img[..., :3] - target) ** 2)
        loss.backward()
        opt.step()
        with torch.no_grad():
            params[:, :3].clamp_(0.01, 1.0)
            params[:, 3].clamp_(0.0, 10.0)
        dev_abs = torch.abs(params[:, :3] - hidden[:, :3])
        err, mean_err = float(torch.max(dev_abs).item()), float(torch.mean(dev_abs).item())
        hist.append((float(loss.item()), err, mean_err))
        if log_every and (it % log_every == 0 or it == steps - 1):
            print(f"adam {it}: loss {hist[-1][0]:.4e} |albedo - hidden| max {err:.4f} mean {mean_err:.4f}",
                  flush=True)
    return params.detach().cpu().numpy(), hist


def _main(argv=None) -> int:
    """``python -m mcrt_amd.diff [--steps N] [--log-every K]``: the inverse
    workload (512 x 512, 32 spp, depth 6; target at 8x the spp from an
    independent frame range), N Adam steps, one JSON line with the timing and
    the albedo error before and after."""
    import argparse
    import json
    import time

    import torch

    from ._lib import Context

    ap = argparse.ArgumentParser(prog="python -m mcrt_amd.diff")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--log-every", type=int, default=0)
    ap.add_argument("--mode", choices=["replay", "tape"], default="replay",
                    help="adjoint form: replay (re-trace, no tape) or tape")
    a = ap.parse_args(argv)
    torch.cuda.init()
    inv = inverse_workload()
    with Context(0) as ctx:
        ctx.upload_scene(inv.work.buffers)
        ctx.set_diff_mode(a.mode)
        target_spp = 8 * inv.work.spp
        optimize(ctx, inv, steps=2, target_spp=target_spp)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.set_profiling(True)
        ctx.reset_stats()
        _, hist = optimize(ctx, inv, steps=a.steps, target_spp=target_spp, log_every=a.log_every)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        fwd_ms, nf = ctx.kernel_stats("diff_primal")
        adj_ms, na = ctx.kernel_stats("diff_adjoint")
        ctx.set_profiling(False)
    print(json.dumps({"mode": a.mode, "steps": a.steps,
                      "primal_ms": round(fwd_ms / max(nf, 1), 3), "adjoint_ms": round(adj_ms / max(na, 1), 3), "total_s": round(el, 3), "ms_per_step": round(el / a.steps * 1e3, 3),
                      "albedo_err_first": {"max": round(hist[0][1], 4), "mean": round(hist[0][2], 4)},
                      "albedo_err_last": {"max": round(hist[-1][1], 4), "mean": round(hist[-1][2], 4)}}),
          flush=True)
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(_main())
