"""Learned denoiser on the HIP BF16-MFMA kernels (no reference counterpart).

:class:`Denoiser` wraps ``mcrt_denoiser_*``: a 6-layer 3x3 conv net
(10 -> 32 x5 -> 3, ReLU, global skip in log space), trained with Adam on
(low-spp, high-spp) pairs that the renderer produces itself.  The bilateral
filter (``mcrt_denoise_bilateral_dev``) remains the fallback.

Training data: seeded variants of the Cornell box and of the displaced
icosphere scene with a random camera, random materials and a random light
position, rendered at 256 x 256 (noisy at `spp_in`, reference at `spp_ref`,
both with the first-hit albedo and G-buffer of the noisy render).

Evaluation (:func:`time_to_rel_mse`): the 1920 x 1080 icosphere frame,
rendered at increasing spp and denoised by either filter; relMSE is taken
on fixed 32 x 32 windows against a 65,536-spp FP64 reference of the same
windows, and the figure of merit is the wall time of render + denoise at the
first spp whose relMSE is at most 1e-2.

    python -m mcrt_amd.learned train --out weights.npy   # on a GPU
    python -m mcrt_amd.learned eval --weights weights.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from dataclasses import dataclass
from typing import List

import numpy as np

from . import camera as cam_mod
from . import configs, scene as scene_mod, scenes
from ._lib import Context, McrtError, MCRT_OK, load

N_PARAMS = 40771
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(PKG_DIR, "weights", "denoiser_c4.npy")
WEIGHTS_NEE_PATH = os.path.join(PKG_DIR, "weights", "denoiser_c4_nee.npy")  # trained on NEE renders
REFERENCE_CROPS = os.path.join(os.path.dirname(PKG_DIR), "tests", "golden", "c4_reference_crops.npz")
# Full-frame reference of the same frame and frames (GPU, 65,536 spp, float16;
# tests/golden/make_c4_full_reference.py), checked against the windows.
REFERENCE_FULL = os.path.join(os.path.dirname(PKG_DIR), "tests", "golden", "c4_full_reference.npz")
EVAL_W, EVAL_H, EVAL_DEPTH = 1920, 1080, 8
INFER_FLOP_PER_PIXEL = 81216  # 2 * 9 * (10*32 + 4*32*32 + 32*3) multiply-adds of the six convs


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

    LOSSES = {"log_l2": 0, "rel_mse": 1}

    def set_loss(self, kind: str) -> None:
        """Training loss: "log_l2" (L2 of log1p, default) or "rel_mse"
        ((x - r)^2 / (r^2 + 1e-2), the evaluation metric)."""
        self.ctx._check(self._lib.mcrt_denoiser_set_loss(self._h, self.LOSSES[kind]))

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
        v, t, _ = scenes.displaced_icosphere(n=59, seed=seed, mat_base=base.name_to_idx["m0"], n_mats=8)
        extra = (v, t)
    loaded = scene_mod.load_scene_from_str(text, f"variant-{seed}", extra_mesh=extra)
    return loaded, scene_mod.build_scene_buffers(loaded)


@dataclass
class Pairs:
    noisy: object   # torch (L, N, H, W, 4): one input per sample count of spp_levels
    albedo: object  # (L, N, H, W, 4)
    gbuf: object    # (L, N, H, W, 4)
    ref: object     # (N, H, W, 4)
    spp_levels: tuple = (4,)


def reference_render(ctx: Context, cam, W: int, H: int, spp_in: int, spp_ref: int, depth: int, noisy, ref,
                     tmp_gbuf, stream) -> None:
    """Mean of frames spp_in .. spp_in+spp_ref-1, independent of the noisy
    input (frames 0 .. spp_in-1).  The renderer's running mean weights frame
    f by 1/(f+1), i.e. it is a plain mean only from frame 0, so render frames
    0 .. spp_in+spp_ref-1 and remove the noisy frames' share:
    ref = ((spp_in+spp_ref) * mean_all - spp_in * noisy) / spp_ref."""
    n_all = spp_in + spp_ref
    if n_all > 65535:
        raise ValueError("reference range exceeds the renderer's frame-count clamp")
    ctx.render_dev(cam, W, H, n_all, 0, depth, None, ref.data_ptr(), tmp_gbuf.data_ptr(), stream)
    ref.mul_(n_all / spp_ref).sub_(noisy, alpha=spp_in / spp_ref)


def render_pairs(ctx: Context, n_variants: int, size: int = 256, spp_in=4, spp_ref: int = 4096,
                 depth: int = 8, seed0: int = 1000, log=None, nee: bool = False) -> Pairs:
    """Training pairs; with `nee`, input and reference are both rendered
    with next-event estimation (same expectation, lower noise).  `spp_in` is
    one sample count or several: each variant then has one input per count
    (the first s frames, 0 .. s-1) and one reference, the mean of the
    spp_ref frames after the largest count, independent of every input."""
    levels = tuple(int(v) for v in (spp_in if isinstance(spp_in, (list, tuple)) else (spp_in,)))
    ctx.set_nee(nee)
    try:
        return _render_pairs(ctx, n_variants, size, levels, spp_ref, depth, seed0, log)
    finally:
        ctx.set_nee(False)


def _render_pairs(ctx, n_variants, size, levels, spp_ref, depth, seed0, log) -> Pairs:
    import torch

    dev = torch.device("cuda:0")
    L = len(levels)
    top = levels.index(max(levels))
    noisy = torch.empty((L, n_variants, size, size, 4), dtype=torch.float32, device=dev)
    albedo = torch.empty_like(noisy)
    gbuf = torch.empty_like(noisy)
    ref = torch.empty((n_variants, size, size, 4), dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    tmp_g = torch.empty((size, size, 4), dtype=torch.float32, device=dev)
    for i in range(n_variants):
        kind = "c1" if i % 2 == 0 else "c2"
        loaded, sb = variant_scene(seed0 + i, kind)
        cam = cam_mod.camera_for(loaded, size, size)
        ctx.upload_scene(sb)
        for l, spp in enumerate(levels):
            ctx.render_aux_dev(cam, size, size, spp, 0, depth, None, noisy[l, i].data_ptr(), gbuf[l, i].data_ptr(),
                               albedo[l, i].data_ptr(), s)
        reference_render(ctx, cam, size, size, levels[top], spp_ref, depth, noisy[top, i], ref[i], tmp_g, s)
        if log and (i % 16 == 0 or i == n_variants - 1):
            torch.cuda.synchronize()
            log(f"pairs {i + 1}/{n_variants}")
    torch.cuda.synchronize()
    return Pairs(noisy, albedo, gbuf, ref, levels)


def train(dn: Denoiser, pairs: Pairs, steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 0,
          log=None, loss: str = "log_l2", warm_steps: int = 0) -> List[float]:
    """Adam steps on random batches; the first `warm_steps` use the log-L2
    loss, the rest `loss`.  The learning rate decays from lr to lr/10."""
    import torch

    L, n = pairs.noisy.shape[0], pairs.noisy.shape[1]
    rng = np.random.default_rng(seed)
    losses = []
    dn.set_loss("log_l2" if warm_steps > 0 else loss)
    for it in range(steps):
        if it == warm_steps and warm_steps > 0:
            dn.set_loss(loss)
        idx = torch.as_tensor(rng.choice(n, size=min(batch, n), replace=False), device=pairs.noisy.device)
        # each batch entry takes the input of a random sample count
        lvl = torch.as_tensor(rng.integers(L, size=len(idx)), device=pairs.noisy.device)
        lr_t = lr * (0.1 ** (it / max(steps, 1)))
        # Augmentation: one of the 8 symmetries of the square tile, shared by
        # the batch (the G-buffer normals stay world-space features).
        k, flip = int(rng.integers(4)), bool(rng.integers(2))

        def aug(t):
            t = t[lvl, idx] if t.dim() == 5 else t[idx]
            if flip:
                t = torch.flip(t, dims=(2,))
            return torch.rot90(t, k, dims=(1, 2)).contiguous() if k else t.contiguous()

        value = dn.train_step(aug(pairs.noisy), aug(pairs.albedo), aug(pairs.gbuf), aug(pairs.ref), lr_t,
                              want_loss=(it % 50 == 0 or it == steps - 1))
        if value is not None:
            losses.append(value)
            if log:
                log(f"train step {it}: loss {value:.4e}")
    return losses


def save_weights(dn: Denoiser, path: str, meta: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.save(path, dn.get_weights())
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load_trained(ctx: Context, path: str = WEIGHTS_PATH) -> Denoiser:
    """A denoiser with the committed trained weights (raises if absent)."""
    w = np.load(path, allow_pickle=False)
    dn = Denoiser(ctx, seed=0)
    dn.set_weights(w)
    return dn


# ---------------------------------------------------------------------------
# Evaluation: wall time to relMSE <= 1e-2 at 1080p
# ---------------------------------------------------------------------------
def rel_mse(x: np.ndarray, y: np.ndarray) -> float:
    """mean over pixels and RGB of (x - y)^2 / (y^2 + 1e-2)."""
    x = x[..., :3].astype(np.float64)
    y = y[..., :3].astype(np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))


def eval_workload():
    return configs.c2(width=EVAL_W, height=EVAL_H, spp=1, depth=EVAL_DEPTH)


def load_reference_crops(path: str = REFERENCE_CROPS):
    d = np.load(path, allow_pickle=False)
    return d["crops"].astype(np.float64), [tuple(int(v) for v in o) for o in d["origins"]], int(d["spp"])


def load_full_reference(path: str = REFERENCE_FULL):
    """(H, W, 3) float32 full-frame reference, or None without the fixture."""
    if not os.path.exists(path):
        return None
    d = np.load(path, allow_pickle=False)
    return d["image"].astype(np.float32)


def crops_rel_mse(img: np.ndarray, crops: np.ndarray, origins) -> float:
    c = crops.shape[1]
    x = np.stack([img[y0:y0 + c, x0:x0 + c, :3] for x0, y0 in origins])
    return rel_mse(x, crops)


def time_to_rel_mse(ctx: Context, wl, method: str, spps, crops, origins, dn: Denoiser = None,
                    threshold: float = 1e-2, reps: int = 3, log=None, full=None):
    """Sweep spp upward; at each, time render (+ denoise) over `reps` runs
    (median, inputs resident, scene uploaded) and take relMSE on the
    reference windows and, with `full` (the full-frame reference), on every
    pixel.  Stops at the first spp at or below `threshold` on the full frame
    (on the windows without it).  method: "none", "bilateral" or "learned",
    optionally prefixed "nee_" (or "nee" alone) for renders with next-event
    estimation."""
    ctx.set_nee(method.startswith("nee"))
    try:
        return _time_to_rel_mse(ctx, wl, method.replace("nee_", "").replace("nee", "none"), spps, crops,
                                origins, dn, threshold, reps, log, method, full)
    finally:
        ctx.set_nee(False)


def _time_to_rel_mse(ctx, wl, method, spps, crops, origins, dn, threshold, reps, log, label, full=None):
    import torch

    dev = torch.device("cuda:0")
    W, H = wl.width, wl.height
    s = torch.cuda.current_stream(dev).cuda_stream
    accum = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    gbuf, albedo, out = torch.empty_like(accum), torch.empty_like(accum), torch.empty_like(accum)
    def run(spp):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if method == "learned":
            ctx.render_aux_dev(wl.camera, W, H, spp, 0, wl.depth, None, accum.data_ptr(), gbuf.data_ptr(),
                               albedo.data_ptr(), s)
            dn.infer(accum[None], albedo[None], gbuf[None], out[None], s)
        else:
            ctx.render_dev(wl.camera, W, H, spp, 0, wl.depth, None, accum.data_ptr(), gbuf.data_ptr(), s)
            if method == "bilateral":
                ctx.denoise_bilateral_dev(None, W, H, accum.data_ptr(), gbuf.data_ptr(), out.data_ptr(), s)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run(1)  # warm-up: first launches and the per-size buffers
    rows, hit = [], None
    for spp in spps:
        # Above 128 spp one run lasts ~0.5 s or more and varies little.
        times = [run(spp) for _ in range(reps if spp <= 128 else 1)]
        img = (accum if method == "none" else out).cpu().numpy()
        err_win = crops_rel_mse(img, crops, origins)
        row = {"spp": spp, "ms": round(float(np.median(times)) * 1e3, 3), "relmse_windows": round(err_win, 6)}
        err = err_win
        if full is not None:
            err = rel_mse(img, full)
            row["relmse_full"] = round(err, 6)
        row["relmse"] = round(err, 6)
        rows.append(row)
        if log:
            log(f"{label} spp {spp}: {row['ms']:.2f} ms relMSE {err:.4e} (windows {err_win:.4e})")
        if err <= threshold:
            hit = row
            break
        # A method whose error has stopped falling (a biased denoiser) will
        # not reach the threshold: stop when the last three points improve
        # on the best before them by less than 10%.
        if len(rows) >= 6:
            before = min(r["relmse"] for r in rows[:-3])
            if min(r["relmse"] for r in rows[-3:]) > 0.9 * before:
                break
    return {"reached": hit, "sweep": rows}


def time_infer(dn: Denoiser, W: int, H: int, reps: int = 10) -> float:
    """Median seconds of one learned-denoiser inference on a W x H frame."""
    import torch

    dev = torch.device("cuda:0")
    x = torch.rand((1, H, W, 4), device=dev)
    g = torch.rand((1, H, W, 4), device=dev)
    out = torch.empty_like(x)
    s = torch.cuda.current_stream(dev).cuda_stream
    dn.infer(x, g, g, out, s)
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dn.infer(x, g, g, out, s)
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    return float(np.median(ts))


def evaluate(ctx: Context, dn: Denoiser, reps: int = 3, log=None, dn_nee: Denoiser = None) -> dict:
    """The sweep for every method; `dn_nee` (weights trained on renders with
    next-event estimation) denoises the NEE renders when given."""
    crops, origins, ref_spp = load_reference_crops()
    full = load_full_reference()
    wl = eval_workload()
    ctx.upload_scene(wl.buffers)
    # Steps of at most ~1/3 so that neither filter is charged for a coarse grid.
    spps = [1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 56, 64, 80, 96, 128, 160, 192,
            256, 320, 384, 512, 640, 768, 1024, 1280, 1536, 2048]
    res = {"frame": f"icosphere_c2 {EVAL_W}x{EVAL_H} depth {EVAL_DEPTH}",
           "reference": (f"every pixel: GPU {ref_spp}-spp reference of the oracle's frames, checked against "
                         f"{len(origins)} FP64 oracle windows" if full is not None else
                         f"{len(origins)} windows {crops.shape[1]}x{crops.shape[2]}, {ref_spp} spp FP64 oracle"),
           "relmse_over": "full frame" if full is not None else "windows",
           "threshold": 1e-2}
    # The metric's comparison (plain renders), then the same with next-event
    # estimation.
    for method in ("learned", "bilateral", "none", "nee_learned", "nee_bilateral", "nee"):
        net = dn_nee if (method == "nee_learned" and dn_nee is not None) else dn
        res[method] = time_to_rel_mse(ctx, wl, method, spps, crops, origins, dn=net, reps=reps, log=log,
                                      full=full)
    res["nee_learned_weights"] = "trained on NEE renders" if dn_nee is not None else "trained on plain renders"
    t_inf = time_infer(dn, EVAL_W, EVAL_H)
    res["infer_ms"] = round(t_inf * 1e3, 3)
    res["infer_tflops"] = round(INFER_FLOP_PER_PIXEL * EVAL_W * EVAL_H / t_inf / 1e12, 2)
    lt = res["learned"]["reached"]
    bt = res["bilateral"]["reached"]
    res["speedup_vs_bilateral"] = (round(bt["ms"] / lt["ms"], 3) if lt and bt else None)
    return res


def _main(argv=None):
    import torch

    ap = argparse.ArgumentParser(prog="python -m mcrt_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--variants", type=int, default=512)
    t.add_argument("--spp-in", default="2,4,8,16", help="input sample count, or several separated by commas")
    t.add_argument("--spp-ref", type=int, default=4096)
    t.add_argument("--steps", type=int, default=192000)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--loss", choices=sorted(Denoiser.LOSSES), default="rel_mse")
    t.add_argument("--warm-steps", type=int, default=2000, help="log-L2 steps before --loss")
    t.add_argument("--nee", action="store_true", help="render inputs and references with next-event estimation")
    t.add_argument("--out", default=None, help="default: the package's weights file for the estimator")
    t.add_argument("--init", default=None, help="continue training from these weights (their .json beside them)")
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS_PATH)
    e.add_argument("--weights-nee", default=WEIGHTS_NEE_PATH)
    e.add_argument("--reps", type=int, default=3)
    ib = sub.add_parser("infer-bench", help="time inference on a 1920 x 1080 frame (committed weights)")
    ib.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    torch.cuda.init()
    with Context(0) as ctx:
        if a.cmd == "train":
            t0 = time.time()
            out = a.out or (WEIGHTS_NEE_PATH if a.nee else WEIGHTS_PATH)
            spp_in = [int(v) for v in str(a.spp_in).split(",")]
            pairs = render_pairs(ctx, a.variants, spp_in=spp_in, spp_ref=a.spp_ref, log=log, nee=a.nee)
            t_pairs = time.time() - t0
            log(f"pairs rendered in {t_pairs:.1f} s")
            dn = Denoiser(ctx, seed=a.seed)
            init_meta = None
            if a.init:
                dn.set_weights(np.load(a.init, allow_pickle=False))
                side = os.path.splitext(a.init)[0] + ".json"
                if os.path.exists(side):
                    with open(side) as f:
                        init_meta = json.load(f)
            t1 = time.time()
            losses = train(dn, pairs, a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log, loss=a.loss,
                           warm_steps=a.warm_steps)
            torch.cuda.synchronize()
            t_train = time.time() - t1
            log(f"trained {a.steps} steps in {t_train:.1f} s")
            save_weights(dn, out, {"variants": a.variants, "size": 256, "spp_in": spp_in,
                                   "spp_ref": a.spp_ref, "nee": a.nee, "steps": a.steps, "batch": a.batch,
                                   "lr": a.lr, "seed": a.seed, "loss": a.loss, "warm_steps": a.warm_steps,
                                   "loss_first": losses[0], "loss_last": losses[-1],
                                   "pairs_s": round(t_pairs, 1), "train_s": round(t_train, 1),
                                   **({"init": os.path.basename(a.init), "init_meta": init_meta} if a.init else {})})
            log(f"weights -> {out}")
        elif a.cmd == "infer-bench":
            dn = load_trained(ctx)
            t = time_infer(dn, EVAL_W, EVAL_H, reps=a.reps)
            print(json.dumps({"frame": f"{EVAL_W}x{EVAL_H}", "infer_ms": round(t * 1e3, 4),
                              "tflops": round(INFER_FLOP_PER_PIXEL * EVAL_W * EVAL_H / t / 1e12, 1)}), flush=True)
        else:
            dn = load_trained(ctx, a.weights)
            dn_nee = load_trained(ctx, a.weights_nee) if os.path.exists(a.weights_nee) else None
            print(json.dumps(evaluate(ctx, dn, reps=a.reps, log=log, dn_nee=dn_nee)), flush=True)


if __name__ == "__main__":
    sys.exit(_main())
