#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

One step = one full progressive render of the workload (every frame of its
spp, every bounce) through ``mcrt_render_dev`` on device-resident buffers.
The scene is uploaded before timing; the timed region holds only GPU work.

Output: one JSON line with Msamples/s, the roofline of the dominant kernel
(``extend``, the closest-hit pass) measured with HIP events on its launch
stream, and the CPU oracle timed on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--workload c3|c2|c1]

Default workload: the voxel terrain at 1920 x 1080, 16 spp, depth 6, with the
bilateral denoiser in every step (the 1080p configuration of the metric).
The 16M-ray closest-hit microbenchmark runs on the displaced-icosphere scene.

Beside the main figure the line carries:
- ``denoise_eval``: wall time of render + denoise until relMSE <= 1e-2 on the
  1920 x 1080 icosphere frame, learned (BF16 MFMA) vs bilateral vs none, and
  the learned denoiser's inference rate against the BF16 MFMA peak;
- ``inverse``: 200 Adam steps of inverse rendering through the path-replay
  adjoint (icosphere scene, 512 x 512, 32 spp, depth 6, 8 albedos).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1080p; time to relMSE≤1e-2, learned vs bilateral denoise"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
# Measured bytes between L2 and memory per k_extend launch (FETCH_SIZE x 2 +
# WRITE_SIZE x 1, factors from a 1 GiB copy in the same rocprofv3 --pmc runs),
# per workload at its default size; profiles/traffic_extend_voxel_c3_1080p.md.
EXTEND_TRAFFIC_BYTES = {"voxel_c3": 12.358e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="c3")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--no-microbench", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-denoise-eval", action="store_true",
                   help="skip the 1080p time-to-relMSE comparison of the learned and bilateral denoisers")
    p.add_argument("--no-inverse", action="store_true",
                   help="skip the 200-step inverse-rendering (path-replay adjoint + Adam) timing")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="first copy MIB MiB device-to-device three times (a plain copy of known size that "
                        "calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU design is "
              "described but not built", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from mcrt_amd import configs
    from mcrt_amd._lib import Context

    log(f"building workload {a.workload}")
    w = configs.by_name(a.workload)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.copy_calibration > 0:
        n = a.copy_calibration << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {a.copy_calibration} MiB read + written")
        del src, dst
    ctx = Context(0)
    ctx.upload_scene(w.buffers)
    accum = torch.empty((w.height, w.width, 4), dtype=torch.float32, device=dev)
    gbuf = torch.empty_like(accum)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    denoised = torch.empty_like(accum) if w.denoise == "bilateral" else None

    def step():
        ctx.render_dev(w.camera, w.width, w.height, w.spp, 0, w.depth, None, accum.data_ptr(),
                       gbuf.data_ptr(), sp)
        if denoised is not None:
            ctx.denoise_bilateral_dev(None, w.width, w.height, accum.data_ptr(), gbuf.data_ptr(),
                                      denoised.data_ptr(), sp)

    log(f"warmup {a.warmup}")
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {a.steps} steps")
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = (t1 - t0) / a.steps
    samples = w.width * w.height * w.spp
    value = samples / sec / 1e6
    segments, _ = ctx.render_counters()
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scenes; no external assets)",
        "config": {
            "workload": w.name,
            "width": w.width,
            "height": w.height,
            "spp": w.spp,
            "max_depth": w.depth,
            "triangles": int(len(w.buffers.triangles)),
            "denoise": w.denoise,
        },
        "mrays_per_s": round(segments / sec / 1e6, 3),
        "segments_per_step": int(segments),
    }

    if not a.no_roofline:
        out["roofline"] = roofline(ctx, w, step, segments)
    if not a.no_microbench:
        log("microbench: 16M rays on the icosphere scene")
        mb = configs.c2(width=64, height=64, spp=1)
        ctx.upload_scene(mb.buffers)
        out["microbench"] = microbench(ctx, mb)
    if not a.no_denoise_eval:
        out["denoise_eval"] = denoise_eval(ctx)
    if not a.no_inverse:
        out["inverse"] = inverse(ctx)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds)
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


def roofline(ctx, w, step, segments):
    """Dominant kernel: extend (closest hit).  Algorithmic bytes per segment
    B_ray = 24 (ray in) + 8 (hit out) + 32 * nodes + 36 * triangles, with the
    node and triangle counts of the reference-order traversal measured by the
    CPU oracle on a 1-spp sample of the same render."""
    import torch

    import oracle

    ctx.set_profiling(True)
    ctx.reset_stats()
    step()
    torch.cuda.synchronize()
    ms, launches = ctx.kernel_stats("extend")
    shade_ms, _ = ctx.kernel_stats("shade")
    ctx.set_profiling(False)
    log("counting reference traversal work with the oracle (1 spp sample)")
    sw, sh = w.width // 4, w.height // 4
    cam = __import__("mcrt_amd.camera", fromlist=["camera_for"]).camera_for(w.loaded, sw, sh)
    _, _, cnt = oracle.render(oracle.Scene(w.buffers), cam, sw, sh, 1, 0, w.depth, counts=True)
    nodes, tris, sph, seg = cnt
    b_ray = 24.0 + 8.0 + 32.0 * nodes / seg + 36.0 * tris / seg + 32.0 * sph / seg
    total_bytes = b_ray * segments
    achieved = total_bytes / (ms * 1e-3) / 1e9
    avg_ms = ms / max(launches, 1)
    traffic = EXTEND_TRAFFIC_BYTES.get(w.name)
    return {
        "kernel": "extend",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_gbs": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
        "traffic_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        "algorithmic_bytes_per_launch": round(total_bytes / max(launches, 1)),
        "avg_launch_ms": round(avg_ms, 4),
        "launches": int(launches),
        "bytes_per_segment": round(b_ray, 1),
        "nodes_per_segment": round(nodes / seg, 2),
        "tris_per_segment": round(tris / seg, 2),
        "extend_ms_per_step": round(ms, 3),
        "shade_ms_per_step": round(shade_ms, 3),
    }


def microbench(ctx, w, n=1 << 24, reps=5):
    """Closest hit of 16M rays (origins uniform in the scene box, directions
    uniform on the sphere) against the workload's scene; ids + t out."""
    import torch

    from mcrt_amd import configs, scenes

    lo, hi = configs.scene_aabb(w.buffers)
    o, d = scenes.microbench_rays(n, lo, hi, seed=2024)
    dev = torch.device("cuda:0")
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    ids = torch.empty(n, dtype=torch.int32, device=dev)
    ts = torch.empty(n, dtype=torch.float32, device=dev)
    sp = torch.cuda.current_stream(dev).cuda_stream

    def go():
        ctx.trace_closest_dev(to.data_ptr(), td.data_ptr(), n, 1e-4, 1e9, ids.data_ptr(),
                              ts.data_ptr(), sp)

    go()
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        go()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    ms, launches = ctx.kernel_stats("trace_closest")
    ctx.set_profiling(False)
    kernel_ms = ms / max(launches, 1)
    hit = float((ids != -1).float().mean().item())
    return {"rays": n, "mrays_per_s": round(n / el / 1e6, 1),
            "kernel_ms": round(kernel_ms, 3), "kernel_mrays_per_s": round(n / (kernel_ms * 1e-3) / 1e6, 1),
            "hit_fraction": round(hit, 4)}


def denoise_eval(ctx):
    """Wall time of render + denoise to relMSE <= 1e-2 on the 1920 x 1080
    icosphere frame: learned (committed weights) vs bilateral vs none; relMSE
    on fixed windows against the 65,536-spp FP64 reference of those windows."""
    from mcrt_amd import learned

    missing = [p for p in (learned.WEIGHTS_PATH, learned.REFERENCE_CROPS) if not os.path.exists(p)]
    if missing:
        return {"skipped": "missing " + ", ".join(os.path.relpath(p, ROOT) for p in missing)}
    log("denoise eval: 1080p icosphere, learned vs bilateral")
    dn = learned.load_trained(ctx)
    dn_nee = learned.load_trained(ctx, learned.WEIGHTS_NEE_PATH) if os.path.exists(learned.WEIGHTS_NEE_PATH) else None
    res = learned.evaluate(ctx, dn, reps=3, log=log, dn_nee=dn_nee)
    dn.close()
    if dn_nee is not None:
        dn_nee.close()
    res["infer_mfma"] = {"bound": "mfma", "achieved": res["infer_tflops"], "peak": MFMA_BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(res["infer_tflops"] / MFMA_BF16_PEAK_TFLOPS, 4)}
    return res


def inverse(ctx, steps=200):
    """Inverse rendering: 200 Adam steps on 8 material albedos through the
    path-replay adjoint, icosphere scene 512 x 512, 32 spp, depth 6."""
    import torch

    from mcrt_amd import diff

    log(f"inverse rendering: {steps} Adam steps")
    inv = diff.inverse_workload()
    ctx.upload_scene(inv.work.buffers)
    w = inv.work
    target_spp = 8 * w.spp
    diff.optimize(ctx, inv, steps=2, target_spp=target_spp)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = diff.optimize(ctx, inv, steps=steps, target_spp=target_spp)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"workload": f"{w.name} {w.width}x{w.height} {w.spp} spp depth {w.depth}, {len(inv.init)} "
                        f"materials x (RGB albedo, emission); target {target_spp} spp, independent frames; "
                        "decorrelated primal / replay frames per step",
            "steps": steps, "total_s": round(el, 3), "ms_per_step": round(el / steps * 1e3, 3),
            # samples of the primal render plus samples of the adjoint (each traced and replayed)
            "msamples_fwd_plus_adj_per_s": round(2 * w.width * w.height * w.spp * steps / el / 1e6, 1),
            "loss_first": round(hist[0][0], 6), "loss_last": round(hist[-1][0], 6),
            "albedo_err_first": {"max": round(hist[0][1], 4), "mean": round(hist[0][2], 4)},
            "albedo_err_last": {"max": round(hist[-1][1], 4), "mean": round(hist[-1][2], 4)}}


def cpu_baseline(w, seconds):
    """The CPU oracle (FP32 C restatement of the reference shader, pthreads)
    rendering whole frames of the same workload until ~`seconds` elapse."""
    import oracle

    threads = min(os.cpu_count() or 1, 16)
    sc = oracle.Scene(w.buffers)
    frames = 0
    t0 = time.perf_counter()
    acc = None
    while True:
        acc, _ = oracle.render(sc, w.camera, w.width, w.height, 1, frames, w.depth, accum_in=acc,
                               threads=threads)
        frames += 1
        el = time.perf_counter() - t0
        if el >= seconds or frames >= w.spp:
            break
    value = w.width * w.height * frames / el / 1e6
    return {
        "value": round(value, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} of {w.spp} frames of {w.name} at {w.width}x{w.height}, depth {w.depth}",
    }


if __name__ == "__main__":
    sys.exit(main())
