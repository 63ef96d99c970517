#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

One step = one full progressive render of the workload (every frame of its
spp, every bounce) through ``mcrt_render_dev`` on device-resident buffers,
plus the bilateral pass.  The scene is uploaded before timing; the timed
region holds only GPU work.

Output: one JSON line.  ``value`` is Msamples/s of the 1920 x 1080 voxel
terrain at 16 spp, depth 6, with the bilateral denoiser (the 1080p
configuration of the metric).  Beside it:

- ``roofline``: the dominant kernel (``extend``, closest hit).  Its launch
  time comes from HIP events on its stream; its HBM bytes per launch
  (FETCH_SIZE, WRITE_SIZE) and its VALU issue and lane counts (SQ counters)
  from ``rocprofv3 --pmc`` passes that this script runs on itself (``--probe``)
  in the same invocation, each counter scaled by what it reads for a copy
  of known size in the same pass;
- ``peaks_measured``: HBM copy bandwidth and dense BF16 MFMA rate measured
  on the device (``mcrt_measure_peaks``);
- ``cornell_c1`` (256 x 256, 16 spp, depth 8) and ``icosphere_c2`` (1024 x
  1024, 64 spp, depth 8) render lines, relMSE against the 65,536-spp FP64
  oracle where its fixture exists;
- ``microbench``: 16M closest-hit rays on the icosphere scene;
- ``denoise_eval``: time to relMSE <= 1e-2, learned vs bilateral (1080p);
- ``inverse``: 200 Adam steps through the path-replay adjoint;
- ``cpu_baseline``: the CPU oracle on the host's cores.

    python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import signal
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1080p; time to relMSE≤1e-2, learned vs bilateral denoise"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16, published
N_XCD = 8
COPY_PROBE_BYTES = 1 << 30
EXTEND_C3 = "k_extend<256, false, true>"   # culling stack (the terrain spills the L2), bounces 1+
EXTEND_C3_FIRST = "k_extend_packet<256, false>"  # the same scene's first bounce, wave-uniform packets
EXTEND_C2 = "k_extend<256, false, false>"  # plain stack (cache-resident scene), bounces 1+
EXTEND_C2_FIRST = "k_extend_packet<256, true>"   # its first bounce: packets, camera rays computed in the kernel
COPY_KERNEL = "k_copy_probe"
VALU_KERNEL = "k_valu_probe"
PMC_PASSES = [
    ["FETCH_SIZE"],
    ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum"],
    ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_ACTIVE_INST_VALU", "SQ_INSTS_VMEM_RD",
     "SQ_WAVE_CYCLES", "GRBM_GUI_ACTIVE", "TA_BUSY_avr"],
]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="c3")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter passes")
    p.add_argument("--no-microbench", action="store_true")
    p.add_argument("--no-lines", action="store_true", help="skip the C1 and C2 render lines")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-denoise-eval", action="store_true",
                   help="skip the 1080p time-to-relMSE comparison of the learned and bilateral denoisers")
    p.add_argument("--no-inverse", action="store_true",
                   help="skip the 200-step inverse-rendering (path-replay adjoint + Adam) timing")
    p.add_argument("--counters-dir", default=None,
                   help="keep the rocprofv3 counter CSVs here (default: a temporary directory)")
    p.add_argument("--probe", default=None, metavar="DIR",
                   help="internal: run the counter probe on the scenes cached in DIR and exit")
    return p.parse_args()


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


# ---------------------------------------------------------------------------
# Scene cache shared with the counter probe (a structured-array .npz, no pickle)
# ---------------------------------------------------------------------------
def save_workload(w, path):
    import numpy as np

    sb = w.buffers
    np.savez(path, **{k: v for k, v in sb.__dict__.items() if isinstance(v, np.ndarray)},
             n_real_spheres=sb.n_real_spheres, n_real_tris=sb.n_real_tris, camera=w.camera,
             size=np.array([w.width, w.height, w.spp, w.depth]), name=np.array(w.name),
             denoise=np.array(w.denoise or ""))


def load_workload(path):
    import numpy as np

    from mcrt_amd import configs
    from mcrt_amd.scene import SceneBuffers

    z = np.load(path, allow_pickle=False)
    fields = {k: z[k] for k in ("spheres", "sphere_orig", "sphere_bvh", "vertices", "triangles", "tri_orig",
                                "mesh_bvh", "materials", "env", "emissive_spheres")}
    sb = SceneBuffers(**fields, n_real_spheres=int(z["n_real_spheres"]), n_real_tris=int(z["n_real_tris"]))
    W, H, spp, depth = (int(v) for v in z["size"])
    return configs.Workload(str(z["name"]), None, sb, z["camera"], W, H, spp, depth, str(z["denoise"]) or None)


def render_step(ctx, w, accum, gbuf, denoised, sp):
    ctx.render_dev(w.camera, w.width, w.height, w.spp, 0, w.depth, None, accum.data_ptr(), gbuf.data_ptr(), sp)
    if denoised is not None:
        ctx.denoise_bilateral_dev(None, w.width, w.height, accum.data_ptr(), gbuf.data_ptr(), denoised.data_ptr(),
                                  sp)


def probe(cache_dir):
    """Counter probe: one copy of known size, then one step of each cached
    workload.  Run under rocprofv3 --pmc by counter_passes()."""
    import torch

    from mcrt_amd._lib import Context, measure_peaks

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # its copy dispatches calibrate FETCH_SIZE / WRITE_SIZE, its FP32 fma
    # dispatches the cycles one VALU instruction holds a SIMD
    measure_peaks(0, COPY_PROBE_BYTES)
    with Context(0) as ctx:
        for name in ("c3", "c2"):
            w = load_workload(os.path.join(cache_dir, f"{name}.npz"))
            ctx.upload_scene(w.buffers)
            accum = torch.empty((w.height, w.width, 4), dtype=torch.float32, device=dev)
            gbuf = torch.empty_like(accum)
            den = torch.empty_like(accum) if w.denoise == "bilateral" else None
            render_step(ctx, w, accum, gbuf, den, torch.cuda.current_stream(dev).cuda_stream)
            torch.cuda.synchronize()
    return 0


def main():
    a = parse()
    if a.probe:
        return probe(a.probe)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU design is "
              "described but not built", file=sys.stderr)
        return 2
    import torch

    from mcrt_amd import configs
    from mcrt_amd._lib import Context

    log(f"building workload {a.workload}")
    w = configs.by_name(a.workload)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = Context(0)
    ctx.upload_scene(w.buffers)
    accum = torch.empty((w.height, w.width, 4), dtype=torch.float32, device=dev)
    gbuf = torch.empty_like(accum)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    denoised = torch.empty_like(accum) if w.denoise == "bilateral" else None

    def step():
        render_step(ctx, w, accum, gbuf, denoised, sp)

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
    if w.name == "voxel_c3":  # the oracle-B windows are of the voxel terrain frame
        relmse_c3 = relmse_windows(w, accum, denoised, "c3_oracle_b.npz")
        if relmse_c3:
            out["relmse_vs_oracle_b"] = relmse_c3

    lines = {}
    if not a.no_lines:
        lines["cornell_c1"] = c1_line(ctx)
        lines["icosphere_c2"] = c2_line(ctx)
        ctx.upload_scene(w.buffers)

    counters = None
    if not a.no_counters and not a.no_roofline:
        counters = counter_passes(w, a.counters_dir)
    peaks = measured_peaks()
    out["peaks_measured"] = peaks
    if not a.no_roofline:
        out["roofline"] = roofline(ctx, w, step, segments, counters, peaks)
    if "icosphere_c2" in lines and counters and counters.get(EXTEND_C2):
        lines["icosphere_c2"]["extend_counters"] = kernel_fractions(counters, EXTEND_C2,
                                                                    lines["icosphere_c2"].get("extend_ms"))
    out.update(lines)
    if not a.no_microbench:
        log("microbench: 16M rays on the icosphere scene")
        mb = configs.c2(width=64, height=64, spp=1)
        ctx.upload_scene(mb.buffers)
        out["microbench"] = microbench(ctx, mb)
    if not a.no_denoise_eval:
        out["denoise_eval"] = denoise_eval(ctx, peaks)
    if not a.no_inverse:
        out["inverse"] = inverse(ctx)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds)
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


# ---------------------------------------------------------------------------
# Counters (rocprofv3 --pmc on this script's --probe)
# ---------------------------------------------------------------------------
def _rocprofv3():
    p = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    return p if os.path.exists(p) else None


def counter_passes(w, keep_dir=None):
    """Per-dispatch means of every counter, by kernel name (template
    arguments included), over the three passes; None when rocprofv3 is
    unavailable or a pass fails."""
    import numpy as np

    from mcrt_amd import configs

    prof = _rocprofv3()
    if not prof:
        log("counters: rocprofv3 not found, skipped")
        return None
    work = tempfile.mkdtemp(prefix="mcrt_bench_")
    out_dir = keep_dir or work
    try:
        save_workload(w, os.path.join(work, "c3.npz"))
        save_workload(configs.c2(), os.path.join(work, "c2.npz"))
        means = {}
        for k, counters in enumerate(PMC_PASSES):
            d = os.path.join(out_dir, f"pass{k}")
            cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "p", "--",
                   sys.executable, os.path.abspath(__file__), "--probe", work]
            log(f"counters: pass {k + 1}/{len(PMC_PASSES)} ({' '.join(counters)})")
            env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
            p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                 start_new_session=True)
            try:
                txt, _ = p.communicate(timeout=240)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                log(f"counters: pass {k + 1} timed out; counters skipped")
                return None
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            if p.returncode != 0 and not files:
                log(f"counters: pass {k + 1} failed (rc {p.returncode}): {txt.decode(errors='replace')[-400:]}")
                return None
            for f in files:
                acc = {}
                for r in csv.DictReader(open(f)):
                    acc.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], []).append(
                        float(r["Counter_Value"]))
                for kern, cs in acc.items():
                    for c, vals in cs.items():
                        means.setdefault(kern, {})[c] = (float(np.mean(vals)), len(vals))
        return group_counter_means(means)
    finally:
        if keep_dir is None:
            shutil.rmtree(work, ignore_errors=True)
        else:
            for f in ("c3.npz", "c2.npz"):
                try:
                    os.remove(os.path.join(work, f))
                except OSError:
                    pass


def group_counter_means(means):
    """Per-dispatch means by kernel name -> by the short names the bench
    reports.  An extension step is two kernels (first bounce, later
    bounces); its means are taken over the dispatches of both, as the event
    timing of the step is."""
    groups = {EXTEND_C3: (EXTEND_C3, EXTEND_C3_FIRST), EXTEND_C2: (EXTEND_C2, EXTEND_C2_FIRST), COPY_KERNEL: (COPY_KERNEL,),
              VALU_KERNEL: (VALU_KERNEL,), "k_shade": ("k_shade",)}
    by_short = {}
    for short, names in groups.items():
        acc = {}
        for kern, cs in means.items():
            if any(n in kern for n in names):
                for c, (m, n) in cs.items():
                    tot, cnt = acc.get(c, (0.0, 0))
                    acc[c] = (tot + m * n, cnt + n)
        if acc:
            by_short[short] = {c: (tot / cnt, cnt) for c, (tot, cnt) in acc.items()}
    return by_short


def traffic_factors(counters):
    """Counter scale factors from the probe copy (1 GiB read + 1 GiB written
    per dispatch): bytes per counted KiB."""
    cp = counters.get(COPY_KERNEL, {})
    if "FETCH_SIZE" not in cp or "WRITE_SIZE" not in cp:
        return None
    return (COPY_PROBE_BYTES / (cp["FETCH_SIZE"][0] * 1024.0), COPY_PROBE_BYTES / (cp["WRITE_SIZE"][0] * 1024.0))


def valu_cycles_per_inst(counters):
    """Cycles one wave64 VALU instruction holds a SIMD, measured on the probe
    of independent FP32 fma chains at 8 waves per SIMD (issue-bound by
    construction): SIMDs x cycles / instructions.  None without its counters."""
    import torch

    v = counters.get(VALU_KERNEL, {})
    if "SQ_INSTS_VALU" not in v or "GRBM_GUI_ACTIVE" not in v:
        return None
    simds = torch.cuda.get_device_properties(0).multi_processor_count * 4
    return simds * (v["GRBM_GUI_ACTIVE"][0] / N_XCD) / max(v["SQ_INSTS_VALU"][0], 1.0)


def kernel_fractions(counters, kern, avg_ms):
    """HBM traffic and VALU figures of one kernel from the counter means."""
    c = counters.get(kern, {})
    f = traffic_factors(counters)
    res = {}
    if f and "FETCH_SIZE" in c and "WRITE_SIZE" in c:
        traffic = c["FETCH_SIZE"][0] * 1024.0 * f[0] + c["WRITE_SIZE"][0] * 1024.0 * f[1]
        res["traffic_bytes"] = round(traffic)
        res["fetch_factor"], res["write_factor"] = round(f[0], 4), round(f[1], 4)
        if avg_ms:
            res["hbm_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
            res["hbm_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c:
        h, m = c["TCC_HIT_sum"][0], c["TCC_MISS_sum"][0]
        res["l2_hit_rate"] = round(h / max(h + m, 1.0), 4)
    if "SQ_INSTS_VALU" in c and "GRBM_GUI_ACTIVE" in c:
        import torch

        simds = torch.cuda.get_device_properties(0).multi_processor_count * 4
        cycles = c["GRBM_GUI_ACTIVE"][0] / N_XCD  # the counter sums the 8 XCDs
        insts = c["SQ_INSTS_VALU"][0]
        # VALU instruction rate against the rate of the FP32 fma probe in the
        # same counter pass (full waves, issue-bound).  Not a busy fraction:
        # it exceeds 1 where instructions run with idle lane groups or an
        # empty exec mask, which issue in fewer cycles than a full-wave fma.
        cpi = valu_cycles_per_inst(counters)
        if cpi:
            res["valu_cycles_per_inst_probe"] = round(cpi, 3)
            res["valu_inst_rate_vs_fma_probe"] = round(insts * cpi / (simds * cycles), 4)
        res["valu_insts"] = round(insts)
        if "SQ_THREAD_CYCLES_VALU" in c:
            lanes = c["SQ_THREAD_CYCLES_VALU"][0] / max(insts, 1.0)
            res["active_lanes_per_valu"] = round(lanes, 2)
            res["lane_efficiency"] = round(lanes / 64.0, 4)
        if "TA_BUSY_avr" in c:
            res["ta_busy_frac"] = round(c["TA_BUSY_avr"][0] / cycles, 4)
        if "SQ_INSTS_VMEM_RD" in c:
            res["vmem_rd_insts"] = round(c["SQ_INSTS_VMEM_RD"][0])
        res["cycles_per_xcd"] = round(cycles)
    return res


def measured_peaks():
    from mcrt_amd._lib import measure_peaks

    log("peaks: copy, MFMA and VALU probes")
    gbs, tflops, tflops16, valu = measure_peaks(0, COPY_PROBE_BYTES)
    return {"hbm_copy_gbs": round(gbs, 1), "hbm_copy_frac_of_published": round(gbs / HBM_PEAK_GBS, 4),
            "mfma_bf16_tflops": round(tflops, 1),
            "mfma_bf16_frac_of_published": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4),
            "mfma_bf16_16x16x32_tflops": round(tflops16, 1),
            "valu_fp32_tflops": round(valu, 1),
            "method": "mcrt_measure_peaks: one-float4-per-thread copy of 1 GiB (read + write counted), "
                      "register-operand MFMA loops on every SIMD (v_mfma_f32_32x32x16_bf16; "
                      "v_mfma_f32_16x16x32_bf16, the convolutions' shape), independent non-packed FP32 fma "
                      "chains at 8 waves per SIMD; HIP events, best of several runs"}


def roofline(ctx, w, step, segments, counters, peaks):
    """Dominant kernel: extend (closest hit).  `achieved` is its HBM traffic
    per launch (counters of the same invocation) over its mean launch time
    (HIP events on its stream).  The reference-order byte model of the
    traversal is reported beside it as a cache-reuse ratio, not as a
    fraction: the 2.16M-triangle scene is served largely from L2 and MALL."""
    import torch

    import oracle

    ctx.set_profiling(True)
    ctx.reset_stats()
    step()
    torch.cuda.synchronize()
    ms, launches = ctx.kernel_stats("extend")
    shade_ms, _ = ctx.kernel_stats("shade")
    ctx.set_profiling(False)
    avg_ms = ms / max(launches, 1)
    log("counting reference traversal work with the oracle (1 spp sample)")
    sw, sh = w.width // 4, w.height // 4
    cam = __import__("mcrt_amd.camera", fromlist=["camera_for"]).camera_for(w.loaded, sw, sh)
    _, _, cnt = oracle.render(oracle.Scene(w.buffers), cam, sw, sh, 1, 0, w.depth, counts=True)
    nodes, tris, sph, seg = cnt
    b_ray = 24.0 + 8.0 + 32.0 * nodes / seg + 36.0 * tris / seg + 32.0 * sph / seg
    alg_gbs = b_ray * segments / (ms * 1e-3) / 1e9
    res = {
        "kernel": "extend",
        "bound": "hbm",
        "achieved": None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": None,
        "traffic": None,
        "avg_launch_ms": round(avg_ms, 4),
        "launches": int(launches),
        "extend_ms_per_step": round(ms, 3),
        "shade_ms_per_step": round(shade_ms, 3),
        "segments_per_launch": round(segments / max(launches, 1)),
        "reference_order_model": {
            "bytes_per_segment": round(b_ray, 1), "nodes_per_segment": round(nodes / seg, 2),
            "tris_per_segment": round(tris / seg, 2), "algorithmic_gbs": round(alg_gbs, 1),
            "algorithmic_over_hbm_peak": round(alg_gbs / HBM_PEAK_GBS, 3),
            "note": "B_ray = 24 + 8 + 32 N_nodes + 36 N_tris with the oracle's reference-order counts; above 1 "
                    "because the traversal set is reused from L2/MALL, so it is a reuse ratio, not a roofline "
                    "fraction"},
    }
    if counters and counters.get(EXTEND_C3 if w.name == "voxel_c3" else EXTEND_C2):
        kf = kernel_fractions(counters, EXTEND_C3 if w.name == "voxel_c3" else EXTEND_C2, avg_ms)
        if "traffic_bytes" in kf:
            res["traffic"] = kf["traffic_bytes"]
            res["achieved"] = kf["hbm_gbs"]
            res["frac"] = kf["hbm_frac"]
            res["frac_of_measured_copy"] = round(kf["hbm_gbs"] / peaks["hbm_copy_gbs"], 4)
            res["bytes_per_segment"] = round(kf["traffic_bytes"] / max(segments / max(launches, 1), 1), 1)
        # The probe renders one step: one extension dispatch per bounce (C3).
        n = counters[EXTEND_C3 if w.name == "voxel_c3" else EXTEND_C2].get("FETCH_SIZE", (0.0, 0))[1]
        kf["dispatches"] = n
        if w.name == "voxel_c3" and n != w.depth:
            kf["dispatches_note"] = f"expected {w.depth} extension dispatches per render, counted {n}"
        res["counters"] = kf
        res["counters_source"] = ("rocprofv3 --pmc passes of `bench.py --probe` in this invocation: "
                                  + " | ".join(" ".join(p) for p in PMC_PASSES))
        res["binding"] = ("not HBM bandwidth: VALU issue over lanes that diverge (counters.valu_inst_rate_vs_fma_probe, "
                          "counters.lane_efficiency) and the latency of dependent node fetches; each 4-wide node "
                          "step slab-tests 4 boxes exactly (about 150 VALU instructions) from three 16-B loads, and "
                          "the vector-memory address unit (counters.ta_busy_frac) is no longer the busiest unit, "
                          "as it was on the binary pairs, where one more 16-B load per node step cost 8.5%")
    return res


# ---------------------------------------------------------------------------
# Render lines and relMSE against the FP64 oracle
# ---------------------------------------------------------------------------
def _rel_mse(x, y):
    import numpy as np

    x = np.asarray(x, np.float64)[..., :3]
    y = np.asarray(y, np.float64)[..., :3]
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))


def relmse_windows(w, accum, denoised, fixture):
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", fixture)
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    acc = accum.cpu().numpy()
    den = denoised.cpu().numpy() if denoised is not None else None
    res = {"reference": f"{len(z['origins'])} windows {z['crops'].shape[1]}x{z['crops'].shape[2]}, "
                        f"{int(z['spp'])} spp FP64 oracle"}
    crops = z["crops"]
    for key, img in (("render", acc), ("denoised", den)):
        if img is None:
            continue
        parts = [img[y:y + crops.shape[1], x:x + crops.shape[2]] for (x, y) in z["origins"]]
        res[key] = round(_rel_mse(np.stack(parts), crops), 6)
    return res


def _time_render(ctx, w, reps, warm=1):
    import torch

    dev = torch.device("cuda:0")
    accum = torch.empty((w.height, w.width, 4), dtype=torch.float32, device=dev)
    gbuf = torch.empty_like(accum)
    sp = torch.cuda.current_stream(dev).cuda_stream
    for _ in range(warm):
        render_step(ctx, w, accum, gbuf, None, sp)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        render_step(ctx, w, accum, gbuf, None, sp)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    seg, _ = ctx.render_counters()
    ctx.set_profiling(True)
    ctx.reset_stats()
    render_step(ctx, w, accum, gbuf, None, sp)
    torch.cuda.synchronize()
    ext_ms, ext_n = ctx.kernel_stats("extend")
    ctx.set_profiling(False)
    return statistics.median(times), seg, accum, gbuf, ext_ms / max(ext_n, 1), ext_n


def c1_line(ctx):
    """Cornell box 256 x 256, 16 spp, depth 8: Msamples/s (median of 20),
    relMSE against the FP64 oracle's full frame, 8-bit PSNR against the FP32
    oracle's render of the same frames."""
    import numpy as np

    import oracle
    from mcrt_amd import configs, headless, regression

    log("line: cornell_c1 256x256 16 spp")
    w = configs.c1()
    ctx.upload_scene(w.buffers)
    t, seg, accum, gbuf, ext_ms, ext_n = _time_render(ctx, w, 20, warm=3)
    res = {"width": w.width, "height": w.height, "spp": w.spp, "max_depth": w.depth,
           "triangles": int(len(w.buffers.triangles)), "ms_per_render": round(t * 1e3, 3),
           "msamples_per_s": round(w.width * w.height * w.spp / t / 1e6, 2),
           "mrays_per_s": round(seg / t / 1e6, 2), "extend_ms": round(ext_ms, 4), "extend_launches": int(ext_n),
           "bound": "launch: 36 triangles and 1M samples in about 1 ms, so kernel launches and persistent-grid "
                    "drains, not memory or VALU, set the time"}
    acc = accum.cpu().numpy()
    ref, _ = oracle.render(oracle.Scene(w.buffers), w.camera, w.width, w.height, w.spp, 0, w.depth)
    cmp = regression.compare_images(headless.to_rgba8(acc), headless.to_rgba8(ref))
    res["vs_oracle_fp32_8bit"] = {"mse": round(cmp.mse, 8), "psnr_db": round(cmp.psnr, 2) if np.isfinite(cmp.psnr)
                                  else "inf"}
    path = os.path.join(ROOT, "tests", "golden", "c1_oracle_b.npz")
    if os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        res["relmse_vs_oracle_b"] = round(_rel_mse(acc, z["image"]), 6)
        res["oracle_b"] = f"full frame, {int(z['spp'])} spp FP64"
    return res


def c2_line(ctx):
    """Displaced icosphere in the Cornell box, 1024 x 1024, 64 spp, depth 8."""
    from mcrt_amd import configs

    log("line: icosphere_c2 1024x1024 64 spp")
    w = configs.c2()
    ctx.upload_scene(w.buffers)
    t, seg, _, _, ext_ms, ext_n = _time_render(ctx, w, 3, warm=1)
    return {"width": w.width, "height": w.height, "spp": w.spp, "max_depth": w.depth,
            "triangles": int(len(w.buffers.triangles)), "ms_per_render": round(t * 1e3, 3),
            "msamples_per_s": round(w.width * w.height * w.spp / t / 1e6, 2),
            "mrays_per_s": round(seg / t / 1e6, 2), "extend_ms": round(ext_ms, 4), "extend_launches": int(ext_n)}


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
        ctx.trace_closest_dev(to.data_ptr(), td.data_ptr(), n, 1e-4, 1e9, ids.data_ptr(), ts.data_ptr(), sp)

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


def denoise_eval(ctx, peaks):
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
                         "unit": "TFLOP/s", "frac": round(res["infer_tflops"] / MFMA_BF16_PEAK_TFLOPS, 4),
                         "frac_of_measured": round(res["infer_tflops"] / peaks["mfma_bf16_tflops"], 4),
                         "frac_of_measured_16x16x32": round(res["infer_tflops"] / peaks["mfma_bf16_16x16x32_tflops"], 4)}
    return res


def inverse(ctx, steps=200):
    """Inverse rendering: 200 Adam steps on 8 material albedos through the
    path-replay adjoint, icosphere scene 512 x 512, 32 spp, depth 6; then
    the adjoint's time per call in each form (replay, tape)."""
    import torch

    from mcrt_amd import diff

    log(f"inverse rendering: {steps} Adam steps")
    inv = diff.inverse_workload()
    ctx.upload_scene(inv.work.buffers)
    w = inv.work
    target_spp = 8 * w.spp
    ctx.set_diff_mode("replay")
    diff.optimize(ctx, inv, steps=2, target_spp=target_spp)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = diff.optimize(ctx, inv, steps=steps, target_spp=target_spp)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    forms = {}
    for mode in ("replay", "tape"):
        ctx.set_diff_mode(mode)
        diff.optimize(ctx, inv, steps=2, target_spp=w.spp)
        torch.cuda.synchronize()
        ctx.set_profiling(True)
        ctx.reset_stats()
        diff.optimize(ctx, inv, steps=20, target_spp=w.spp)
        torch.cuda.synchronize()
        fms, fn = ctx.kernel_stats("diff_primal")
        ams, an = ctx.kernel_stats("diff_adjoint")
        ctx.set_profiling(False)
        tape_bytes = 0
        if mode == "tape":
            from mcrt_amd._lib import diff_tape_bytes

            tape_bytes = diff_tape_bytes(w.width, w.height, w.spp, w.depth)
        forms[mode] = {"adjoint_ms": round(ams / max(an, 1), 3), "primal_ms": round(fms / max(fn, 1), 3),
                       "tape_bytes": tape_bytes}
    ctx.set_diff_mode("replay")
    return {"workload": f"{w.name} {w.width}x{w.height} {w.spp} spp depth {w.depth}, {len(inv.init)} "
                        f"materials x (RGB albedo, emission); target {target_spp} spp, independent frames; "
                        "decorrelated primal / replay frames per step",
            "adjoint": "replay (re-trace, no tape)",
            "steps": steps, "total_s": round(el, 3), "ms_per_step": round(el / steps * 1e3, 3),
            # samples of the primal render plus samples of the adjoint (each traced once)
            "msamples_fwd_plus_adj_per_s": round(2 * w.width * w.height * w.spp * steps / el / 1e6, 1),
            "loss_first": round(hist[0][0], 6), "loss_last": round(hist[-1][0], 6),
            "albedo_err_first": {"max": round(hist[0][1], 4), "mean": round(hist[0][2], 4)},
            "albedo_err_last": {"max": round(hist[-1][1], 4), "mean": round(hist[-1][2], 4)},
            "adjoint_forms": forms}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, seconds):
    """The CPU oracle (FP32 C restatement of the reference shader, pthreads)
    rendering whole frames of the same workload: median of 3 timed runs of
    equal frame counts sized to ~`seconds` in total; the host SAH build of
    the same mesh is timed on its own."""
    import oracle
    from mcrt_amd import build_mesh_bvh

    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    # One GPU's share of a GPU box's host is 16 cores whatever the machine
    # reports; a host with fewer available uses all of them.
    threads = max(1, min(avail, 16))
    sc = oracle.Scene(w.buffers)
    t0 = time.perf_counter()
    oracle.render(sc, w.camera, w.width, w.height, 1, 0, w.depth, threads=threads)
    t1 = time.perf_counter() - t0
    frames = max(1, int(round(seconds / 3.0 / t1)))
    frames = min(frames, w.spp)
    runs = []
    for r in range(3):
        t0 = time.perf_counter()
        oracle.render(sc, w.camera, w.width, w.height, frames, 1 + r * frames, w.depth, threads=threads)
        runs.append(time.perf_counter() - t0)
    el = statistics.median(runs)
    value = w.width * w.height * frames / el / 1e6
    t0 = time.perf_counter()
    if w.loaded is not None and len(w.loaded.mesh_triangles):
        build_mesh_bvh(w.loaded.mesh_vertices, w.loaded.mesh_triangles)
    bvh_s = time.perf_counter() - t0
    return {
        "value": round(value, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of 3 runs of {frames} frame(s) of {w.name} at {w.width}x{w.height}, depth {w.depth} "
                  f"(render only; runs {', '.join(f'{x:.2f}' for x in runs)} s)",
        "host": {"nproc": nproc, "available_cores": avail, "cpu_model": _cpu_model(),
                 "threads_note": "16 = one GPU's share of the GPU box's host cores"},
        "bvh_build_s": round(bvh_s, 3),
        "bvh_build": f"host C++ binned SAH of the {len(w.buffers.triangles)}-triangle mesh, single thread",
        "label": "CPU restatement of the reference shader, not the reference binary",
    }


if __name__ == "__main__":
    sys.exit(main())
