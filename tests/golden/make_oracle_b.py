#!/usr/bin/env python3
"""Converged FP64 references ("oracle B") of the plain-render workloads.

* ``c1``: the Cornell box (36 triangles) at 256 x 256, depth 8, every pixel.
* ``c3``: the voxel terrain (2.16M triangles) at 1920 x 1080, depth 6, on
  eight 32 x 32 windows (a full frame would take days on a CPU).

Each reference is the reference estimator (path_trace.wgsl:583-662) with FP64
accumulation of 65,536 samples per pixel from frame 2^24 on, a frame range
that no GPU render of the tests or the benchmark uses, so the reference is
independent of the images it judges.  Output: tests/golden/<workload>_oracle_b.npz
(float32 radiance, window origins, settings), used for relMSE.

    python tests/golden/make_oracle_b.py c1 [--spp 65536] [--threads N]
    python tests/golden/make_oracle_b.py c3
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mcrt_amd import configs  # noqa: E402

FIRST = 1 << 24
C3_CROP = 32
C3_WINDOWS = [(x - C3_CROP // 2, y - C3_CROP // 2) for y in (270, 810) for x in (240, 720, 1200, 1680)]


def windows_for(name, w):
    """(x0, y0, cw, ch) list: the whole image for c1, fixed 32 x 32 windows for c3."""
    if name == "c1":
        # bands of 4 rows so progress is printed as the frame fills in
        return [(0, y, w.width, 4) for y in range(0, w.height, 4)]
    return [(x, y, C3_CROP, C3_CROP) for (x, y) in C3_WINDOWS]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["c1", "c3"])
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--out", default=None)
    ap.add_argument("--cache", default=None,
                    help="directory for finished windows, so an interrupted run resumes where it stopped")
    ap.add_argument("--only", default=None,
                    help="comma-separated window indices: compute these into --cache and stop (no fixture)")
    a = ap.parse_args()
    w = configs.by_name(a.workload)
    out_path = a.out or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                     f"{a.workload}_oracle_b.npz")
    sc = oracle.Scene(w.buffers)
    wins = windows_for(a.workload, w)
    parts = []
    t0 = time.time()
    if a.cache:
        os.makedirs(a.cache, exist_ok=True)
    only = None if a.only is None else {int(v) for v in a.only.split(",")}
    if only is not None and not a.cache:
        ap.error("--only needs --cache")
    for i, (x0, y0, cw, ch) in enumerate(wins):
        if only is not None and i not in only:
            continue
        cached = os.path.join(a.cache, f"{a.workload}_{a.spp}_{i}.npy") if a.cache else None
        if cached and os.path.exists(cached):
            parts.append(np.load(cached))
            continue
        parts.append(oracle.reference64_window(sc, w.camera, w.width, w.height, x0, y0, cw, ch, a.spp,
                                               FIRST, w.depth, threads=a.threads))
        if cached:
            np.save(cached, parts[-1])
        print(f"window {i + 1}/{len(wins)} at ({x0},{y0}) done, {time.time() - t0:.0f}s", flush=True)
    if only is not None:
        return
    if a.workload == "c1":
        image = np.concatenate(parts, axis=0).astype(np.float32)
        np.savez_compressed(out_path, image=image, spp=a.spp, first_frame=FIRST, depth=w.depth,
                            width=w.width, height=w.height)
    else:
        crops = np.stack(parts).astype(np.float32)
        np.savez_compressed(out_path, crops=crops, origins=np.array(C3_WINDOWS, np.int32), spp=a.spp,
                            first_frame=FIRST, depth=w.depth, width=w.width, height=w.height)
    print(out_path)


if __name__ == "__main__":
    main()
