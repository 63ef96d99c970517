#!/usr/bin/env python3
"""Converged reference crops of the 1080p icosphere frame ("oracle B").

Renders eight 32 x 32 windows of the displaced-icosphere scene (the C2
scene) at 1920 x 1080, depth 8, with the CPU oracle's reference estimator
and FP64 accumulation, 65,536 samples per pixel from frame 2^24 on (a range
no GPU render in the tests or the benchmark uses).  The result is the
fixture tests/golden/c4_reference_crops.npz used for relMSE.

    python tests/golden/make_c4_reference.py [--spp 65536] [--threads N]
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

W, H, DEPTH, CROP = 1920, 1080, 8, 32
FIRST = 1 << 24
CROPS = [(x - CROP // 2, y - CROP // 2) for y in (270, 810) for x in (240, 720, 1200, 1680)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "c4_reference_crops.npz"))
    a = ap.parse_args()
    w = configs.c2(width=W, height=H, spp=1, depth=DEPTH)
    sc = oracle.Scene(w.buffers)
    out = np.empty((len(CROPS), CROP, CROP, 3), np.float64)
    t0 = time.time()
    for i, (x0, y0) in enumerate(CROPS):
        out[i] = oracle.reference64_window(sc, w.camera, W, H, x0, y0, CROP, CROP, a.spp, FIRST, DEPTH,
                                           threads=a.threads)
        print(f"crop {i} at ({x0},{y0}) done, {time.time() - t0:.0f}s", flush=True)
    np.savez_compressed(a.out, crops=out.astype(np.float32), origins=np.array(CROPS, np.int32),
                        spp=a.spp, first_frame=FIRST, depth=DEPTH, width=W, height=H)
    print(a.out)


if __name__ == "__main__":
    main()
