#!/usr/bin/env python3
"""Full-frame converged reference of the 1080p icosphere evaluation frame.

Renders the C4 evaluation frame (the displaced-icosphere scene at 1920 x
1080, depth 8) on the GPU with the reference estimator, 65,536 samples per
pixel from frame 2^24 on: the frames of the CPU oracle's FP64 window
reference (tests/golden/c4_reference_crops.npz, make_c4_reference.py), so
on those windows the two are the same sample set and must agree to
rounding.  Uses mcrt_diff_render_dev with no parameter materials, whose
image is the plain mean of its frames (the render's running mean weights
frames from 65,535 on by 1/65,536 and is no mean there); chunk means are
summed in FP64.  Output: tests/golden/c4_full_reference.npz, float16
radiance (H, W, 3) plus the check against the window reference.

    python tests/golden/make_c4_full_reference.py [--spp 65536] [--chunk 1024]    # on an MI355X
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

FIRST = 1 << 24


def main():
    import torch

    from mcrt_amd import learned
    from mcrt_amd._lib import Context

    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "c4_full_reference.npz"))
    a = ap.parse_args()
    assert a.spp % a.chunk == 0
    wl = learned.eval_workload()
    W, H = wl.width, wl.height
    dev = torch.device("cuda:0")
    img = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
    pmap = np.full(64, -1, np.int32)
    params = np.zeros((1, 4), np.float32)
    n = a.spp // a.chunk
    t0 = time.time()
    with Context(0) as ctx:
        ctx.upload_scene(wl.buffers)
        s = torch.cuda.current_stream(dev).cuda_stream
        for k in range(n):
            ctx.diff_render_dev(wl.camera, W, H, a.chunk, FIRST + k * a.chunk, wl.depth, params[:0], pmap,
                                img.data_ptr(), s)
            acc += img[..., :3].double()
            if k % 8 == 7 or k == n - 1:
                torch.cuda.synchronize()
                print(f"chunk {k + 1}/{n} ({(k + 1) * a.chunk} spp), {time.time() - t0:.0f} s", flush=True)
    ref = (acc / n).cpu().numpy()
    crops, origins, crop_spp = learned.load_reference_crops()
    check = {}
    if crop_spp == a.spp:
        check["relmse_vs_oracle_windows"] = learned.crops_rel_mse(ref, crops, origins)
        c = crops.shape[1]
        win = np.stack([ref[y0:y0 + c, x0:x0 + c] for x0, y0 in origins])
        check["max_rel_diff_windows"] = float(np.max(np.abs(win - crops) / (np.abs(crops) + 1e-3)))
    print(check, flush=True)
    np.savez_compressed(a.out, image=ref.astype(np.float16), spp=a.spp, first_frame=FIRST, depth=wl.depth,
                        width=W, height=H, **{k: np.float64(v) for k, v in check.items()})
    print(a.out)


if __name__ == "__main__":
    main()
