"""Converged FP64 references ("oracle B") of the plain renders, against the oracle.

tests/golden/c1_oracle_b.npz holds the full 256 x 256 Cornell frame and
tests/golden/c3_oracle_b.npz eight 32 x 32 windows of the 1080p voxel
terrain, each 65,536 samples per pixel from frame 2^24 with FP64
accumulation (tests/golden/make_oracle_b.py).  Recomputing a few pixels at
the full sample count must give the stored values exactly, and the light
seen directly is colour x strength exactly.
"""
import os

import numpy as np
import pytest

import oracle
from mcrt_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return np.load(path, allow_pickle=False)


def test_c1_full_frame_fixture_recomputes_exactly():
    z = _load("c1_oracle_b.npz")
    img = z["image"]
    assert img.shape == (256, 256, 3) and int(z["spp"]) == 65536 and int(z["first_frame"]) == 1 << 24
    assert np.all(np.isfinite(img)) and img.min() > 0.0
    w = configs.c1()
    sc = oracle.Scene(w.buffers)
    for (x0, y0) in ((0, 0), (131, 77)):
        win = oracle.reference64_window(sc, w.camera, 256, 256, x0, y0, 2, 2, 65536, 1 << 24, 8, threads=4)
        assert np.array_equal(win.astype(np.float32), img[y0:y0 + 2, x0:x0 + 2]), (x0, y0)
    # the area light seen directly: colour (1, 1, 1) x strength 15, every sample
    assert np.any(np.all(img == np.float32(15.0), axis=-1))


def test_c1_low_spp_render_is_near_the_converged_frame():
    z = _load("c1_oracle_b.npz")
    w = configs.c1(spp=16)
    acc, _ = oracle.render(oracle.Scene(w.buffers), w.camera, 256, 256, 16, 0, 8)
    y = z["image"].astype(np.float64)
    rel = np.mean((acc[..., :3] - y) ** 2 / (y * y + 1e-2))
    assert 1e-3 < rel < 1.0  # noisy at 16 spp, but the same image


def test_c3_window_fixture_recomputes_exactly():
    z = _load("c3_oracle_b.npz")
    crops, origins = z["crops"], z["origins"]
    assert crops.shape == (8, 32, 32, 3) and int(z["spp"]) == 65536
    w = configs.c3()
    sc = oracle.Scene(w.buffers)
    for k in (3, 6):
        x0, y0 = (int(v) for v in origins[k])
        win = oracle.reference64_window(sc, w.camera, w.width, w.height, x0 + 30, y0 + 31, 2, 1, 65536, 1 << 24,
                                        w.depth, threads=4)
        assert np.array_equal(win.astype(np.float32), crops[k, 31:32, 30:32]), k
