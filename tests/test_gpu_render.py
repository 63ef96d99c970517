"""GPU render and denoise parity against the CPU oracle (same PCG stream).

Stated per-pixel tolerance (the reference's image cannot be produced here;
the oracle is its FP32 restatement):
  * G-buffer (first-hit normal + depth of the last frame) identical on
    >= 99.9% of pixels;
  * HDR radiance within 1e-4 relative (+1e-6 absolute) on >= 99% of pixels —
    the remainder is paths that diverge after a last-ulp difference of a
    transcendental (sin/cos/atan2/asin/pow of the device library vs glibc);
  * 8-bit tonemapped image within the reference's own regression gate
    (src/regression.rs:308-315: MSE <= 0.002, PSNR >= 38 dB) by a wide margin.
"""
import os

import numpy as np
import pytest

import oracle
from mcrt_amd import camera as cam_mod, configs, headless, regression, scene as scene_mod
from mcrt_amd import texture, types as T

pytestmark = pytest.mark.gpu
SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def _parity(acc, gb, ref, rgb, frac_px=0.99, rel=1e-4):
    same_g = np.mean(np.all(gb == rgb, axis=-1))
    diff = np.abs(acc[..., :3] - ref[..., :3])
    ok_px = np.all(diff <= rel * np.abs(ref[..., :3]) + 1e-6, axis=-1)
    r = regression.compare_images(headless.to_rgba8(acc), headless.to_rgba8(ref))
    return same_g, float(np.mean(ok_px)), r


def test_cornell_c1_full_config(ctx):
    w = configs.c1()
    print("render c1 256x256 16spp", flush=True)
    ctx.upload_scene(w.buffers)
    acc, gb = ctx.render(w.camera, w.width, w.height, w.spp, 0, w.depth)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, w.width, w.height, w.spp, 0, w.depth)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    print(f"gbuf {same_g:.5f} px {ok:.5f} mse {r.mse:.3e} psnr {r.psnr:.1f}")
    assert np.all(np.isfinite(acc)) and np.all(acc[..., 3] == 1.0)
    assert same_g >= 0.999
    assert ok >= 0.99
    assert r.mse <= 0.002 and r.psnr >= 38.0
    assert r.psnr >= 50.0


def test_progressive_resume_is_exact(ctx):
    """Frames 0..7 then 8..15 resumed from the running mean equal 0..15."""
    w = configs.c1(width=96, height=80)
    ctx.upload_scene(w.buffers)
    full, gfull = ctx.render(w.camera, w.width, w.height, 16, 0, 8)
    half, _ = ctx.render(w.camera, w.width, w.height, 8, 0, 8)
    rest, grest = ctx.render(w.camera, w.width, w.height, 8, 8, 8, accum_in=half)
    assert np.array_equal(full.view(np.uint32), rest.view(np.uint32))
    assert np.array_equal(gfull.view(np.uint32), grest.view(np.uint32))


def test_running_mean_weight_freezes_after_65535(ctx):
    """Weight 1/(min(f, 65535) + 1) (path_trace.wgsl:657): frames past 65535
    keep weight 1/65536 — compared with the oracle at the same frame indices."""
    w = configs.c1(width=32, height=32)
    ctx.upload_scene(w.buffers)
    prev = np.full((32, 32, 4), 0.25, np.float32)
    acc, gb = ctx.render(w.camera, 32, 32, 3, 65534, 8, accum_in=prev)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, 32, 32, 3, 65534, 8, accum_in=prev)
    same_g, ok, _ = _parity(acc, gb, ref, rgb)
    assert same_g >= 0.99 and ok >= 0.99


def test_stream_halves_are_schedule_free(ctx):
    """A cache-resident scene traces a batch as two halves on two streams:
    frame halves for 4+ frames, pixel halves for 1-3.  A 5-frame render (frame
    halves) must equal five chained 1-frame renders (pixel halves) bit for bit,
    at an odd pixel count, and the 1-frame path must match the oracle."""
    W, H = 37, 23
    w = configs.c2(width=W, height=H, spp=5)
    ctx.upload_scene(w.buffers)
    print("render c2 37x23 5spp vs 5 x 1spp", flush=True)
    full, gfull = ctx.render(w.camera, W, H, 5, 0, 8)
    acc = None
    for f in range(5):
        acc, g = ctx.render(w.camera, W, H, 1, f, 8, accum_in=acc)
    assert np.array_equal(full.view(np.uint32), acc.view(np.uint32))
    assert np.array_equal(gfull, g)
    one, g1 = ctx.render(w.camera, W, H, 1, 0, 8)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, W, H, 1, 0, 8)
    same_g, ok, _ = _parity(one, g1, ref, rgb)
    assert same_g >= 0.999 and ok >= 0.99


@pytest.mark.parametrize("size", [(1, 1), (1, 3), (3, 1), (2, 2)])
def test_tiny_images(ctx, size):
    """Renders of a handful of pixels: pixel halves need two pixels, and a
    single pixel stays on one stream.  First hits match the oracle exactly,
    and a 5-frame render (frame halves) equals five chained 1-frame renders
    bit for bit."""
    W, H = size
    w = configs.c2(width=W, height=H, spp=5)
    ctx.upload_scene(w.buffers)
    print(f"render c2 {W}x{H}", flush=True)
    full, gfull = ctx.render(w.camera, W, H, 5, 0, 8)
    acc = None
    for f in range(5):
        acc, g = ctx.render(w.camera, W, H, 1, f, 8, accum_in=acc)
    assert np.array_equal(full.view(np.uint32), acc.view(np.uint32))
    _, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, W, H, 5, 0, 8)
    assert np.array_equal(gfull, rgb) and np.array_equal(g, rgb)


def test_thin_lens_camera_matches_oracle(ctx):
    """A pinhole camera's first bounce recomputes its camera rays from the
    queue slot; a camera with a lens (nonzero defocus vectors) keeps the rays
    that k_generate stores.  Both must match the oracle.  With a lens the
    primary ray origin itself goes through sin/cos of the lens angle, so a
    last-ulp difference of the device library moves a first hit, and the
    G-buffer bar is the radiance bar (99%) there; pinhole keeps 99.9%."""
    W, H = 48, 40
    w = configs.c2(width=W, height=H, spp=3)
    lens = w.camera.copy()
    lens["defocus_u"][:3] = (0.05, 0.0, 0.0)
    lens["defocus_v"][:3] = (0.0, 0.05, 0.0)
    ctx.upload_scene(w.buffers)
    for name, cam, g_bar in (("thin lens", lens, 0.99), ("pinhole", w.camera, 0.999)):
        print(f"render c2 48x40 3spp {name}", flush=True)
        acc, gb = ctx.render(cam, W, H, 3, 0, 8)
        ref, rgb = oracle.render(oracle.Scene(w.buffers), cam, W, H, 3, 0, 8)
        same_g, ok, _ = _parity(acc, gb, ref, rgb)
        print(f"  gbuf {same_g:.5f} px {ok:.5f}", flush=True)
        assert same_g >= g_bar and ok >= 0.99, name


@pytest.mark.parametrize("depth", [1, 2, 6])
def test_bounce_limits(ctx, depth):
    w = configs.c1(width=64, height=64)
    ctx.upload_scene(w.buffers)
    acc, gb = ctx.render(w.camera, 64, 64, 4, 0, depth)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, 64, 64, 4, 0, depth)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    assert same_g >= 0.999 and ok >= 0.99 and r.psnr >= 45


@pytest.mark.parametrize("name", ["diffuse_sphere", "glass_sphere", "mirror_sphere", "emissive_box",
                                  "cornell_box_mini", "skymap_sphere"])
def test_reference_test_scenes(ctx, name, tmp_path):
    path = os.path.join(SCENES, name + ".yaml")
    loaded = scene_mod.load_scene_from_yaml(path)
    env = None
    if loaded.env_map_path:
        # The reference generates this sky with tests/assets/gen_test_skymap.py.
        sky = tmp_path / "sky.hdr"
        texture.write_hdr(str(sky), texture.generate_skymap(1024, 512))
        env = texture.try_load_hdr(str(sky))
    sb = scene_mod.build_scene_buffers(loaded, env=env)
    cam = cam_mod.camera_for(loaded, 96, 96)
    ctx.upload_scene(sb)
    acc, gb = ctx.render(cam, 96, 96, 16, 0, 8)
    ref, rgb = oracle.render(oracle.Scene(sb), cam, 96, 96, 16, 0, 8)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    th = regression.load_thresholds(path)
    print(f"{name}: gbuf {same_g:.4f} px {ok:.4f} mse {r.mse:.2e} psnr {r.psnr:.1f}")
    assert same_g >= 0.995 and ok >= 0.98
    assert th.passes(r)


def test_icosphere_c2_small(ctx):
    w = configs.c2(width=128, height=128, spp=4)
    ctx.upload_scene(w.buffers)
    acc, gb = ctx.render(w.camera, 128, 128, 4, 0, 8)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, 128, 128, 4, 0, 8)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    print(f"c2: gbuf {same_g:.4f} px {ok:.4f} psnr {r.psnr:.1f}")
    assert same_g >= 0.999 and ok >= 0.98 and r.psnr >= 45


@pytest.fixture(scope="module")
def c3_small():
    return configs.c3(width=192, height=108, spp=2)


def test_voxel_terrain_c3_small(ctx, c3_small):
    """Voxel terrain (2.16M triangles, sun sky) at reduced resolution."""
    w = c3_small
    ctx.upload_scene(w.buffers)
    acc, gb = ctx.render(w.camera, w.width, w.height, w.spp, 0, w.depth)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, w.width, w.height, w.spp, 0, w.depth)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    print(f"c3: gbuf {same_g:.4f} px {ok:.4f} psnr {r.psnr:.1f}")
    assert same_g >= 0.999 and ok >= 0.98 and r.mse <= 0.002 and r.psnr >= 38
    den = ctx.denoise_bilateral(acc, gb)
    oden = oracle.denoise(acc, gb, T.default_denoise_params())
    np.testing.assert_allclose(den[..., :3], oden[..., :3], rtol=2e-5, atol=1e-5)


def test_voxel_terrain_with_spheres(ctx, c3_spheres):
    """Spheres in the voxel terrain: the culling-stack extension kernel with a
    sphere phase (quantized sphere pairs) before the mesh phase (4-wide
    nodes), every bounce (no first-bounce packets with spheres)."""
    w, sb = c3_spheres
    ctx.upload_scene(sb)
    acc, gb = ctx.render(w.camera, w.width, w.height, w.spp, 0, w.depth)
    ref, rgb = oracle.render(oracle.Scene(sb), w.camera, w.width, w.height, w.spp, 0, w.depth)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    print(f"c3 + spheres: gbuf {same_g:.4f} px {ok:.4f} psnr {r.psnr:.1f}")
    assert same_g >= 0.999 and ok >= 0.98, (same_g, ok)


def test_voxel_terrain_pixel_major_queue(ctx):
    """Paths are queued pixel-major (a pixel's frames adjacent): at an odd
    size with 3 frames per batch every (pixel, frame) path must still be
    rendered exactly once (G-buffer and image against the oracle)."""
    w = configs.c3(width=70, height=45, spp=3)
    ctx.upload_scene(w.buffers)
    acc, gb = ctx.render(w.camera, w.width, w.height, 3, 0, w.depth)
    ref, rgb = oracle.render(oracle.Scene(w.buffers), w.camera, w.width, w.height, 3, 0, w.depth)
    same_g, ok, r = _parity(acc, gb, ref, rgb)
    assert same_g >= 0.999 and ok >= 0.98, (same_g, ok)


def test_segment_counter_matches_oracle(ctx):
    w = configs.c1(width=64, height=64)
    ctx.upload_scene(w.buffers)
    ctx.render(w.camera, 64, 64, 4, 0, 8)
    seg, smp = ctx.render_counters()
    _, _, cnt = oracle.render(oracle.Scene(w.buffers), w.camera, 64, 64, 4, 0, 8, counts=True)
    assert smp == 64 * 64 * 4
    assert abs(seg - cnt[3]) <= 1e-3 * cnt[3]


def test_segment_counter_sums_over_batches(ctx):
    """A render is split into batches of at most 2^26 paths: at 2048 x 2048,
    17 spp runs as 16 + 1 frames, and its segment count must be the sum of
    the two renders of those frame ranges."""
    w = configs.c1(width=2048, height=2048)
    ctx.upload_scene(w.buffers)
    ctx.render(w.camera, 2048, 2048, 17, 0, 8, want_gbuf=False)
    seg, smp = ctx.render_counters()
    ctx.render(w.camera, 2048, 2048, 16, 0, 8, want_gbuf=False)
    seg_a, _ = ctx.render_counters()
    ctx.render(w.camera, 2048, 2048, 1, 16, 8, want_gbuf=False)
    seg_b, _ = ctx.render_counters()
    assert smp == 2048 * 2048 * 17
    assert seg_a > 16 * 2048 * 2048 and seg_b > 2048 * 2048
    assert seg == seg_a + seg_b


def test_bilateral_matches_oracle(ctx):
    w = configs.c1(width=128, height=96)
    ctx.upload_scene(w.buffers)
    acc, gb = ctx.render(w.camera, 128, 96, 4, 0, 8)
    # Whole exponents take the kernel's repeated-squaring path (0 included:
    # pow(x, 0) = 1 even at x = 0); 7.5 and 0.3 take the general pow.
    # Radii 8 and below are staged in LDS; 9, 12 and 20 read taps from memory;
    # a negative radius runs no tap and copies the pixel, as the shader's
    # empty loops do.
    for radius, ss, sd, npow in ((3, 2.0, 0.5, 8.0), (1, 1.0, 0.25, 2.0), (5, 3.0, 1.0, 16.0), (0, 2.0, 0.5, 8.0),
                                 (2, 1.5, 0.75, 7.5), (3, 2.0, 0.5, 0.3), (3, 2.0, 0.5, 0.0), (8, 4.0, 0.5, 8.0),
                                 (9, 4.0, 0.5, 8.0), (12, 5.0, 0.4, 8.0), (20, 3.0, 0.5, 2.5), (-1, 2.0, 0.5, 8.0),
                                 (-7, 2.0, 0.5, 8.0)):
        p = T.default_denoise_params()
        p["radius"], p["sigma_s"], p["sigma_d"], p["normal_pow"] = radius, ss, sd, npow
        g = ctx.denoise_bilateral(acc, gb, p)
        o = oracle.denoise(acc, gb, p)
        np.testing.assert_allclose(g[..., :3], o[..., :3], rtol=2e-5, atol=1e-6)
        assert np.all(g[..., 3] == 1.0)


def test_bilateral_radius_limit_and_stream_order(ctx):
    """Radii above 256 are refused with a message; parameter changes between
    calls need no device-wide synchronisation (each call is stream-ordered)."""
    import torch

    acc = np.random.default_rng(3).random((40, 48, 4)).astype(np.float32)
    gb = np.zeros((40, 48, 4), np.float32)
    gb[..., 2] = 1.0
    gb[..., 3] = 5.0
    p = T.default_denoise_params()
    p["radius"] = 257
    with pytest.raises(Exception) as e:
        ctx.denoise_bilateral(acc, gb, p)
    assert "256" in str(e.value)
    dev = torch.device("cuda:0")
    ta, tg = torch.from_numpy(acc).to(dev), torch.from_numpy(gb).to(dev)
    outs = [torch.empty_like(ta) for _ in range(3)]
    s = torch.cuda.Stream(dev)
    with torch.cuda.stream(s):
        for o, (r, sg) in zip(outs, ((3, 2.0), (1, 0.5), (10, 3.0))):
            q = T.default_denoise_params()
            q["radius"], q["sigma_s"] = r, sg
            ctx.denoise_bilateral_dev(q, 48, 40, ta.data_ptr(), tg.data_ptr(), o.data_ptr(), s.cuda_stream)
    s.synchronize()
    for o, (r, sg) in zip(outs, ((3, 2.0), (1, 0.5), (10, 3.0))):
        q = T.default_denoise_params()
        q["radius"], q["sigma_s"] = r, sg
        np.testing.assert_allclose(o.cpu().numpy()[..., :3], oracle.denoise(acc, gb, q)[..., :3],
                                   rtol=2e-5, atol=1e-6)


def test_bilateral_sky_passthrough_quirk(ctx):
    """Sky texels (depth 1e9, zero normal) have all-zero weights and return the
    raw pixel (denoise.wgsl:61-62, :115-117)."""
    acc = np.random.default_rng(0).random((20, 24, 4)).astype(np.float32)
    gb = np.zeros((20, 24, 4), np.float32)
    gb[..., 3] = 1e9
    out = ctx.denoise_bilateral(acc, gb)
    assert np.array_equal(out[..., :3], acc[..., :3])


def test_headless_cli_end_to_end(tmp_path):
    """`python -m mcrt_amd --headless scene.yaml --output o.png ...` (the
    mirror of src/main.rs:56-80): exit status, PNG size, the plain render
    against the oracle's tonemapped image, and the denoiser / NEE extensions."""
    from PIL import Image

    scene = os.path.join(SCENES, "cornell_box_mini.yaml")
    out = str(tmp_path / "plain.png")
    assert headless.main(["--headless", scene, "--output", out, "--width", "48", "--height", "32",
                          "--spp", "4"]) == 0
    png = np.asarray(Image.open(out))
    assert png.shape == (32, 48, 4)
    loaded = scene_mod.load_scene_from_yaml(scene)
    sb = scene_mod.build_scene_buffers(loaded)
    ref, _ = oracle.render(oracle.Scene(sb), cam_mod.camera_for(loaded, 48, 32), 48, 32, 4, 0, 8)
    r = regression.compare_images(png, headless.to_rgba8(ref))
    assert r.mse <= 0.002 and r.psnr >= 38.0
    box = os.path.join(SCENES, "emissive_box.yaml")
    for extra in (["--denoise", "bilateral"], ["--denoise", "learned"], ["--nee"],
                  ["--nee", "--denoise", "learned"]):
        out = str(tmp_path / ("_".join(x.strip("-") for x in extra) + ".png"))
        assert headless.main(["--headless", box, "--output", out, "--width", "40", "--height", "24",
                              "--spp", "2"] + extra) == 0, extra
        assert np.asarray(Image.open(out)).shape == (24, 40, 4)
    assert headless.main(["--headless", box, "--output", str(tmp_path / "x.png"), "--denoise", "nope"]) == 1
