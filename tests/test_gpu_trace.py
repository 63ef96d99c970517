"""GPU closest-hit parity (mcrt_trace_closest through the C-ABI) against the
CPU oracle.

Bar: primitive ids bit-exact and hit t bit-exact (the kernel evaluates the
shader's slab and Möller–Trumbore expressions with the same FP32 roundings,
so the stated tolerance of 1e-5 relative is met with zero error).
"""
import os

import numpy as np
import pytest

import oracle
from mcrt_amd import configs, scene as scene_mod, scenes
from mcrt_amd import types as T

pytestmark = pytest.mark.gpu
SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def _cmp(ctx, sb, o, d, tmin=1e-4, tmax=1e9, mode="exact"):
    ctx.upload_scene(sb)
    gid, gt = ctx.trace_closest(o, d, tmin, tmax)
    oid, ot = oracle.trace(oracle.Scene(sb), o, d, tmin, tmax, mode=mode)
    bad = np.flatnonzero((gid != oid) | (gt.view(np.uint32) != ot.view(np.uint32)))
    return gid, gt, oid, ot, bad


@pytest.fixture(scope="module")
def c2():
    return configs.c2(width=64, height=64, spp=1)


def test_c2_one_million_rays_bit_exact(ctx, c2):
    lo, hi = configs.scene_aabb(c2.buffers)
    o, d = scenes.microbench_rays(1 << 20, lo, hi, seed=1234)
    print("trace c2 1M rays", flush=True)
    gid, gt, oid, ot, bad = _cmp(ctx, c2.buffers, o, d)
    assert len(bad) == 0, f"{len(bad)} mismatches, first {bad[:5]}: gpu {gid[bad[:5]]} {gt[bad[:5]]} oracle {oid[bad[:5]]} {ot[bad[:5]]}"
    assert np.mean(gid != T.ID_MISS) > 0.5


def test_c2_matches_brute_force_sample(ctx, c2):
    lo, hi = configs.scene_aabb(c2.buffers)
    o, d = scenes.microbench_rays(2000, lo, hi, seed=99)
    ctx.upload_scene(c2.buffers)
    gid, gt = ctx.trace_closest(o, d)
    bid, bt = oracle.trace(oracle.Scene(c2.buffers), o, d, mode="brute")
    assert np.array_equal(gid, bid)
    assert np.array_equal(gt.view(np.uint32), bt.view(np.uint32))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 4097])
def test_ragged_batches(ctx, c2, n):
    lo, hi = configs.scene_aabb(c2.buffers)
    o, d = scenes.microbench_rays(n, lo, hi, seed=n)
    _, _, _, _, bad = _cmp(ctx, c2.buffers, o, d)
    assert len(bad) == 0


def test_empty_batch(ctx, c2):
    ctx.upload_scene(c2.buffers)
    ids, ts = ctx.trace_closest(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert ids.shape == (0,) and ts.shape == (0,)


def test_axis_aligned_and_interval_edges(ctx):
    w = configs.c1(width=8, height=8, spp=1)
    rng = np.random.default_rng(5)
    o = rng.uniform([-4, 0.5, -4], [4, 9.5, 4], (3000, 3)).astype(np.float32)
    axes = np.eye(3, dtype=np.float32)
    d = np.concatenate([axes, -axes] * 500)[:3000]
    for tmin, tmax in ((1e-4, 1e9), (0.5, 3.0), (0.0, 1e9)):
        _, _, _, _, bad = _cmp(ctx, w.buffers, o, d, tmin, tmax)
        assert len(bad) == 0, (tmin, tmax, bad[:5])


@pytest.mark.parametrize("name", ["diffuse_sphere", "glass_sphere", "mirror_sphere", "emissive_box",
                                  "cornell_box_mini"])
def test_reference_scenes_with_spheres(ctx, name):
    loaded = scene_mod.load_scene_from_yaml(os.path.join(SCENES, name + ".yaml"))
    sb = scene_mod.build_scene_buffers(loaded)
    lo = np.array([-6, -1, -6], np.float32)
    hi = np.array([6, 8, 6], np.float32)
    o, d = scenes.microbench_rays(20000, lo, hi, seed=3)
    gid, _, oid, _, bad = _cmp(ctx, sb, o, d)
    assert len(bad) == 0, bad[:5]
    bid, bt = oracle.trace(oracle.Scene(sb), o, d, mode="brute")
    assert np.array_equal(gid, bid)
    if len(loaded.spheres):
        assert np.any(gid & T.ID_SPHERE_BIT) and np.any(gid != T.ID_MISS)


def test_no_mesh_stub_scene(ctx):
    text = "materials:\n  w: {type: lambertian}\nobjects:\n  - {type: sphere, center: [0,0,0], radius: 1, material: w}\n"
    loaded = scene_mod.load_scene_from_str(text)
    sb = scene_mod.build_scene_buffers(loaded)
    o = np.array([[0, 0, 5], [0, 5, 0], [3, 3, 3]], np.float32)
    d = np.array([[0, 0, -1], [0, 1, 0], [-1, -1, -1]], np.float32)
    gid, gt, oid, ot, bad = _cmp(ctx, sb, o, d)
    assert len(bad) == 0
    assert gid[0] == T.ID_SPHERE_BIT and abs(gt[0] - 4.0) < 1e-6
    assert gid[1] == T.ID_MISS and gt[1] == -1.0


def test_sixteen_million_rays_properties(ctx, c2):
    """Full microbench size: deterministic across runs, hits carry t in range,
    and a 1/64 strided sample is bit-exact against the oracle."""
    lo, hi = configs.scene_aabb(c2.buffers)
    n = 1 << 24
    o, d = scenes.microbench_rays(n, lo, hi, seed=2024)
    ctx.upload_scene(c2.buffers)
    print("trace c2 16M rays", flush=True)
    gid, gt = ctx.trace_closest(o, d)
    gid2, gt2 = ctx.trace_closest(o, d)
    assert np.array_equal(gid, gid2) and np.array_equal(gt.view(np.uint32), gt2.view(np.uint32))
    hit = gid != T.ID_MISS
    assert np.all((gt[hit] >= 1e-4) & (gt[hit] < 1e9)) and np.all(gt[~hit] == -1.0)
    assert np.all(gid[hit] < len(c2.buffers.triangles))
    sel = np.arange(0, n, 64)
    oid, ot = oracle.trace(oracle.Scene(c2.buffers), o[sel], d[sel], mode="exact")
    assert np.array_equal(gid[sel], oid)
    assert np.array_equal(gt[sel].view(np.uint32), ot.view(np.uint32))


def test_every_slot_traced_once(ctx, c2):
    """Slots are handed out as one static chunk per wave, then pulled in
    chunks from a counter.  An odd count above the static share must leave no
    output slot unwritten (outputs start as a sentinel on the device), and
    the head, the tail and a strided sample match the oracle."""
    import torch

    lo, hi = configs.scene_aabb(c2.buffers)
    n = (1 << 24) + 17
    o, d = scenes.microbench_rays(n, lo, hi, seed=77)
    ctx.upload_scene(c2.buffers)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    ids = torch.full((n,), 0x7FFFFFF0, dtype=torch.int32, device=dev)
    ts = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    print("trace c2 16M+17 rays into sentinels", flush=True)
    ctx.trace_closest_dev(to.data_ptr(), td.data_ptr(), n, 1e-4, 1e9, ids.data_ptr(), ts.data_ptr(),
                          torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    gid = ids.cpu().numpy().view(np.uint32)
    gt = ts.cpu().numpy()
    assert not np.any(gid == 0x7FFFFFF0) and not np.any(np.isnan(gt))
    sel = np.concatenate([np.arange(0, 1000), np.arange(n - 1000, n), np.arange(1000, n - 1000, 997)])
    oid, ot = oracle.trace(oracle.Scene(c2.buffers), o[sel], d[sel], mode="exact")
    assert np.array_equal(gid[sel], oid)
    assert np.array_equal(gt[sel].view(np.uint32), ot.view(np.uint32))


def test_c3_large_scene_bit_exact(ctx):
    """The 2.16M-triangle voxel terrain (a traversal set far beyond the L2):
    ids and t must be the oracle's, bit for bit, for incoherent rays and for
    rays exactly parallel to the voxel faces (zero direction components: the
    slab form's infinite inverse)."""
    w = configs.c3(width=64, height=36)
    lo, hi = configs.scene_aabb(w.buffers)
    o, d = scenes.microbench_rays(1 << 18, lo, hi, seed=77)
    rng = np.random.default_rng(5)
    n_ax = 1 << 15
    oa = (lo + rng.random((n_ax, 3)) * (hi - lo)).astype(np.float32)
    axes = np.eye(3, dtype=np.float32)[rng.integers(0, 3, n_ax)] * rng.choice([-1.0, 1.0], (n_ax, 1)).astype(np.float32)
    o = np.concatenate([o, oa])
    d = np.concatenate([d, axes])
    print("trace c3 large scene", flush=True)
    gid, gt, oid, ot, bad = _cmp(ctx, w.buffers, o, d)
    assert len(bad) == 0, f"{len(bad)} mismatches, first {bad[:5]}: gpu {gid[bad[:5]]} {gt[bad[:5]]} oracle {oid[bad[:5]]} {ot[bad[:5]]}"
    assert np.mean(gid != T.ID_MISS) > 0.1


def test_c3_with_spheres_bit_exact(ctx, c3_spheres):
    """Spheres and the 2.16M-triangle terrain: ids and t equal the oracle's
    bit for bit, for incoherent and axis-parallel rays."""
    w, sb = c3_spheres
    lo, hi = configs.scene_aabb(sb)
    o, d = scenes.microbench_rays(1 << 17, lo, hi, seed=21)
    rng = np.random.default_rng(8)
    n_ax = 1 << 13
    oa = (lo + rng.random((n_ax, 3)) * (hi - lo)).astype(np.float32)
    axes = np.eye(3, dtype=np.float32)[rng.integers(0, 3, n_ax)] * rng.choice([-1.0, 1.0], (n_ax, 1)).astype(np.float32)
    o = np.concatenate([o, oa])
    d = np.concatenate([d, axes])
    print("trace c3 + 300 spheres", flush=True)
    gid, gt, oid, ot, bad = _cmp(ctx, sb, o, d)
    assert len(bad) == 0, f"{len(bad)} mismatches, first {bad[:5]}: gpu {gid[bad[:5]]} {gt[bad[:5]]} oracle {oid[bad[:5]]} {ot[bad[:5]]}"
    assert np.mean(gid != T.ID_MISS) > 0.1
    assert np.sum((oid != T.ID_MISS) & ((oid & T.ID_SPHERE_BIT) != 0)) > 100  # spheres are hit too
