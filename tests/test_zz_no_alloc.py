"""mcrt_trace_closest_noalloc_dev: the closest-hit query with its scratch in
a static device array instead of library allocations.

Bar: ids and t bitwise equal to mcrt_trace_closest_dev on the same rays, for
two calls in a row (the second smaller, so stale scratch from the first would
show), on a cache-resident scene (plain stack) and on one whose traversal
set spills the L2 (culling stack), both deep enough to use the overflow
columns.
"""
import numpy as np
import pytest

from mcrt_amd import configs, scene as scene_mod, scenes

pytestmark = pytest.mark.gpu


def _big_icosphere():
    text = "materials:\n  w: {type: lambertian}\n"
    v, t, _ = scenes.displaced_icosphere(n=170, seed=4, n_mats=1)
    loaded = scene_mod.load_scene_from_str(text, extra_mesh=(v, t))
    return scene_mod.build_scene_buffers(loaded)


def _trace_both(ctx, o, d):
    import torch

    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    n = len(o)
    ids = [torch.full((n,), 7, dtype=torch.int32, device=dev) for _ in range(2)]
    ts = [torch.full((n,), 7.0, dtype=torch.float32, device=dev) for _ in range(2)]
    sp = torch.cuda.current_stream(dev).cuda_stream
    ctx.trace_closest_dev(to.data_ptr(), td.data_ptr(), n, 1e-4, 1e9, ids[0].data_ptr(), ts[0].data_ptr(), sp)
    ctx.trace_closest_noalloc_dev(to.data_ptr(), td.data_ptr(), n, 1e-4, 1e9, ids[1].data_ptr(),
                                  ts[1].data_ptr(), sp)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in ids], [x.cpu().numpy() for x in ts]


@pytest.mark.parametrize("which", ["icosphere_c2", "icosphere_512k"])
def test_noalloc_trace_equals_trace_dev(ctx, which):
    sb = configs.c2(width=64, height=64, spp=1).buffers if which == "icosphere_c2" else _big_icosphere()
    print(f"noalloc trace: {which}, {len(sb.triangles)} triangles", flush=True)
    ctx.upload_scene(sb)
    lo, hi = configs.scene_aabb(sb)
    for k, n in enumerate((1 << 20, 300_001)):
        o, d = scenes.microbench_rays(n, lo, hi, seed=77 + k)
        ids, ts = _trace_both(ctx, o, d)
        assert np.array_equal(ids[0], ids[1]), (which, n, np.flatnonzero(ids[0] != ids[1])[:5])
        assert np.array_equal(ts[0].view(np.uint32), ts[1].view(np.uint32)), (which, n)
        assert np.mean(ids[1] != -1) > 0.3


def test_noalloc_trace_rejects_missing_scene_and_null_pointers():
    from mcrt_amd._lib import Context, McrtError

    with Context(0) as c:
        with pytest.raises(McrtError) as e:
            c.trace_closest_noalloc_dev(1, 1, 4, 1e-4, 1e9, 1, 1, 0)
        assert e.value.code == -6  # MCRT_ERR_NO_SCENE
        c.upload_scene(configs.c1(width=8, height=8).buffers)
        with pytest.raises(McrtError) as e:
            c.trace_closest_noalloc_dev(None, None, 4, 1e-4, 1e9, None, None, 0)
        assert e.value.code == -2  # MCRT_ERR_BAD_ARG
