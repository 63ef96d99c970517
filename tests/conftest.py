import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
SCENES = os.path.join(GOLDEN, "scenes")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def ctx():
    from mcrt_amd._lib import Context

    c = Context(0)
    yield c
    c.close()


def c3_with_spheres(width=64, height=36, spp=2, n=300, seed=3):
    """The voxel terrain with n spheres of radius 0.5-4.5 scattered through
    its box: a scene whose traversal set spills the L2 and that has spheres,
    so the culling-stack kernels walk the sphere tree's quantized pairs and
    then the mesh tree's 4-wide nodes in one traversal."""
    import dataclasses

    import numpy as np

    from mcrt_amd import _lib, configs, types as T

    w = configs.c3(width=width, height=height, spp=spp)
    lo, hi = configs.scene_aabb(w.buffers)
    rng = np.random.default_rng(seed)
    sph = np.zeros(n, T.GpuSphere)
    sph["center_r"][:, :3] = lo + rng.random((n, 3)) * (hi - lo)
    sph["center_r"][:, 3] = 0.5 + 4.0 * rng.random(n)
    sph["mat_and_pad"][:, 0] = rng.integers(0, int(w.buffers.materials["mat_count"]), n)
    snodes, sord, sorig = _lib.build_sphere_bvh(sph)
    sb = dataclasses.replace(w.buffers, spheres=sord, sphere_orig=sorig, sphere_bvh=snodes, n_real_spheres=n)
    return w, sb


@pytest.fixture(scope="session")
def c3_spheres():
    return c3_with_spheres()
