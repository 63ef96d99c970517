"""The converged 1080p reference windows used for relMSE, against the oracle.

The fixture (tests/golden/c4_reference_crops.npz, written by
tests/golden/make_c4_reference.py) holds eight 32 x 32 windows of the
1920 x 1080 displaced-icosphere frame at 65,536 spp.  Recomputing a 2 x 2
corner of two windows at the full sample count must give the stored values,
and the relMSE definition is checked on closed forms.
"""
import os

import numpy as np
import pytest

import oracle
from mcrt_amd import learned

FIXTURE = learned.REFERENCE_CROPS


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(FIXTURE):
        pytest.skip("reference windows not generated")
    d = np.load(FIXTURE, allow_pickle=False)
    return d


def test_fixture_shape_and_metadata(ref):
    assert ref["crops"].shape == (8, 32, 32, 3)
    assert int(ref["spp"]) == 65536 and int(ref["depth"]) == learned.EVAL_DEPTH
    assert (int(ref["width"]), int(ref["height"])) == (learned.EVAL_W, learned.EVAL_H)
    assert np.all(np.isfinite(ref["crops"])) and np.all(ref["crops"] >= 0)


@pytest.mark.parametrize("k", [0, 5])
def test_fixture_corner_recomputes_exactly(ref, k):
    wl = learned.eval_workload()
    sc = oracle.Scene(wl.buffers)
    x0, y0 = (int(v) for v in ref["origins"][k])
    got = oracle.reference64_window(sc, wl.camera, learned.EVAL_W, learned.EVAL_H, x0 + 3, y0 + 7, 2, 2,
                                    int(ref["spp"]), int(ref["first_frame"]), int(ref["depth"]), threads=4)
    np.testing.assert_allclose(got.astype(np.float32), ref["crops"][k, 7:9, 3:5], rtol=1e-6, atol=0)


def test_rel_mse_closed_forms():
    y = np.full((4, 4, 3), 0.3)
    assert learned.rel_mse(y, y) == 0.0
    # (0.1)^2 / (0.09 + 0.01) = 0.1
    assert abs(learned.rel_mse(y + 0.1, y) - 0.1) < 1e-12
    # Black reference: error over the 1e-2 floor.
    z = np.zeros((2, 2, 3))
    assert abs(learned.rel_mse(z + 0.05, z) - 0.25) < 1e-12
    # Alpha channel is ignored.
    a = np.concatenate([y, np.full((4, 4, 1), 7.0)], -1)
    assert learned.rel_mse(a, np.concatenate([y, np.zeros((4, 4, 1))], -1)) == 0.0


def test_full_frame_reference_matches_the_oracle_windows():
    """The GPU full-frame reference (65,536 spp of the same frames, float16)
    agrees with the FP64 oracle windows up to float16 rounding and rare path
    divergences: it was rendered from the same PCG frames."""
    from mcrt_amd import learned

    full = learned.load_full_reference()
    assert full is not None and full.shape == (1080, 1920, 3)
    assert np.all(np.isfinite(full)) and full.min() >= 0.0
    crops, origins, spp = learned.load_reference_crops()
    z = np.load(learned.REFERENCE_FULL, allow_pickle=False)
    assert int(z["spp"]) == spp == 65536 and int(z["first_frame"]) == 1 << 24
    assert learned.crops_rel_mse(full, crops, origins) < 2e-6
    # as rendered, before the float16 store
    assert float(z["relmse_vs_oracle_windows"]) < 1e-9
