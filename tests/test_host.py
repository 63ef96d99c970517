"""Host-side logic and the C-ABI library surface (no GPU compute)."""
import os
import re

import numpy as np
import pytest

import oracle
from mcrt_amd import _lib, headless, regression, texture
from mcrt_amd import types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mcrt_abi.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(mcrt_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.exported_symbols()) == sorted(
        n for n in names if n in _lib.exported_symbols())


def test_create_without_device_reports_no_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.McrtError) as e:
        _lib.Context(0)
    assert e.value.code == -1 and "adapter" in str(e.value)


def test_pod_sizes():
    assert (T.GpuSphere.itemsize, T.GpuBvhNode.itemsize, T.GpuVertex.itemsize, T.GpuTriangle.itemsize,
            T.GpuMaterial.itemsize, T.GpuMaterialData.itemsize, T.CameraUniform.itemsize,
            T.DenoiseParams.itemsize) == (32, 48, 48, 16, 48, 3088, 112, 16)


def test_tonemap_matches_oracle():
    vals = np.array([0, 1e-6, 0.01, 0.1, 0.5, 1, 2, 7.5, 100, 1e6, -1, np.nan, np.inf], np.float32)
    got = headless.tonemap_to_u8(vals)
    ref = [oracle.tonemap_u8(float(v)) for v in vals]
    assert list(got) == ref


def test_compare_images_closed_forms():
    a = np.zeros((4, 5, 4), np.uint8)
    r = regression.compare_images(a, a)
    assert r.mse == 0 and r.psnr == float("inf") and r.max_delta == 0
    b = a.copy()
    b[..., :3] = 255
    b[..., 3] = 7  # alpha ignored
    r = regression.compare_images(a, b)
    assert r.mse == 1.0 and r.psnr == 0.0 and r.max_delta == 1.0
    c = a.copy()
    c[0, 0, 0] = 255
    r = regression.compare_images(a, c)
    assert abs(r.mse - 1 / 60) < 1e-12
    with pytest.raises(regression.DimensionMismatch):
        regression.compare_images(a, np.zeros((5, 4, 4), np.uint8))


def test_thresholds_from_sidecar():
    scenes = os.path.join(ROOT, "tests", "golden", "scenes")
    t = regression.load_thresholds(os.path.join(scenes, "emissive_box.yaml"))
    assert (t.mse, t.psnr) == (0.005, 35.0)
    t = regression.load_thresholds("/nonexistent.yaml")
    assert (t.mse, t.psnr) == (0.002, 38.0)


def test_sidecar_precedence(tmp_path):
    scenes = os.path.join(ROOT, "tests", "golden", "scenes")
    assert headless.load_sidecar(os.path.join(scenes, "glass_sphere.yaml")) == {"width": 256, "height": 256, "spp": 256}
    assert headless.load_sidecar(str(tmp_path / "none.yaml")) == {}
    assert headless.get_next_after(["--headless", "a.yaml", "--spp"], "--spp") is None
    assert headless.get_next_after(["--width", "64"], "--width") == "64"


def test_cli_errors_without_output(capsys):
    assert headless.main(["--headless", "x.yaml"]) == 1
    assert "--output" in capsys.readouterr().err
    assert headless.main(["--headless"]) == 1


def test_hdr_round_trip(tmp_path):
    sky = texture.generate_skymap(64, 32)
    p = tmp_path / "s.hdr"
    texture.write_hdr(str(p), sky)
    back = texture.read_hdr(str(p))
    assert back.shape == sky.shape
    # RGBE keeps 8 bits of mantissa relative to each pixel's largest channel.
    err = np.abs(back - sky).max(axis=-1)
    assert np.all(err <= sky.max(axis=-1) / 128.0 + 1e-30)
    env = texture.try_load_hdr(str(p))
    assert env.shape == (512, 1024, 4) and np.all(env[..., 3] == 1)


def test_skymap_closed_form():
    sky = texture.generate_skymap(1024, 512, sun_elevation_deg=35.0, sun_azimuth_deg=30.0)
    # Sun peak: 80000 * sun colour at the pixel nearest the sun direction.
    u = int(round((30.0 / 360.0 + 0.5) * 1024))
    v = int(round((0.5 - 35.0 / 180.0) * 512))
    sun = np.array([(0xFF / 255) ** 2.2, (0xEE / 255) ** 2.2, (0xDD / 255) ** 2.2])
    peak = sky[v - 2:v + 3, u - 2:u + 3].reshape(-1, 3).max(axis=0)
    assert np.all(peak > 0.8 * 80000 * sun)
    assert np.allclose(sky[-1, 0], 0.1 * (0x33 / 255) ** 2.2, rtol=1e-5)  # ground fill
