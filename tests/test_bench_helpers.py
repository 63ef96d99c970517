"""bench.py's host helpers (no GPU): the counter arithmetic behind the
roofline fields, the scene cache the counter probe loads, and the one-GPU
contract."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_kernel_fractions_closed_forms(monkeypatch):
    # A copy of exactly 1 GiB that FETCH_SIZE counts at half (the part's
    # behaviour) gives factors 2 and 1.  The FP32 fma probe issuing one wave64
    # instruction per 2 cycles on every SIMD calibrates 2 cycles per
    # instruction; a kernel issuing at that rate with all lanes active then has
    # an instruction rate of 1 against the probe and lane efficiency 1.
    class Props:
        multi_processor_count = 256

    import torch

    monkeypatch.setattr(torch.cuda, "get_device_properties", lambda i: Props())
    gib_kib = (1 << 30) / 1024.0
    cycles = 1_000_000
    counters = {
        bench.COPY_KERNEL: {"FETCH_SIZE": (gib_kib / 2, 6), "WRITE_SIZE": (gib_kib, 6)},
        bench.VALU_KERNEL: {"GRBM_GUI_ACTIVE": (8.0 * 3000, 4), "SQ_INSTS_VALU": (1024 * 3000 / 2.0, 4)},
        "k": {"FETCH_SIZE": (1000.0, 1), "WRITE_SIZE": (500.0, 1), "TCC_HIT_sum": (3.0, 1),
              "TCC_MISS_sum": (1.0, 1), "GRBM_GUI_ACTIVE": (8.0 * cycles, 1),
              "SQ_INSTS_VALU": (1024 * cycles / 2.0, 1), "SQ_THREAD_CYCLES_VALU": (64 * 1024 * cycles / 2.0, 1),
              "TA_BUSY_avr": (cycles / 2.0, 1), "SQ_INSTS_VMEM_RD": (7.0, 1)},
    }
    f = bench.traffic_factors(counters)
    assert f == pytest.approx((2.0, 1.0))
    r = bench.kernel_fractions(counters, "k", avg_ms=1.0)
    assert r["traffic_bytes"] == round(1000 * 1024 * 2 + 500 * 1024)
    assert r["hbm_gbs"] == pytest.approx(r["traffic_bytes"] / 1e-3 / 1e9, abs=0.05)  # rounded to 0.1
    assert r["l2_hit_rate"] == 0.75
    assert r["valu_cycles_per_inst_probe"] == 2.0
    assert r["valu_inst_rate_vs_fma_probe"] == 1.0 and r["lane_efficiency"] == 1.0
    assert r["ta_busy_frac"] == 0.5 and r["cycles_per_xcd"] == cycles
    # without the probe's counters no issue fraction is claimed
    del counters[bench.VALU_KERNEL]
    r = bench.kernel_fractions(counters, "k", avg_ms=1.0)
    assert "valu_inst_rate_vs_fma_probe" not in r and r["lane_efficiency"] == 1.0


def test_missing_copy_calibration_gives_no_traffic():
    r = bench.kernel_fractions({"k": {"FETCH_SIZE": (1.0, 1), "WRITE_SIZE": (1.0, 1)}}, "k", 1.0)
    assert "traffic_bytes" not in r


def test_probe_scene_cache_round_trip(tmp_path):
    from mcrt_amd import configs

    w = configs.c1(width=32, height=24, spp=3)
    p = str(tmp_path / "w.npz")
    bench.save_workload(w, p)
    v = bench.load_workload(p)
    for k, a in w.buffers.__dict__.items():
        b = getattr(v.buffers, k)
        assert (np.array_equal(a, b) if isinstance(a, np.ndarray) else a == b), k
    assert (v.width, v.height, v.spp, v.depth, v.name) == (32, 24, 3, 8, w.name)
    assert np.array_equal(v.camera, w.camera)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_c3_extension_counters_average_over_both_kernels():
    """A render runs the first bounce as k_extend_packet and the later bounces
    as k_extend: the step's per-launch means weigh every dispatch of both
    (C3: <256, false> + <256, false, true>; C2: <256, true> + <256, false,
    false>), and the two scenes stay apart."""
    means = {
        "void mcrt::k_extend_packet<256, false>(mcrt::ExtArgs)": {"FETCH_SIZE": (100.0, 1)},
        "void mcrt::k_extend<256, false, true>(mcrt::ExtArgs)": {"FETCH_SIZE": (400.0, 5)},
        "void mcrt::k_extend_packet<256, true>(mcrt::ExtArgs)": {"FETCH_SIZE": (3.0, 2)},
        "void mcrt::k_extend<256, false, false>(mcrt::ExtArgs)": {"FETCH_SIZE": (7.0, 14)},
    }
    g = bench.group_counter_means(means)
    assert g[bench.EXTEND_C3]["FETCH_SIZE"] == ((100.0 + 5 * 400.0) / 6, 6)
    assert g[bench.EXTEND_C2]["FETCH_SIZE"] == ((2 * 3.0 + 14 * 7.0) / 16, 16)
