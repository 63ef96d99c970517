"""Quantized BVH child pairs contain the reference boxes (no GPU).

Scenes whose traversal set spills the L2 traverse 32-B child pairs
(mcrt_amd/csrc/node_quant.h): per axis an FP32 grid origin and a power-of-two
step, and one byte per plane decoded as fma(q, step, origin).  The encoder
picks each byte with that same expression, so a decoded box contains its
reference box and the reference slab test (path_trace.wgsl:248-256), monotone
in each plane, accepts every ray interval it accepts on the reference box.
tests/node_quant_check.cpp (built with hipcc as a host program) encodes
adversarial pairs (flat, tiny, large, negative, denormal, far from the
origin, sharing planes) and counts decoded planes on the inner side of a
reference plane, planes further than one grid step (or two ulps where the
step is finer than FP32 spacing) from it, and non-finite boxes accepted.
Moving one lower plane a step inward after encoding must be caught.
"""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("node_quant") / "node_quant_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-o", exe,
                    os.path.join(HERE, "node_quant_check.cpp")], check=True, capture_output=True)
    return exe


def run(exe, cases, seed, plant=False):
    out = subprocess.run([exe, str(cases), str(seed)] + (["plant"] if plant else []), check=True,
                         capture_output=True, text=True).stdout
    m = re.search(r"cases (\d+) fail (\d+) outside (\d+) loose (\d+) nonfin (\d+)", out)
    return tuple(int(v) for v in m.groups())


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_decoded_boxes_contain_the_reference_boxes(checker, seed):
    cases, fail, outside, loose, nonfin = run(checker, 200_000, seed)
    assert fail == 0 and outside == 0 and loose == 0 and nonfin == 0


def test_an_inward_plane_is_caught(checker):
    cases, fail, outside, loose, nonfin = run(checker, 20_000, 4, plant=True)
    assert outside > cases // 2
