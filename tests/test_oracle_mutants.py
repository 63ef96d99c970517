"""The known-answer pins catch plausible restatement mistakes (no GPU).

Each case plants one mistake in a copy of the CPU oracle's source, builds it
into a temporary library, swaps it in, and checks that the pin written for
that part of the oracle fails.  The mistakes are the ones a restatement of
path_trace.wgsl / denoise.wgsl could make: sigma^2 for 2 sigma^2
(denoise.wgsl:65), the neighbour's depth as the divisor (:100-103), |cos| for
max(0, cos) (:95), swapped atan2 operands or an unflipped v in the env lookup
(path_trace.wgsl:233-242), an unflipped camera v or a skipped lens draw
(:640-651), and an off-by-one freeze of the running mean (:58, :657).
"""
import ctypes
import os
import subprocess

import pytest

import oracle

import test_oracle as pins

SRC = os.path.join(os.path.dirname(oracle.__file__), "mcrt_oracle.c")

# (name, original text, mistake, pin tests that must fail)
MUTANTS = [
    ("spatial_sigma_squared", "float inv2ss = 1.0f / (2.0f * dp->sigma_s * dp->sigma_s);",
     "float inv2ss = 1.0f / (dp->sigma_s * dp->sigma_s);",
     [lambda: pins.test_bilateral_impulse_response_is_the_normalised_gaussian(3, 2.0)]),
    ("depth_over_neighbour", "float drel = (dpth - dq) / fmaxf(dpth, 1e-4f);",
     "float drel = (dpth - dq) / fmaxf(dq, 1e-4f);",
     [pins.test_bilateral_depth_term_is_relative_to_the_centre_depth]),
    ("depth_absolute", "float drel = (dpth - dq) / fmaxf(dpth, 1e-4f);",
     "float drel = (dpth - dq);",
     [pins.test_bilateral_depth_term_is_relative_to_the_centre_depth]),
    ("normal_abs_cos", "powf(fmaxf(0.0f, dot(np, nq)), dp->normal_pow)",
     "powf(fabsf(dot(np, nq)), dp->normal_pow)",
     [lambda: pins.test_bilateral_normal_term_is_clamped_cosine_power(8.0)]),
    ("env_atan2_swapped", "float phi = atan2f(dir.z, dir.x);", "float phi = atan2f(dir.x, dir.z);",
     [lambda: pins.test_env_lookup_oblique_directions(pins.make_coded_env_scene())]),
    ("env_v_unflipped", "float v = 0.5f - theta / PI_F;", "float v = 0.5f + theta / PI_F;",
     [lambda: pins.test_env_lookup_oblique_directions(pins.make_coded_env_scene())]),
    ("camera_v_unflipped", "float v = 1.0f - ((float)y + oracle_rand_f32(&rng)) / (float)H;",
     "float v = ((float)y + oracle_rand_f32(&rng)) / (float)H;",
     [pins.test_camera_jitter_draws_fix_the_first_hit_depth]),
    ("no_lens_draws", "    random_in_unit_disk(&rng, &lx, &ly);\n    v3 lens",
     "    lx = 0.0f; ly = 0.0f;\n    v3 lens",
     [pins.test_camera_lens_draws_offset_the_origin,
      pins.test_first_scatter_uses_draws_five_and_six]),
    ("freeze_off_by_one", "uint32_t fc = f < MCRT_MAX_FRAME_COUNT ? f : MCRT_MAX_FRAME_COUNT;",
     "uint32_t fc = f < 65534u ? f : 65534u;",
     [pins.test_running_mean_weight_freezes_at_65535]),
]


@pytest.fixture
def swap_oracle():
    """Loads a given library as the oracle for the duration of a test."""
    saved_lib, saved_path = oracle._lib, oracle.LIB

    def use(path):
        oracle._lib = None
        oracle.LIB = path
        oracle.lib()

    yield use
    oracle._lib, oracle.LIB = saved_lib, saved_path


@pytest.mark.parametrize("name,orig,bad,checks", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_fails_on_planted_mistake(tmp_path, swap_oracle, name, orig, bad, checks):
    src = open(SRC).read()
    assert src.count(orig) >= 1, f"mutation site of {name} not found"
    mutated = tmp_path / "mutant.c"
    mutated.write_text(src.replace(orig, bad, 1))
    so = str(tmp_path / "libmutant.so")
    hdr = os.path.dirname(SRC)
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-pthread", "-std=c11",
                    "-I", hdr, "-o", so, str(mutated), "-lm"], check=True)
    swap_oracle(so)
    assert isinstance(oracle._lib, ctypes.CDLL) and oracle.LIB == so
    for check in checks:
        with pytest.raises(AssertionError):
            check()
