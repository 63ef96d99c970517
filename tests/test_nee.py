"""Next-event estimation (the build's opt-in estimator, no reference
counterpart): CPU pins of the oracle's statement of it.

- Closed form: a Lambertian floor of albedo rho under a sphere light of
  radiance Le and radius R centred at height h, black sky, depth 2: the
  floor's radiance straight below the light is rho * Le * (R / h)^2 (a
  uniformly emitting sphere irradiates like a point of intensity
  pi R^2 Le at its centre).  NEE and the plain estimator both converge to it.
- Unbiased: on the Cornell box (a triangle light) the NEE image converges to
  the converged FP64 reference of the plain estimator.
- Lower variance: at equal spp its relMSE is well below the plain render's.
A mistake in a pdf, a MIS weight or a cosine (a dropped 1/pi, |cos| at the
light, the emitter choice probability) moves the mean and fails the first
two; a broken light-sample direction fails all three.
"""
import numpy as np

import oracle
from mcrt_amd import camera as cam_mod, configs, learned, scene as scene_mod

RHO, LE, R, H = 0.5, 4.0, 0.5, 2.0


def _black_env():
    e = np.zeros((512, 1024, 4), np.float32)
    e[..., 3] = 1
    return e


def sphere_light_scene():
    text = ("camera: {look_from: [0, 1.0, 3.0], look_at: [0, 0, 0], vfov: 2}\n"
            "materials:\n"
            f"  floor: {{type: lambertian, albedo: [{RHO}, {RHO}, {RHO}]}}\n"
            f"  light: {{type: emissive, albedo: [1, 1, 1], emission_strength: {LE}}}\n"
            "objects:\n"
            "  - {type: plane, normal: [0, 1, 0], half_extents: [20, 20], material: floor}\n"
            f"  - {{type: sphere, center: [0, {H}, 0], radius: {R}, material: light}}\n")
    loaded = scene_mod.load_scene_from_str(text)
    sb = scene_mod.build_scene_buffers(loaded, env=_black_env())
    return loaded, sb


def test_sphere_light_closed_form():
    loaded, sb = sphere_light_scene()
    cam = cam_mod.camera_for(loaded, 8, 8)
    sc = oracle.Scene(sb)
    expect = RHO * LE * (R / H) ** 2
    nee, gb = oracle.render_nee(sc, cam, 8, 8, 256, 0, 2, threads=4)
    assert np.all(gb[..., 1] > 0.99)  # every pixel sees the floor
    m = float(nee[..., :3].mean())
    print(f"NEE mean {m:.5f}, closed form {expect:.5f}")
    assert abs(m - expect) < 0.01 * expect
    plain, _ = oracle.render(sc, cam, 8, 8, 1024, 0, 2, threads=4)
    mp = float(plain[..., :3].mean())
    print(f"plain mean {mp:.5f}")
    assert abs(mp - expect) < 0.05 * expect
    # NEE is far less noisy per pixel than BSDF-only sampling at equal spp.
    plain256, _ = oracle.render(sc, cam, 8, 8, 256, 0, 2, threads=4)
    assert nee[..., :3].std() < 0.25 * plain256[..., :3].std()


def square_light_scene(facing_down=True, culled=False):
    n = -1 if facing_down else 1
    cull = ", backface_culling: true" if culled else ""
    text = ("camera: {look_from: [0, 1.0, 3.0], look_at: [0, 0, 0], vfov: 2}\n"
            "materials:\n"
            f"  floor: {{type: lambertian, albedo: [{RHO}, {RHO}, {RHO}]}}\n"
            f"  light: {{type: emissive, albedo: [1, 1, 1], emission_strength: {LE}}}\n"
            "objects:\n"
            "  - {type: plane, normal: [0, 1, 0], half_extents: [20, 20], material: floor}\n"
            f"  - {{type: plane, normal: [0, {n}, 0], half_extents: [0.5, 0.5], material: light{cull},"
            f" transform: {{translate: [0, {H}, 0]}}}}\n")
    loaded = scene_mod.load_scene_from_str(text)
    return loaded, scene_mod.build_scene_buffers(loaded, env=_black_env())


def square_light_closed_form(half=0.5, h=H):
    """rho * Le * (form factor of a centred parallel square), as 4 corner
    rectangles: F = [X/sqrt(1+X^2) atan(Y/sqrt(1+X^2)) + (X <-> Y)] / 2pi."""
    x = half / h
    q = np.sqrt(1 + x * x)
    f = (2 * x / q * np.arctan(x / q)) / (2 * np.pi)
    return RHO * LE * 4 * f


def test_square_light_closed_form_and_one_sided_lights():
    expect = square_light_closed_form()
    for culled in (False, True):
        loaded, sb = square_light_scene(facing_down=True, culled=culled)
        cam = cam_mod.camera_for(loaded, 8, 8)
        sc = oracle.Scene(sb)
        nee, _ = oracle.render_nee(sc, cam, 8, 8, 512, 0, 2, threads=4)
        plain, _ = oracle.render(sc, cam, 8, 8, 2048, 0, 2, threads=4)
        m, mp = float(nee[..., :3].mean()), float(plain[..., :3].mean())
        print(f"culled={culled}: NEE {m:.5f} plain {mp:.5f} closed form {expect:.5f}")
        assert abs(m - expect) < 0.01 * expect
        assert abs(mp - expect) < 0.05 * expect
    # A one-sided light facing away from the floor sends it nothing, in both
    # estimators (the back face is culled for BSDF rays and light samples).
    loaded, sb = square_light_scene(facing_down=False, culled=True)
    cam = cam_mod.camera_for(loaded, 8, 8)
    sc = oracle.Scene(sb)
    nee, _ = oracle.render_nee(sc, cam, 8, 8, 64, 0, 2, threads=4)
    plain, _ = oracle.render(sc, cam, 8, 8, 64, 0, 2, threads=4)
    assert np.all(nee[..., :3] == 0.0) and np.all(plain[..., :3] == 0.0)


def test_cornell_nee_converges_with_less_noise():
    # The reference (4096 spp of the plain estimator) carries relMSE ~2e-3
    # of its own, which bounds how far the comparison can go.
    w = configs.c1(width=24, height=24)
    sc = oracle.Scene(w.buffers)
    ref = oracle.reference64_window(sc, w.camera, 24, 24, 0, 0, 24, 24, 4096, 1 << 20, 8, threads=8)
    nee64, _ = oracle.render_nee(sc, w.camera, 24, 24, 64, 0, 8, threads=8)
    plain64, _ = oracle.render(sc, w.camera, 24, 24, 64, 0, 8, threads=8)
    e_nee = learned.rel_mse(nee64, ref)
    e_plain = learned.rel_mse(plain64, ref)
    print(f"relMSE at 64 spp: NEE {e_nee:.4e}, plain {e_plain:.4e}")
    assert e_nee < 0.15 * e_plain
    nee1k, _ = oracle.render_nee(sc, w.camera, 24, 24, 1024, 0, 8, threads=8)
    e1k = learned.rel_mse(nee1k, ref)
    print(f"NEE at 1024 spp: {e1k:.4e}")
    assert e1k < 0.3 * e_nee and e1k < 4e-3
    r = float(ref.mean())
    assert abs(float(nee1k[..., :3].mean()) - r) < 0.01 * r
