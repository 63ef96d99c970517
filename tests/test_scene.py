"""Scene YAML loader and camera (no GPU).

Known answers follow the reference's own unit tests (src/scene.rs:1383-1723,
src/camera.rs:137-221) and closed forms.
"""
import math
import os

import numpy as np
import pytest

from mcrt_amd import camera as cam_mod, scene as S, types as T

F = np.float32
SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")
M = "materials:\n  m: { type: lambertian, albedo: [1.0, 1.0, 1.0] }\n"


def test_plane_two_triangles_normal_and_corners():
    y = M + "objects:\n  - {type: plane, normal: [0, 1, 0], half_extents: [5, 3], material: m}\n"
    s = S.load_scene_from_str(y)
    assert len(s.mesh_vertices) == 4 and len(s.mesh_triangles) == 2
    assert np.all(s.mesh_vertices["normal"][:, :3] == [0, 1, 0])
    p = s.mesh_vertices["position"][:, :3]
    # Gram-Schmidt: helper x, t1 = normalize(x cross n) = +z, t2 = n cross t1 = +x.
    np.testing.assert_array_equal(p, [[-3, 0, -5], [-3, 0, 5], [3, 0, 5], [3, 0, -5]])
    assert list(s.mesh_triangles["v"][0]) == [0, 1, 2] and list(s.mesh_triangles["v"][1]) == [0, 2, 3]
    assert np.all(s.mesh_triangles["mat_idx"] == 0)


def test_plane_backface_flag_in_bit_31():
    y = M + ("objects:\n  - {type: plane, normal: [0, 0, 1], half_extents: [1, 1], material: m,"
             " backface_culling: true}\n  - {type: plane, normal: [0, 0, 1], half_extents: [1, 1], material: m}\n")
    s = S.load_scene_from_str(y)
    mi = s.mesh_triangles["mat_idx"]
    assert np.all(mi[:2] == (1 << 31)) and np.all(mi[2:] == 0)


def test_box_faces_spans_and_normals():
    y = M + "objects:\n  - {type: box, half_extents: [1.0, 2.0, 3.0], material: m}\n"
    s = S.load_scene_from_str(y)
    v = s.mesh_vertices
    assert len(v) == 24 and len(s.mesh_triangles) == 12
    for n, span in (((1, 0, 0), (0, 4, 6)), ((-1, 0, 0), (0, 4, 6)), ((0, 1, 0), (2, 0, 6)),
                    ((0, -1, 0), (2, 0, 6)), ((0, 0, 1), (2, 4, 0)), ((0, 0, -1), (2, 4, 0))):
        f = v[np.all(v["normal"][:, :3] == n, axis=1)]
        assert len(f) == 4
        p = f["position"][:, :3]
        np.testing.assert_allclose(p.max(0) - p.min(0), span, atol=1e-5)


def test_box_rotate_y_90_maps_plus_x_to_minus_z():
    y = M + "objects:\n  - {type: box, half_extents: [1, 1, 1], material: m, transform: {rotate: [0, 90, 0]}}\n"
    v = S.load_scene_from_str(y).mesh_vertices
    f = v[np.abs(v["normal"][:, 2] + 1) < 1e-5]
    assert len(f) == 4
    np.testing.assert_allclose(f["position"][:, 2], -1.0, atol=1e-5)
    np.testing.assert_allclose(np.linalg.norm(v["normal"][:, :3], axis=1), 1.0, atol=1e-6)


def test_euler_order_is_rz_ry_rx():
    m = np.array(S.euler_to_mat3(90, 0, 0), np.float64)
    np.testing.assert_allclose(m @ [0, 1, 0], [0, 0, 1], atol=1e-6)
    m = np.array(S.euler_to_mat3(90, 90, 0), np.float64)  # R = Ry(90) Rx(90)
    ry = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]])
    rx = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])
    np.testing.assert_allclose(m, ry @ rx, atol=1e-6)


def test_material_defaults_and_packing():
    y = ("materials:\n  a: {type: lambertian}\n  b: {type: metal, albedo: [0.1, 0.2, 0.3], fuzz: 0.25}\n"
         "  c: {type: dielectric}\n  d: {type: emissive, albedo: [1, 0.5, 0.25], emission_strength: 15}\n"
         "  e: {type: emissive}\n")
    s = S.load_scene_from_str(y)
    m = s.materials["materials"]
    assert s.materials["mat_count"] == 5
    assert list(m[0]["type_pad"]) == [0, 0, 0, 0] and np.all(m[0]["albedo_fuzz"] == [1, 1, 1, 0])
    assert m[1]["type_pad"][0] == 1 and np.allclose(m[1]["albedo_fuzz"], [0.1, 0.2, 0.3, 0.25])
    assert m[2]["type_pad"][0] == 2 and m[2]["ior_pad"][0] == F(1.5)
    assert m[3]["type_pad"][0] == 3 and m[3]["ior_pad"][0] == 15 and np.all(m[3]["albedo_fuzz"][:3] == [1, 0.5, 0.25])
    assert m[4]["ior_pad"][0] == 1.0


def test_palette_order_world_materials_first_and_inline_appended():
    y = ("world_materials:\n  stone: {type: lambertian}\n  dup: {type: metal}\n"
         "materials:\n  dup: {type: dielectric}\n  w: {type: lambertian}\n"
         "objects:\n  - {type: sphere, center: [0,0,0], radius: 1, material: {type: metal, fuzz: 0.5}}\n"
         "  - {type: sphere, center: [0,0,0], radius: 1, material: w}\n")
    s = S.load_scene_from_str(y)
    assert s.name_to_idx == {"stone": 0, "dup": 1, "w": 2}
    assert s.materials["materials"][1]["type_pad"][0] == T.MAT_METAL  # first occurrence wins
    assert s.materials["mat_count"] == 4 and s.spheres["mat_and_pad"][0, 0] == 3
    assert s.material_names == ["stone", "dup", "w", "inline-3"]


def test_errors():
    with pytest.raises(S.SceneError) as e:
        S.load_scene_from_str(M + "objects:\n  - {type: sphere, center: [0,0,0], radius: 1, material: nope}\n")
    assert e.value.kind == "UnknownMaterial"
    many = "materials:\n" + "".join(f"  m{i}: {{type: lambertian}}\n" for i in range(65))
    with pytest.raises(S.SceneError) as e:
        S.load_scene_from_str(many)
    assert e.value.kind == "TooManyMaterials"
    with pytest.raises(S.SceneError) as e:
        S.load_scene_from_str("objects: [ {type: torus} ]")
    assert e.value.kind == "Parse"
    with pytest.raises(S.SceneError) as e:
        S.load_scene_from_str("materials:\n  a: {type: lambertian}\n  a: {type: metal}\n")
    assert e.value.kind == "Parse"
    with pytest.raises(S.SceneError) as e:
        S.load_scene_from_yaml("/nonexistent/scene.yaml")
    assert e.value.kind == "Io"


def test_stub_mesh_and_emissive_list():
    y = ("materials:\n  L: {type: emissive}\n  w: {type: lambertian}\nobjects:\n"
         "  - {type: sphere, center: [0,0,0], radius: 1, material: L}\n"
         "  - {type: sphere, center: [3,0,0], radius: 1, material: w}\n")
    s = S.load_scene_from_str(y)
    assert len(s.mesh_vertices) == 1 and np.all(s.mesh_vertices["position"][0] == [1e9, 1e9, 1e9, 1])
    assert len(s.mesh_triangles) == 1 and list(s.mesh_triangles["v"][0]) == [0, 0, 0]
    assert len(s.emissive_spheres) == 1 and s.materials["n_emissive"] == 1
    sb = S.build_scene_buffers(s)
    assert len(sb.spheres) == 2 and sorted(sb.sphere_orig) == [0, 1]


def test_reference_fixture_scenes_load():
    for f in sorted(os.listdir(SCENES)):
        if f.endswith(".yaml"):
            s = S.load_scene_from_yaml(os.path.join(SCENES, f))
            assert s.camera is not None and s.materials["mat_count"] >= 2


# ---------------------------------------------------------------------------
# Camera (src/camera.rs:76-127)
# ---------------------------------------------------------------------------
def _v(c, k):
    return np.asarray(c[k][:3], np.float64)


def test_camera_centre_ray_hits_look_at_and_extents():
    c = cam_mod.compute_camera(400, 200, (0, 1, 3.5), (0, 0, 0), 60.0, 0.0, math.sqrt(1 + 12.25))
    centre = _v(c, "lower_left") + 0.5 * _v(c, "horizontal") + 0.5 * _v(c, "vertical")
    np.testing.assert_allclose(centre, [0, 0, 0], atol=1e-5)
    fd = math.sqrt(13.25)
    vh = 2 * math.tan(math.radians(30))
    np.testing.assert_allclose(np.linalg.norm(_v(c, "vertical")), fd * vh, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(_v(c, "horizontal")), 2 * fd * vh, rtol=1e-6)
    assert np.all(c["defocus_u"] == 0) and np.all(c["defocus_v"] == 0)
    assert abs(_v(c, "horizontal") @ _v(c, "vertical")) < 1e-4


def test_camera_gimbal_guard_and_defocus():
    c = cam_mod.compute_camera(100, 100, (0, 5, 0), (0, 0, 0), 45.0, 0.5, 5.0)
    assert np.all(np.isfinite(c["horizontal"])) and np.linalg.norm(_v(c, "horizontal")) > 0
    np.testing.assert_allclose(np.linalg.norm(_v(c, "defocus_u")), 0.25, rtol=1e-6)


def test_camera_for_defaults():
    s = S.load_scene_from_str(M)
    c = cam_mod.camera_for(s, 64, 64)
    np.testing.assert_allclose(c["origin"][:3], [0, 1, 3.5])
    y = "camera: {look_from: [0, 0, 4], look_at: [0, 0, 0]}\n"
    s = S.load_scene_from_str(y)
    assert s.camera.vfov == 60.0 and s.camera.aperture == 0.0 and s.camera.focus_dist == 4.0
