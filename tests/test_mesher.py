"""Greedy voxel mesher (mcrt_mesh_voxel_chunks) against the reference's known
answers and brute-force invariants (no GPU; the mesher is host code).

Known answers are those of the reference's own tests,
src/world/mesher.rs:429-507 (air chunk, single voxel, merged slab, culled
shared face, merge offsets), here with their exact triangle counts.  The
brute-force checks rasterise every quad back onto the voxel faces it covers
and compare with a numpy count of exposed faces (an opaque voxel whose
neighbour in the face direction is not opaque; outside the chunk grid is air,
as the reference's missing neighbour, :255-290).
"""
import numpy as np
import pytest

from mcrt_amd import world
from mcrt_amd._lib import mesh_voxel_chunks

S = world.VOXEL_SIZE
OPAQUE = np.zeros(256, np.uint8)
OPAQUE[1:5] = 1          # ids 1-4 opaque, id 5 present but not opaque (like water)
MAT = np.arange(256, dtype=np.uint32) * 3 + 7  # distinct material slot per id


def grid(cx=1, cz=1):
    return np.zeros((cx, cz, 16, 256, 16), np.uint16)


def mesh(blocks, cx0=0, cz0=0):
    return mesh_voxel_chunks(blocks, cx0, cz0, S, OPAQUE, MAT)


def quads(verts, tris):
    """Per quad: axis, sign, voxel-unit corners (4, 3), material."""
    assert len(tris) % 2 == 0 and len(verts) == 2 * len(tris)
    out = []
    for q in range(len(tris) // 2):
        p = verts["position"][4 * q:4 * q + 4, :3].astype(np.float64) / S
        n = verts["normal"][4 * q, :3]
        axis = int(np.argmax(np.abs(n)))
        out.append((axis, int(np.sign(n[axis])), np.rint(p).astype(np.int64), int(tris[2 * q]["mat_idx"])))
    return out


# ---------------------------------------------------------------------------
# Reference known answers (src/world/mesher.rs:429-507)
# ---------------------------------------------------------------------------
def test_air_chunk_produces_no_triangles():
    v, t = mesh(grid())
    assert len(v) == 0 and len(t) == 0


def test_single_voxel_produces_12_triangles_of_the_unit_cube():
    b = grid()
    b[0, 0, 8, 8, 8] = 2
    v, t = mesh(b)
    assert len(t) == 12 and len(v) == 24
    lo, hi = v["position"][:, :3].min(0), v["position"][:, :3].max(0)
    np.testing.assert_array_equal(lo, np.float32(8 * S))
    np.testing.assert_array_equal(hi, np.float32(9 * S))
    seen = sorted((a, s) for a, s, _, _ in quads(v, t))
    assert seen == [(a, s) for a in range(3) for s in (-1, 1)]
    assert np.all(t["mat_idx"] == MAT[2])


def test_flat_2x2_slab_merges_every_face():
    b = grid()
    b[0, 0, 0:2, 0, 0:2] = 1
    v, t = mesh(b)
    # top and bottom merge to one quad each, each side is one 2 x 1 quad
    assert len(t) == 12 < 24
    areas = sorted(int(round(np.linalg.norm(np.cross(q[2][1] - q[2][0], q[2][3] - q[2][0]))))
                   for q in quads(v, t))
    assert areas == [2, 2, 2, 2, 4, 4]


def test_adjacent_blocks_cull_shared_face():
    b = grid()
    b[0, 0, 0, 0, 0] = 1
    b[0, 0, 1, 0, 0] = 1
    v, t = mesh(b)
    assert len(t) == 12  # 6 merged quads: no face between the two blocks
    xs = [q for q in quads(v, t) if q[0] == 0]
    planes = sorted(int(q[2][0][0]) for q in xs)
    assert planes == [0, 2]  # only the outer +-X faces, at x = 0 and x = 2


def test_merged_chunks_offset_indices_and_positions():
    b = grid(2, 1)
    b[0, 0, 0, 0, 0] = 1
    b[1, 0, 0, 0, 0] = 1
    v, t = mesh(b)
    assert len(t) == 24 and len(v) == 48
    assert t["v"].max() + 1 == len(v)
    assert t["v"][12:].min() == 24  # the second chunk's triangles start after the first's 24 vertices
    np.testing.assert_array_equal(v["position"][24:, :3] - v["position"][:24, :3],
                                  np.tile(np.float32([16 * S, 0, 0]), (24, 1)))


def test_neighbour_chunk_culls_across_the_border():
    b = grid(2, 1)
    b[0, 0, 15, 5, 3] = 1
    b[1, 0, 0, 5, 3] = 1
    v, t = mesh(b)
    # each chunk keeps five faces of its voxel; the shared x = 16 face is culled
    assert len(t) == 20
    assert not any(q[0] == 0 and q[2][0][0] == 16 for q in quads(v, t))


def test_non_opaque_blocks_are_skipped_and_expose_faces():
    b = grid()
    b[0, 0, 4, 4, 4] = 1
    b[0, 0, 5, 4, 4] = 5  # not opaque: no mesh of its own, the +X face of (4,4,4) stays
    v, t = mesh(b)
    assert len(t) == 12


def test_chunk_origin_uses_grid_offset():
    b = grid()
    b[0, 0, 0, 0, 0] = 1
    v, _ = mesh(b, cx0=-3, cz0=2)
    lo = v["position"][:, :3].min(0)
    np.testing.assert_array_equal(lo, np.float32([-3 * 16 * S, 0, 2 * 16 * S]))


# ---------------------------------------------------------------------------
# Brute force: every exposed face is covered exactly once
# ---------------------------------------------------------------------------
def exposed_faces(blocks):
    """(6, X, Y, Z) bool: voxel is opaque and its neighbour in direction
    (+X, -X, +Y, -Y, +Z, -Z) is not; outside the grid is air."""
    cx, cz = blocks.shape[:2]
    w = blocks.transpose(0, 2, 3, 1, 4).reshape(cx * 16, 256, cz * 16)
    op = OPAQUE[np.minimum(w, 255)].astype(bool)
    out = np.zeros((6,) + op.shape, bool)
    for axis in range(3):
        pad = [(0, 0)] * 3
        pad[axis] = (1, 1)
        p = np.pad(op, pad)
        sl_next = [slice(None)] * 3
        sl_prev = [slice(None)] * 3
        sl_next[axis] = slice(2, None)
        sl_prev[axis] = slice(0, -2)
        out[2 * axis] = op & ~p[tuple(sl_next)]
        out[2 * axis + 1] = op & ~p[tuple(sl_prev)]
    return out, w


def rasterise(blocks, v, t, cx0=0, cz0=0):
    cx, cz = blocks.shape[:2]
    cover = np.zeros((6, cx * 16, 256, cz * 16), np.int32)
    _, w = exposed_faces(blocks)
    verts_vox = v["position"][:, :3].astype(np.float64) / S
    for qi, (axis, sign, p, mat) in enumerate(quads(v, t)):
        p = p - np.array([cx0 * 16, 0, cz0 * 16])
        lo, hi = p.min(0), p.max(0)
        assert lo[axis] == hi[axis]  # planar, axis-aligned
        layer = lo[axis] - (1 if sign > 0 else 0)
        sl = [slice(lo[k], hi[k]) for k in range(3)]
        sl[axis] = slice(layer, layer + 1)
        d = 2 * axis + (0 if sign > 0 else 1)
        cover[d][tuple(sl)] += 1
        ids = np.unique(w[tuple(sl)])
        assert len(ids) == 1 and MAT[ids[0]] == mat  # one block id per merged quad
        # winding: emit_quad's corner order (mesher.rs:359-394) makes the
        # geometric normal of each triangle sign * (u_axis x v_axis): outward
        # for X and Z faces, inward for Y faces (u = X, v = Z).
        for tri in (t["v"][2 * qi], t["v"][2 * qi + 1]):
            a, b_, c = (verts_vox[i] for i in tri)
            g = np.cross(b_ - a, c - a)
            assert np.sign(g[axis]) == sign * (-1 if axis == 1 else 1)
    return cover


@pytest.mark.parametrize("seed", [1, 2])
def test_random_world_faces_covered_exactly_once(seed):
    rng = np.random.default_rng(seed)
    b = grid(2, 2)
    # sparse blobs in the lower 48 layers with four block types and some
    # non-opaque voxels, plus one solid column touching the top of the chunk
    dense = rng.random((2, 2, 16, 48, 16))
    ids = rng.integers(1, 6, (2, 2, 16, 48, 16)).astype(np.uint16)
    b[:, :, :, :48, :] = np.where(dense < 0.35, ids, 0)
    b[1, 0, 3, :, 7] = 2
    v, t = mesh(b, cx0=-1, cz0=4)
    exp, _ = exposed_faces(b)
    cover = rasterise(b, v, t, cx0=-1, cz0=4)
    assert np.array_equal(cover, exp.astype(np.int32))
    # greedy merging emits fewer quads than exposed faces
    assert len(t) // 2 < exp.sum()


def test_c3_terrain_area_per_direction_matches_exposed_faces():
    blocks = world.terrain_blocks(32, 32, 11)
    opq = np.zeros(256, np.uint8)
    mat = np.zeros(256, np.uint32)
    for bid in world.BLOCK_MATERIALS:
        opq[bid] = 1
        mat[bid] = bid
    v, t = mesh_voxel_chunks(blocks, -16, -16, S, opq, mat)
    assert len(t) == 2161464
    global OPAQUE
    saved = OPAQUE
    try:
        OPAQUE = opq
        exp, _ = exposed_faces(blocks)
    finally:
        OPAQUE = saved
    p = v["position"][:, :3].astype(np.float64).reshape(-1, 4, 3) / S
    n = v["normal"][::4, :3]
    area = np.abs(np.cross(p[:, 1] - p[:, 0], p[:, 3] - p[:, 0])).sum(1)
    for d in range(6):
        axis, sign = d // 2, (1 if d % 2 == 0 else -1)
        sel = n[:, axis] == sign
        assert int(round(area[sel].sum())) == int(exp[d].sum()), d
    del blocks
