"""The C++ SAH builders of the C-ABI (host-only, no GPU).

Pins: structural invariants of the reference layout (src/bvh.rs:543-677
test designs), and identity with a literal sort-based restatement of
build_recursive (src/mesh.rs:129-275) written here in plain Python: the
library partitions by bin instead of sorting, and must produce the same
nodes and the same leaf order.
"""
import numpy as np
import pytest

from mcrt_amd import build_mesh_bvh, build_sphere_bvh, scenes
from mcrt_amd import types as T

F = np.float32


def _random_tris(n, seed, spread=10.0, size=0.5):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 3))
    p = (c[:, None, :] + rng.uniform(-size, size, (n, 3, 3))).reshape(-1, 3)
    v = np.zeros(3 * n, T.GpuVertex)
    v["position"][:, :3] = p.astype(np.float32)
    v["position"][:, 3] = 1
    t = np.zeros(n, T.GpuTriangle)
    t["v"] = np.arange(3 * n).reshape(n, 3)
    t["mat_idx"] = np.arange(n) % 5
    return v, t


# ---------------------------------------------------------------------------
# Literal restatement of the reference's sort-based build (small n only)
# ---------------------------------------------------------------------------
def _ref_build(verts, tris):
    def tri_box(t):
        p = verts["position"][t["v"], :3]
        mn = p.min(axis=0).astype(F)
        mx = p.max(axis=0).astype(F)
        eps = F(1e-4)
        return np.minimum(mn, mx - eps), np.maximum(mx, mn + eps)

    prims = []
    for i, t in enumerate(tris):
        mn, mx = tri_box(t)
        prims.append((mn, mx, (mn + mx) * F(0.5), i))
    nodes, ordered = [], []

    def half_area(mn, mx):
        d = mx - mn
        return d[0] * d[1] + d[1] * d[2] + d[0] * d[2]

    def union(boxes):
        mn = np.full(3, np.inf, F)
        mx = np.full(3, -np.inf, F)
        for a, b in boxes:
            mn = np.minimum(mn, a)
            mx = np.maximum(mx, b)
        return mn, mx

    def best_split(ps, cmn, cmx, nmn, nmx):
        parent = max(half_area(nmn, nmx), F(1e-8))
        best, found, axis_b, split_b = F(len(ps)), False, 0, 0
        for axis in range(3):
            ln = cmx[axis] - cmn[axis]
            if ln < F(1e-8):
                continue
            bins = [[] for _ in range(8)]
            for p in ps:
                tt = (p[2][axis] - cmn[axis]) / ln
                bins[min(int(tt * F(8)), 7)].append(p)
            lb, lc = [], []
            acc = []
            for b in range(8):
                acc = acc + [(p[0], p[1]) for p in bins[b]]
                lb.append(union(acc))
                lc.append(len(acc))
            rb_list, rc = [], 0
            for s in range(6, -1, -1):
                rb_list = rb_list + [(p[0], p[1]) for p in bins[s + 1]]
                rc += len(bins[s + 1])
                if lc[s] == 0 or rc == 0:
                    continue
                rmn, rmx = union(rb_list)
                cost = (half_area(*lb[s]) * F(lc[s]) + half_area(rmn, rmx) * F(rc)) / parent
                if cost < best:
                    best, found, axis_b, split_b = cost, True, axis, lc[s]
        if found and 0 < split_b < len(ps):
            return axis_b, split_b
        return None

    def rec(ps):
        ni = len(nodes)
        nodes.append(None)
        mn, mx = union([(p[0], p[1]) for p in ps])
        if len(ps) > 4:
            cmn, cmx = union([(p[2], p[2]) for p in ps])
            s = best_split(ps, cmn, cmx, mn, mx)
        else:
            s = None
        if s is None:
            nodes[ni] = (mn, mx, len(ordered), len(ps))
            ordered.extend(p[3] for p in ps)
            return
        axis, k = s
        ps = sorted(ps, key=lambda p: (p[2][axis], p[3]))
        rec(ps[:k])
        right = len(nodes)
        rec(ps[k:])
        nodes[ni] = (mn, mx, right, 0)

    rec(prims)
    return nodes, ordered


@pytest.mark.parametrize("n,seed", [(7, 1), (60, 2), (300, 3)])
def test_matches_literal_sort_based_build(n, seed):
    v, t = _random_tris(n, seed)
    nodes, ordered, orig = build_mesh_bvh(v, t)
    rn, ro = _ref_build(v, t)
    assert len(nodes) == len(rn)
    for nd, (mn, mx, off, cnt) in zip(nodes, rn):
        assert np.array_equal(nd["aabb_min"][:3], mn) and np.array_equal(nd["aabb_max"][:3], mx)
        assert nd["right_or_offset"] == off and nd["prim_count"] == cnt
    assert list(orig) == ro
    assert np.array_equal(ordered, t[orig])


def _check_structure(nodes, n_prims, prim_boxes):
    n = len(nodes)
    assert 1 <= n <= max(2 * n_prims - 1, 1)
    seen = np.zeros(n_prims, int)
    stack = [0]
    visited = 0
    while stack:
        i = stack.pop()
        visited += 1
        nd = nodes[i]
        lo, hi = nd["aabb_min"][:3], nd["aabb_max"][:3]
        if nd["prim_count"] == 0:
            l, r = i + 1, nd["right_or_offset"]
            assert l < n and i + 1 < r < n
            for c in (l, r):
                assert np.all(nodes[c]["aabb_min"][:3] >= lo) and np.all(nodes[c]["aabb_max"][:3] <= hi)
            stack += [r, l]
        else:
            s, c = nd["right_or_offset"], nd["prim_count"]
            assert s + c <= n_prims
            seen[s:s + c] += 1
            bl, bh = prim_boxes
            assert np.all(bl[s:s + c] >= lo) and np.all(bh[s:s + c] <= hi)
    assert visited == n and np.all(seen == 1)


def test_icosphere_structure_and_determinism():
    v, t, _ = scenes.displaced_icosphere(n=20, seed=4)
    nodes, ordered, orig = build_mesh_bvh(v, t)
    p = v["position"][ordered["v"], :3]
    mn, mx = p.min(axis=1), p.max(axis=1)
    eps = np.float32(1e-4)
    _check_structure(nodes, len(t), (np.minimum(mn, mx - eps), np.maximum(mx, mn + eps)))
    assert sorted(orig) == list(range(len(t)))
    leaf = nodes["prim_count"][nodes["prim_count"] > 0]
    assert leaf.max() <= 4
    n2, o2, g2 = build_mesh_bvh(v, t)
    assert np.array_equal(nodes, n2) and np.array_equal(orig, g2)


def test_flat_triangles_get_padded_boxes():
    v, t = _random_tris(1, 0)
    v["position"][:, 2] = -5.0
    nodes, _, _ = build_mesh_bvh(v, t)
    assert nodes[0]["aabb_max"][2] - nodes[0]["aabb_min"][2] > 1e-5


def test_degenerate_centroids_make_one_big_leaf():
    v, t = _random_tris(1, 0)
    t = np.repeat(t, 9)
    nodes, ordered, orig = build_mesh_bvh(v, t)
    assert len(nodes) == 1 and nodes[0]["prim_count"] == 9 and list(orig) == list(range(9))


def test_empty_and_single():
    v, t = _random_tris(1, 0)
    nodes, ordered, orig = build_mesh_bvh(v, t[:0])
    assert len(nodes) == 0
    nodes, ordered, orig = build_mesh_bvh(v, t)
    assert len(nodes) == 1 and nodes[0]["prim_count"] == 1


def test_bad_vertex_index_is_rejected():
    from mcrt_amd import McrtError

    v, t = _random_tris(2, 0)
    t[1]["v"][2] = 1000
    with pytest.raises(McrtError):
        build_mesh_bvh(v, t)


def test_sphere_bvh_structure():
    rng = np.random.default_rng(9)
    s = np.zeros(200, T.GpuSphere)
    s["center_r"][:, :3] = rng.uniform(-20, 20, (200, 3))
    s["center_r"][:, 3] = rng.uniform(0.1, 1.0, 200)
    s["mat_and_pad"][:, 0] = np.arange(200) % 3
    nodes, ordered, orig = build_sphere_bvh(s)
    c = ordered["center_r"]
    _check_structure(nodes, 200, (c[:, :3] - c[:, 3:], c[:, :3] + c[:, 3:]))
    assert np.array_equal(ordered, s[orig])


# ---------------------------------------------------------------------------
# Tree validation shared with mcrt_upload_scene (mcrt_validate_bvh)
# ---------------------------------------------------------------------------
def _tree(spec):
    """spec: per node ('i', right) for internal or ('l', first, count)."""
    nodes = np.zeros(len(spec), T.GpuBvhNode)
    for i, s in enumerate(spec):
        if s[0] == "i":
            nodes[i]["right_or_offset"], nodes[i]["prim_count"] = s[1], 0
        else:
            nodes[i]["right_or_offset"], nodes[i]["prim_count"] = s[1], s[2]
        nodes[i]["aabb_min"][:3], nodes[i]["aabb_max"][:3] = -1, 1
    return nodes


def test_validate_accepts_builder_trees_and_stubs():
    from mcrt_amd import _lib

    v, t, _ = scenes.displaced_icosphere(n=6, seed=1, n_mats=1)
    nodes, _, _ = build_mesh_bvh(v, t)
    _lib.validate_bvh(nodes, len(t))
    _lib.validate_bvh(np.zeros(0, T.GpuBvhNode), 0)           # empty
    _lib.validate_bvh(np.zeros(1, T.GpuBvhNode), 0)           # the reference's zeroed stub node
    _lib.validate_bvh(_tree([("l", 0, 3)]), 3)                 # a single leaf root
    _lib.validate_bvh(_tree([("i", 2), ("l", 0, 2), ("l", 2, 1)]), 3)


@pytest.mark.parametrize("spec,n_prims,reason", [
    # node 2 is internal but no parent reaches it (the case that used to
    # index past the converted-node vector)
    ([("i", 3), ("l", 0, 1), ("i", 4), ("l", 1, 1), ("l", 2, 1)], 3, "not reachable"),
    # node 3 is the right child of both 0 and 1
    ([("i", 3), ("i", 3), ("l", 0, 1), ("l", 1, 1)], 2, "two parents"),
    # leaves overlap
    ([("i", 2), ("l", 0, 2), ("l", 1, 2)], 3, "two leaves"),
    # primitive 3 in no leaf
    ([("i", 2), ("l", 0, 2), ("l", 2, 1)], 4, "in no leaf"),
    # leaf beyond the primitive count
    ([("i", 2), ("l", 0, 2), ("l", 2, 5)], 4, "beyond"),
    # right child outside the layout / not after the left child
    ([("i", 7), ("l", 0, 1), ("l", 1, 1)], 2, "outside the pre-order"),
    ([("i", 1), ("l", 0, 1), ("l", 1, 1)], 2, "outside the pre-order"),
    # an internal node as the last node has no left child
    ([("i", 2), ("l", 0, 1), ("i", 3)], 1, "outside the pre-order"),
])
def test_validate_rejects_malformed_trees(spec, n_prims, reason):
    from mcrt_amd import _lib

    with pytest.raises(_lib.McrtError) as e:
        _lib.validate_bvh(_tree(spec), n_prims)
    assert e.value.code == -2  # MCRT_ERR_BAD_ARG
    assert reason in str(e.value), str(e.value)
