// bvh_refine.h — host-side leaf refinement of a reference mesh BVH, used by
// mcrt_upload_scene before the tree is converted to device child pairs.
// Host code only (no HIP), so the tests build it with the host compiler.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/mcrt_abi.h"

namespace mcrt {

constexpr uint32_t kNoParent = 0xFFFFFFFFu;

// Leaf refinement of a validated mesh tree.  Every leaf of 3 to kRefineMax
// triangles becomes a subtree whose leaves hold at most 2 (a voxel face's
// two triangles share one box): an exact SAH sweep over the leaf's
// triangles, split while a leaf holds more than 2 or a split is cheaper than
// the leaf, unless the two halves would get equal boxes (then both are
// visited whenever one is), over the same padded triangle boxes as the reference builder
// (src/bvh.rs:190-204).  Each new box is widened by a margin of 2^-18 of the
// scene's largest root-box coordinate or extent, far above the rounding of
// the slab test, and then clipped to the given leaf's box.  A new box
// therefore lies inside the given leaf box (it accepts no ray interval the
// given box rejects), and a hit point that is not on the given box's own
// boundary lies inside it by the margin, so the slab test does not reject it
// at a new plane through a vertex.  Internal nodes and leaves of at most 2
// triangles are kept as given.  The result is again
// a reference pre-order tree; perm[j] is the given leaf-order index of the
// triangle at refined index j (triangles move only within their leaf).
// Closest-hit results do not depend on the tree: a child is visited when the
// slab test passes and its entry does not exceed best t, and the hit is the
// lowest (t, original id).
constexpr uint32_t kRefineMax = 8;

struct LeafTri {
    float mn[3], mx[3], c[3];
    uint32_t idx;
};

inline float half_area3(const float mn[3], const float mx[3]) {
    const float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return d0 * d1 + d1 * d2 + d0 * d2;
}

// Union of the padded triangle boxes of t[0, n) (the SAH's parent box) and
// the stored box: the union widened by the margin and clipped to the leaf.
inline void leaf_box(const LeafTri *t, size_t n, const GpuBvhNode &leaf, float margin, float tmn[3],
                     float tmx[3], float mn[3], float mx[3]) {
    for (int k = 0; k < 3; ++k) {
        tmn[k] = INFINITY;
        tmx[k] = -INFINITY;
    }
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            tmn[k] = std::fmin(tmn[k], t[i].mn[k]);
            tmx[k] = std::fmax(tmx[k], t[i].mx[k]);
        }
    for (int k = 0; k < 3; ++k) {
        mn[k] = std::fmax(tmn[k] - margin, leaf.aabb_min[k]);
        mx[k] = std::fmin(tmx[k] + margin, leaf.aabb_max[k]);
    }
}

inline void emit_node(std::vector<GpuBvhNode> &out, const float mn[3], const float mx[3], uint32_t r,
               uint32_t count) {
    GpuBvhNode g{};
    for (int k = 0; k < 3; ++k) {
        g.aabb_min[k] = mn[k];
        g.aabb_max[k] = mx[k];
    }
    g.right_or_offset = r;
    g.prim_count = count;
    out.push_back(g);
}

// Subtree over t[0, n) (n <= kRefineMax, so the recursion is shallow).
inline void refine_leaf(LeafTri *t, size_t n, const GpuBvhNode &leaf, float margin,
                        std::vector<GpuBvhNode> &out, std::vector<uint32_t> &perm) {
    float tmn[3], tmx[3], mn[3], mx[3];
    leaf_box(t, n, leaf, margin, tmn, tmx, mn, mx);
    float best = INFINITY;
    int bax = -1;
    size_t bk = 0;
    if (n > 1) {
        const float parent = std::fmax(half_area3(tmn, tmx), 1e-8f);
        for (int ax = 0; ax < 3; ++ax) {
            std::sort(t, t + n, [ax](const LeafTri &a, const LeafTri &b) {
                return a.c[ax] < b.c[ax] || (a.c[ax] == b.c[ax] && a.idx < b.idx);
            });
            float ra[kRefineMax];
            float rmn[3] = {INFINITY, INFINITY, INFINITY}, rmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (size_t i = n; i-- > 1;) {
                for (int k = 0; k < 3; ++k) {
                    rmn[k] = std::fmin(rmn[k], t[i].mn[k]);
                    rmx[k] = std::fmax(rmx[k], t[i].mx[k]);
                }
                ra[i] = half_area3(rmn, rmx);
            }
            float lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (size_t i = 0; i + 1 < n; ++i) {
                for (int k = 0; k < 3; ++k) {
                    lmn[k] = std::fmin(lmn[k], t[i].mn[k]);
                    lmx[k] = std::fmax(lmx[k], t[i].mx[k]);
                }
                const float cost =
                    (half_area3(lmn, lmx) * (float)(i + 1) + ra[i + 1] * (float)(n - i - 1)) / parent;
                if (cost < best) {
                    best = cost;
                    bax = ax;
                    bk = i + 1;
                }
            }
        }
    }
    bool leaf_here = bax < 0 || (n <= 2 && !(best < (float)n));
    if (!leaf_here) {
        std::sort(t, t + n, [bax](const LeafTri &a, const LeafTri &b) {
            return a.c[bax] < b.c[bax] || (a.c[bax] == b.c[bax] && a.idx < b.idx);
        });
        // Two children with equal stored boxes (the margin and the clip can
        // make them equal) are both visited whenever one is: one leaf then.
        float a0[3], a1[3], b0[3], b1[3], u[3], v[3];
        leaf_box(t, bk, leaf, margin, u, v, a0, a1);
        leaf_box(t + bk, n - bk, leaf, margin, u, v, b0, b1);
        leaf_here = !std::memcmp(a0, b0, sizeof a0) && !std::memcmp(a1, b1, sizeof a1);
    }
    if (leaf_here) {
        emit_node(out, mn, mx, (uint32_t)perm.size(), (uint32_t)n);
        for (size_t i = 0; i < n; ++i) perm.push_back(t[i].idx);
        return;
    }
    const size_t self = out.size();
    emit_node(out, mn, mx, 0u, 0u);
    refine_leaf(t, bk, leaf, margin, out, perm);
    out[self].right_or_offset = (uint32_t)out.size();
    refine_leaf(t + bk, n - bk, leaf, margin, out, perm);
}

// False (out untouched) when no leaf holds more than 2 triangles.  The
// margin is 2^margin_exp of the scene scale.
inline bool refine_mesh_tree(const GpuBvhNode *nodes, uint32_t n, const GpuVertex *verts,
                             const GpuTriangle *tris, uint32_t n_tris, std::vector<GpuBvhNode> &out,
                             std::vector<uint32_t> &perm, int margin_exp = -18) {
    if (n == 0 || (n == 1 && nodes[0].prim_count == 0)) return false;
    bool any = false;
    for (uint32_t i = 0; i < n && !any; ++i)
        any = nodes[i].prim_count > 2 && nodes[i].prim_count <= kRefineMax;
    if (!any) return false;
    float scale = 0.0f;
    for (int k = 0; k < 3; ++k)
        scale = std::fmax(scale, std::fmax(std::fmax(std::fabs(nodes[0].aabb_min[k]), std::fabs(nodes[0].aabb_max[k])),
                                           nodes[0].aabb_max[k] - nodes[0].aabb_min[k]));
    if (!std::isfinite(scale)) return false;
    const float margin = std::ldexp(scale, margin_exp);
    out.clear();
    out.reserve((size_t)n * 2);
    perm.clear();
    perm.reserve(n_tris);
    std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, kNoParent}};  // (node, new parent awaiting its right index)
    LeafTri lt[kRefineMax];
    while (!todo.empty()) {
        const auto [i, parent] = todo.back();
        todo.pop_back();
        if (parent != kNoParent) out[parent].right_or_offset = (uint32_t)out.size();
        const GpuBvhNode &g = nodes[i];
        if (g.prim_count == 0) {
            out.push_back(g);
            todo.push_back({g.right_or_offset, (uint32_t)out.size() - 1u});
            todo.push_back({i + 1, kNoParent});
        } else if (g.prim_count <= 2 || g.prim_count > kRefineMax) {
            out.push_back(g);
            out.back().right_or_offset = (uint32_t)perm.size();
            for (uint32_t k = 0; k < g.prim_count; ++k) perm.push_back(g.right_or_offset + k);
        } else {
            for (uint32_t k = 0; k < g.prim_count; ++k) {
                const uint32_t p = g.right_or_offset + k;
                LeafTri &x = lt[k];
                for (int a = 0; a < 3; ++a) {
                    x.mn[a] = INFINITY;
                    x.mx[a] = -INFINITY;
                }
                for (int v = 0; v < 3; ++v) {
                    const float *q = verts[tris[p].v[v]].position;
                    for (int a = 0; a < 3; ++a) {
                        x.mn[a] = std::fmin(x.mn[a], q[a]);
                        x.mx[a] = std::fmax(x.mx[a], q[a]);
                    }
                }
                for (int a = 0; a < 3; ++a) {  // padded as src/bvh.rs:190-204
                    const float lo = x.mn[a], hi = x.mx[a];
                    x.mn[a] = std::fmin(lo, hi - 1e-4f);
                    x.mx[a] = std::fmax(hi, lo + 1e-4f);
                    x.c[a] = (x.mn[a] + x.mx[a]) * 0.5f;
                }
                x.idx = p;
            }
            refine_leaf(lt, g.prim_count, g, margin, out, perm);
        }
    }
    return true;
}

}  // namespace mcrt
