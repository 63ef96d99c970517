// bvh_build.cpp — host-side binned-SAH BVH builders of the C-ABI.
//
// Semantics follow the reference builders src/mesh.rs:82-291 (triangles) and
// src/bvh.rs:239-463 (spheres): 8 bins per axis, leaf forced at <= 4
// primitives, split accepted only when strictly cheaper than the leaf cost N,
// cost (A_L*N_L + A_R*N_R) / A_P over half-areas, pre-order layout with
// left child = parent + 1.  Triangle boxes are padded by 1e-4 on flat axes
// (src/bvh.rs:190-204, src/mesh.rs:282-291); sphere boxes are not.
//
// Differences in mechanics, not in result:
//  - The reference sorts the whole slice by centroid at every split
//    (src/mesh.rs:153-157) and then cuts it at the left count.  The prims of
//    bins 0..=split are exactly that sorted prefix (bins are monotone in the
//    centroid), so a linear partition by bin yields the same two sets.  Only
//    the order inside a leaf depends on the sort; a leaf therefore sorts its
//    own prims by the parent's split axis, which reproduces the reference's
//    leaf order in O(n log n) total instead of O(n log^2 n).
//  - Centroid ties are broken by original index, where the reference's
//    sort_unstable_by leaves the order unspecified.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mcrt_abi.h"

namespace {

constexpr int kSahBins = 8;     // src/bvh.rs:38
constexpr size_t kLeafMax = 4;  // src/bvh.rs:32

struct Aabb {
    float mn[3], mx[3];
    static Aabb empty() {
        Aabb a;
        for (int i = 0; i < 3; ++i) {
            a.mn[i] = INFINITY;
            a.mx[i] = -INFINITY;
        }
        return a;
    }
    // Rust f32::min/max ignore a NaN operand, as fminf/fmaxf do.
    Aabb onion(const Aabb &o) const {
        Aabb r;
        for (int i = 0; i < 3; ++i) {
            r.mn[i] = std::fmin(mn[i], o.mn[i]);
            r.mx[i] = std::fmax(mx[i], o.mx[i]);
        }
        return r;
    }
    float half_area() const {  // src/bvh.rs:164-171, same operation order
        float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
        return d0 * d1 + d1 * d2 + d0 * d2;
    }
    void centroid(float c[3]) const {
        for (int i = 0; i < 3; ++i) c[i] = (mn[i] + mx[i]) * 0.5f;
    }
    Aabb padded() const {  // src/bvh.rs:190-204
        const float eps = 1e-4f;
        Aabb r;
        for (int i = 0; i < 3; ++i) {
            r.mn[i] = std::fmin(mn[i], mx[i] - eps);
            r.mx[i] = std::fmax(mx[i], mn[i] + eps);
        }
        return r;
    }
};

struct Prim {
    Aabb bounds;
    float c[3];
    uint32_t idx;  // original primitive index
};

// Rust `(t * 8.0) as usize` saturates: NaN and negatives -> 0.
inline int bin_of(float t) {
    float x = t * (float)kSahBins;
    int b = 0;
    if (x > 0.0f) b = x >= 2147483647.0f ? kSahBins - 1 : (int)x;
    return std::min(b, kSahBins - 1);
}

struct Split {
    int axis = 0;
    int bin = 0;       // prims with bin <= this go left
    size_t left = 0;   // left count
    bool found = false;
};

// find_best_split, src/mesh.rs:206-275 (identical to src/bvh.rs:388-463).
Split find_best_split(const Prim *p, size_t n, const Aabb &cb, const Aabb &nb) {
    Split s;
    float parent = std::fmax(nb.half_area(), 1e-8f);
    float best = (float)n;
    for (int axis = 0; axis < 3; ++axis) {
        float len = cb.mx[axis] - cb.mn[axis];
        if (len < 1e-8f) continue;
        Aabb bb[kSahBins];
        uint32_t bc[kSahBins] = {0};
        for (int b = 0; b < kSahBins; ++b) bb[b] = Aabb::empty();
        for (size_t i = 0; i < n; ++i) {
            float t = (p[i].c[axis] - cb.mn[axis]) / len;
            int b = bin_of(t);
            bb[b] = bb[b].onion(p[i].bounds);
            bc[b] += 1;
        }
        Aabb lb[kSahBins];
        uint32_t lc[kSahBins];
        lb[0] = bb[0];
        lc[0] = bc[0];
        for (int i = 1; i < kSahBins; ++i) {
            lb[i] = lb[i - 1].onion(bb[i]);
            lc[i] = lc[i - 1] + bc[i];
        }
        Aabb rb = Aabb::empty();
        uint32_t rc = 0;
        for (int split = kSahBins - 2; split >= 0; --split) {
            rb = rb.onion(bb[split + 1]);
            rc += bc[split + 1];
            uint32_t l = lc[split], r = rc;
            if (l == 0 || r == 0) continue;
            float cost = (lb[split].half_area() * (float)l + rb.half_area() * (float)r) / parent;
            if (cost < best) {
                best = cost;
                s.axis = axis;
                s.bin = split;
                s.left = l;
                s.found = true;
            }
        }
    }
    if (!(s.found && s.left > 0 && s.left < n)) s.found = false;
    return s;
}

template <class T>
struct Builder {
    std::vector<Prim> &prims;
    const T *src;
    GpuBvhNode *nodes;
    uint32_t n_nodes = 0;
    T *ordered;
    uint32_t *orig;
    uint32_t n_ordered = 0;

    void set_box(GpuBvhNode &nd, const Aabb &b) {
        for (int i = 0; i < 3; ++i) {
            nd.aabb_min[i] = b.mn[i];
            nd.aabb_max[i] = b.mx[i];
        }
        nd.aabb_min[3] = 0.0f;
        nd.aabb_max[3] = 0.0f;
    }

    void make_leaf(size_t lo, size_t hi, const Aabb &b, int parent_axis, uint32_t ni) {
        if (parent_axis >= 0) {
            std::sort(prims.begin() + lo, prims.begin() + hi,
                      [parent_axis](const Prim &a, const Prim &c) {
                          float x = a.c[parent_axis], y = c.c[parent_axis];
                          if (x < y) return true;
                          if (y < x) return false;
                          return a.idx < c.idx;
                      });
        }
        GpuBvhNode &nd = nodes[ni];
        std::memset(&nd, 0, sizeof nd);
        set_box(nd, b);
        nd.right_or_offset = n_ordered;
        nd.prim_count = (uint32_t)(hi - lo);
        for (size_t i = lo; i < hi; ++i) {
            ordered[n_ordered] = src[prims[i].idx];
            orig[n_ordered] = prims[i].idx;
            ++n_ordered;
        }
    }

    // build_recursive, src/mesh.rs:129-183.  The recursion depth equals the
    // tree depth (the reference recurses the same way).
    void build(size_t lo, size_t hi, int parent_axis) {
        uint32_t ni = n_nodes++;
        Aabb b = Aabb::empty();
        for (size_t i = lo; i < hi; ++i) b = b.onion(prims[i].bounds);
        size_t n = hi - lo;
        if (n <= kLeafMax) {
            make_leaf(lo, hi, b, parent_axis, ni);
            return;
        }
        Aabb cb = Aabb::empty();
        for (size_t i = lo; i < hi; ++i) {
            Aabb pt;
            for (int k = 0; k < 3; ++k) pt.mn[k] = pt.mx[k] = prims[i].c[k];
            cb = cb.onion(pt);
        }
        Split s = find_best_split(prims.data() + lo, n, cb, b);
        if (!s.found) {
            make_leaf(lo, hi, b, parent_axis, ni);
            return;
        }
        const int axis = s.axis;
        const float cmin = cb.mn[axis];
        const float len = cb.mx[axis] - cb.mn[axis];
        const int sbin = s.bin;
        auto mid = std::stable_partition(prims.begin() + lo, prims.begin() + hi,
                                         [=](const Prim &p) {
                                             return bin_of((p.c[axis] - cmin) / len) <= sbin;
                                         });
        size_t m = (size_t)(mid - prims.begin());
        build(lo, m, axis);
        uint32_t right = n_nodes;
        build(m, hi, axis);
        GpuBvhNode &nd = nodes[ni];
        std::memset(&nd, 0, sizeof nd);
        set_box(nd, b);
        nd.right_or_offset = right;
        nd.prim_count = 0;
    }
};

inline bool finite3(const float *p) {
    return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]);
}

}  // namespace

extern "C" int mcrt_build_mesh_bvh(const GpuVertex *verts, uint32_t n_verts,
                                   const GpuTriangle *tris, uint32_t n_tris,
                                   GpuBvhNode *out_nodes, uint32_t *out_n_nodes,
                                   GpuTriangle *out_ordered, uint32_t *out_orig_index) {
    if (!out_n_nodes) return MCRT_ERR_BAD_ARG;
    *out_n_nodes = 0;
    if (n_tris == 0) return MCRT_OK;
    if (!verts || !tris || !out_nodes || !out_ordered || !out_orig_index)
        return MCRT_ERR_BAD_ARG;
    std::vector<Prim> prims(n_tris);
    for (uint32_t i = 0; i < n_tris; ++i) {
        const GpuTriangle &t = tris[i];
        if (t.v[0] >= n_verts || t.v[1] >= n_verts || t.v[2] >= n_verts)
            return MCRT_ERR_BAD_ARG;
        Aabb a = Aabb::empty();  // Aabb::from_points_3, src/bvh.rs:116-127
        for (int k = 0; k < 3; ++k) {
            const float *p = verts[t.v[k]].position;
            for (int c = 0; c < 3; ++c) {
                a.mn[c] = std::fmin(a.mn[c], p[c]);
                a.mx[c] = std::fmax(a.mx[c], p[c]);
            }
        }
        prims[i].bounds = a.padded();
        prims[i].bounds.centroid(prims[i].c);
        prims[i].idx = i;
    }
    Builder<GpuTriangle> b{prims, tris, out_nodes, 0, out_ordered, out_orig_index, 0};
    b.build(0, n_tris, -1);
    *out_n_nodes = b.n_nodes;
    return MCRT_OK;
}

extern "C" int mcrt_build_sphere_bvh(const GpuSphere *spheres, uint32_t n_spheres,
                                     GpuBvhNode *out_nodes, uint32_t *out_n_nodes,
                                     GpuSphere *out_ordered, uint32_t *out_orig_index) {
    if (!out_n_nodes) return MCRT_ERR_BAD_ARG;
    *out_n_nodes = 0;
    if (n_spheres == 0) return MCRT_OK;
    if (!spheres || !out_nodes || !out_ordered || !out_orig_index) return MCRT_ERR_BAD_ARG;
    std::vector<Prim> prims(n_spheres);
    for (uint32_t i = 0; i < n_spheres; ++i) {
        const float *c = spheres[i].center_r;
        float r = c[3];
        if (!finite3(c) || !(r >= 0.0f) || !std::isfinite(r)) return MCRT_ERR_BAD_ARG;
        Aabb a;  // Aabb::from_sphere, src/bvh.rs:131-140
        for (int k = 0; k < 3; ++k) {
            a.mn[k] = c[k] - r;
            a.mx[k] = c[k] + r;
        }
        prims[i].bounds = a;
        a.centroid(prims[i].c);
        prims[i].idx = i;
    }
    Builder<GpuSphere> b{prims, spheres, out_nodes, 0, out_ordered, out_orig_index, 0};
    b.build(0, n_spheres, -1);
    *out_n_nodes = b.n_nodes;
    return MCRT_OK;
}
