// Checks mcrt_amd/csrc/bvh_refine.h on the CPU (host compiler only).
//
// For seeded scenes (axis-aligned voxel quads of two triangles, random
// triangles, flat and coincident triangles) the reference tree comes from
// this library's builder (mcrt_build_mesh_bvh), is refined, and is checked:
//   - pre-order shape: left child at i + 1, right child after it, every node
//     reached once, leaves disjoint and covering every triangle;
//   - leaves of 3..kRefineMax triangles become leaves of at most 2 (a larger
//     one, counted as "big", only where a split would give two equal boxes:
//     none on voxel quads, rare elsewhere), and every triangle stays inside
//     the range of its given leaf;
//   - no internal node has two children with the same box (a split that
//     separates a voxel face's two triangles only adds a node step);
//   - each leaf box contains its triangles' vertices and lies inside the box
//     of its given leaf, each child box lies in its parent's;
//   - closest hits (lowest (t, original index)) through the refined tree
//     equal those through the given tree, for rays aimed at triangles
//     (centroids, vertices, edge points) and random rays; and where the given
//     tree's hit agrees with brute force over all triangles (it need not for
//     a ray through a box corner, where the slab test rounds), so does the
//     refined tree's.  Rays aimed exactly at vertices are the hard case: a
//     refined leaf box has planes through vertices that the given box does
//     not.
// Usage: bvh_refine_check SEED [plant|nomargin]; "plant" moves one refined
// leaf box off its triangles, and "nomargin" refines with no margin (new
// planes through vertices); the hit check must catch both.
// Prints "cases C fail F shape S contain K hits H dup D big B leaves L voxbig V"
// (fail counts rays whose refined result differs from the given tree's, dup
// internal nodes with two equal child boxes, big refined leaves of more than
// 2 triangles from given leaves of 3..kRefineMax, voxbig those on voxel
// quads).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../mcrt_amd/csrc/bvh_refine.h"

using namespace mcrt;

namespace {

struct Scene {
    std::vector<GpuVertex> v;
    std::vector<GpuTriangle> t;
    void tri(const float a[3], const float b[3], const float c[3]) {
        const float *p[3] = {a, b, c};
        GpuTriangle g{};
        for (int k = 0; k < 3; ++k) {
            GpuVertex x{};
            for (int j = 0; j < 3; ++j) x.position[j] = p[k][j];
            g.v[k] = (uint32_t)v.size();
            v.push_back(x);
        }
        t.push_back(g);
    }
};

Scene make_scene(int kind, std::mt19937 &rng) {
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    Scene s;
    if (kind == 0) {  // voxel-like: unit quads on a heightfield, tops and sides
        const int n = 24;
        for (int x = 0; x < n; ++x)
            for (int z = 0; z < n; ++z) {
                const float h = std::floor(4.0f * U(rng));
                const float a[3] = {(float)x, h, (float)z}, b[3] = {(float)x + 1, h, (float)z},
                            c[3] = {(float)x + 1, h, (float)z + 1}, d[3] = {(float)x, h, (float)z + 1};
                s.tri(a, b, c);
                s.tri(a, c, d);
                if (U(rng) < 0.5f) {  // a side face
                    const float e[3] = {(float)x, h - 1, (float)z}, f[3] = {(float)x + 1, h - 1, (float)z};
                    s.tri(e, f, b);
                    s.tri(e, b, a);
                }
            }
    } else if (kind == 1) {  // random small triangles in a box
        for (int i = 0; i < 1500; ++i) {
            float p[3][3];
            const float c[3] = {10 * U(rng), 10 * U(rng), 10 * U(rng)};
            for (auto &q : p)
                for (int k = 0; k < 3; ++k) q[k] = c[k] + U(rng) - 0.5f;
            s.tri(p[0], p[1], p[2]);
        }
    } else {  // flat, coincident and sliver triangles
        for (int i = 0; i < 600; ++i) {
            const float y = std::floor(3 * U(rng));
            const float a[3] = {4 * U(rng), y, 4 * U(rng)}, b[3] = {4 * U(rng), y, 4 * U(rng)},
                        c[3] = {4 * U(rng), y, 4 * U(rng)};
            s.tri(a, b, c);
            if (i % 5 == 0) s.tri(a, b, c);                             // coincident copy
            if (i % 7 == 0) {
                const float d[3] = {a[0] + 1e-6f, a[1], a[2] + 3.0f};  // sliver
                s.tri(a, d, b);
            }
        }
    }
    return s;
}

struct Ray {
    float o[3], d[3], inv[3];
};

// Moller-Trumbore, t in [tmin, tmax).
bool mt(const Scene &s, const GpuTriangle &g, const Ray &r, float tmin, float tmax, float &t) {
    const float *p0 = s.v[g.v[0]].position, *p1 = s.v[g.v[1]].position, *p2 = s.v[g.v[2]].position;
    float e1[3], e2[3], sv[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = p1[k] - p0[k];
        e2[k] = p2[k] - p0[k];
        sv[k] = r.o[k] - p0[k];
    }
    const float h[3] = {r.d[1] * e2[2] - r.d[2] * e2[1], r.d[2] * e2[0] - r.d[0] * e2[2],
                        r.d[0] * e2[1] - r.d[1] * e2[0]};
    const float a = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
    if (std::fabs(a) < 1e-8f) return false;
    const float f = 1.0f / a;
    const float u = f * (sv[0] * h[0] + sv[1] * h[1] + sv[2] * h[2]);
    if (u < 0.0f || u > 1.0f) return false;
    const float q[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2],
                        sv[0] * e1[1] - sv[1] * e1[0]};
    const float v = f * (r.d[0] * q[0] + r.d[1] * q[1] + r.d[2] * q[2]);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]);
    return t >= tmin && t < tmax;
}

bool slab(const GpuBvhNode &n, const Ray &r, float tmin, float tmax, float &entry) {
    float t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        const float a = (n.aabb_min[k] - r.o[k]) * r.inv[k], b = (n.aabb_max[k] - r.o[k]) * r.inv[k];
        t0 = std::fmax(t0, std::fmin(a, b));
        t1 = std::fmin(t1, std::fmax(a, b));
    }
    entry = t0;
    return t0 <= t1;
}

// Closest hit: lowest (t, original index); returns the original index or ~0u.
uint32_t brute(const Scene &s, const std::vector<GpuTriangle> &tris, const std::vector<uint32_t> &orig,
               const Ray &r, float &best) {
    uint32_t id = ~0u;
    best = INFINITY;
    for (size_t i = 0; i < tris.size(); ++i) {
        float t;
        if (mt(s, tris[i], r, 1e-4f, INFINITY, t) && (t < best || (t == best && orig[i] < id))) {
            best = t;
            id = orig[i];
        }
    }
    return id;
}

uint32_t traverse(const Scene &s, const std::vector<GpuBvhNode> &nodes, const std::vector<GpuTriangle> &tris,
                  const std::vector<uint32_t> &orig, const Ray &r, float &best) {
    uint32_t id = ~0u;
    best = INFINITY;
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        // The traversal rule: the slab test passes on the query interval and
        // the entry does not exceed best t * (1 + 2^-10).
        float e;
        if (!slab(nodes[i], r, 1e-4f, INFINITY, e) || !(e <= best * 1.0009765625f)) continue;
        if (nodes[i].prim_count == 0) {
            st.push_back(nodes[i].right_or_offset);
            st.push_back(i + 1);
            continue;
        }
        for (uint32_t k = 0; k < nodes[i].prim_count; ++k) {
            const uint32_t p = nodes[i].right_or_offset + k;
            float t;
            if (mt(s, tris[p], r, 1e-4f, INFINITY, t) &&
                (t < best || (t == best && orig[p] < id))) {
                best = t;
                id = orig[p];
            }
        }
    }
    return id;
}

}  // namespace

int main(int argc, char **argv) {
    const int seed = argc > 1 ? atoi(argv[1]) : 1;
    const bool plant = argc > 2 && !strcmp(argv[2], "plant");
    const bool nomargin = argc > 2 && !strcmp(argv[2], "nomargin");
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    long cases = 0, fail = 0, shape = 0, contain = 0, hits = 0, dup = 0, big = 0, leaves = 0, voxbig = 0;
    for (int kind = 0; kind < 3; ++kind) {
        Scene s = make_scene(kind, rng);
        const uint32_t nt = (uint32_t)s.t.size();
        std::vector<GpuBvhNode> given(2 * (size_t)nt);
        std::vector<GpuTriangle> ord(nt);
        std::vector<uint32_t> orig(nt);
        uint32_t nn = 0;
        if (mcrt_build_mesh_bvh(s.v.data(), (uint32_t)s.v.size(), s.t.data(), nt, given.data(), &nn, ord.data(),
                                orig.data()) != MCRT_OK) {
            printf("build failed\n");
            return 1;
        }
        given.resize(nn);
        std::vector<GpuBvhNode> rn;
        std::vector<uint32_t> perm;
        if (!refine_mesh_tree(given.data(), nn, s.v.data(), ord.data(), nt, rn, perm, nomargin ? -1000 : -18)) {
            printf("scene %d: nothing to refine\n", kind);
            ++shape;
            continue;
        }
        // shape: pre-order, reachability, leaves partition [0, nt)
        std::vector<uint8_t> seen(rn.size(), 0), covered(nt, 0);
        std::vector<uint32_t> todo{0u};
        while (!todo.empty()) {
            const uint32_t i = todo.back();
            todo.pop_back();
            if (i >= rn.size() || seen[i]++) {
                ++shape;
                continue;
            }
            if (rn[i].prim_count == 0) {
                if (rn[i].right_or_offset <= i + 1) ++shape;
                todo.push_back(rn[i].right_or_offset);
                todo.push_back(i + 1);
                const uint32_t r2 = rn[i].right_or_offset;
                if (r2 < rn.size() && !memcmp(rn[i + 1].aabb_min, rn[r2].aabb_min, 12) &&
                    !memcmp(rn[i + 1].aabb_max, rn[r2].aabb_max, 12))
                    ++dup;
                for (const uint32_t c : {i + 1, rn[i].right_or_offset})
                    for (int k = 0; k < 3 && c < rn.size(); ++k)
                        if (rn[c].aabb_min[k] < rn[i].aabb_min[k] || rn[c].aabb_max[k] > rn[i].aabb_max[k])
                            ++contain;
            } else {
                for (uint32_t k = 0; k < rn[i].prim_count; ++k) {
                    const uint32_t p = rn[i].right_or_offset + k;
                    if (p >= nt || covered[p]++) ++shape;
                    if (p >= nt) continue;
                    for (int v = 0; v < 3; ++v)
                        for (int a = 0; a < 3; ++a) {
                            const float x = s.v[ord[perm[p]].v[v]].position[a];
                            if (x < rn[i].aabb_min[a] || x > rn[i].aabb_max[a]) ++contain;
                        }
                }
            }
        }
        for (size_t i = 0; i < rn.size(); ++i) shape += !seen[i];
        for (uint32_t p = 0; p < nt; ++p) shape += !covered[p];
        // leaf sizes and leaf membership
        std::vector<uint32_t> leaf_of(nt, ~0u);
        for (uint32_t i = 0; i < nn; ++i)
            for (uint32_t k = 0; given[i].prim_count && k < given[i].prim_count; ++k)
                leaf_of[given[i].right_or_offset + k] = i;
        for (size_t i = 0; i < rn.size(); ++i) {
            if (rn[i].prim_count == 0) continue;
            const uint32_t g = leaf_of[perm[rn[i].right_or_offset]];
            const uint32_t gc = given[g].prim_count;
            for (int k = 0; k < 3; ++k)
                if (rn[i].aabb_min[k] < given[g].aabb_min[k] || rn[i].aabb_max[k] > given[g].aabb_max[k]) ++contain;
            ++leaves;
            if (gc >= 3 && gc <= kRefineMax && rn[i].prim_count > 2) {
                ++big;
                voxbig += kind == 0;
            }
            for (uint32_t k = 0; k < rn[i].prim_count; ++k)
                if (leaf_of[perm[rn[i].right_or_offset + k]] != g) ++shape;
        }
        std::vector<GpuTriangle> rt(nt);
        std::vector<uint32_t> ro(nt);
        for (uint32_t j = 0; j < nt; ++j) {
            rt[j] = ord[perm[j]];
            ro[j] = orig[perm[j]];
        }
        size_t pl = 0;  // the first leaf
        while (rn[pl].prim_count == 0) ++pl;
        if (plant) {  // move that leaf's box off its triangles
            rn[pl].aabb_min[1] = rn[pl].aabb_max[1] + 1.0f;
            rn[pl].aabb_max[1] += 2.0f;
        }
        // hits
        for (int q = 0; q < 4000; ++q) {
            Ray r;
            const uint32_t tgt = (uint32_t)(U(rng) * nt) % nt;
            const GpuTriangle &g = plant && q % 2 ? rt[rn[pl].right_or_offset] : s.t[tgt];
            float aim[3];
            const int mode = q % 4;
            const float w0 = U(rng), w1 = U(rng) * (1 - w0);
            for (int k = 0; k < 3; ++k) {
                const float *p0 = s.v[g.v[0]].position, *p1 = s.v[g.v[1]].position, *p2 = s.v[g.v[2]].position;
                if (mode == 0) aim[k] = (p0[k] + p1[k] + p2[k]) / 3.0f;
                else if (mode == 1) aim[k] = p0[k];
                else if (mode == 2) aim[k] = p0[k] + w0 * (p1[k] - p0[k]);
                else aim[k] = p0[k] + w0 * (p1[k] - p0[k]) + w1 * (p2[k] - p0[k]);
            }
            for (int k = 0; k < 3; ++k) r.o[k] = aim[k] + 20.0f * (U(rng) - 0.5f) + (k == 1 ? 12.0f : 0.0f);
            for (int k = 0; k < 3; ++k) r.d[k] = aim[k] - r.o[k];
            if (q % 16 == 15)
                for (int k = 0; k < 3; ++k) r.d[k] = U(rng) - 0.5f;
            if (q % 32 == 7) r.d[0] = 0.0f;  // axis-parallel component
            for (int k = 0; k < 3; ++k) r.inv[k] = 1.0f / r.d[k];
            float tb, tt;
            const uint32_t a = brute(s, rt, ro, r, tb);
            const uint32_t b = traverse(s, rn, rt, ro, r, tt);
            float tg;
            const uint32_t c = traverse(s, given, ord, orig, r, tg);
            ++cases;
            hits += a != ~0u;
            if (b != c || (c != ~0u && tt != tg)) ++fail;
        }
    }
    printf("cases %ld fail %ld shape %ld contain %ld hits %ld dup %ld big %ld leaves %ld voxbig %ld\n", cases, fail,
           shape, contain, hits, dup, big, leaves, voxbig);
    return 0;
}
