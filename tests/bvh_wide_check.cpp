// Checks mcrt_amd/csrc/bvh_wide.h on the CPU (built with hipcc as a host
// program: the header shares node_quant.h with the device code).
//
// For seeded scenes (voxel quads on a heightfield, random triangles, flat and
// coincident triangles) the tree comes from this library's builder, is
// refined (bvh_refine.h) and converted to binary child pairs with shuffled
// pair indices and the 2-triangle flag on 2-triangle leaves, as at upload.
// The 4-wide plan, the leaf-record order and the 48-B nodes are then built,
// and checked:
//   - the record order is a permutation, every binary leaf reference points
//     at its own records in the new order, and the leaves of a wide node
//     follow one another;
//   - every wide child's decoded box contains the binary box it came from,
//     its reference (wide index, or record and 2-triangle flag) is the
//     binary one, and unused slots carry meta 0;
//   - closest hits (lowest (t, original index)) of an exact emulation of the
//     kernel's wide node step (dev_common.h node_step_wide: decode, slab
//     test, cull limit, sort on packed keys, pushes farthest first, culling
//     pops) equal those
//     of the binary traversal over the exact pairs with the kernel's rule,
//     where the binary traversal agrees with brute force over all triangles
//     (it need not for a ray through a box corner, where the slab test
//     rounds); and the stack never holds more than the plan's bound.
// Usage: bvh_wide_check SEED [plant]; "plant" moves one leaf child's record
// offset by one in its wide node, which the hit check must catch.
// Prints "cases C fail F diff D diffq Q order O box B ref R stack S nodes N full U"
// (fail: wide != binary where binary == brute force; diff: wide != binary;
// diffq: the quantized 32-B binary pairs != the exact pairs, for comparison;
// order / box / ref: structural faults; stack: rays whose stack exceeded the
// bound; full: wide nodes with 4 children).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../mcrt_amd/csrc/bvh_refine.h"
#include "../mcrt_amd/csrc/bvh_wide.h"

using namespace mcrt;

namespace {

constexpr uint32_t kLeaf = 0x80000000u, kPair = 0x40000000u, kNo = 0xFFFFFFFFu;
constexpr float kSlack = 1.0009765625f;

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
    if (kind == 0) {  // voxel faces: a heightfield's tops and some sides, two triangles each
        const int n = 40;
        for (int x = 0; x < n; ++x)
            for (int z = 0; z < n; ++z) {
                const float h = std::floor(6.0f * U(rng));
                const float a[3] = {(float)x, h, (float)z}, b[3] = {(float)x + 1, h, (float)z},
                            c[3] = {(float)x + 1, h, (float)z + 1}, d[3] = {(float)x, h, (float)z + 1};
                s.tri(a, b, c);
                s.tri(a, c, d);
                if (U(rng) < 0.5f) {
                    const float e[3] = {(float)x, h - 1, (float)z}, f[3] = {(float)x + 1, h - 1, (float)z};
                    s.tri(e, f, b);
                    s.tri(e, b, a);
                }
            }
    } else if (kind == 1) {  // random small triangles
        for (int i = 0; i < 3000; ++i) {
            float p[3][3];
            const float c[3] = {10 * U(rng), 10 * U(rng), 10 * U(rng)};
            for (auto &q : p)
                for (int k = 0; k < 3; ++k) q[k] = c[k] + U(rng) - 0.5f;
            s.tri(p[0], p[1], p[2]);
        }
    } else {  // flat, coincident and sliver triangles
        for (int i = 0; i < 1200; ++i) {
            const float y = std::floor(3 * U(rng));
            const float a[3] = {4 * U(rng), y, 4 * U(rng)}, b[3] = {4 * U(rng), y, 4 * U(rng)},
                        c[3] = {4 * U(rng), y, 4 * U(rng)};
            s.tri(a, b, c);
            if (i % 5 == 0) s.tri(a, b, c);
            if (i % 7 == 0) {
                const float d[3] = {a[0] + 1e-6f, a[1], a[2] + 3.0f};
                s.tri(a, d, b);
            }
        }
    }
    return s;
}

struct Ray {
    float o[3], d[3], inv[3];
};

struct Hit {
    float t = 0.0f;
    uint32_t id = kNo, prim = kNo;
};

// Moller-Trumbore with precomputed FP32 edges, t in [tmin, tmax).
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

// The kernel's slab test (dev_common.h slab2).
void slab(const float lo[3], const float hi[3], const Ray &r, float &te, float &tx) {
    float a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        const float t0 = (lo[k] - r.o[k]) * r.inv[k], t1 = (hi[k] - r.o[k]) * r.inv[k];
        a[k] = std::fmin(t0, t1);
        b[k] = std::fmax(t0, t1);
    }
    te = std::fmax(std::fmax(a[0], a[1]), a[2]);
    tx = std::fmin(std::fmin(b[0], b[1]), b[2]);
}

float next_up(float lim) {
    uint32_t u;
    std::memcpy(&u, &lim, 4);
    ++u;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Tris {
    const Scene *s;
    std::vector<GpuTriangle> t;
    std::vector<uint32_t> orig;
    std::vector<uint8_t> last;
    void leaf(uint32_t i, const Ray &r, float tmin, float tmax, Hit &h) const {
        for (;; ++i) {
            float t;
            if (mt(*s, this->t[i], r, tmin, tmax, t) && (h.prim == kNo || t < h.t || (t == h.t && orig[i] < h.id))) {
                h.t = t;
                h.id = orig[i];
                h.prim = i;
            }
            if (last[i]) break;
        }
    }
};

const float kTmin = 1e-4f, kTmax = 1e30f;
long g_empty_pass = 0;

bool root_pass(const float lo[3], const float hi[3], const Ray &r) {
    float te, tx;
    slab(lo, hi, r, te, tx);
    return tx > std::fmax(te, kTmin) && te < kTmax;
}

// Binary traversal with the kernel's rule (node_step), over the exact pairs,
// or (qw non-empty) over their quantized 32-B form (node_quant.h).
Hit binary(const WidePairs &bp, uint32_t root, const float rlo[3], const float rhi[3], const Tris &T, const Ray &r,
           const std::vector<uint32_t> &qw = {}) {
    Hit h;
    h.t = kTmax;
    if (!root_pass(rlo, rhi, r)) return h;
    std::vector<std::pair<uint32_t, float>> st;
    uint32_t cur = root;
    while (cur != kNo) {
        if (cur & kLeaf) {
            T.leaf(cur & ~(kLeaf | kPair), r, kTmin, kTmax, h);
            cur = kNo;
        } else {
            const float lim = h.t * kSlack, bound = std::fmin(kTmax, next_up(lim));
            float lo[2][3], hi[2][3], te[2], tx[2];
            bool hit[2];
            for (int k = 0; k < 2; ++k) {
                if (qw.empty())
                    bp.box(2 * cur + k, lo[k], hi[k]);
                else
                    qnode_decode(&qw[(size_t)cur * kQuantWords], k, lo[k], hi[k]);
                slab(lo[k], hi[k], r, te[k], tx[k]);
                hit[k] = tx[k] > std::fmax(te[k], kTmin) && te[k] < bound;
            }
            const uint32_t c0 = bp.ref(2 * cur), c1 = bp.ref(2 * cur + 1);
            if (hit[0] && hit[1]) {
                const bool sw = te[1] < te[0];
                st.push_back({sw ? c0 : c1, sw ? te[0] : te[1]});
                cur = sw ? c1 : c0;
                continue;
            }
            if (hit[0] || hit[1]) {
                cur = hit[0] ? c0 : c1;
                continue;
            }
            cur = kNo;
        }
        const float lim = h.t * kSlack;
        while (!st.empty()) {
            auto e = st.back();
            st.pop_back();
            if (!(e.second > lim)) {
                cur = e.first;
                break;
            }
        }
    }
    return h;
}

// Emulation of the kernel's wide node step over the 48-B nodes.
Hit wide(const std::vector<uint32_t> &w, const float rlo[3], const float rhi[3], const Tris &T, const Ray &r,
         size_t &max_sp) {
    Hit h;
    h.t = kTmax;
    if (!root_pass(rlo, rhi, r)) return h;
    std::vector<std::pair<uint32_t, float>> st;
    uint32_t cur = 0;
    const uint32_t n_wide = (uint32_t)(w.size() / kWideWords);
    while (cur != kNo) {
        if (cur & kLeaf) {
            const uint32_t i = cur & ~(kLeaf | kPair);
            if (i >= T.t.size()) {
                h.id = kNo - 1;  // a record outside the array: reported as a failure
                return h;
            }
            T.leaf(i, r, kTmin, kTmax, h);
            cur = kNo;
        } else {
            if (cur >= n_wide) {
                h.id = kNo - 1;
                return h;
            }
            const uint32_t *n = &w[(size_t)cur * kWideWords];
            const float lim = h.t * kSlack, bound = std::fmin(kTmax, next_up(lim));
            uint32_t key[4];
            for (uint32_t k = 0; k < 4; ++k) {
                float lo[3], hi[3], te, tx;
                wide_decode(n, (int)k, lo, hi);
                slab(lo, hi, r, te, tx);
                const uint32_t m = (n[11] >> (8 * k)) & 0xFFu;
                const bool pass = tx > std::fmax(te, kTmin) && te < bound;
                if (pass && m == 0u) g_empty_pass++;  // an unused slot must never pass
                const float c = std::fmax(std::fmax(te, kTmin), 0.0f);
                uint32_t bits;
                std::memcpy(&bits, &c, 4);
                key[k] = pass ? ((bits & ~3u) | k) : 0xFFFFFFFFu;
            }
            auto cas = [&](int a, int b) {
                const uint32_t lo = std::min(key[a], key[b]), hi = std::max(key[a], key[b]);
                key[a] = lo;
                key[b] = hi;
            };
            cas(0, 1);
            cas(2, 3);
            cas(0, 2);
            cas(1, 3);
            cas(1, 2);
            auto ref = [&](uint32_t kk) {
                const uint32_t m = (n[11] >> (8 * (kk & 3u))) & 0xFFu;
                return (m & 0x80u) ? (kLeaf | ((m & 0x40u) << 24) | (n[10] + (m & 0x3Fu))) : (n[9] + (m & 3u));
            };
            auto dist = [](uint32_t kk) {
                const uint32_t b = kk & ~3u;
                float f;
                std::memcpy(&f, &b, 4);
                return f;
            };
            if (key[0] != 0xFFFFFFFFu) {
                for (int k = 3; k >= 1; --k)
                    if (key[k] != 0xFFFFFFFFu) st.push_back({ref(key[k]), dist(key[k])});
                max_sp = std::max(max_sp, st.size());
                cur = ref(key[0]);
                continue;
            }
            cur = kNo;
        }
        const float lim = h.t * kSlack;
        while (!st.empty()) {
            auto e = st.back();
            st.pop_back();
            if (!(e.second > lim)) {
                cur = e.first;
                break;
            }
        }
    }
    return h;
}

Hit brute(const Tris &T, const Ray &r) {
    Hit h;
    h.t = kTmax;
    for (uint32_t i = 0; i < T.t.size(); ++i) {
        float t;
        if (mt(*T.s, T.t[i], r, kTmin, kTmax, t) && (h.prim == kNo || t < h.t || (t == h.t && T.orig[i] < h.id))) {
            h.t = t;
            h.id = T.orig[i];
            h.prim = i;
        }
    }
    return h;
}

bool same(const Hit &a, const Hit &b) {
    return a.id == b.id && (a.id == kNo || std::memcmp(&a.t, &b.t, 4) == 0);
}

}  // namespace

int main(int argc, char **argv) {
    const int seed = argc > 1 ? atoi(argv[1]) : 1;
    const bool plant = argc > 2 && !strcmp(argv[2], "plant");
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    long cases = 0, fail = 0, diff = 0, diffq = 0, order = 0, boxf = 0, reff = 0, stackf = 0, nodes = 0, full = 0;
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
        if (!refine_mesh_tree(given.data(), nn, s.v.data(), ord.data(), nt, rn, perm)) {
            rn = given;
            perm.resize(nt);
            for (uint32_t i = 0; i < nt; ++i) perm[i] = i;
        }
        Tris T;
        T.s = &s;
        T.t.resize(nt);
        T.orig.resize(nt);
        T.last.assign(nt, 0);
        for (uint32_t j = 0; j < nt; ++j) {
            T.t[j] = ord[perm[j]];
            T.orig[j] = orig[perm[j]];
        }
        // binary pairs, internal nodes numbered in a shuffled order
        std::vector<uint32_t> internal, pidx(rn.size(), kNo);
        for (uint32_t i = 0; i < rn.size(); ++i)
            if (rn[i].prim_count == 0) internal.push_back(i);
            else T.last[rn[i].right_or_offset + rn[i].prim_count - 1] = 1;
        std::vector<uint32_t> order_ix(internal.size());
        for (uint32_t i = 0; i < order_ix.size(); ++i) order_ix[i] = i;
        std::shuffle(order_ix.begin(), order_ix.end(), rng);
        for (uint32_t i = 0; i < internal.size(); ++i) pidx[internal[i]] = order_ix[i];
        auto bref = [&](uint32_t i) -> uint32_t {
            if (rn[i].prim_count == 0) return pidx[i];
            return kLeaf | (rn[i].prim_count == 2 ? kPair : 0u) | rn[i].right_or_offset;
        };
        std::vector<float4> pairs(4 * internal.size());
        for (uint32_t i : internal) {
            const GpuBvhNode &a = rn[i + 1], &b = rn[rn[i].right_or_offset];
            float4 *q = &pairs[4 * (size_t)pidx[i]];
            q[0] = make_float4(a.aabb_min[0], a.aabb_max[0], a.aabb_min[1], a.aabb_max[1]);
            q[1] = make_float4(b.aabb_min[0], b.aabb_max[0], b.aabb_min[1], b.aabb_max[1]);
            q[2] = make_float4(a.aabb_min[2], a.aabb_max[2], b.aabb_min[2], b.aabb_max[2]);
            uint32_t r0 = bref(i + 1), r1 = bref(rn[i].right_or_offset);
            float f0, f1;
            std::memcpy(&f0, &r0, 4);
            std::memcpy(&f1, &r1, 4);
            q[3] = make_float4(f0, f1, 0.0f, 0.0f);
        }
        const uint32_t root = bref(0);
        const float *rlo = rn[0].aabb_min, *rhi = rn[0].aabb_max;
        WidePairs bp{pairs.data(), internal.size(), kLeaf, kPair, kNo};
        // leaf records before the reorder, for the membership check
        std::vector<std::vector<uint32_t>> before;  // per binary slot with a leaf: original ids
        std::vector<uint32_t> leaf_slots;
        for (uint32_t sl = 0; sl < 2 * internal.size(); ++sl) {
            const uint32_t r = bp.ref(sl);
            if (!(r & kLeaf)) continue;
            std::vector<uint32_t> ids;
            for (uint32_t i = r & ~(kLeaf | kPair);; ++i) {
                ids.push_back(T.orig[i]);
                if (T.last[i]) break;
            }
            before.push_back(ids);
            leaf_slots.push_back(sl);
        }
        WidePlan plan;
        std::vector<uint32_t> wperm, wstart, words;
        if (!wide_plan(bp, root, plan) || !wide_leaf_order(bp, plan, T.last.data(), nt, wperm, wstart)) {
            printf("scene %d: no plan\n", kind);
            ++order;
            continue;
        }
        // apply the order
        {
            std::vector<uint8_t> seen(nt, 0);
            for (uint32_t j = 0; j < nt; ++j)
                if (wperm[j] >= nt || seen[wperm[j]]++) ++order;
            if (wperm.size() != nt) ++order;
            Tris W = T;
            for (uint32_t j = 0; j < nt; ++j) {
                W.t[j] = T.t[wperm[j]];
                W.orig[j] = T.orig[wperm[j]];
                W.last[j] = T.last[wperm[j]];
            }
            for (size_t i = 0; i < internal.size(); ++i) {
                float4 &q = pairs[4 * i + 3];
                for (float *f : {&q.x, &q.y}) {
                    uint32_t r;
                    std::memcpy(&r, f, 4);
                    if (!(r & kLeaf)) continue;
                    const uint32_t ns = wstart[r & ~(kLeaf | kPair)];
                    if (ns == kNo) {
                        ++order;
                        continue;
                    }
                    r = (r & (kLeaf | kPair)) | ns;
                    std::memcpy(f, &r, 4);
                }
            }
            T = W;
        }
        for (size_t k = 0; k < leaf_slots.size(); ++k) {  // each leaf still holds its own records
            const uint32_t r = bp.ref(leaf_slots[k]);
            uint32_t i = r & ~(kLeaf | kPair);
            for (size_t j = 0; j < before[k].size(); ++j, ++i)
                if (i >= nt || T.orig[i] != before[k][j] || (bool)T.last[i] != (j + 1 == before[k].size())) {
                    ++order;
                    break;
                }
        }
        if (!wide_emit(bp, plan, words)) {
            printf("scene %d: emit failed\n", kind);
            ++order;
            continue;
        }
        nodes += (long)plan.nodes.size();
        std::vector<uint32_t> qpairs(internal.size() * kQuantWords);  // the 32-B pairs, for comparison
        for (size_t i = 0; i < internal.size() && !qpairs.empty(); ++i) {
            float lo0[3], hi0[3], lo1[3], hi1[3];
            bp.box(2 * (uint32_t)i, lo0, hi0);
            bp.box(2 * (uint32_t)i + 1, lo1, hi1);
            if (!qnode_encode(lo0, hi0, lo1, hi1, bp.ref(2 * (uint32_t)i), bp.ref(2 * (uint32_t)i + 1),
                              &qpairs[i * kQuantWords]))
                qpairs.clear();
        }
        // structure: boxes contain the binary boxes, references match
        for (size_t w = 0; w < plan.nodes.size(); ++w) {
            const WidePlan::Node &nd = plan.nodes[w];
            const uint32_t *n = &words[w * kWideWords];
            full += nd.n == 4u;
            uint32_t leaf_end = kNo;
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t m = (n[11] >> (8 * k)) & 0xFFu;
                if (k >= nd.n) {
                    reff += m != 0u;
                    continue;
                }
                float lo[3], hi[3], blo[3], bhi[3];
                wide_decode(n, (int)k, lo, hi);
                bp.box(nd.slot[k], blo, bhi);
                for (int a = 0; a < 3; ++a)
                    if (!(lo[a] <= blo[a]) || !(hi[a] >= bhi[a])) ++boxf;
                const uint32_t r = bp.ref(nd.slot[k]);
                const uint32_t got =
                    (m & 0x80u) ? (kLeaf | ((m & 0x40u) << 24) | (n[10] + (m & 0x3Fu))) : (n[9] + (m & 3u));
                if (m == 0u) ++reff;
                if (r & kLeaf) {
                    if (got != r) ++reff;
                    // leaves of a node follow one another
                    const uint32_t st0 = r & ~(kLeaf | kPair);
                    if (leaf_end != kNo && st0 != leaf_end) ++order;
                    uint32_t e = st0;
                    while (!T.last[e]) ++e;
                    leaf_end = e + 1;
                } else if (got >= plan.nodes.size() || plan.src[got] != r) {
                    ++reff;
                }
            }
        }
        uint32_t planted = kNo;  // the record whose leaf child was moved
        if (plant) {  // one leaf child's record offset moved by one
            for (size_t w = 0; w < plan.nodes.size(); ++w) {
                uint32_t *n = &words[w * kWideWords];
                bool done = false;
                for (int k = 0; k < 4 && !done; ++k) {
                    const uint32_t m = (n[11] >> (8 * k)) & 0xFFu;
                    if ((m & 0x80u) && (m & 0x3Fu) == 0u && n[10] + 1u < nt) {
                        n[11] += 1u << (8 * k);
                        planted = n[10];
                        done = true;
                    }
                }
                if (done) break;
            }
        }
        // hits
        for (int q = 0; q < 4000; ++q) {
            Ray r;
            const GpuTriangle &g = planted != kNo && q % 2 ? T.t[planted] : s.t[(uint32_t)(U(rng) * nt) % nt];
            float aim[3];
            const int mode = q % 5;
            const float w0 = U(rng), w1 = U(rng) * (1 - w0);
            for (int k = 0; k < 3; ++k) {
                const float *p0 = s.v[g.v[0]].position, *p1 = s.v[g.v[1]].position, *p2 = s.v[g.v[2]].position;
                if (mode == 0) aim[k] = (p0[k] + p1[k] + p2[k]) / 3.0f;
                else if (mode == 1) aim[k] = p0[k];
                else if (mode == 2) aim[k] = p0[k] + w0 * (p1[k] - p0[k]);
                else aim[k] = p0[k] + w0 * (p1[k] - p0[k]) + w1 * (p2[k] - p0[k]);
            }
            for (int k = 0; k < 3; ++k) r.o[k] = aim[k] + 30.0f * (U(rng) - 0.5f) + (k == 1 ? 15.0f : 0.0f);
            for (int k = 0; k < 3; ++k) r.d[k] = aim[k] - r.o[k];
            if (mode == 4) {  // an axis-parallel ray (zero direction components)
                const int ax = (int)(U(rng) * 3) % 3;
                for (int k = 0; k < 3; ++k) r.d[k] = k == ax ? (U(rng) < 0.5f ? -1.0f : 1.0f) : 0.0f;
                for (int k = 0; k < 3; ++k) r.o[k] = aim[k] - 20.0f * r.d[k];
            }
            for (int k = 0; k < 3; ++k) r.inv[k] = 1.0f / r.d[k];
            ++cases;
            size_t max_sp = 0;
            const Hit hb = binary(bp, root, rlo, rhi, T, r);
            const Hit hw = wide(words, rlo, rhi, T, r, max_sp);
            if (!qpairs.empty() && !same(hb, binary(bp, root, rlo, rhi, T, r, qpairs))) ++diffq;
            if (max_sp > plan.stack_need) ++stackf;
            if (!same(hb, hw)) {
                ++diff;
                if (same(hb, brute(T, r))) ++fail;
            }
        }
    }
    printf("cases %ld fail %ld diff %ld diffq %ld order %ld box %ld ref %ld stack %ld nodes %ld full %ld empty %ld\n",
           cases, fail, diff, diffq, order, boxf, reff, stackf, nodes, full, g_empty_pass);
    return 0;
}
