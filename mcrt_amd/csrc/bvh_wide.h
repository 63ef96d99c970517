// bvh_wide.h — 4-wide quantized BVH nodes (48 B) for the culling-stack
// traversal of scenes whose traversal set spills the L2 (host side; the
// device decoder is dev_common.h node_step_wide).
//
// A wide node replaces a binary internal node and part of the subtree below
// it: starting from its two children, the interior child with the largest
// surface area is replaced by its own two children until the node holds 4
// children or only leaves.  One node step then tests up to 4 boxes from
// three 16-B loads, where the binary pairs take two loads per 2 boxes and
// one more node step per level; the traversal's load instructions per node
// step are what bound it on these scenes.
//
// Layout, 12 words:
//   w0..w2   per axis the grid word of node_quant.h (origin bits, step exponent);
//   w3..w8   24 plane bytes, child k at bytes 6k..6k+5 (lo.x, hi.x, lo.y,
//            hi.y, lo.z, hi.z), 4 to a word, low byte first;
//   w9       base_node: wide index of the first interior child;
//   w10      base_leaf: record index of the first leaf child;
//   w11      one meta byte per child k (byte k):
//              0                     no child (its planes are all byte 0:
//                                    a point box, which a slab test never
//                                    passes; the kernel also masks it)
//              0x10 | j              interior child: wide node base_node + j
//              0x80 | quad << 6 | j  leaf child at record base_leaf + j
//                                    (j < 64; quad = the kLeafPair flag)
// The interior children of a node take consecutive wide indices, and the
// leaf records are reordered at upload so that the leaves of a node follow
// one another; the planes are the binary pairs' boxes on a shared grid, so
// they contain them, and the traversal visits a superset of the binary
// traversal's leaves (closest-hit results are decided in the leaves).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <vector>

#include "node_quant.h"

namespace mcrt {

constexpr uint32_t kWideWords = 12;  // 48 B per wide node
constexpr uint32_t kWideLeafOff = 63;

// The binary child pairs, 4 float4 per internal node:
//   q0 = child 0 (lo.x, hi.x, lo.y, hi.y), q1 = child 1 the same,
//   q2 = (lo.z, hi.z of child 0, lo.z, hi.z of child 1), q3 = (ref0, ref1) bits.
// A child slot is pair index * 2 + side.
struct WidePairs {
    const float4 *q;
    size_t n_pairs;
    uint32_t leaf_bit, pair_bit, none;
    uint32_t ref(uint32_t slot) const {
        uint32_t r;
        const float4 &w = q[4 * (size_t)(slot >> 1) + 3];
        std::memcpy(&r, (slot & 1u) ? &w.y : &w.x, 4);
        return r;
    }
    void box(uint32_t slot, float lo[3], float hi[3]) const {
        const float4 *p = q + 4 * (size_t)(slot >> 1);
        const float4 &xy = p[slot & 1u];
        lo[0] = xy.x;
        hi[0] = xy.y;
        lo[1] = xy.z;
        hi[1] = xy.w;
        lo[2] = (slot & 1u) ? p[2].z : p[2].x;
        hi[2] = (slot & 1u) ? p[2].w : p[2].y;
    }
    bool interior(uint32_t r) const { return r != none && !(r & leaf_bit); }
};

struct WidePlan {
    struct Node {
        uint32_t n = 0;
        uint32_t slot[4] = {0, 0, 0, 0};  // binary child slots
    };
    std::vector<Node> nodes;             // by wide index
    std::vector<uint32_t> src;           // the binary pair each wide node replaces
    std::vector<uint32_t> first_child;   // wide index of a node's first interior child
    std::vector<uint32_t> visit;         // wide indices in depth-first visit order
    uint32_t stack_need = 0;             // most stack entries a traversal holds
};

inline float wide_half_area(const float lo[3], const float hi[3]) {
    const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    return x * y + y * z + z * x;
}

// Groups the binary tree under the internal root into wide nodes.  Which
// binary nodes become wide nodes is chosen to minimise the summed surface
// area of the wide nodes (the expected number of wide-node visits of a ray
// under the surface-area heuristic), each holding at most 4 children, by a
// dynamic program over the binary tree (the collapse of Ylitie, Karras and
// Laine, "Efficient incoherent ray traversal on GPUs through compressed wide
// BVHs", HPG 2017, with leaves kept as they are):
//   f(s, k)  least area of the wide nodes below child slot s when s's subtree
//            takes at most k child slots of its wide parent: 0 for a leaf;
//            for an internal node, area(s) + g(s, 4) as a wide node of its
//            own (one slot), or g(s, k) with its children lifted (k >= 2);
//   g(s, m)  min over a + b = m of f(left, a) + f(right, b).
// Indices are given depth-first: visiting a node gives its interior children
// the next consecutive indices, and its first child's subtree is visited
// before the second's, so a path down the tree reads nearby nodes.  False
// where a reference points outside the pairs or the tree is not a tree.
inline bool wide_plan(const WidePairs &bp, uint32_t root, WidePlan &out) {
    out = WidePlan{};
    if (!bp.interior(root) || root >= bp.n_pairs) return false;
    // Post-order over the internal nodes reachable from the root.
    const size_t np = bp.n_pairs;
    std::vector<uint8_t> seen(np, 0);
    std::vector<uint32_t> post;
    {
        std::vector<std::pair<uint32_t, bool>> st{{root, false}};
        while (!st.empty()) {
            auto [p, done] = st.back();
            st.pop_back();
            if (done) {
                post.push_back(p);
                continue;
            }
            if (seen[p]) return false;
            seen[p] = 1;
            st.push_back({p, true});
            for (uint32_t side = 0; side < 2; ++side) {
                const uint32_t r = bp.ref(2u * p + side);
                if (!bp.interior(r)) continue;
                if (r >= np) return false;
                st.push_back({r, false});
            }
        }
    }
    // f per child slot (k = 1..4), g and its split per pair (m = 2..4),
    // children before parents.
    std::vector<float> f(2 * np * 4, 0.0f), g(np * 3, 0.0f);
    std::vector<uint8_t> split(np * 3, 1);
    auto F = [&](uint32_t slot, uint32_t k) -> float & { return f[(size_t)slot * 4 + (k - 1)]; };
    auto G = [&](uint32_t p, uint32_t m) -> float & { return g[(size_t)p * 3 + (m - 2)]; };
    for (const uint32_t p : post) {
        for (uint32_t side = 0; side < 2; ++side) {
            const uint32_t slot = 2u * p + side, r = bp.ref(slot);
            if (!bp.interior(r)) continue;  // a leaf: f = 0
            float lo[3], hi[3];
            bp.box(slot, lo, hi);
            const float own = std::max(wide_half_area(lo, hi), 0.0f) + G(r, 4);
            F(slot, 1) = own;
            for (uint32_t k = 2; k <= 4; ++k) F(slot, k) = std::min(own, G(r, k));
        }
        for (uint32_t m = 2; m <= 4; ++m) {
            uint32_t arg = 1;
            float best = F(2u * p, 1) + F(2u * p + 1u, m - 1);
            for (uint32_t a = 2; a < m; ++a) {
                const float c = F(2u * p, a) + F(2u * p + 1u, m - a);
                if (c < best) {
                    best = c;
                    arg = a;
                }
            }
            G(p, m) = best;
            split[(size_t)p * 3 + (m - 2)] = (uint8_t)arg;
        }
    }
    // The child slots of a wide node over pair p with m slots to fill: a
    // child slot stays a slot (a leaf, or a wide node of its own) or is
    // replaced by its children's slots, as the program chose.
    auto gather = [&](uint32_t p, uint32_t m, WidePlan::Node &nd) {
        std::vector<std::pair<uint32_t, uint32_t>> todo{{p, m}};  // (pair, slots)
        while (!todo.empty()) {
            auto [q, k] = todo.back();
            todo.pop_back();
            const uint32_t a = split[(size_t)q * 3 + (k - 2)];
            for (uint32_t side = 0; side < 2; ++side) {
                const uint32_t slot = 2u * q + side, kk = side ? k - a : a;
                const uint32_t r = bp.ref(slot);
                if (!bp.interior(r) || kk == 1 || F(slot, 1) <= G(r, kk))
                    nd.slot[nd.n++] = slot;
                else
                    todo.push_back({r, kk});
            }
        }
    };
    std::vector<uint32_t> need{0u};  // stack entries held on entering each node
    out.src.push_back(root);
    out.nodes.resize(1);
    out.first_child.resize(1);
    std::vector<uint32_t> todo{0u};
    while (!todo.empty()) {
        const uint32_t w = todo.back();
        todo.pop_back();
        out.visit.push_back(w);
        WidePlan::Node nd;
        gather(out.src[w], 4, nd);
        const uint32_t held = need[w] + nd.n - 1u;  // a node step pushes all but the nearest hit child
        out.stack_need = std::max(out.stack_need, held);
        const uint32_t first = (uint32_t)out.src.size();
        out.first_child[w] = first;
        for (uint32_t k = 0; k < nd.n; ++k) {
            const uint32_t r = bp.ref(nd.slot[k]);
            if (!bp.interior(r)) continue;
            out.src.push_back(r);
            need.push_back(held);
        }
        const uint32_t end = (uint32_t)out.src.size();
        out.nodes.resize(end);
        out.first_child.resize(end);
        out.nodes[w] = nd;
        for (uint32_t c = end; c > first; --c) todo.push_back(c - 1u);
    }
    out.stack_need = std::max(out.stack_need, 1u);
    return true;
}

// Leaf-record order in which the leaves of every wide node follow one
// another (nodes in depth-first visit order, a node's leaves in slot order),
// then the records no leaf references.  perm[new] = old; start[old leaf start] = new
// start (kNone elsewhere).  last[i] != 0 marks the last record of a leaf.
// False where a leaf runs past the records or two leaves share a record.
inline bool wide_leaf_order(const WidePairs &bp, const WidePlan &p, const uint8_t *last, uint32_t n_prims,
                            std::vector<uint32_t> &perm, std::vector<uint32_t> &start) {
    perm.clear();
    perm.reserve(n_prims);
    start.assign(n_prims, bp.none);
    std::vector<uint8_t> placed(n_prims, 0);
    for (const uint32_t w : p.visit)
        for (uint32_t k = 0, n = p.nodes[w].n; k < n; ++k) {
            const WidePlan::Node &nd = p.nodes[w];
            const uint32_t r = bp.ref(nd.slot[k]);
            if (r == bp.none || !(r & bp.leaf_bit)) continue;
            const uint32_t s = r & ~(bp.leaf_bit | bp.pair_bit);
            if (s >= n_prims || start[s] != bp.none) return false;
            start[s] = (uint32_t)perm.size();
            for (uint32_t i = s;; ++i) {
                if (i >= n_prims || placed[i]) return false;
                placed[i] = 1;
                perm.push_back(i);
                if (last[i]) break;
            }
        }
    for (uint32_t i = 0; i < n_prims; ++i)
        if (!placed[i]) perm.push_back(i);
    return true;
}

// The wide nodes of a plan over binary pairs whose leaf references already
// follow the order of wide_leaf_order (12 words each).  False where a box
// cannot be encoded or a leaf lies beyond kWideLeafOff records from its
// node's first leaf.
inline bool wide_emit(const WidePairs &bp, const WidePlan &p, std::vector<uint32_t> &words) {
    words.assign(p.nodes.size() * kWideWords, 0u);
    for (size_t w = 0; w < p.nodes.size(); ++w) {
        const WidePlan::Node &nd = p.nodes[w];
        float lo[4][3], hi[4][3];
        uint32_t bytes[24] = {};
        uint32_t meta = 0, base_leaf = bp.none, n_int = 0;
        for (uint32_t k = 0; k < nd.n; ++k) {
            bp.box(nd.slot[k], lo[k], hi[k]);
            const uint32_t r = bp.ref(nd.slot[k]);
            if (r == bp.none) return false;
            if (r & bp.leaf_bit) {
                const uint32_t s = r & ~(bp.leaf_bit | bp.pair_bit);
                if (base_leaf == bp.none) base_leaf = s;
                if (s < base_leaf || s - base_leaf > kWideLeafOff) return false;
                meta |= (0x80u | ((r & bp.pair_bit) ? 0x40u : 0u) | (s - base_leaf)) << (8 * k);
            } else {
                meta |= (0x10u | n_int++) << (8 * k);
            }
        }
        uint32_t *o = &words[w * kWideWords];
        if (!qgrid_encode(lo, hi, (int)nd.n, o, bytes)) return false;
        for (uint32_t k = nd.n; k < 4u; ++k)
            for (int b = 0; b < 6; ++b) bytes[6 * k + b] = 0u;
        for (int j = 0; j < 6; ++j)
            o[3 + j] = bytes[4 * j] | (bytes[4 * j + 1] << 8) | (bytes[4 * j + 2] << 16) | (bytes[4 * j + 3] << 24);
        o[9] = n_int ? p.first_child[w] : 0u;
        o[10] = base_leaf == bp.none ? 0u : base_leaf;
        o[11] = meta;
    }
    return true;
}

// Decodes child k (0..3) of a wide node: lo[3], hi[3].
__host__ __device__ inline void wide_decode(const uint32_t w[kWideWords], int k, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        const float st = qnode_step(w[a]), o = qnode_origin(w[a]);
        const int b = 6 * k + 2 * a;
        lo[a] = qnode_plane(w[3 + b / 4], b % 4, st, o);
        hi[a] = qnode_plane(w[3 + (b + 1) / 4], (b + 1) % 4, st, o);
    }
}

}  // namespace mcrt
