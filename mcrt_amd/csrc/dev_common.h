// dev_common.h — device-side scene layout, arithmetic and the persistent
// closest-hit traversal shared by the render, trace and adjoint kernels of
// libmcrt_hip.so (gfx950, wave64).
//
// The arithmetic restates src/shaders/path_trace.wgsl operation by operation
// (the file is compiled with -ffp-contract=off, so nothing is fused), which
// makes intersection results bit-identical to a CPU evaluation of the same
// expressions.  Vector helpers evaluate left to right: dot = x*x' + y*y' +
// z*z', normalize(v) = v * (1 / sqrt(dot(v, v))).  Division and sqrt are the
// correctly rounded HIP defaults.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mcrt_abi.h"
#include "node_quant.h"

namespace mcrt {

// ---------------------------------------------------------------------------
// Device scene layout
// ---------------------------------------------------------------------------
// BVH nodes are stored as child pairs: one 64-B node per internal node of the
// reference pre-order tree, holding both children's boxes, so one node fetch
// slab-tests two boxes and orders them.
//   q0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   q1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   q2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   q3 = (ref0, ref1, -, -) as uint bits
// Scenes whose traversal set spills the L2 (DevScene::stack_cull) store the
// same pairs quantized to 32 B instead (node_quant.h): half the bytes per
// node step, boxes that contain the reference boxes.
// A child reference is a device node index, or kLeafBit | first primitive
// index in leaf order.  A leaf's primitives run until the one whose
// kPrimLast flag is set; in the mesh tree of a scene on the culling-stack
// kernels, kLeafPair also marks a quad leaf (tri_leaf): two triangles that
// share a vertex and an edge, both held in the first 64-B record.
constexpr uint32_t kNodeCache = 64;  // 64-B child pairs of the top levels in LDS (4 KB per block)
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafPair = 0x40000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kPrimLast = 1u;
constexpr uint32_t kPrimCull = 2u;
constexpr uint32_t kQuadSwap = 4u;  // quad record: the leaf's first triangle is (a, c, d), the second (a, b, c)

// Triangle (leaf order, DevScene::tri_stride float4 apart: 48 B for scenes
// that fit the L2, 64 B with a padding float4 otherwise), used by traversal:
//   t0 = (v0.xyz, orig_id bits)  t1 = (e1.xyz, flags bits)
//   t2 = (e2.xyz, -)
// e1 = v1 - v0 and e2 = v2 - v0 are formed on the host in FP32, the same
// rounding the shader applies at :302-303.
// Triangle shading record (leaf order, DevScene::tshade_stride float4 apart:
// 48 B for scenes that fit the L2, 64 B with a padding float4 otherwise, so
// that each record gathered from memory is one aligned 64-B sector), used for
// the back-face test and by the shade pass:
//   n0 = (normal of v0, mat_idx bits incl. the cull flag)
//   n1 = (normal of v1, -)  n2 = (normal of v2, -)
// Sphere (32 B, leaf order):
//   s0 = (centre.xyz, radius)  s1 = (mat, orig_id, flags, -) as uint bits

struct DevScene {
    const float4 *mnodes;
    const float4 *mnodes_full;  // the pairs at 64 B (packet traversal): mnodes, or a copy beside quantized ones
    const float4 *mtris;
    const float4 *tshade;
    uint32_t tri_stride;     // float4 per triangle record: 3, or 4 (padded) for scenes that spill the L2
    uint32_t tshade_stride;  // float4 per shading record: 3, or 4 (padded) for scenes that spill the L2
    uint32_t mroot;  // kNone = no mesh
    float mroot_lo[3], mroot_hi[3];
    const float4 *snodes;
    const float4 *sph;
    uint32_t sroot;  // kNone = no spheres
    float sroot_lo[3], sroot_hi[3];
    const GpuMaterial *mats;
    const float4 *env;
    uint32_t env_w, env_h;
    uint32_t stack_depth;     // traversal stack entries needed per lane
    uint32_t refill;          // idle lanes (1..64) at which a traversal wave loads new rays
    uint32_t mtop;            // plain-stack kernels: mesh nodes [0, mtop) are kept in LDS (<= kNodeCache)
    uint32_t stack_cull;      // traversal kernels with the culling stack (TStack<true>), and
                              // quantized 32-B child pairs in mnodes / snodes (node_quant.h)
    const uint4 *mwide;       // culling-stack kernels: the mesh tree as 4-wide 48-B nodes (bvh_wide.h)
    uint32_t mwroot;          // their root: wide node 0, a leaf reference, or kNone
    // Emitters for next-event estimation: emissive triangles and spheres,
    // chosen with probability proportional to area x luminance.
    //   emit[2e] = (prob, kind | leaf-order index bits, area, -),
    //   emit[2e+1] = (Le.rgb, -); emit_cdf[e] = sum of prob up to e.
    // tri_emit / sph_emit map a leaf-order primitive to its emitter or kNone.
    const float4 *emit;
    const float *emit_cdf;
    const uint32_t *tri_emit, *sph_emit;
    uint32_t n_emit;
};

// ---------------------------------------------------------------------------
// Vector helpers (explicit operation order)
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize(V3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return inv * a;
}
__device__ __forceinline__ V3 xyz(float4 q) { return v3(q.x, q.y, q.z); }
__device__ __forceinline__ V3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }

// WGSL f32 -> u32 saturates (negative and NaN -> 0).
__device__ __forceinline__ uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// ---------------------------------------------------------------------------
// PCG, path_trace.wgsl:194-204, seed :636-637
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t rot = s >> 28u;
    s = ((s >> (rot + 4u)) ^ s) * 277803737u;
    return (s >> 22u) ^ s;
}
__device__ __forceinline__ float rand_f32(uint32_t &st) {
    st = pcg_hash(st);
    return __uint_as_float((st >> 9u) | 0x3f800000u) - 1.0f;
}

constexpr float kTau = 6.283185307179586f;
constexpr float kPi = 3.141592653589793f;

// ---------------------------------------------------------------------------
// Slab test on a box, path_trace.wgsl:248-256.  inv = 1/dir computed once
// per ray (the value the shader recomputes per call).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void slab2(V3 o, V3 inv, float lox, float hix, float loy,
                                      float hiy, float loz, float hiz, float &te,
                                      float &tx) {
    float t0x = (lox - o.x) * inv.x, t1x = (hix - o.x) * inv.x;
    float t0y = (loy - o.y) * inv.y, t1y = (hiy - o.y) * inv.y;
    float t0z = (loz - o.z) * inv.z, t1z = (hiz - o.z) * inv.z;
    te = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    tx = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
}

// ---------------------------------------------------------------------------
// Möller–Trumbore, path_trace.wgsl:300-317, with edges precomputed.
// Returns true for a hit with t in [t_min, t_max) and t > 0 (:444).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool mt_hit(V3 o, V3 d, V3 v0, V3 e1, V3 e2, float t_min,
                                       float t_max, float &t, float &u, float &v) {
    V3 h = cross(d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-8f) return false;
    float f = 1.0f / a;
    V3 s = o - v0;
    u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot(e2, q);
    if (t < t_min || t >= t_max) return false;
    return t > 0.0f;
}

// Sphere roots, path_trace.wgsl:262-280.  Returns t or -1.
__device__ __forceinline__ float sphere_t(V3 o, V3 d, float4 s0, float t_min, float t_max) {
    V3 c = xyz(s0);
    float r = s0.w;
    V3 oc = o - c;
    float a = dot(d, d);
    float h = dot(oc, d);
    float cc = dot(oc, oc) - r * r;
    float disc = h * h - a * cc;
    if (disc < 0.0f) return -1.0f;
    float sq = sqrtf(disc);
    float t = (-h - sq) / a;
    if (t <= t_min || t >= t_max) {
        t = (-h + sq) / a;
        if (t <= t_min || t >= t_max) return -1.0f;
    }
    return t;
}

// Interpolated shading normal of a leaf-order triangle at (u, v),
// path_trace.wgsl:448-450.  mat receives mat_idx with its cull flag.
__device__ __forceinline__ V3 tri_normal(const DevScene &sc, uint32_t ordered, float u, float v,
                                         uint32_t &mat) {
    const float4 *r = sc.tshade + sc.tshade_stride * ordered;
    float4 a = r[0], b = r[1], c = r[2];
    mat = __float_as_uint(a.w);
    float w = 1.0f - u - v;
    return normalize((w * xyz(a) + u * xyz(b)) + v * xyz(c));
}

// ---------------------------------------------------------------------------
// Persistent closest-hit traversal.
//
// Visit rule (identical to the oracle's order-free closest hit): a child is
// entered when the reference slab test passes against the query interval
// [t_min, t_max) and its entry distance does not exceed best_t*(1+2^-10).
// Accept rule: lowest (t, original id) lexicographically.  Both rules are
// independent of visit order, so near-first ordering, the stack layout and
// the scheduling of rays onto lanes leave every result unchanged.
//
// Scheduling: every wave's first chunk is fixed by its index, min(kChunk,
// ceil(count / waves)) slots, so a launch starts without touching the shared
// counter and a queue of up to kChunk slots per wave needs no atomic at all.
// After that a wave pulls chunks of kChunk slots with one atomic.  (The
// counter is a single device-scope word that every wave of the grid would
// otherwise hit at once on entry; those pulls queue behind each other.)  A
// wave hands its slots to idle lanes whenever at least DevScene::refill lanes
// are idle, so lanes whose rays finished early are refilled instead of idling
// until the slowest ray of the wave ends.  Inside an iteration every lane first walks
// internal nodes until it reaches a leaf (or finishes), then the leaves are
// intersected together (the "while-while" order).
//
// Stack: the first kShort entries of each lane live in LDS
// (lds[i * kBlock + tid]); deeper entries go to a per-lane column of a
// global overflow buffer (ovf[(i - kShort) * words * ovf_stride + gtid]).
// ---------------------------------------------------------------------------
constexpr float kCullSlack = 1.0009765625f;  // 1 + 2^-10
constexpr uint32_t kChunk = 256;

struct Best {
    float t;
    uint32_t id;    // original id (for the tie rule / reporting)
    uint32_t prim;  // leaf-order index; kLeafBit marks a sphere
    float u, v;
};

__device__ __forceinline__ bool better(float t, uint32_t id, const Best &b) {
    return t < b.t || (t == b.t && id < b.id);
}

// Traversal stack of one lane.  kCull = false: entries are node references,
// 16 in LDS.  kCull = true: an entry also holds the entry distance of its box,
// measured when it was pushed, and a pop skips entries that start beyond the
// lane's current cull limit without fetching them (the rule that culls a
// child at its parent, applied with the closer hit found since); 10 entries
// in LDS keep the block's LDS at the plain stack's occupancy.
template <bool kCull>
struct TStack {
    static constexpr uint32_t kShort = kCull ? 10u : 16u;
    static constexpr uint32_t kWords = kCull ? 2u : 1u;  // 32-bit words per entry
    uint32_t *lds;
    uint32_t *ovf;
    uint32_t ovf_stride, tid, gtid, block;
    // Reads and writes go through explicitly LDS- and global-qualified
    // pointers.  Through plain pointers the compiler merges the LDS and the
    // overflow access into one flat access, which routes every LDS push or pop
    // through the vector-memory path (its address unit is the busiest unit of
    // the traversal) and makes each node step wait on both counters.
    __device__ __forceinline__ void set(uint32_t i, uint32_t v, float te) const {
        if (i < kShort) {
            __attribute__((address_space(3))) uint32_t *l = (__attribute__((address_space(3))) uint32_t *)lds;
            l[i * block + tid] = v;
            if (kCull) l[(kShort + i) * block + tid] = __float_as_uint(te);
        } else {
            __attribute__((address_space(1))) uint32_t *g = (__attribute__((address_space(1))) uint32_t *)ovf;
            const size_t e = (size_t)(i - kShort) * kWords * ovf_stride + gtid;
            g[e] = v;
            if (kCull) g[e + ovf_stride] = __float_as_uint(te);
        }
    }
    __device__ __forceinline__ uint32_t get(uint32_t i, float &te) const {
        if (i < kShort) {
            const __attribute__((address_space(3))) uint32_t *l = (const __attribute__((address_space(3))) uint32_t *)lds;
            if (kCull) te = __uint_as_float(l[(kShort + i) * block + tid]);
            return l[i * block + tid];
        }
        const __attribute__((address_space(1))) uint32_t *g = (const __attribute__((address_space(1))) uint32_t *)ovf;
        const size_t e = (size_t)(i - kShort) * kWords * ovf_stride + gtid;
        if (kCull) te = __uint_as_float(g[e + ovf_stride]);
        return g[e];
    }
    // Top entry whose box may still hold a hit (entry distance within lim),
    // or kNone when the stack runs out.
    __device__ __forceinline__ uint32_t pop(uint32_t &sp, float lim) const {
        while (sp) {
            float te = 0.0f;
            const uint32_t v = get(--sp, te);
            if (!kCull || !(te > lim)) return v;
        }
        return kNone;
    }
};
using Stack = TStack<false>;

// Root entry of one BVH for a ray: the root box test of the reference
// (the root node is popped and slab-tested first, :375-379).
// kWide: the mesh root is that of the 4-wide tree.
template <bool kWide = false>
__device__ __forceinline__ uint32_t enter_root(const DevScene &sc, bool sphere, V3 o, V3 inv,
                                               float t_min, float t_max) {
    const uint32_t root = sphere ? sc.sroot : (kWide ? sc.mwroot : sc.mroot);
    if (root == kNone) return kNone;
    const float *lo = sphere ? sc.sroot_lo : sc.mroot_lo;
    const float *hi = sphere ? sc.sroot_hi : sc.mroot_hi;
    float te, tx;
    slab2(o, inv, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], te, tx);
    return (tx > fmaxf(te, t_min) && te < t_max) ? root : kNone;
}

// One internal-node step: slab-test both children, descend near-first.
// kQuant: the node is a quantized 32-B pair (two 16-B loads instead of four);
// its decoded boxes contain the reference boxes, and the slab test on them is
// the reference's.
template <bool kQuant, class S>
__device__ __forceinline__ void node_step(const float4 *nodes, V3 o, V3 inv, float t_min,
                                          float t_max, float best_t, uint32_t &cur, uint32_t &sp,
                                          const S &st, const float4 *cache = nullptr, uint32_t ncache = 0) {
    float te0, tx0, te1, tx1;
    uint32_t c0, c1;
    if constexpr (kQuant) {
        const uint4 *n = reinterpret_cast<const uint4 *>(nodes) + 2u * cur;
        const uint4 wa = n[0], wb = n[1];
        const float sx = qnode_step(wa.x), ox = qnode_origin(wa.x);
        const float sy = qnode_step(wa.y), oy = qnode_origin(wa.y);
        const float sz = qnode_step(wa.z), oz = qnode_origin(wa.z);
        slab2(o, inv, qnode_plane(wa.w, 0, sx, ox), qnode_plane(wa.w, 1, sx, ox), qnode_plane(wa.w, 2, sy, oy),
              qnode_plane(wa.w, 3, sy, oy), qnode_plane(wb.x, 0, sz, oz), qnode_plane(wb.x, 1, sz, oz), te0, tx0);
        slab2(o, inv, qnode_plane(wb.x, 2, sx, ox), qnode_plane(wb.x, 3, sx, ox), qnode_plane(wb.y, 0, sy, oy),
              qnode_plane(wb.y, 1, sy, oy), qnode_plane(wb.y, 2, sz, oz), qnode_plane(wb.y, 3, sz, oz), te1, tx1);
        c0 = wb.z;
        c1 = wb.w;
    } else {
        typedef float nf4 __attribute__((ext_vector_type(4)));
        nf4 r0, r1, r2, r3;
        if (cur < ncache) {
            const __attribute__((address_space(3))) nf4 *n =
                (const __attribute__((address_space(3))) nf4 *)cache + 4u * cur;
            r0 = n[0];
            r1 = n[1];
            r2 = n[2];
            r3 = n[3];
        } else {
            const __attribute__((address_space(1))) nf4 *n = (const __attribute__((address_space(1))) nf4 *)nodes + 4u * cur;
            r0 = n[0];
            r1 = n[1];
            r2 = n[2];
            r3 = n[3];
        }
        const float4 q0 = make_float4(r0.x, r0.y, r0.z, r0.w), q1 = make_float4(r1.x, r1.y, r1.z, r1.w);
        const float4 q2 = make_float4(r2.x, r2.y, r2.z, r2.w), q3 = make_float4(r3.x, r3.y, r3.z, r3.w);
        slab2(o, inv, q0.x, q0.y, q0.z, q0.w, q2.x, q2.y, te0, tx0);
        slab2(o, inv, q1.x, q1.y, q1.z, q1.w, q2.z, q2.w, te1, tx1);
        c0 = __float_as_uint(q3.x);
        c1 = __float_as_uint(q3.y);
    }
    const float lim = best_t * kCullSlack;
    // te < t_max (the reference test) and !(te > lim) (the cull) as one
    // compare: te < min(t_max, next float above lim) holds exactly when both
    // do, and is false for a NaN te, which the reference test rejects too.
    const float bound = fminf(t_max, __uint_as_float(__float_as_uint(lim) + 1u));
    bool h0 = tx0 > fmaxf(te0, t_min) && te0 < bound;
    bool h1 = tx1 > fmaxf(te1, t_min) && te1 < bound;
    if (h0 && h1) {
        uint32_t nearc = c0, farc = c1;
        if (te1 < te0) {
            nearc = c1;
            farc = c0;
        }
        st.set(sp, farc, te1 < te0 ? te0 : te1);
        ++sp;
        cur = nearc;
    } else if (h0) {
        cur = c0;
    } else if (h1) {
        cur = c1;
    } else {
        cur = st.pop(sp, lim);
    }
}

// One 4-wide node step (culling stack; bvh_wide.h layout): slab-test the up
// to 4 children on their shared grid, descend into the nearest child that
// passes, and push the others farthest first, each with its entry distance.
// A child passes where the reference slab test passes on its decoded box
// (which contains the binary child box it came from) and the entry is within
// the cull limit, the same rule as node_step.
template <class S>
__device__ __forceinline__ void node_step_wide(const uint4 *wide, V3 o, V3 inv, float t_min, float t_max,
                                               float best_t, uint32_t &cur, uint32_t &sp, const S &st) {
    const uint4 *n = wide + 3u * cur;
    const uint4 wa = n[0], wb = n[1], wc = n[2];
    const float sx = qnode_step(wa.x), ox = qnode_origin(wa.x);
    const float sy = qnode_step(wa.y), oy = qnode_origin(wa.y);
    const float sz = qnode_step(wa.z), oz = qnode_origin(wa.z);
    float te0, tx0, te1, tx1, te2, tx2, te3, tx3;
    slab2(o, inv, qnode_plane(wa.w, 0, sx, ox), qnode_plane(wa.w, 1, sx, ox), qnode_plane(wa.w, 2, sy, oy),
          qnode_plane(wa.w, 3, sy, oy), qnode_plane(wb.x, 0, sz, oz), qnode_plane(wb.x, 1, sz, oz), te0, tx0);
    slab2(o, inv, qnode_plane(wb.x, 2, sx, ox), qnode_plane(wb.x, 3, sx, ox), qnode_plane(wb.y, 0, sy, oy),
          qnode_plane(wb.y, 1, sy, oy), qnode_plane(wb.y, 2, sz, oz), qnode_plane(wb.y, 3, sz, oz), te1, tx1);
    slab2(o, inv, qnode_plane(wb.z, 0, sx, ox), qnode_plane(wb.z, 1, sx, ox), qnode_plane(wb.z, 2, sy, oy),
          qnode_plane(wb.z, 3, sy, oy), qnode_plane(wb.w, 0, sz, oz), qnode_plane(wb.w, 1, sz, oz), te2, tx2);
    slab2(o, inv, qnode_plane(wb.w, 2, sx, ox), qnode_plane(wb.w, 3, sx, ox), qnode_plane(wc.x, 0, sy, oy),
          qnode_plane(wc.x, 1, sy, oy), qnode_plane(wc.x, 2, sz, oz), qnode_plane(wc.x, 3, sz, oz), te3, tx3);
    const float lim = best_t * kCullSlack;
    const float bound = fminf(t_max, __uint_as_float(__float_as_uint(lim) + 1u));
    const uint32_t meta = wc.w;
    // Sort keys: for a passing child, the bits of its entry distance raised
    // to at least t_min and +0 (non-negative, so the bits order as unsigned
    // integers), with the slot in the 2 low bits; all ones for the others.
    // The distance a key carries (low bits cleared) is at most max(entry,
    // t_min, 0), so a box that holds the closest hit t* (entry <= t*,
    // t* >= t_min, t* > 0) is never culled by it.  An unused slot is a point
    // box (every plane byte 0), whose exit never exceeds its entry: it never
    // passes, with no test of its meta byte (such a test became a branch
    // around each child, and the node's first 16-B load was split in two).
    auto key = [&](float te, float tx, uint32_t k) {
        const bool h = tx > fmaxf(te, t_min) && te < bound;
        return h ? ((__float_as_uint(fmaxf(fmaxf(te, t_min), 0.0f)) & ~3u) | k) : 0xFFFFFFFFu;
    };
    uint32_t k0 = key(te0, tx0, 0), k1 = key(te1, tx1, 1), k2 = key(te2, tx2, 2), k3 = key(te3, tx3, 3);
    auto cas = [](uint32_t &a, uint32_t &b) {
        const uint32_t lo = min(a, b), hi = max(a, b);
        a = lo;
        b = hi;
    };
    cas(k0, k1);
    cas(k2, k3);
    cas(k0, k2);
    cas(k1, k3);
    cas(k1, k2);
    auto ref = [&](uint32_t key) {
        const uint32_t m = (meta >> (8u * (key & 3u))) & 0xFFu;
        return (m & 0x80u) ? (kLeafBit | ((m & 0x40u) << 24) | (wc.z + (m & 0x3Fu))) : (wc.y + (m & 3u));
    };
    auto dist = [](uint32_t key) { return __uint_as_float(key & ~3u); };
    if (k0 == 0xFFFFFFFFu) {
        cur = st.pop(sp, lim);
        return;
    }
    if (k3 != 0xFFFFFFFFu) st.set(sp++, ref(k3), dist(k3));
    if (k2 != 0xFFFFFFFFu) st.set(sp++, ref(k2), dist(k2));
    if (k1 != 0xFFFFFFFFu) st.set(sp++, ref(k1), dist(k1));
    cur = ref(k0);
}

__device__ __forceinline__ void tri_test(const DevScene &sc, uint32_t i, float4 a, float4 b, float4 c, V3 o,
                                         V3 d, float t_min, float t_max, Best &best) {
    float t, u, v;
    if (mt_hit(o, d, xyz(a), xyz(b), xyz(c), t_min, t_max, t, u, v)) {
        uint32_t id = __float_as_uint(a.w);
        if (best.prim == kNone || better(t, id, best)) {
            bool ok = true;
            if (__float_as_uint(b.w) & kPrimCull) {  // :438-452, culled candidates do not shrink t
                uint32_t mat;
                V3 nn = tri_normal(sc, i, u, v, mat);
                ok = dot(d, nn) < 0.0f;
            }
            if (ok) {
                best.t = t;
                best.id = id;
                best.prim = i;
                best.u = u;
                best.v = v;
            }
        }
    }
}

// Leaf at primitive i.  pair: the leaf is a quad, two triangles (a, b, c)
// and (a, c, d) in either order, and the first record (64 B) holds both:
// its spare slots carry the second triangle's id (third float4, w) and its
// remaining edge and flags (fourth float4).  The shared edge c - a is the
// same FP32 value in both triangles' own records, so both tests see exactly
// their own records' operands, from 4 loads instead of 6.  kQuads = false
// compiles without that path (kernels of scenes whose trees never set
// kLeafPair: it cost their registers).
template <bool kQuads = true>
__device__ __forceinline__ void tri_leaf(const DevScene &sc, uint32_t i, bool pair, V3 o, V3 d, float t_min,
                                         float t_max, Best &best) {
    if (kQuads && pair) {
        const float4 *tp = sc.mtris + sc.tri_stride * i;
        const float4 a = tp[0], b = tp[1], c = tp[2], q = tp[3];
        tri_test(sc, i, a, b, c, o, d, t_min, t_max, best);
        const bool sw = (__float_as_uint(q.w) & kQuadSwap) != 0u;
        const float4 a1 = make_float4(a.x, a.y, a.z, c.w);
        const float4 b1 = sw ? q : make_float4(c.x, c.y, c.z, q.w);
        const float4 c1 = sw ? b : q;
        tri_test(sc, i + 1u, a1, b1, c1, o, d, t_min, t_max, best);
        return;
    }
    for (;;) {
        const float4 *tp = sc.mtris + sc.tri_stride * i;
        float4 a = tp[0], b = tp[1], c = tp[2];
        tri_test(sc, i, a, b, c, o, d, t_min, t_max, best);
        if (__float_as_uint(b.w) & kPrimLast) break;
        ++i;
    }
}

__device__ __forceinline__ void sphere_leaf(const DevScene &sc, uint32_t i, V3 o, V3 d,
                                            float t_min, float t_max, Best &best) {
    for (;;) {
        const float4 *p = sc.sph + 2u * i;
        float4 s0 = p[0], s1 = p[1];
        uint32_t flags = __float_as_uint(s1.z);
        float t = sphere_t(o, d, s0, t_min, t_max);
        if (t > 0.0f) {
            uint32_t id = kLeafBit | __float_as_uint(s1.y);
            if (best.prim == kNone || better(t, id, best)) {
                best.t = t;
                best.id = id;
                best.prim = kLeafBit | i;
                best.u = 0.0f;
                best.v = 0.0f;
            }
        }
        if (flags & kPrimLast) break;
        ++i;
    }
}

// Closest hit of one ray per lane (world_hit semantics and visit rule of
// trace_persistent), for megakernels that keep a whole path in one thread.
// Called by every active lane of a wave together (it votes with ballots);
// its loop has trace_persistent's while-while order without the refill:
// lanes walk internal nodes, postponing the first leaf they reach, until
// every lane holds a leaf or is done, then the leaves are intersected
// together, so lanes at leaves and lanes at nodes do not take turns.
template <bool kSpheres>
__device__ __forceinline__ Best trace_one(const DevScene &sc, V3 o, V3 d, float t_min, float t_max,
                                          const Stack &st) {
    const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    Best best, sbest;
    best.t = t_max;
    best.id = best.prim = kNone;
    best.u = best.v = 0.0f;
    sbest = best;
    bool sphere_phase = kSpheres && sc.sroot != kNone;
    float t_cur = t_max;
    uint32_t cur = enter_root(sc, sphere_phase, o, inv, t_min, t_cur);
    uint32_t sp = 0, leaf = kNone;
    bool done = false;
    for (;;) {
        for (;;) {
            if (!done && leaf == kNone && cur != kNone && (cur & kLeafBit)) {
                leaf = cur;
                cur = st.pop(sp, best.t * kCullSlack);
            }
            const bool in = !done && cur != kNone && !(cur & kLeafBit);
            if (__ballot(in && leaf == kNone) == 0) break;
            if (in) {
                const float4 *nodes = (kSpheres && sphere_phase) ? sc.snodes : sc.mnodes;
                if (sc.stack_cull)
                    node_step<true>(nodes, o, inv, t_min, t_cur, best.t, cur, sp, st);
                else
                    node_step<false>(nodes, o, inv, t_min, t_cur, best.t, cur, sp, st);
            }
        }
        if (!done && leaf != kNone) {
            if (kSpheres && sphere_phase)
                sphere_leaf(sc, leaf & ~kLeafBit, o, d, t_min, t_cur, best);
            else
                tri_leaf(sc, leaf & ~(kLeafBit | kLeafPair), (leaf & kLeafPair) != 0u, o, d, t_min, t_cur, best);
            leaf = kNone;
        }
        if (!done && cur == kNone && leaf == kNone) {
            if (kSpheres && sphere_phase) {  // spheres -> mesh (world_hit, :471-479)
                sbest = best;
                sphere_phase = false;
                t_cur = sbest.prim != kNone ? sbest.t : t_max;
                best.prim = kNone;
                best.id = kNone;
                best.t = t_cur;
                cur = enter_root(sc, false, o, inv, t_min, t_cur);
            }
            if (!sphere_phase && cur == kNone) done = true;
        }
        if (__ballot(!done) == 0) break;
    }
    return (!kSpheres || best.prim != kNone) ? best : sbest;
}

// IO policy: count(), work(), load(slot, o, d, t_min, t_max), store(slot, best),
// and kAnyHit with any_hit(slot).  An any-hit slot is a visibility query (a
// shadow ray): its ray ends at the first leaf that yields a hit in
// [t_min, t_max), and a sphere hit skips the mesh phase.  Up to the first hit
// the traversal is the closest-hit one (best t is still t_max), so a miss is
// reported exactly when the closest-hit traversal reports one; the hit stored
// is some hit, not the closest.  Policies without kAnyHit compile none of it
// (a per-lane flag in the extension kernels cost them 1%).
// kSpheres = false compiles the variant for scenes without spheres (no
// sphere phase state), launched when the sphere BVH is empty; kCull selects
// the culling stack (lds holds TStack<kCull>::kShort * kWords words per lane).
template <int kBlock, bool kSpheres, bool kCull, class IO>
__device__ __forceinline__ void trace_persistent(const DevScene &sc, const IO &io, uint32_t *lds,
                                                 uint32_t *ovf, uint32_t ovf_stride,
                                                 float4 *cache = nullptr) {
    uint32_t ncache = 0;
    if (!kCull && cache) {  // the breadth-first top levels of the mesh tree
        ncache = sc.mtop;
        for (uint32_t i = threadIdx.x; i < 4u * ncache; i += kBlock) cache[i] = sc.mnodes[i];
        __syncthreads();
    }
    const uint32_t lane = __lane_id();
    TStack<kCull> st{lds, ovf, ovf_stride, threadIdx.x, blockIdx.x * kBlock + threadIdx.x, (uint32_t)kBlock};
    const uint32_t total = io.count();
    const uint32_t waves = gridDim.x * (uint32_t)(kBlock / 64);
    const uint32_t wave_id = blockIdx.x * (uint32_t)(kBlock / 64) + threadIdx.x / 64u;
    const uint32_t first = min(kChunk, (total + waves - 1u) / waves);  // static first chunk
    const uint32_t dyn0 = first * waves;                                // first pulled slot
    uint32_t wbase = min(wave_id * first, total);                       // wave-local slot range
    uint32_t wend = min(wbase + first, total);
    bool more = dyn0 < total;
    // lane state
    uint32_t slot = kNone;
    V3 o = v3(0, 0, 0), d = v3(0, 0, 0), inv = v3(0, 0, 0);
    float t_min = 0.0f, t_q = 0.0f, t_cur = 0.0f;  // query and current-phase upper bounds
    uint32_t cur = kNone, sp = 0;
    uint32_t leaf = kNone;  // postponed leaf
    bool sphere_phase = false, any = false;
    Best best, sbest;
    best.t = 0.0f;
    best.id = best.prim = kNone;
    best.u = best.v = 0.0f;
    sbest = best;
    for (;;) {
        // ---- refill idle lanes --------------------------------------------
        const uint64_t idle = __ballot(slot == kNone);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle >= sc.refill) {
            if (wbase >= wend && more) {
                uint32_t g = 0;
                if (lane == 0) g = dyn0 + atomicAdd(io.work(), kChunk);
                g = __shfl(g, 0);
                if (g < total) {
                    wbase = g;
                    wend = min(g + kChunk, total);
                } else {
                    more = false;
                }
            }
            const uint32_t avail = wend - wbase;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            if (slot == kNone && rank < avail) {
                slot = wbase + rank;
                io.load(slot, o, d, t_min, t_q);
                if constexpr (IO::kAnyHit) any = io.any_hit(slot);
                inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
                best.prim = kNone;
                best.id = kNone;
                best.t = t_q;
                if (kSpheres) sbest = best;
                sphere_phase = kSpheres && sc.sroot != kNone;
                t_cur = t_q;
                cur = enter_root<kCull>(sc, sphere_phase, o, inv, t_min, t_cur);
                sp = 0;
                leaf = kNone;
            }
            wbase += min(avail, nidle);
        }
        if (__ballot(slot != kNone) == 0) {
            if (!more && wbase >= wend) break;
            continue;
        }
        // ---- internal nodes until every lane holds a leaf or is done -------
        // A lane that reaches a leaf postpones it and keeps descending while
        // other lanes still search for theirs; its extra steps cull with an
        // older (larger) best t, so they only visit more nodes, and they keep
        // the lane busy instead of idle in the wave.
        for (;;) {
            if (slot != kNone && leaf == kNone && cur != kNone && (cur & kLeafBit)) {
                leaf = cur;
                cur = st.pop(sp, best.t * kCullSlack);
            }
            const bool in = slot != kNone && cur != kNone && !(cur & kLeafBit);
            if (__ballot(in && leaf == kNone) == 0) break;
            if (in) {
                if (kCull && !(kSpheres && sphere_phase)) {
                    node_step_wide(sc.mwide, o, inv, t_min, t_cur, best.t, cur, sp, st);
                } else {
                    const float4 *nodes = (kSpheres && sphere_phase) ? sc.snodes : sc.mnodes;
                    node_step<kCull>(nodes, o, inv, t_min, t_cur, best.t, cur, sp, st, cache,
                                     (kSpheres && sphere_phase) ? 0u : ncache);
                }
            }
        }
        // ---- leaves ----------------------------------------------------------
        if (slot != kNone && leaf != kNone) {
            if (kSpheres && sphere_phase)
                sphere_leaf(sc, leaf & ~kLeafBit, o, d, t_min, t_cur, best);
            else
                tri_leaf<kCull>(sc, leaf & ~(kLeafBit | kLeafPair), (leaf & kLeafPair) != 0u, o, d, t_min, t_cur, best);
            leaf = kNone;
            if (IO::kAnyHit && any && best.prim != kNone) {  // a visibility query is answered
                cur = kNone;
                sp = 0;
            }
        }
        // ---- phase end: spheres -> mesh (world_hit, :471-479) or finish ----
        if (slot != kNone && cur == kNone && leaf == kNone) {
            if (kSpheres && sphere_phase) {
                sbest = best;
                sphere_phase = false;
                if (!(IO::kAnyHit && any && sbest.prim != kNone)) {
                    t_cur = sbest.prim != kNone ? sbest.t : t_q;
                    best.prim = kNone;
                    best.id = kNone;
                    best.t = t_cur;
                    cur = enter_root<kCull>(sc, false, o, inv, t_min, t_cur);
                }
            }
            if (!sphere_phase && cur == kNone) {
                io.store(slot, (!kSpheres || best.prim != kNone) ? best : sbest);
                slot = kNone;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Packet traversal for coherent queues (first-bounce camera rays: a wave's
// 64 slots are 4 neighbouring pixels x 16 frames).  The wave walks one
// wave-uniform path through the tree: a child is entered when any lane's
// slab test passes it with the lane's own cull limit, so every lane visits a
// superset of the nodes its own traversal visits and keeps the lowest
// (t, original id); results equal the per-lane traversal's.  Nodes and
// triangles are wave-uniform and come in through scalar loads (no
// vector-memory address work): a node is the exact 64-B pair, one
// s_load_dwordx16, with no plane decode (DevScene::mnodes_full; scenes with
// quantized pairs upload the 64-B pairs beside them).  The stack is wave-uniform and held one entry
// per lane in two register pairs (128 entries; trees are at most 96 deep).
// Each stack entry keeps the smallest entry distance over the lanes that
// hit the far child, so a pop is skipped when no lane's limit reaches it.
// Sphere-free scenes, first bounce.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
    for (int s = 32; s >= 1; s >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, s));
    return v;
}

template <int kBlock, class IO>
__device__ __forceinline__ void trace_packets(const DevScene &sc, const IO &io) {
    typedef float pf4 __attribute__((ext_vector_type(4)));
    const __attribute__((address_space(4))) pf4 *nf = (const __attribute__((address_space(4))) pf4 *)sc.mnodes_full;
    const __attribute__((address_space(4))) pf4 *tr = (const __attribute__((address_space(4))) pf4 *)sc.mtris;
    const uint32_t lane = __lane_id();
    const uint32_t total = io.count();
    const uint32_t waves = gridDim.x * (uint32_t)(kBlock / 64);
    const uint32_t wave_id = blockIdx.x * (uint32_t)(kBlock / 64) + threadIdx.x / 64u;
    const uint32_t first = min(kChunk, (total + waves - 1u) / waves);
    const uint32_t dyn0 = first * waves;
    uint32_t cbase = min(wave_id * first, total), cend = min(cbase + first, total);
    bool more = dyn0 < total;
    for (;;) {
        if (cbase >= cend) {
            if (!more) break;
            uint32_t g = 0;
            if (lane == 0) g = dyn0 + atomicAdd(io.work(), kChunk);
            g = __builtin_amdgcn_readfirstlane(g);
            if (g >= total) break;
            cbase = g;
            cend = min(g + kChunk, total);
        }
        const uint32_t slot = cbase + lane;
        const bool valid = slot < cend;
        cbase = min(cbase + 64u, cend);
        V3 o = v3(0, 0, 0), d = v3(1, 1, 1);
        float t_min = 0.0f, t_max = 0.0f;
        if (valid) io.load(slot, o, d, t_min, t_max);
        const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        Best best;
        best.t = t_max;
        best.id = best.prim = kNone;
        best.u = best.v = 0.0f;
        const bool act = valid && enter_root(sc, false, o, inv, t_min, t_max) != kNone;
        uint32_t cur = __ballot(act) ? sc.mroot : kNone;
        uint32_t sp = 0;
        uint32_t sr0 = 0, sr1 = 0, st0 = 0, st1 = 0;  // stack entry k in lane k % 64 of pair k / 64
        while (cur != kNone) {
            if (!(cur & kLeafBit)) {
                const __attribute__((address_space(4))) pf4 *qf = nf + 4u * cur;
                const pf4 q0 = qf[0], q1 = qf[1], q2 = qf[2], q3 = qf[3];
                float te0, tx0, te1, tx1;
                slab2(o, inv, q0.x, q0.y, q0.z, q0.w, q2.x, q2.y, te0, tx0);
                slab2(o, inv, q1.x, q1.y, q1.z, q1.w, q2.z, q2.w, te1, tx1);
                const float lim = best.t * kCullSlack;
                const float bound = fminf(t_max, __uint_as_float(__float_as_uint(lim) + 1u));
                const bool h0 = act && tx0 > fmaxf(te0, t_min) && te0 < bound;
                const bool h1 = act && tx1 > fmaxf(te1, t_min) && te1 < bound;
                const uint64_t m0 = __ballot(h0), m1 = __ballot(h1);
                const uint32_t c0 = __float_as_uint(q3.x), c1 = __float_as_uint(q3.y);
                if (m0 && m1) {
                    const uint64_t both = m0 & m1;
                    const uint64_t first1 = __ballot(h0 && h1 && te1 < te0);
                    const bool near1 = both ? 2u * (uint32_t)__popcll(first1) > (uint32_t)__popcll(both)
                                            : __popcll(m1) > __popcll(m0);
                    const bool hf = near1 ? h0 : h1;
                    const float tef = fmaxf(near1 ? te0 : te1, t_min);
                    const uint32_t fte = __builtin_amdgcn_readfirstlane(wave_umin(hf ? __float_as_uint(tef) : 0x7F800000u));
                    const uint32_t far = near1 ? c0 : c1;
                    if (sp < 64u) {
                        sr0 = lane == sp ? far : sr0;
                        st0 = lane == sp ? fte : st0;
                    } else {
                        sr1 = lane == sp - 64u ? far : sr1;
                        st1 = lane == sp - 64u ? fte : st1;
                    }
                    ++sp;
                    cur = near1 ? c1 : c0;
                    continue;
                }
                if (m0) {
                    cur = c0;
                    continue;
                }
                if (m1) {
                    cur = c1;
                    continue;
                }
            } else {
                uint32_t i = cur & ~(kLeafBit | kLeafPair);
                const bool pair = (cur & kLeafPair) != 0u;
                for (uint32_t k = 0;; ++k, ++i) {
                    i = __builtin_amdgcn_readfirstlane(i);  // wave-uniform: scalar loads
                    const __attribute__((address_space(4))) pf4 *tp = tr + sc.tri_stride * i;
                    const pf4 a = tp[0], b = tp[1], c = tp[2];
                    if (act)
                        tri_test(sc, i, make_float4(a.x, a.y, a.z, a.w), make_float4(b.x, b.y, b.z, b.w),
                                 make_float4(c.x, c.y, c.z, c.w), o, d, t_min, t_max, best);
                    if (pair ? k == 1u : (__float_as_uint(b.w) & kPrimLast) != 0u) break;
                }
            }
            cur = kNone;
            while (sp > 0u) {
                --sp;
                const uint32_t r = sp < 64u ? __builtin_amdgcn_readlane(sr0, sp) : __builtin_amdgcn_readlane(sr1, sp - 64u);
                const uint32_t te = sp < 64u ? __builtin_amdgcn_readlane(st0, sp) : __builtin_amdgcn_readlane(st1, sp - 64u);
                if (__ballot(act && __uint_as_float(te) <= best.t * kCullSlack)) {
                    cur = r;
                    break;
                }
            }
        }
        if (valid) io.store(slot, best);
    }
}

}  // namespace mcrt
