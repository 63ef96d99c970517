// dev_common.h — device-side scene layout and arithmetic shared by the
// render, trace and adjoint kernels of libmcrt_hip.so (gfx950, wave64).
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
// A child reference is a device node index, or kLeafBit | first primitive
// index in leaf order.  A leaf's primitives run until the one whose
// kPrimLast flag is set.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kPrimLast = 1u;
constexpr uint32_t kPrimCull = 2u;

// Triangle (48 B, leaf order):
//   t0 = (v0.xyz, orig_id bits)  t1 = (e1.xyz, flags bits)
//   t2 = (e2.xyz, ordered index bits)
// e1 = v1 - v0 and e2 = v2 - v0 are formed on the host in FP32, the same
// rounding the shader applies at :302-303.
// Sphere (32 B, leaf order):
//   s0 = (centre.xyz, radius)  s1 = (mat, orig_id, flags, -) as uint bits

struct DevScene {
    const float4 *mnodes;
    const float4 *mtris;
    const GpuVertex *verts;
    const uint4 *tris_ref;  // ordered GpuTriangle
    uint32_t mroot;         // kNone = no mesh
    float mroot_lo[3], mroot_hi[3];
    const float4 *snodes;
    const float4 *sph;
    uint32_t sroot;  // kNone = no spheres
    float sroot_lo[3], sroot_hi[3];
    const GpuMaterial *mats;
    const float4 *env;
    uint32_t env_w, env_h;
    uint32_t stack_depth;  // LDS stack entries per lane
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
// Slab test on a child box, path_trace.wgsl:248-256.  inv = 1/dir computed
// once per ray (same value the shader recomputes per call).
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

// Interpolated shading normal of an ordered triangle at (u, v),
// path_trace.wgsl:448-450.
__device__ __forceinline__ V3 tri_normal(const DevScene &sc, uint32_t ordered, float u, float v,
                                         uint4 &tri) {
    tri = sc.tris_ref[ordered];
    const GpuVertex &va = sc.verts[tri.x];
    const GpuVertex &vb = sc.verts[tri.y];
    const GpuVertex &vc = sc.verts[tri.z];
    float w = 1.0f - u - v;
    return normalize((w * ld3(va.normal) + u * ld3(vb.normal)) + v * ld3(vc.normal));
}

// ---------------------------------------------------------------------------
// Closest hit over one child-pair BVH.
//
// Visit rule (identical to the oracle's order-free closest hit): a child is
// entered when the reference slab test passes against the query interval
// [t_min, t_max) and its entry distance does not exceed best_t*(1+2^-10).
// Accept rule: lowest (t, original id) lexicographically.  Both rules are
// independent of visit order, so near-first ordering and the LDS stack leave
// the result unchanged.
//
// The stack lives in LDS, one column per lane: stack[i * kBlock + tid].
// ---------------------------------------------------------------------------
constexpr float kCullSlack = 1.0009765625f;  // 1 + 2^-10

struct Best {
    float t;
    uint32_t id;    // original id (for the tie rule / reporting)
    uint32_t prim;  // leaf-order index (shading)
    float u, v;
};

__device__ __forceinline__ bool better(float t, uint32_t id, const Best &b) {
    return t < b.t || (t == b.t && id < b.id);
}

template <int kBlock, bool kSphere>
__device__ __forceinline__ void traverse(const DevScene &sc, V3 o, V3 d, V3 inv, float t_min,
                                         float t_max, Best &best, uint32_t *stack,
                                         uint32_t &work_nodes) {
    const uint32_t root = kSphere ? sc.sroot : sc.mroot;
    if (root == kNone) return;
    {
        const float *lo = kSphere ? sc.sroot_lo : sc.mroot_lo;
        const float *hi = kSphere ? sc.sroot_hi : sc.mroot_hi;
        float te, tx;
        slab2(o, inv, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], te, tx);
        if (!(tx > fmaxf(te, t_min) && te < t_max)) return;
    }
    const float4 *nodes = kSphere ? sc.snodes : sc.mnodes;
    const uint32_t tid = threadIdx.x;
    uint32_t sp = 0;
    uint32_t cur = root;
    for (;;) {
        if (!(cur & kLeafBit)) {
            ++work_nodes;
            const float4 *n = nodes + 4u * cur;
            float4 q0 = n[0], q1 = n[1], q2 = n[2], q3 = n[3];
            float te0, tx0, te1, tx1;
            slab2(o, inv, q0.x, q0.y, q0.z, q0.w, q2.x, q2.y, te0, tx0);
            slab2(o, inv, q1.x, q1.y, q1.z, q1.w, q2.z, q2.w, te1, tx1);
            const float lim = best.t * kCullSlack;
            bool h0 = tx0 > fmaxf(te0, t_min) && te0 < t_max && !(te0 > lim);
            bool h1 = tx1 > fmaxf(te1, t_min) && te1 < t_max && !(te1 > lim);
            uint32_t c0 = __float_as_uint(q3.x), c1 = __float_as_uint(q3.y);
            if (h0 && h1) {
                uint32_t nearc = c0, farc = c1;
                if (te1 < te0) {
                    nearc = c1;
                    farc = c0;
                }
                stack[sp * kBlock + tid] = farc;
                ++sp;
                cur = nearc;
                continue;
            }
            if (h0) {
                cur = c0;
                continue;
            }
            if (h1) {
                cur = c1;
                continue;
            }
        } else {
            uint32_t i = cur & ~kLeafBit;
            if (!kSphere) {
                for (;;) {
                    const float4 *tp = sc.mtris + 3u * i;
                    float4 a = tp[0], b = tp[1], c = tp[2];
                    uint32_t flags = __float_as_uint(b.w);
                    float t, u, v;
                    if (mt_hit(o, d, xyz(a), xyz(b), xyz(c), t_min, t_max, t, u, v)) {
                        uint32_t id = __float_as_uint(a.w);
                        if (best.prim == kNone || better(t, id, best)) {
                            bool ok = true;
                            uint32_t ordered = __float_as_uint(c.w);
                            if (flags & kPrimCull) {
                                uint4 tri;
                                V3 nn = tri_normal(sc, ordered, u, v, tri);
                                ok = dot(d, nn) < 0.0f;
                            }
                            if (ok) {
                                best.t = t;
                                best.id = id;
                                best.prim = ordered;
                                best.u = u;
                                best.v = v;
                            }
                        }
                    }
                    if (flags & kPrimLast) break;
                    ++i;
                }
            } else {
                for (;;) {
                    const float4 *spp = sc.sph + 2u * i;
                    float4 s0 = spp[0], s1 = spp[1];
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
        }
        if (sp == 0) break;
        --sp;
        cur = stack[sp * kBlock + tid];
    }
}

// world_hit, path_trace.wgsl:471-479: spheres first, then the mesh against
// [t_min, sphere t) — the sphere wins a tie.
template <int kBlock>
__device__ __forceinline__ Best world_hit(const DevScene &sc, V3 o, V3 d, float t_min,
                                          float t_max, uint32_t *stack, uint32_t &work) {
    V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    Best sb;
    sb.t = t_max;
    sb.id = kNone;
    sb.prim = kNone;
    sb.u = sb.v = 0.0f;
    traverse<kBlock, true>(sc, o, d, inv, t_min, t_max, sb, stack, work);
    float mesh_tmax = sb.prim != kNone ? sb.t : t_max;
    Best mb;
    mb.t = mesh_tmax;
    mb.id = kNone;
    mb.prim = kNone;
    mb.u = mb.v = 0.0f;
    traverse<kBlock, false>(sc, o, d, inv, t_min, mesh_tmax, mb, stack, work);
    if (mb.prim != kNone) return mb;
    return sb;
}

}  // namespace mcrt
