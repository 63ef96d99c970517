// nee.h — next-event estimation (NEE) with multiple importance sampling,
// opt-in.  No counterpart in the reference: path_trace.wgsl samples only the
// BSDF (:583-620) and reserves binding 12 (the emissive-sphere list,
// src/gpu_layout.rs:51-52, src/scene.rs:912-923) for this estimator.
//
// At every Lambertian vertex that continues the path, one emitter is chosen
// with probability proportional to area x luminance, one point (triangle:
// uniform in area) or direction (sphere: uniform in the cone it subtends) is
// sampled, and a shadow ray carries the contribution
//   thr * albedo/pi * Le * cos_x * w_light / pdf_light
// with the power heuristic w_light = pdf_light^2 / (pdf_light^2 + pdf_bsdf^2)
// against the cosine-weighted BSDF sampling (pdf_bsdf = cos_x / pi).  An
// emitter reached by BSDF sampling after a Lambertian vertex is weighted by
// w_bsdf = pdf_bsdf^2 / (pdf_bsdf^2 + pdf_light^2); after the camera or a
// specular event its weight is 1.  Emission is two-sided (:608-610), so the
// triangle pdf uses |cos| at the light.  Light samples draw from their own
// PCG stream (hashed from the path state), so the BSDF path is the reference
// estimator's draw for draw.
#pragma once

#include "shading.h"

namespace mcrt {

constexpr uint32_t kEmitSphere = 0x80000000u;
constexpr uint32_t kLightStream = 0x68E31DA4u;
constexpr float kInvPi = 0.31830988618379067f;

struct LightSample {
    V3 dir;
    float dist;  // distance to the sampled point on the emitter
    V3 Le;
    float pdf;   // solid-angle pdf, emitter choice included
};

__device__ __forceinline__ float power_heuristic(float a, float b) {
    const float a2 = a * a;
    return a2 / (a2 + b * b);
}

// Orthonormal basis around a unit vector (Duff et al., "Building an
// Orthonormal Basis, Revisited", JCGT 6(1), 2017).
__device__ __forceinline__ void onb(V3 n, V3 &t1, V3 &t2) {
    const float sg = copysignf(1.0f, n.z);
    const float a = -1.0f / (sg + n.z);
    const float b = n.x * n.y * a;
    t1 = v3(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x);
    t2 = v3(b, sg + n.y * n.y * a, -n.y);
}

// 1 - cos(theta_max) of the cone a sphere of squared radius r2 subtends at
// squared distance d2 (> r2), without cancellation.
__device__ __forceinline__ float cone_one_minus_cos(float r2, float d2) {
    const float s2 = r2 / d2;
    return s2 / (1.0f + sqrtf(1.0f - s2));
}

__device__ __forceinline__ uint32_t pick_emitter(const DevScene &sc, float u) {
    uint32_t lo = 0, hi = sc.n_emit - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (u < sc.emit_cdf[mid])
            hi = mid;
        else
            lo = mid + 1u;
    }
    return lo;
}

// Solid-angle pdf with which sample_light produces the direction d from x
// that reaches emitter e at distance t.
__device__ __forceinline__ float light_pdf(const DevScene &sc, uint32_t e, V3 x, V3 d, float t) {
    const float4 r0 = sc.emit[2u * e];
    const uint32_t ref = __float_as_uint(r0.y);
    if (ref & kEmitSphere) {
        const float4 s0 = sc.sph[2u * (ref & ~kEmitSphere)];
        const V3 w = xyz(s0) - x;
        const float d2 = dot(w, w), r2 = s0.w * s0.w;
        if (!(d2 > r2)) return 0.0f;
        return r0.x / (kTau * cone_one_minus_cos(r2, d2));
    }
    const float4 *tp = sc.mtris + sc.tri_stride * ref;
    const V3 ng = cross(xyz(tp[1]), xyz(tp[2]));
    const float cosl = fabsf(dot(ng, d)) / sqrtf(dot(ng, ng));
    if (!(cosl > 1e-7f)) return 0.0f;
    return r0.x * (t * t) / (r0.z * cosl);
}

__device__ __forceinline__ bool sample_light(const DevScene &sc, V3 x, uint32_t &lrng, LightSample &ls) {
    const float u0 = rand_f32(lrng), u1 = rand_f32(lrng), u2 = rand_f32(lrng);
    const uint32_t e = pick_emitter(sc, u0);
    const float4 r0 = sc.emit[2u * e], r1 = sc.emit[2u * e + 1u];
    ls.Le = xyz(r1);
    const uint32_t ref = __float_as_uint(r0.y);
    if (ref & kEmitSphere) {
        const float4 s0 = sc.sph[2u * (ref & ~kEmitSphere)];
        const V3 w = xyz(s0) - x;
        const float d2 = dot(w, w), r2 = s0.w * s0.w;
        if (!(d2 > r2)) return false;
        const float omc = cone_one_minus_cos(r2, d2);
        const float ct = 1.0f - u1 * omc;
        const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const float phi = kTau * u2;
        const V3 wn = (1.0f / sqrtf(d2)) * w;
        V3 t1, t2;
        onb(wn, t1, t2);
        ls.dir = normalize((st * cosf(phi)) * t1 + (st * sinf(phi)) * t2 + ct * wn);
        const float b = dot(ls.dir, w);
        ls.dist = b - sqrtf(fmaxf(r2 - (d2 - b * b), 0.0f));
        ls.pdf = r0.x / (kTau * omc);
        return ls.dist > 0.0f && ls.pdf > 0.0f && ls.pdf < 3.0e38f;
    }
    const float4 *tp = sc.mtris + sc.tri_stride * ref;
    const V3 v0 = xyz(tp[0]), e1 = xyz(tp[1]), e2 = xyz(tp[2]);
    const float su = sqrtf(u1);
    const float bu = su * (1.0f - u2), bv = su * u2;
    const V3 y = v0 + bu * e1 + bv * e2;
    const V3 w = y - x;
    const float d2 = dot(w, w);
    if (!(d2 > 0.0f)) return false;
    ls.dist = sqrtf(d2);
    ls.dir = (1.0f / ls.dist) * w;
    if (__float_as_uint(tp[1].w) & kPrimCull) {
        // A back-face-culled emitter is invisible from behind (:438-452): a
        // sample whose direction meets its back face contributes nothing.
        uint32_t mat;
        if (!(dot(ls.dir, tri_normal(sc, ref, bu, bv, mat)) < 0.0f)) return false;
    }
    const V3 ng = cross(e1, e2);
    const float cosl = fabsf(dot(ng, ls.dir)) / sqrtf(dot(ng, ng));
    if (!(cosl > 1e-7f)) return false;
    ls.pdf = r0.x * d2 / (r0.z * cosl);
    return ls.pdf < 3.0e38f;
}

// Emitter index of a hit primitive (kLeafBit marks a sphere), or kNone.
__device__ __forceinline__ uint32_t emitter_of(const DevScene &sc, uint32_t prim) {
    return (prim & kLeafBit) ? sc.sph_emit[prim & ~kLeafBit] : sc.tri_emit[prim];
}

}  // namespace mcrt
