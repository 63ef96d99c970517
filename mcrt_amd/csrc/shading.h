// shading.h — camera rays, sampling, environment lookup, surface records and
// material scatter of src/shaders/path_trace.wgsl, shared by the wavefront
// render kernels and the differentiable (adjoint) kernels.
#pragma once

#include "dev_common.h"

namespace mcrt {

// ---------------------------------------------------------------------------
// Sampling helpers, path_trace.wgsl:210-227
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 random_unit_sphere(uint32_t &rng) {
    float cos_phi = 1.0f - 2.0f * rand_f32(rng);
    float sin_phi = sqrtf(fmaxf(0.0f, 1.0f - cos_phi * cos_phi));
    float theta = kTau * rand_f32(rng);
    return v3(sin_phi * cosf(theta), sin_phi * sinf(theta), cos_phi);
}

__device__ __forceinline__ V3 cosine_scatter(uint32_t &rng, V3 n) {
    V3 s = n + random_unit_sphere(rng);
    if (dot(s, s) < 1e-6f) return n;
    return normalize(s);
}

__device__ __forceinline__ V3 reflect_vec(V3 v, V3 n) { return v - (2.0f * dot(v, n)) * n; }

__device__ __forceinline__ V3 refract_vec(V3 v, V3 n, float eta) {
    float cos_theta = fminf(dot(-v, n), 1.0f);
    V3 r_perp = eta * (v + cos_theta * n);
    V3 r_para = (-sqrtf(fmaxf(0.0f, 1.0f - dot(r_perp, r_perp)))) * n;
    return r_perp + r_para;
}

__device__ __forceinline__ float schlick(float cos_theta, float eta) {
    float r0 = (1.0f - eta) / (1.0f + eta);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * powf(1.0f - cos_theta, 5.0f);
}

// sample_env_map, path_trace.wgsl:233-242.
__device__ __forceinline__ V3 sample_env(const DevScene &sc, V3 dir) {
    float phi = atan2f(dir.z, dir.x);
    float theta = asinf(fminf(fmaxf(dir.y, -1.0f), 1.0f));
    float u = phi / kTau + 0.5f;
    float v = 0.5f - theta / kPi;
    uint32_t ix = f2u_sat(u * (float)sc.env_w) % sc.env_w;
    uint32_t iy = f2u_sat(v * (float)sc.env_h) % sc.env_h;
    return xyz(sc.env[(size_t)iy * sc.env_w + ix]);
}

// cs_main's seed, sub-pixel jitter and thin-lens ray (:633-652).  The lens
// disk always consumes two draws, even for a pinhole (:223-227).
__device__ __forceinline__ void camera_ray(const CameraUniform &c, uint32_t width, uint32_t height,
                                           uint32_t x, uint32_t y, uint32_t frame, uint32_t &rng,
                                           V3 &ro, V3 &rd) {
    rng = pcg_hash(pcg_hash(x + y * width) ^ (frame * 2654435761u));
    float u = ((float)x + rand_f32(rng)) / (float)width;
    float v = 1.0f - ((float)y + rand_f32(rng)) / (float)height;
    V3 focal = (ld3(c.lower_left) + u * ld3(c.horizontal)) + v * ld3(c.vertical);
    float r = sqrtf(rand_f32(rng));
    float th = kTau * rand_f32(rng);
    float lx = r * cosf(th), ly = r * sinf(th);
    V3 lens = lx * ld3(c.defocus_u) + ly * ld3(c.defocus_v);
    ro = ld3(c.origin) + lens;
    rd = normalize(focal - ro);
}

// camera_ray for a pinhole camera: every defocus component is zero and no
// origin component is -0.  The lens offset lx * defocus_u + ly * defocus_v is
// then a zero for every finite (lx, ly), and origin + that zero is the origin
// bit for bit, so the ray is the one camera_ray returns without the lens
// trigonometry.  The two lens draws are still consumed.
__device__ __forceinline__ void pinhole_ray(const CameraUniform &c, uint32_t width, uint32_t height, uint32_t x,
                                            uint32_t y, uint32_t frame, uint32_t &rng, V3 &ro, V3 &rd) {
    rng = pcg_hash(pcg_hash(x + y * width) ^ (frame * 2654435761u));
    float u = ((float)x + rand_f32(rng)) / (float)width;
    float v = 1.0f - ((float)y + rand_f32(rng)) / (float)height;
    V3 focal = (ld3(c.lower_left) + u * ld3(c.horizontal)) + v * ld3(c.vertical);
    rand_f32(rng);
    rand_f32(rng);
    ro = ld3(c.origin);
    rd = normalize(focal - ro);
}

// ---------------------------------------------------------------------------
// Hit record of a closest hit (:282-287 spheres, :445-459 triangles)
// ---------------------------------------------------------------------------
struct Surface {
    V3 p, n;
    bool front;
    uint32_t mat;
};

__device__ __forceinline__ Surface surface(const DevScene &sc, V3 o, V3 d, float t, float u,
                                           float v, uint32_t prim) {
    Surface s;
    s.p = o + t * d;
    if (prim & kLeafBit) {
        float4 s0 = sc.sph[2u * (prim & ~kLeafBit)];
        float4 s1 = sc.sph[2u * (prim & ~kLeafBit) + 1u];
        V3 c = xyz(s0);
        float r = s0.w;
        V3 outn = v3((s.p.x - c.x) / r, (s.p.y - c.y) / r, (s.p.z - c.z) / r);
        s.front = dot(d, outn) < 0.0f;
        s.n = s.front ? outn : -outn;
        s.mat = __float_as_uint(s1.x);
    } else {
        uint32_t mat;
        V3 n = tri_normal(sc, prim, u, v, mat);
        s.front = dot(d, n) < 0.0f;
        s.n = s.front ? n : -n;
        s.mat = mat & 0x7FFFFFFFu;
    }
    return s;
}

// ---------------------------------------------------------------------------
// scatter, path_trace.wgsl:521-573
// ---------------------------------------------------------------------------
struct Scattered {
    V3 dir, att;
    bool absorbed;
};

__device__ __forceinline__ Scattered scatter_material(const GpuMaterial &m, V3 d, const Surface &s,
                                                      uint32_t &rng) {
    Scattered r;
    r.absorbed = false;
    const uint32_t type = m.type_pad[0];
    const V3 albedo = ld3(m.albedo_fuzz);
    if (type == MCRT_MAT_METAL) {  // :530-541
        float fuzz = m.albedo_fuzz[3];
        V3 refl = reflect_vec(normalize(d), s.n);
        V3 pert = refl + fuzz * random_unit_sphere(rng);
        r.dir = normalize(dot(pert, pert) >= 1e-6f ? pert : refl);
        r.att = albedo;
        r.absorbed = dot(r.dir, s.n) <= 0.0f;
    } else if (type == MCRT_MAT_DIELECTRIC) {  // :543-564
        r.att = v3(1.0f, 1.0f, 1.0f);
        float ior = m.ior_pad[0];
        float eta = s.front ? 1.0f / ior : ior;
        float cos_theta = fminf(dot(-d, s.n), 1.0f);
        float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
        bool cannot = eta * sin_theta > 1.0f;
        float sc = (!s.front && !cannot)
                       ? sqrtf(fmaxf(0.0f, 1.0f - eta * eta * (1.0f - cos_theta * cos_theta)))
                       : cos_theta;
        float refl = schlick(sc, eta);
        if (cannot || refl > rand_f32(rng))
            r.dir = reflect_vec(d, s.n);
        else
            r.dir = normalize(refract_vec(d, s.n, eta));
    } else {  // lambertian :521-528
        r.dir = cosine_scatter(rng, s.n);
        r.att = albedo;
    }
    return r;
}

}  // namespace mcrt
