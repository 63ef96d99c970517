// render_kernels.hip — HIP kernels of the render hot path for gfx950.
//
// The headless render of the reference is one cs_main megakernel dispatch per
// sample (src/headless.rs:271-289, path_trace.wgsl:626-662).  Here it is a
// wavefront pipeline over a batch of F frames x W*H pixels of paths:
//
//   k_generate    camera rays + PCG seeds for every (frame, pixel) path
//   k_extend      closest hit for every queued path (persistent, wave64
//                 chunks pulled from an atomic counter, LDS traversal stack)
//   k_shade       G-buffer, emission/miss termination, material scatter,
//                 wave-aggregated enqueue of surviving paths
//   k_accumulate  running mean over the batch's frames in frame order
//
// Every path draws from its own PCG state in the shader's order, and the
// accumulation applies mix(prev, s, 1/(min(f,65535)+1)) frame by frame, so
// the result does not depend on how paths are scheduled.
#include "dev_common.h"
#include "kernels.h"

namespace mcrt {

// Path state, structure of arrays of float4 (16-B coalesced accesses):
//   A = (origin.xyz, path id bits)   B = (dir.xyz, rng bits)
//   C = (throughput.xyz, bounce bits)
// Hit record H = (t, u, v, prim bits) with prim = kNone on a miss.

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

// ---------------------------------------------------------------------------
// k_generate: cs_main's seed, jitter and thin-lens ray (:633-652)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_generate(GenArgs a) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= a.n_paths) return;
    uint32_t pix = p % a.wh;
    uint32_t fi = p / a.wh;
    uint32_t frame = a.frame0 + fi;
    uint32_t x = pix % a.width, y = pix / a.width;
    uint32_t rng = pcg_hash(pcg_hash(x + y * a.width) ^ (frame * 2654435761u));
    float u = ((float)x + rand_f32(rng)) / (float)a.width;
    float v = 1.0f - ((float)y + rand_f32(rng)) / (float)a.height;
    const CameraUniform &c = a.cam;
    V3 focal = (ld3(c.lower_left) + u * ld3(c.horizontal)) + v * ld3(c.vertical);
    // random_in_unit_disk (:223-227): two draws even for a pinhole.
    float r = sqrtf(rand_f32(rng));
    float th = kTau * rand_f32(rng);
    float lx = r * cosf(th), ly = r * sinf(th);
    V3 lens = lx * ld3(c.defocus_u) + ly * ld3(c.defocus_v);
    V3 ro = ld3(c.origin) + lens;
    V3 rd = normalize(focal - ro);
    a.A[p] = make_float4(ro.x, ro.y, ro.z, __uint_as_float(p));
    a.B[p] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(rng));
    a.C[p] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
}

// ---------------------------------------------------------------------------
// Persistent work distribution: each wave takes 64 consecutive slots per
// atomic.  All lanes of a wave leave the loop together.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t grab64(uint32_t *work) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(work, 64u);
    return __shfl(base, 0);
}

// ---------------------------------------------------------------------------
// k_extend: closest hit of every queued path (world_hit at [T_MIN, T_MAX))
// ---------------------------------------------------------------------------
struct ExtIO {
    const ExtArgs &a;
    __device__ uint32_t count() const { return *a.count; }
    __device__ uint32_t *work() const { return a.work; }
    __device__ void load(uint32_t slot, V3 &o, V3 &d, float &tmin, float &tmax) const {
        float4 A = a.A[slot], B = a.B[slot];
        o = xyz(A);
        d = xyz(B);
        tmin = MCRT_T_MIN;
        tmax = MCRT_T_MAX;
    }
    __device__ void store(uint32_t slot, const Best &b) const {
        a.H[slot] = b.prim != kNone ? make_float4(b.t, b.u, b.v, __uint_as_float(b.prim))
                                    : make_float4(-1.0f, 0.0f, 0.0f, __uint_as_float(kNone));
    }
};

template <int kBlock, bool kSpheres>
__global__ __launch_bounds__(kBlock) void k_extend(ExtArgs a) {
    __shared__ uint32_t lds[kShortStack * kBlock];
    trace_persistent<kBlock, kSpheres>(a.scene, ExtIO{a}, lds, a.ovf, a.ovf_stride);
}

// ---------------------------------------------------------------------------
// k_shade: one bounce of path_color (:589-617) for every queued path
// ---------------------------------------------------------------------------
struct Surface {
    V3 p, n;
    bool front;
    uint32_t mat;
};

__device__ __forceinline__ Surface surface(const DevScene &sc, V3 o, V3 d, float4 h) {
    Surface s;
    uint32_t prim = __float_as_uint(h.w);
    float t = h.x;
    s.p = o + t * d;
    if (prim & kLeafBit) {  // sphere, :282-287
        float4 s0 = sc.sph[2u * (prim & ~kLeafBit)];
        float4 s1 = sc.sph[2u * (prim & ~kLeafBit) + 1u];
        V3 c = xyz(s0);
        float r = s0.w;
        V3 outn = v3((s.p.x - c.x) / r, (s.p.y - c.y) / r, (s.p.z - c.z) / r);
        s.front = dot(d, outn) < 0.0f;
        s.n = s.front ? outn : -outn;
        s.mat = __float_as_uint(s1.x);
    } else {  // triangle, :445-458
        uint32_t mat;
        V3 n = tri_normal(sc, prim, h.y, h.z, mat);
        s.front = dot(d, n) < 0.0f;
        s.n = s.front ? n : -n;
        s.mat = mat & 0x7FFFFFFFu;
    }
    return s;
}

__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
    const uint32_t count = *a.count_in;
    for (;;) {
        uint32_t base = grab64(a.work);
        if (base >= count) break;
        uint32_t slot = base + __lane_id();
        bool active = slot < count;
        bool alive = false;
        float4 nA, nB, nC;
        if (active) {
            float4 A = a.A[slot], B = a.B[slot], C = a.C[slot], H = a.H[slot];
            V3 o = xyz(A), d = xyz(B), thr = xyz(C);
            uint32_t pid = __float_as_uint(A.w);
            uint32_t rng = __float_as_uint(B.w);
            uint32_t bounce = __float_as_uint(C.w);
            bool first = bounce == 0u;
            uint32_t pix = pid % a.wh;
            bool gb = first && a.gbuf != nullptr && (pid / a.wh) == a.gbuf_frame_index;
            V3 radiance = v3(0.0f, 0.0f, 0.0f);
            bool done = true;
            if (__float_as_uint(H.w) == kNone) {  // miss, :592-597
                if (gb) a.gbuf[pix] = make_float4(0.0f, 0.0f, 0.0f, MCRT_T_MAX);
                radiance = thr * sample_env(a.scene, d);
            } else {
                Surface s = surface(a.scene, o, d, H);
                if (gb) a.gbuf[pix] = make_float4(s.n.x, s.n.y, s.n.z, H.x);
                uint32_t mi = s.mat < 63u ? s.mat : 63u;
                const GpuMaterial &m = a.scene.mats[mi];
                uint32_t type = m.type_pad[0];
                V3 albedo = ld3(m.albedo_fuzz);
                if (type == MCRT_MAT_EMISSIVE) {  // :608-610
                    radiance = m.ior_pad[0] * (thr * albedo);
                } else {
                    V3 dir, att;
                    bool absorbed = false;
                    if (type == MCRT_MAT_METAL) {  // :530-541
                        float fuzz = m.albedo_fuzz[3];
                        V3 refl = reflect_vec(normalize(d), s.n);
                        V3 pert = refl + fuzz * random_unit_sphere(rng);
                        dir = normalize(dot(pert, pert) >= 1e-6f ? pert : refl);
                        att = albedo;
                        absorbed = dot(dir, s.n) <= 0.0f;
                    } else if (type == MCRT_MAT_DIELECTRIC) {  // :543-564
                        att = v3(1.0f, 1.0f, 1.0f);
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
                            dir = reflect_vec(d, s.n);
                        else
                            dir = normalize(refract_vec(d, s.n, eta));
                    } else {  // lambertian :521-528
                        dir = cosine_scatter(rng, s.n);
                        att = albedo;
                    }
                    if (!absorbed) {
                        thr = thr * att;
                        uint32_t nb = bounce + 1u;
                        if (nb < a.max_depth) {
                            done = false;
                            nA = make_float4(s.p.x, s.p.y, s.p.z, A.w);
                            nB = make_float4(dir.x, dir.y, dir.z, __uint_as_float(rng));
                            nC = make_float4(thr.x, thr.y, thr.z, __uint_as_float(nb));
                        }
                    }
                }
            }
            if (done) a.samples[pid] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
            alive = !done;
        }
        // Wave-aggregated enqueue: one atomic per wave.
        uint64_t mask = __ballot(alive);
        uint32_t n_alive = (uint32_t)__popcll(mask);
        if (n_alive) {
            uint32_t obase = 0;
            if (__lane_id() == 0) obase = atomicAdd(a.count_out, n_alive);
            obase = __shfl(obase, 0);
            uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            if (alive) {
                uint32_t o = obase + below;
                a.A_out[o] = nA;
                a.B_out[o] = nB;
                a.C_out[o] = nC;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_accumulate: running mean (:657-661) over the batch's frames in order
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mixf(float x, float y, float w) { return x * (1.0f - w) + y * w; }

__global__ __launch_bounds__(256) void k_accumulate(AccArgs a) {
    uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= a.wh) return;
    float4 acc = a.accum_in ? a.accum_in[pix] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < a.n_frames; ++i) {
        uint32_t f = a.frame0 + i;
        uint32_t fc = f < MCRT_MAX_FRAME_COUNT ? f : MCRT_MAX_FRAME_COUNT;
        float w = 1.0f / (float)(fc + 1u);
        float4 s = a.samples[(size_t)i * a.wh + pix];
        acc.x = mixf(acc.x, s.x, w);
        acc.y = mixf(acc.y, s.y, w);
        acc.z = mixf(acc.z, s.z, w);
    }
    a.accum_out[pix] = make_float4(acc.x, acc.y, acc.z, 1.0f);
}

// ---------------------------------------------------------------------------
// k_trace: standalone closest-hit query (mcrt_trace_closest)
// ---------------------------------------------------------------------------
struct TraceIO {
    const TraceArgs &a;
    __device__ uint32_t count() const { return a.n; }
    __device__ uint32_t *work() const { return a.work; }
    __device__ void load(uint32_t i, V3 &o, V3 &d, float &tmin, float &tmax) const {
        o = v3(a.origins[3u * i], a.origins[3u * i + 1u], a.origins[3u * i + 2u]);
        d = v3(a.dirs[3u * i], a.dirs[3u * i + 1u], a.dirs[3u * i + 2u]);
        tmin = a.tmin;
        tmax = a.tmax;
    }
    __device__ void store(uint32_t i, const Best &b) const {
        a.out_id[i] = b.prim != kNone ? b.id : MCRT_ID_MISS;
        a.out_t[i] = b.prim != kNone ? b.t : -1.0f;
    }
};

template <int kBlock, bool kSpheres>
__global__ __launch_bounds__(kBlock) void k_trace(TraceArgs a) {
    __shared__ uint32_t lds[kShortStack * kBlock];
    trace_persistent<kBlock, kSpheres>(a.scene, TraceIO{a}, lds, a.ovf, a.ovf_stride);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
constexpr int kTraceBlock = 256;

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_generate(const GenArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_generate, dim3((a.n_paths + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_extend(const ExtArgs &a, uint32_t grid, hipStream_t s) {
    if (a.scene.sroot != kNone)
        hipLaunchKernelGGL((k_extend<kTraceBlock, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_extend<kTraceBlock, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_shade(const ShadeArgs &a, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_accumulate(const AccArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3((a.wh + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_trace(const TraceArgs &a, uint32_t grid, hipStream_t s) {
    if (a.scene.sroot != kNone)
        hipLaunchKernelGGL((k_trace<kTraceBlock, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_trace<kTraceBlock, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t occupancy(int which, bool spheres, int *blocks_per_cu) {
    const void *f = nullptr;
    if (which == 0)
        f = spheres ? reinterpret_cast<const void *>(&k_extend<kTraceBlock, true>)
                    : reinterpret_cast<const void *>(&k_extend<kTraceBlock, false>);
    else if (which == 1)
        f = reinterpret_cast<const void *>(&k_shade);
    else
        f = spheres ? reinterpret_cast<const void *>(&k_trace<kTraceBlock, true>)
                    : reinterpret_cast<const void *>(&k_trace<kTraceBlock, false>);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, which == 1 ? 256 : kTraceBlock, 0);
}

uint32_t short_stack() { return kShortStack; }

int trace_block() { return kTraceBlock; }

}  // namespace mcrt
