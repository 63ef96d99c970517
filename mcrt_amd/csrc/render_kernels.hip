// render_kernels.hip — HIP kernels of the render hot path for gfx950.
//
// The headless render of the reference is one cs_main megakernel dispatch per
// sample (src/headless.rs:271-289, path_trace.wgsl:626-662).  Here it is a
// wavefront pipeline over a batch of F frames x W*H pixels of paths:
//
//   k_generate    camera rays + PCG seeds for every (frame, pixel) path
//                 (for a pinhole camera without next-event estimation the
//                 first k_shade, and in cache-resident scenes the first
//                 k_extend, recompute them from the queue slot instead)
//   k_extend      closest hit for every queued path (persistent, wave64, a
//                 static first chunk per wave then slots pulled from an
//                 atomic counter, lane refill, LDS short stack; dev_common.h)
//   k_shade       G-buffer, emission/miss termination, material scatter,
//                 block-aggregated enqueue of surviving paths
//   k_accumulate  running mean over the batch's frames in frame order
//
// Every path draws from its own PCG state in the shader's order, and the
// accumulation applies mix(prev, s, 1/(min(f,65535)+1)) frame by frame, so
// the result does not depend on how paths are scheduled.
#include "kernels.h"
#include "nee.h"
#include "shading.h"

namespace mcrt {

// Path state, structure of arrays of float4 (16-B coalesced accesses):
//   A = (origin.xyz, path id bits)   B = (dir.xyz, rng bits)
//   C = (throughput.xyz, bounce bits)
// Hit record H = (t, u, v, prim bits) with prim = kNone on a miss.

// ---------------------------------------------------------------------------
// k_generate: cs_main's seed, jitter and thin-lens ray (:633-652)
// ---------------------------------------------------------------------------
// Queue order is pixel-major: the batch's frames of one pixel are adjacent,
// so a wave starts from a few neighbouring pixels, and the bounce queues keep
// that order through compaction.  The path id stays frame * wh + row-major
// pixel.
__device__ __forceinline__ void cam_path(const CameraUniform &cam, uint32_t width, uint32_t height, uint32_t wh,
                                         uint32_t frame0, uint32_t nf, uint32_t fo, uint32_t po, uint32_t p,
                                         uint32_t &pid, uint32_t &rng, V3 &ro, V3 &rd) {
    const uint32_t pix = po + p / nf, fi = p % nf;
    pid = (fo + fi) * wh + pix;
    camera_ray(cam, width, height, pix % width, pix / width, frame0 + fi, rng, ro, rd);
}

__device__ __forceinline__ void cam_path(const CamGen &g, uint32_t p, uint32_t &pid, uint32_t &rng, V3 &ro, V3 &rd) {
    cam_path(g.cam, g.width, g.height, g.wh, g.frame0, g.nf, g.fo, g.po, p, pid, rng, ro, rd);
}

__global__ __launch_bounds__(256) void k_generate(GenArgs a) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= a.n_paths) return;
    uint32_t pid, rng;
    V3 ro, rd;
    cam_path(a.cam, a.width, a.height, a.wh, a.frame0, a.n_paths / a.pn, a.fo, a.po, p, pid, rng, ro, rd);
    a.A[p] = make_float4(ro.x, ro.y, ro.z, __uint_as_float(pid));
    a.B[p] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(rng));
    if (a.C) a.C[p] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
    if (a.samples) a.samples[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (a.P) a.P[p] = 0.0f;
}

// ---------------------------------------------------------------------------
// k_extend: closest hit of every queued path (world_hit at [T_MIN, T_MAX))
// ---------------------------------------------------------------------------
// kGen: the launch may recompute first-bounce camera rays (a.gen.on).  Only
// the first-bounce packet kernel of cache-resident, sphere-free scenes does
// (k_extend_packet<.., true>); the per-lane kernels read stored rays (the
// camera arithmetic at their refill point cost the 1080p voxel terrain 2%).
template <bool kGen>
struct ExtIO {
    const ExtArgs &a;
    __device__ uint32_t count() const { return *a.count; }
    __device__ uint32_t *work() const { return a.work; }
    __device__ void load(uint32_t slot, V3 &o, V3 &d, float &tmin, float &tmax) const {
        if (kGen && a.gen.on) {  // first bounce of a pinhole camera
            const CamGen &g = a.gen;
            const uint32_t pix = g.po + slot / g.nf, fi = slot % g.nf;
            uint32_t rng;
            pinhole_ray(g.cam, g.width, g.height, pix % g.width, pix / g.width, g.frame0 + fi, rng, o, d);
        } else {
            float4 A = a.A[slot], B = a.B[slot];
            o = xyz(A);
            d = xyz(B);
        }
        tmin = MCRT_T_MIN;
        tmax = MCRT_T_MAX;
    }
    static constexpr bool kAnyHit = false;
    __device__ bool any_hit(uint32_t) const { return false; }
    __device__ void store(uint32_t slot, const Best &b) const {
        a.H[slot] = b.prim != kNone ? make_float4(b.t, b.u, b.v, __uint_as_float(b.prim))
                                    : make_float4(-1.0f, 0.0f, 0.0f, __uint_as_float(kNone));
    }
};

// Extension rays [0, n_ext) followed by the previous bounce's shadow rays in
// one launch (next-event estimation: one persistent drain per bounce instead
// of two).  A shadow ray adds its contribution to its path's sample when
// nothing lies before the light, as k_occlude does.
struct ExtOccIO {
    const ExtArgs &a;
    uint32_t n_ext;
    __device__ uint32_t count() const { return n_ext + *a.occ_count; }
    __device__ uint32_t *work() const { return a.work; }
    __device__ void load(uint32_t slot, V3 &o, V3 &d, float &tmin, float &tmax) const {
        tmin = MCRT_T_MIN;
        if (slot < n_ext) {
            float4 A = a.A[slot], B = a.B[slot];
            o = xyz(A);
            d = xyz(B);
            tmax = MCRT_T_MAX;
        } else {
            const float4 so = a.S_o[slot - n_ext], sd = a.S_d[slot - n_ext];
            o = xyz(so);
            d = xyz(sd);
            tmax = sd.w;
        }
    }
    static constexpr bool kAnyHit = true;
    __device__ bool any_hit(uint32_t slot) const { return slot >= n_ext; }  // the shadow rays
    __device__ void store(uint32_t slot, const Best &b) const {
        if (slot < n_ext) {
            a.H[slot] = b.prim != kNone ? make_float4(b.t, b.u, b.v, __uint_as_float(b.prim))
                                        : make_float4(-1.0f, 0.0f, 0.0f, __uint_as_float(kNone));
        } else if (b.prim == kNone) {
            const uint32_t i = slot - n_ext;
            const uint32_t pid = __float_as_uint(a.S_o[i].w);
            const float4 c = a.S_c[i], s = a.samples[pid];
            a.samples[pid] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, 0.0f);
        }
    }
};

// Like k_extend, the sphere-free variants run 8 waves per SIMD (64 VGPRs, no
// spill; the 1080p icosphere frame with next-event estimation: 1 spp 3.98 ->
// 3.88 ms, 4 spp 13.05 -> 11.96 ms against the compiler's 7).
template <int kBlock, bool kSpheres, bool kCull>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kSpheres ? 1 : 8, 8)))
void k_extend_occ(ExtArgs a) {
    __shared__ uint32_t lds[TStack<kCull>::kShort * TStack<kCull>::kWords * kBlock];
    __shared__ float4 ncache[kCull ? 1 : kNodeCache * 4];
    trace_persistent<kBlock, kSpheres, kCull>(a.scene, ExtOccIO{a, *a.count}, lds, a.ovf, a.ovf_stride, kCull ? nullptr : ncache);
}

// The sphere-free variants fit 64 VGPRs and run 8 waves per SIMD: the plain
// stack (cache-resident scenes; icosphere Cornell +1.2% over 7) and the
// culling stack over quantized pairs (voxel terrain 1236 -> 1266
// Msamples/s; with 64-B pairs, 8 waves had cost it 3%: more rays in flight
// competed for the cache).  The sphere variants keep the compiler's choice
// (forcing 8 there spills).
template <int kBlock, bool kSpheres, bool kCull>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kSpheres ? 1 : 8, 8)))
void k_extend(ExtArgs a) {
    __shared__ uint32_t lds[TStack<kCull>::kShort * TStack<kCull>::kWords * kBlock];
    __shared__ float4 ncache[kCull ? 1 : kNodeCache * 4];
    trace_persistent<kBlock, kSpheres, kCull>(a.scene, ExtIO<false>{a}, lds, a.ovf, a.ovf_stride, kCull ? nullptr : ncache);
}

// ---------------------------------------------------------------------------
// k_shade: one bounce of path_color (:589-617) for every queued path.
// One thread per queued slot; the grid covers the largest possible queue and
// blocks past the live count return at once.  Survivors are compacted into
// the next queue with one atomic per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kShadeBlock) void k_shade(ShadeArgs a) {
    constexpr uint32_t kWaves = kShadeBlock / 64u;
    __shared__ uint32_t s_wave[kWaves], s_swave[kWaves];
    __shared__ uint32_t s_base, s_sbase;
    const uint32_t count = *a.count_in;
    const uint32_t base = blockIdx.x * kShadeBlock;
    if (base >= count) return;
    const uint32_t slot = base + threadIdx.x;
    bool alive = false, shadow = false;
    float4 nA, nB, nC, sO, sD, sC;
    float n_pdf = 0.0f;
    if (slot < count) {
        float4 A, B, C;
        if (a.gen.on) {  // camera ray of a first-bounce queue, as k_generate writes it
            uint32_t pid0, rng0;
            V3 ro, rd;
            cam_path(a.gen, slot, pid0, rng0, ro, rd);
            A = make_float4(ro.x, ro.y, ro.z, __uint_as_float(pid0));
            B = make_float4(rd.x, rd.y, rd.z, __uint_as_float(rng0));
            C = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
        } else {
            A = a.A[slot];
            B = a.B[slot];
            C = a.C[slot];
        }
        const float4 H = a.H[slot];
        V3 o = xyz(A), d = xyz(B), thr = xyz(C);
        uint32_t pid = __float_as_uint(A.w);
        uint32_t rng = __float_as_uint(B.w);
        uint32_t bounce = __float_as_uint(C.w);
        float4 *gbuf = nullptr;
        uint32_t pix = 0;
        if (bounce == 0u && (a.gbuf != nullptr || a.albedo != nullptr) && pid / a.wh == a.gbuf_frame_index) {
            gbuf = a.gbuf;
            pix = pid % a.wh;
        }
        const bool aux = gbuf != nullptr || (bounce == 0u && a.albedo != nullptr && pid / a.wh == a.gbuf_frame_index);
        V3 radiance = v3(0.0f, 0.0f, 0.0f);
        bool done = true;
        if (__float_as_uint(H.w) == kNone) {  // miss, :592-597
            if (gbuf) gbuf[pix] = make_float4(0.0f, 0.0f, 0.0f, MCRT_T_MAX);
            if (aux && a.albedo) a.albedo[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            radiance = thr * sample_env(a.scene, d);
        } else {
            Surface s = surface(a.scene, o, d, H.x, H.y, H.z, __float_as_uint(H.w));
            if (gbuf) gbuf[pix] = make_float4(s.n.x, s.n.y, s.n.z, H.x);  // :599-602
            const GpuMaterial &m = a.scene.mats[s.mat < 63u ? s.mat : 63u];
            if (aux && a.albedo) {  // first-hit albedo (dielectric: 1)
                V3 al = m.type_pad[0] == MCRT_MAT_DIELECTRIC ? v3(1.0f, 1.0f, 1.0f) : ld3(m.albedo_fuzz);
                a.albedo[pix] = make_float4(al.x, al.y, al.z, 1.0f);
            }
            if (m.type_pad[0] == MCRT_MAT_EMISSIVE) {  // :608-610
                radiance = m.ior_pad[0] * (thr * ld3(m.albedo_fuzz));
                if (a.nee) {  // MIS weight of a BSDF-sampled emitter hit
                    const float pb = a.P[slot];
                    const uint32_t e =
                        (pb > 0.0f && a.scene.n_emit) ? emitter_of(a.scene, __float_as_uint(H.w)) : kNone;
                    if (e != kNone) radiance = power_heuristic(pb, light_pdf(a.scene, e, o, d, H.x)) * radiance;
                }
            } else {
                Scattered sc = scatter_material(m, d, s, rng);
                if (!sc.absorbed) {
                    thr = thr * sc.att;
                    uint32_t nb = bounce + 1u;
                    if (nb < a.max_depth) {
                        done = false;
                        nA = make_float4(s.p.x, s.p.y, s.p.z, A.w);
                        nB = make_float4(sc.dir.x, sc.dir.y, sc.dir.z, __uint_as_float(rng));
                        nC = make_float4(thr.x, thr.y, thr.z, __uint_as_float(nb));
                        const bool lambert = m.type_pad[0] != MCRT_MAT_METAL &&
                                             m.type_pad[0] != MCRT_MAT_DIELECTRIC;
                        if (a.nee && lambert) {
                            n_pdf = fmaxf(dot(s.n, sc.dir), 0.0f) * kInvPi;
                            LightSample ls;
                            uint32_t lrng = pcg_hash(rng ^ kLightStream);
                            if (a.scene.n_emit && sample_light(a.scene, s.p, lrng, ls)) {
                                const float cosx = dot(s.n, ls.dir);
                                if (cosx > 0.0f) {
                                    // thr already holds thr * albedo.
                                    const float wgt = power_heuristic(ls.pdf, cosx * kInvPi) * cosx * kInvPi / ls.pdf;
                                    const V3 c = wgt * (thr * ls.Le);
                                    shadow = true;
                                    sO = make_float4(s.p.x, s.p.y, s.p.z, A.w);
                                    sD = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, ls.dist * (1.0f - 1e-3f));
                                    sC = make_float4(c.x, c.y, c.z, 0.0f);
                                }
                            }
                        }
                    }
                }
            }
        }
        if (done) {
            if (!a.nee) {
                a.samples[pid] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
            } else if (radiance.x != 0.0f || radiance.y != 0.0f || radiance.z != 0.0f) {
                float4 acc = a.samples[pid];
                a.samples[pid] = make_float4(acc.x + radiance.x, acc.y + radiance.y, acc.z + radiance.z, 0.0f);
            }
        }
        alive = !done;
    }
    // Block-aggregated enqueue.
    const uint64_t mask = __ballot(alive);
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) s_wave[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kWaves; ++w) total += s_wave[w];
        s_base = total ? atomicAdd(a.count_out, total) : 0u;
    }
    __syncthreads();
    if (alive) {
        uint32_t o = s_base;
        for (uint32_t w = 0; w < wave; ++w) o += s_wave[w];
        o += __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        a.A_out[o] = nA;
        a.B_out[o] = nB;
        a.C_out[o] = nC;
        if (a.nee) a.P_out[o] = n_pdf;
    }
    if (!a.nee) return;
    // Shadow rays, block-aggregated the same way.
    const uint64_t smask = __ballot(shadow);
    if (__lane_id() == 0) s_swave[wave] = (uint32_t)__popcll(smask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kWaves; ++w) total += s_swave[w];
        s_sbase = total ? atomicAdd(a.shadow_count, total) : 0u;
    }
    __syncthreads();
    if (shadow) {
        uint32_t o = s_sbase;
        for (uint32_t w = 0; w < wave; ++w) o += s_swave[w];
        o += __builtin_amdgcn_mbcnt_hi((uint32_t)(smask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)smask, 0u));
        a.S_o[o] = sO;
        a.S_d[o] = sD;
        a.S_c[o] = sC;
    }
}

// ---------------------------------------------------------------------------
// k_occlude: next-event shadow rays; a visible light sample adds its
// contribution to its path's sample (one shadow ray per path per bounce).
// ---------------------------------------------------------------------------
struct OccIO {
    const OccArgs &a;
    __device__ uint32_t count() const { return *a.count; }
    __device__ uint32_t *work() const { return a.work; }
    __device__ void load(uint32_t i, V3 &o, V3 &d, float &tmin, float &tmax) const {
        const float4 so = a.S_o[i], sd = a.S_d[i];
        o = xyz(so);
        d = xyz(sd);
        tmin = MCRT_T_MIN;
        tmax = sd.w;
    }
    static constexpr bool kAnyHit = true;
    __device__ bool any_hit(uint32_t) const { return true; }
    __device__ void store(uint32_t i, const Best &b) const {
        if (b.prim != kNone) return;
        const uint32_t pid = __float_as_uint(a.S_o[i].w);
        const float4 c = a.S_c[i], s = a.samples[pid];
        a.samples[pid] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, 0.0f);
    }
};

template <int kBlock, bool kSpheres, bool kCull>
__global__ __launch_bounds__(kBlock) void k_occlude(OccArgs a) {
    __shared__ uint32_t lds[TStack<kCull>::kShort * TStack<kCull>::kWords * kBlock];
    __shared__ float4 ncache[kCull ? 1 : kNodeCache * 4];
    trace_persistent<kBlock, kSpheres, kCull>(a.scene, OccIO{a}, lds, a.ovf, a.ovf_stride, kCull ? nullptr : ncache);
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

// k_extend_packet: first bounce of a sphere-free scene, traced as
// wave-uniform packets (trace_packets).  kGen: the launch recomputes pinhole
// camera rays (a.gen.on) instead of reading them.
template <int kBlock, bool kGen>
__global__ __launch_bounds__(kBlock) void k_extend_packet(ExtArgs a) {
    trace_packets<kBlock>(a.scene, ExtIO<kGen>{a});
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
    static constexpr bool kAnyHit = false;
    __device__ bool any_hit(uint32_t) const { return false; }
    __device__ void store(uint32_t i, const Best &b) const {
        a.out_id[i] = b.prim != kNone ? b.id : MCRT_ID_MISS;
        a.out_t[i] = b.prim != kNone ? b.t : -1.0f;
    }
};

template <int kBlock, bool kSpheres, bool kCull>
__global__ __launch_bounds__(kBlock) void k_trace(TraceArgs a) {
    __shared__ uint32_t lds[TStack<kCull>::kShort * TStack<kCull>::kWords * kBlock];
    __shared__ float4 ncache[kCull ? 1 : kNodeCache * 4];
    trace_persistent<kBlock, kSpheres, kCull>(a.scene, TraceIO{a}, lds, a.ovf, a.ovf_stride, kCull ? nullptr : ncache);
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

// Traversal kernel variant for a scene: with or without the sphere phase,
// with the plain or the culling stack.
template <int kBlock, bool kSpheres, bool kCull>
struct ExtendK {
    static const void *fn() { return reinterpret_cast<const void *>(&k_extend<kBlock, kSpheres, kCull>); }
};
template <int kBlock, bool kSpheres, bool kCull>
struct ExtendOccK {
    static const void *fn() { return reinterpret_cast<const void *>(&k_extend_occ<kBlock, kSpheres, kCull>); }
};
template <int kBlock, bool kSpheres, bool kCull>
struct TraceK {
    static const void *fn() { return reinterpret_cast<const void *>(&k_trace<kBlock, kSpheres, kCull>); }
};
template <int kBlock, bool kSpheres, bool kCull>
struct OccludeK {
    static const void *fn() { return reinterpret_cast<const void *>(&k_occlude<kBlock, kSpheres, kCull>); }
};

template <template <int, bool, bool> class K>
const void *traversal_kernel(const DevScene &sc) {
    const bool sph = sc.sroot != kNone;
    if (sc.stack_cull) return sph ? K<kTraceBlock, true, true>::fn() : K<kTraceBlock, false, true>::fn();
    return sph ? K<kTraceBlock, true, false>::fn() : K<kTraceBlock, false, false>::fn();
}

template <class Args>
hipError_t launch_traversal(const void *fn, const Args &a, uint32_t grid, hipStream_t s) {
    Args copy = a;
    void *args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(kTraceBlock), args, 0, s);
}

hipError_t launch_extend_packet(const ExtArgs &a, uint32_t grid, hipStream_t s) {
    const void *fn = a.gen.on ? reinterpret_cast<const void *>(&k_extend_packet<kTraceBlock, true>)
                              : reinterpret_cast<const void *>(&k_extend_packet<kTraceBlock, false>);
    return launch_traversal(fn, a, grid, s);
}

hipError_t launch_extend(const ExtArgs &a, uint32_t grid, hipStream_t s) {
    if (a.occ_count) return launch_traversal(traversal_kernel<ExtendOccK>(a.scene), a, grid, s);
    return launch_traversal(traversal_kernel<ExtendK>(a.scene), a, grid, s);
}

// `grid` = blocks covering the largest queue of the batch.
hipError_t launch_shade(const ShadeArgs &a, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kShadeBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_accumulate(const AccArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3((a.wh + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_trace(const TraceArgs &a, uint32_t grid, hipStream_t s) {
    return launch_traversal(traversal_kernel<TraceK>(a.scene), a, grid, s);
}

hipError_t launch_occlude(const OccArgs &a, uint32_t grid, hipStream_t s) {
    return launch_traversal(traversal_kernel<OccludeK>(a.scene), a, grid, s);
}

hipError_t occupancy(int which, const DevScene &sc, int *blocks_per_cu) {
    const void *f = nullptr;
    if (which == 0)
        f = traversal_kernel<ExtendK>(sc);
    else if (which == 1)
        f = reinterpret_cast<const void *>(&k_shade);
    else if (which == 3)
        f = traversal_kernel<OccludeK>(sc);
    else
        f = traversal_kernel<TraceK>(sc);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, which == 1 ? (int)kShadeBlock : kTraceBlock, 0);
}

void stack_layout(bool cull, uint32_t *short_entries, uint32_t *words_per_entry) {
    *short_entries = cull ? TStack<true>::kShort : TStack<false>::kShort;
    *words_per_entry = cull ? TStack<true>::kWords : TStack<false>::kWords;
}

int trace_block() { return kTraceBlock; }

}  // namespace mcrt
