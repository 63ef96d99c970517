// adjoint_kernels.hip — differentiable render and path-replay adjoint
// (gfx950).  No counterpart in the reference; the estimator is the reference
// path tracer (path_trace.wgsl:583-662, BSDF sampling, no NEE) with one
// extension: up to kMaxParams "parameter materials", each a Lambertian with
// RGB albedo a_p and a white emission E_p added at every vertex that hits it.
// With E_p = 0 a parameter material is exactly the reference's Lambertian.
//
// For one path with vertices k = 0..n-1 the radiance is
//   L = sum_k T_k * E_k + T_n * L_end,   T_0 = 1,  T_{k+1} = T_k * att_k,
// and the outgoing radiance R_k = E_k + att_k * R_{k+1} obeys R_0 = L.
// Direction sampling never depends on a_p or E_p, so each sample's path is
// fixed by its PCG stream and
//   dL/dE_p    = sum over vertices on p of T_k                 (white)
//   dL/da_p[c] = sum over vertices on p of T_k[c] * R_{k+1}[c].
//
// Two adjoint forms, selected per context (mcrt_diff_set_mode):
//
// Replay (default; after Vicini, Speierer, Jakob, "Path Replay
// Backpropagation", ACM TOG 40(4), 2021): the adjoint re-traces each sample
// from its PCG stream, exactly once, and keeps no tape.  Where PRB walks
// forward with the remaining radiance L - sum_{j<=k} T_j E_j and divides by
// the BSDF value at each vertex (ill-conditioned at small albedos, undefined
// at zero), this estimator has at most kMaxParams parameters, so the replay
// carries their derivatives forward instead: D_p = dT/da_p, times the pixel's
// adjoint delta, updated per vertex as
//   g_p    += E_k * D_p            (every p; E_k the radiance added at k)
//   g_p.w  += delta . T_k          (p = parameter of vertex k)
//   D_p     = D_p * att_k  (+ delta * T_k for p = parameter of vertex k)
// which sums exactly the terms T_k * R_{k+1} above, with no division and
// without the primal radiance L.
//
// Tape: the trace writes each vertex's (E_k, parameter), (att_k, continues)
// and T_k to a per-thread tape in device memory (48 B per vertex, about 1.2
// GB for the 512^2 inverse workload), then walks it backwards,
// R = E_k + att_k * R, adding delta * T_k * R_{k+1} at parameter vertices:
// also division-free.  Same terms, other summation order: the two forms
// agree to FP32 rounding.
//
// Layout: one thread per (frame group, pixel), the group's frames looped
// inside the thread; threads of one wave hold neighbouring pixels of the same
// group.  A small image is split into several frame groups so that the grid
// holds about 2^22 threads (several waves per SIMD slot, which balances the
// waves' uneven path lengths).  Gradients sit in registers (kMaxParams x 4),
// are summed per wave by a fixed butterfly, per block in wave order and then
// over blocks in block order, so results are reproducible.  The primal sums
// each group's frames, then k_diff_image adds the groups in order.
#include "kernels.h"
#include "shading.h"

namespace mcrt {

// Evaluates one vertex of the extended estimator.  Returns false when the
// path ends here (miss, emissive light, absorbed); `emit` is the radiance
// added at this vertex, `att` the attenuation to the next one, `param` the
// parameter index or -1.
struct Vertex {
    V3 emit, att, dir, p;
    int param;
    bool cont;
};

template <bool kSpheres>
__device__ __forceinline__ Vertex diff_vertex(const DiffArgs &a, V3 o, V3 d, uint32_t &rng,
                                              const Stack &st) {
    Vertex vx;
    vx.param = -1;
    vx.cont = false;
    vx.att = v3(0.0f, 0.0f, 0.0f);
    Best h = trace_one<kSpheres>(a.scene, o, d, MCRT_T_MIN, MCRT_T_MAX, st);
    if (h.prim == kNone) {  // miss: environment (:592-597)
        vx.emit = sample_env(a.scene, d);
        return vx;
    }
    Surface s = surface(a.scene, o, d, h.t, h.u, h.v, h.prim);
    vx.p = s.p;
    const uint32_t mi = s.mat < 63u ? s.mat : 63u;
    const int p = a.pmap[mi];
    if (p >= 0) {
        const float4 q = a.params[p];
        vx.param = p;
        vx.emit = v3(q.w, q.w, q.w);
        vx.dir = cosine_scatter(rng, s.n);  // :521-528
        vx.att = v3(q.x, q.y, q.z);
        vx.cont = true;
        return vx;
    }
    const GpuMaterial &m = a.scene.mats[mi];
    if (m.type_pad[0] == MCRT_MAT_EMISSIVE) {  // :608-610
        vx.emit = m.ior_pad[0] * ld3(m.albedo_fuzz);
        return vx;
    }
    vx.emit = v3(0.0f, 0.0f, 0.0f);
    Scattered sc = scatter_material(m, d, s, rng);
    if (sc.absorbed) return vx;
    vx.dir = sc.dir;
    vx.att = sc.att;
    vx.cont = true;
    return vx;
}

template <bool kSpheres>
__device__ __forceinline__ V3 diff_path_radiance(const DiffArgs &a, V3 o, V3 d, uint32_t rng,
                                                 const Stack &st) {
    V3 T = v3(1.0f, 1.0f, 1.0f), L = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t b = 0; b < a.max_depth; ++b) {
        Vertex vx = diff_vertex<kSpheres>(a, o, d, rng, st);
        L = L + T * vx.emit;
        if (!vx.cont) break;
        T = T * vx.att;
        o = vx.p;
        d = vx.dir;
    }
    return L;
}

// The kernels are compiled for 8 waves per SIMD, more than the compiler's
// own choice (primal 5 waves at 89 VGPRs, adjoint 4 at 122), spilling part of
// the 32 gradient accumulators and the traversal state: more waves hide the
// latency of dependent node fetches better than the spills cost (the 512^2
// inverse step went 26.8 -> 21.5 ms, each step up in waves faster than the
// last).  The replay adjoint also carries the 24 parameter derivatives; with
// the while-while traversal of trace_one it takes 10.1 ms per 512^2 x 32 spp
// call at 8 waves, 10.4 at 7 and 10.2 at 6 (the primal 7.8 ms at 8, 8.2 at 7,
// 8.8 at 6).
// kMode: 0 primal, 1 replay adjoint, 2 tape adjoint.
template <int kBlock, bool kSpheres, int kMode>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_diff(DiffArgs a) {
    constexpr bool kAdjoint = kMode != 0;
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t lds[Stack::kShort * Stack::kWords * kBlock];
    __shared__ float red[kAdjoint ? kMaxParams * 4 * kWaves : 1];
    const uint32_t tid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t wh = a.width * a.height;
    const uint32_t grp = tid / wh, pix = tid - grp * wh;
    const uint32_t per = (a.spp + a.groups - 1u) / a.groups;  // frames per group
    const uint32_t i0 = grp * per, i1 = min(a.spp, i0 + per);
    Stack st{lds, a.ovf, a.ovf_stride, threadIdx.x, tid, (uint32_t)kBlock};
    float g[kMaxParams][4];
#pragma unroll
    for (int q = 0; q < kMaxParams; ++q) g[q][0] = g[q][1] = g[q][2] = g[q][3] = 0.0f;
    if (grp < a.groups) {
        const uint32_t x = pix % a.width, y = pix / a.width;
        V3 sum = v3(0.0f, 0.0f, 0.0f);
        V3 delta = v3(0.0f, 0.0f, 0.0f);
        if (kAdjoint) {
            float4 dI = a.d_image[pix];
            const float inv = 1.0f / (float)a.spp;  // dI/dL_s of the mean
            delta = v3(dI.x * inv, dI.y * inv, dI.z * inv);
        }
        for (uint32_t i = i0; i < i1; ++i) {
            uint32_t rng0;
            V3 o, d;
            camera_ray(a.cam, a.width, a.height, x, y, a.frame0 + i, rng0, o, d);
            if (!kAdjoint) {
                sum = sum + diff_path_radiance<kSpheres>(a, o, d, rng0, st);
                continue;
            }
            if (kMode == 1) {
                // Replay: one trace, parameter derivatives carried forward.
                float D[kMaxParams][3];
#pragma unroll
                for (int q = 0; q < kMaxParams; ++q) D[q][0] = D[q][1] = D[q][2] = 0.0f;
                V3 T = v3(1.0f, 1.0f, 1.0f);
                uint32_t rng = rng0;
                for (uint32_t b = 0; b < a.max_depth; ++b) {
                    Vertex vx = diff_vertex<kSpheres>(a, o, d, rng, st);
#pragma unroll
                    for (int q = 0; q < kMaxParams; ++q) {
                        g[q][0] += vx.emit.x * D[q][0];
                        g[q][1] += vx.emit.y * D[q][1];
                        g[q][2] += vx.emit.z * D[q][2];
                    }
                    if (!vx.cont) break;
                    const V3 dT = delta * T;
#pragma unroll
                    for (int q = 0; q < kMaxParams; ++q) {
                        D[q][0] *= vx.att.x;
                        D[q][1] *= vx.att.y;
                        D[q][2] *= vx.att.z;
                        if (q == vx.param) {
                            D[q][0] += dT.x;
                            D[q][1] += dT.y;
                            D[q][2] += dT.z;
                            g[q][3] += (dT.x + dT.y) + dT.z;
                        }
                    }
                    T = T * vx.att;
                    o = vx.p;
                    d = vx.dir;
                }
                continue;
            }
            // Tape: trace once, keeping each vertex on this thread's tape.
            uint32_t n = 0;
            {
                V3 T = v3(1.0f, 1.0f, 1.0f);
                uint32_t rng = rng0;
                for (uint32_t b = 0; b < a.max_depth; ++b) {
                    Vertex vx = diff_vertex<kSpheres>(a, o, d, rng, st);
                    float4 *t = a.tape + 3u * ((size_t)b * a.tape_stride + tid);
                    t[0] = make_float4(vx.emit.x, vx.emit.y, vx.emit.z, (float)vx.param);  // -1..7, exact
                    t[1] = make_float4(vx.att.x, vx.att.y, vx.att.z, 0.0f);
                    t[2] = make_float4(T.x, T.y, T.z, 0.0f);
                    n = b + 1u;
                    if (!vx.cont) break;
                    T = T * vx.att;
                    o = vx.p;
                    d = vx.dir;
                }
            }
            // Walk the tape backwards: R is the radiance arriving at vertex k
            // from beyond it (0 past the last vertex; att is 0 where the path
            // ends, so R restarts at that vertex's emission).
            V3 R = v3(0.0f, 0.0f, 0.0f);
            for (uint32_t b = n; b-- > 0u;) {
                const float4 *t = a.tape + 3u * ((size_t)b * a.tape_stride + tid);
                const float4 t0 = t[0], t1 = t[1], t2 = t[2];
                const int param = (int)t0.w;
                if (param >= 0) {
                    const V3 dT = delta * xyz(t2);
#pragma unroll
                    for (int q = 0; q < kMaxParams; ++q) {
                        if (q == param) {
                            g[q][0] += dT.x * R.x;
                            g[q][1] += dT.y * R.y;
                            g[q][2] += dT.z * R.z;
                            g[q][3] += (dT.x + dT.y) + dT.z;
                        }
                    }
                }
                R = xyz(t0) + xyz(t1) * R;
            }
        }
        if (!kAdjoint) {
            if (a.groups == 1u) {
                const float inv = 1.0f / (float)a.spp;
                a.image[pix] = make_float4(sum.x * inv, sum.y * inv, sum.z * inv, 1.0f);
            } else {
                a.image_part[tid] = make_float4(sum.x, sum.y, sum.z, 0.0f);
            }
        }
    }
    if (kAdjoint) {
        // Fixed-order reduction: a butterfly over the wave's lanes, then
        // thread j adds column j of the block's waves in wave order.
        const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < kMaxParams; ++q)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float v = g[q][c];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
                if (__lane_id() == 0) red[wave * (kMaxParams * 4) + q * 4 + c] = v;
            }
        __syncthreads();
        if (threadIdx.x < kMaxParams * 4) {
            float s = 0.0f;
            for (int w = 0; w < kWaves; ++w) s += red[w * (kMaxParams * 4) + threadIdx.x];
            a.partial[(size_t)blockIdx.x * (kMaxParams * 4) + threadIdx.x] = s;
        }
    }
}

// Primal with frame groups: the mean of each pixel's group sums, added in
// group order.
__global__ void k_diff_image(const float4 *part, uint32_t wh, uint32_t groups, uint32_t spp, float4 *image) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= wh) return;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    for (uint32_t g = 0; g < groups; ++g) {
        const float4 v = part[(size_t)g * wh + pix];
        x += v.x;
        y += v.y;
        z += v.z;
    }
    const float inv = 1.0f / (float)spp;
    image[pix] = make_float4(x * inv, y * inv, z * inv, 1.0f);
}

// Sum of the per-block partial gradients: block j reduces output j, thread t
// adding blocks t, t + 256, ... in order, then a fixed tree over the threads,
// so the result is the same on every run.
__global__ __launch_bounds__(256) void k_diff_sum(const float *partial, uint32_t n_blocks, float *out) {
    __shared__ float s[256];
    const uint32_t j = blockIdx.x;
    float acc = 0.0f;
    for (uint32_t b = threadIdx.x; b < n_blocks; b += 256u) acc += partial[(size_t)b * (kMaxParams * 4) + j];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = 128u; w > 0u; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = s[0];
}

constexpr int kDiffBlock = 256;

// Frame groups per pixel: enough that the grid holds about 2^22 threads,
// at most one frame each.
uint32_t diff_groups(uint32_t wh, uint32_t spp) {
    const uint32_t g = (1u << 22) / wh;
    return g < 1u ? 1u : (g > spp ? spp : g);
}

uint32_t diff_grid(uint32_t wh, uint32_t groups) {
    return (uint32_t)(((uint64_t)wh * groups + kDiffBlock - 1) / kDiffBlock);
}

hipError_t launch_diff(const DiffArgs &a, bool adjoint, hipStream_t s) {
    const uint32_t wh = a.width * a.height;
    const uint32_t grid = diff_grid(wh, a.groups);
    const bool sph = a.scene.sroot != kNone;
    if (adjoint) {
        if (a.tape) {
            if (sph)
                hipLaunchKernelGGL((k_diff<kDiffBlock, true, 2>), dim3(grid), dim3(kDiffBlock), 0, s, a);
            else
                hipLaunchKernelGGL((k_diff<kDiffBlock, false, 2>), dim3(grid), dim3(kDiffBlock), 0, s, a);
        } else {
            if (sph)
                hipLaunchKernelGGL((k_diff<kDiffBlock, true, 1>), dim3(grid), dim3(kDiffBlock), 0, s, a);
            else
                hipLaunchKernelGGL((k_diff<kDiffBlock, false, 1>), dim3(grid), dim3(kDiffBlock), 0, s, a);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (a.n_params > 0u)
            hipLaunchKernelGGL(k_diff_sum, dim3(a.n_params * 4u), dim3(256), 0, s, a.partial, grid, a.grad);
    } else {
        if (sph)
            hipLaunchKernelGGL((k_diff<kDiffBlock, true, 0>), dim3(grid), dim3(kDiffBlock), 0, s, a);
        else
            hipLaunchKernelGGL((k_diff<kDiffBlock, false, 0>), dim3(grid), dim3(kDiffBlock), 0, s, a);
        if (a.groups > 1u) {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_diff_image, dim3((wh + 255u) / 256u), dim3(256), 0, s, a.image_part, wh, a.groups,
                               a.spp, a.image);
        }
    }
    return hipGetLastError();
}

}  // namespace mcrt
