// denoise_kernels.hip — joint-bilateral denoiser for gfx950.
//
// Restates cs_denoise (src/shaders/denoise.wgsl:52-120).  The reference
// issues 2 x 49 texture loads per pixel and recomputes the spatial Gaussian
// per tap; here a 16x16 tile plus its R-pixel halo (clamped at the borders,
// as :74-76) is staged once in LDS, and the spatial weights exp(-dist_sq *
// inv2ss) are computed once per block into a (2R+1)^2 LDS table.  Radii
// beyond the LDS tile (R > 8) take k_bilateral_wide, which reads the taps
// from memory; a negative radius runs no tap, so every pixel keeps its
// value (the shader's loops are empty and its weight sum is 0).
//
// The tile is held as one float array per channel: a tap reads depth first
// (the sky test) and the other channels only if it is kept, and 4-B reads at
// consecutive addresses hit 64 distinct banks.  (Staged as float4, the
// depth-only read strides 16 B across lanes: a 4-way bank conflict per tap.)
//
// Taps are visited in the shader's order (dy outer, dx inner) and every
// weight is formed as w_s * w_n * w_d, so the sums round as in the shader.
// The per-tap transcendentals dominate the kernel: the normal term uses
// repeated squaring when the exponent is a small whole number (the default
// is 8), and the depth term the hardware exp2; both stay within a few ulp of
// pow and exp.
#include "kernels.h"

namespace mcrt {

// x^n for a whole n by binary exponentiation (n uniform across the wave).
__device__ __forceinline__ float ipow(float x, uint32_t n) {
    float r = 1.0f;
    while (n) {
        if (n & 1u) r *= x;
        x *= x;
        n >>= 1;
    }
    return r;
}

__device__ __forceinline__ float fast_exp(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}

constexpr int kTile = 16;
constexpr int kMaxR = 8;  // LDS sized for R <= 8 (the reference default is 3)
constexpr int kSpan = kTile + 2 * kMaxR;

__global__ __launch_bounds__(256) void k_bilateral(BilateralArgs a) {
    __shared__ float s_r[kSpan * kSpan], s_g[kSpan * kSpan], s_b[kSpan * kSpan];
    __shared__ float s_nx[kSpan * kSpan], s_ny[kSpan * kSpan], s_nz[kSpan * kSpan], s_d[kSpan * kSpan];
    __shared__ float s_w[(2 * kMaxR + 1) * (2 * kMaxR + 1)];
    const int R = a.radius;
    const int span = kTile + 2 * R;
    const int x0 = blockIdx.x * kTile - R, y0 = blockIdx.y * kTile - R;
    const int tid = threadIdx.y * kTile + threadIdx.x;
    const int taps = (2 * R + 1) * (2 * R + 1);
    for (int k = tid; k < taps; k += 256) {  // spatial weights, :88-89
        const int dy = k / (2 * R + 1) - R, dx = k % (2 * R + 1) - R;
        s_w[k] = expf(-(float)(dx * dx + dy * dy) * a.inv2ss);
    }
    for (int i = tid; i < span * span; i += 256) {
        int ly = i / span, lx = i - ly * span;
        int gx = min(max(x0 + lx, 0), a.width - 1);
        int gy = min(max(y0 + ly, 0), a.height - 1);
        size_t g = (size_t)gy * a.width + gx;
        const float4 c = a.accum[g], n = a.gbuf[g];
        s_r[i] = c.x;
        s_g[i] = c.y;
        s_b[i] = c.z;
        s_nx[i] = n.x;
        s_ny[i] = n.y;
        s_nz[i] = n.z;
        s_d[i] = n.w;
    }
    __syncthreads();
    const int px = blockIdx.x * kTile + threadIdx.x;
    const int py = blockIdx.y * kTile + threadIdx.y;
    if (px >= a.width || py >= a.height) return;
    const int cl = (threadIdx.y + R) * span + (threadIdx.x + R);
    const float pnx = s_nx[cl], pny = s_ny[cl], pnz = s_nz[cl], pd = s_d[cl];
    const bool on_sky = pd < 0.0f;
    const float dden = fmaxf(pd, 1e-4f);
    const float npow = a.normal_pow;
    const bool whole = npow >= 0.0f && npow <= 64.0f && npow == floorf(npow);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, wsum = 0.0f;
    int k = 0;
    for (int dy = -R; dy <= R; ++dy) {
        for (int dx = -R; dx <= R; ++dx, ++k) {
            // Clamped neighbour: the LDS tile already holds the clamped texel
            // at every halo position, so the local offset is exact.
            const int l = cl + dy * span + dx;
            const float qd = s_d[l];
            const bool q_sky = qd < 0.0f;
            if (on_sky != q_sky) continue;
            const float w_s = s_w[k];
            const float nd = pnx * s_nx[l] + pny * s_ny[l] + pnz * s_nz[l];
            const float nc = fmaxf(0.0f, nd);
            const float w_n = on_sky ? 1.0f : (whole ? ipow(nc, (uint32_t)npow) : powf(nc, npow));
            const float drel = (pd - qd) / dden;
            const float w_d = on_sky ? 1.0f : fast_exp(-(drel * drel) * a.inv2ds);
            const float w = w_s * w_n * w_d;
            cr = cr + w * s_r[l];
            cg = cg + w * s_g[l];
            cb = cb + w * s_b[l];
            wsum = wsum + w;
        }
    }
    float4 o;
    if (wsum > 0.0f) {
        o = make_float4(cr / wsum, cg / wsum, cb / wsum, 1.0f);
    } else {
        o = make_float4(s_r[cl], s_g[cl], s_b[cl], 1.0f);
    }
    a.out[(size_t)py * a.width + px] = o;
}

// Any radius, taps read from memory with clamped addresses (:74-76) and the
// spatial weight computed per tap; the same weight expression and tap order
// as k_bilateral.  For radii beyond the LDS tile, and negative radii (no tap:
// the pixel is copied, as the shader's empty loops leave its weight sum 0).
__global__ __launch_bounds__(256) void k_bilateral_wide(BilateralArgs a) {
    const int px = blockIdx.x * kTile + threadIdx.x;
    const int py = blockIdx.y * kTile + threadIdx.y;
    if (px >= a.width || py >= a.height) return;
    const int R = a.radius;
    const size_t p = (size_t)py * a.width + px;
    const float4 g = a.gbuf[p];
    const bool on_sky = g.w < 0.0f;
    const float dden = fmaxf(g.w, 1e-4f);
    const float npow = a.normal_pow;
    const bool whole = npow >= 0.0f && npow <= 64.0f && npow == floorf(npow);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, wsum = 0.0f;
    for (int dy = -R; dy <= R; ++dy) {
        const int qy = min(max(py + dy, 0), a.height - 1);
        for (int dx = -R; dx <= R; ++dx) {
            const int qx = min(max(px + dx, 0), a.width - 1);
            const size_t q = (size_t)qy * a.width + qx;
            const float4 gq = a.gbuf[q];
            if (on_sky != (gq.w < 0.0f)) continue;
            const float w_s = expf(-(float)(dx * dx + dy * dy) * a.inv2ss);
            const float nc = fmaxf(0.0f, g.x * gq.x + g.y * gq.y + g.z * gq.z);
            const float w_n = on_sky ? 1.0f : (whole ? ipow(nc, (uint32_t)npow) : powf(nc, npow));
            const float drel = (g.w - gq.w) / dden;
            const float w_d = on_sky ? 1.0f : fast_exp(-(drel * drel) * a.inv2ds);
            const float w = w_s * w_n * w_d;
            const float4 c = a.accum[q];
            cr = cr + w * c.x;
            cg = cg + w * c.y;
            cb = cb + w * c.z;
            wsum = wsum + w;
        }
    }
    float4 o;
    if (wsum > 0.0f) {
        o = make_float4(cr / wsum, cg / wsum, cb / wsum, 1.0f);
    } else {
        const float4 c = a.accum[p];
        o = make_float4(c.x, c.y, c.z, 1.0f);
    }
    a.out[p] = o;
}

hipError_t launch_bilateral(const BilateralArgs &a, hipStream_t s) {
    if (a.radius > kMaxBilateralRadius) return hipErrorInvalidValue;
    dim3 grid((a.width + kTile - 1) / kTile, (a.height + kTile - 1) / kTile);
    if (a.radius >= 0 && a.radius <= kMaxR)
        hipLaunchKernelGGL(k_bilateral, grid, dim3(kTile, kTile), 0, s, a);
    else
        hipLaunchKernelGGL(k_bilateral_wide, grid, dim3(kTile, kTile), 0, s, a);
    return hipGetLastError();
}

}  // namespace mcrt
