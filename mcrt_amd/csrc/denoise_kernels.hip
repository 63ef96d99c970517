// denoise_kernels.hip — joint-bilateral denoiser for gfx950.
//
// Restates cs_denoise (src/shaders/denoise.wgsl:52-120).  The reference
// issues 2 x 49 texture loads per pixel and recomputes the spatial Gaussian
// per tap; here a 16x16 tile plus its R-pixel halo (clamped at the borders,
// as :74-76) is staged once in LDS, and the spatial weights come from a
// (2R+1)^2 table computed on the host with the same expression
// exp(-dist_sq * inv2ss).
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
    const int R = a.radius;
    const int span = kTile + 2 * R;
    const int x0 = blockIdx.x * kTile - R, y0 = blockIdx.y * kTile - R;
    const int tid = threadIdx.y * kTile + threadIdx.x;
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
            const float w_s = a.spatial[k];
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

hipError_t launch_bilateral(const BilateralArgs &a, hipStream_t s) {
    if (a.radius < 0 || a.radius > kMaxR) return hipErrorInvalidValue;
    dim3 grid((a.width + kTile - 1) / kTile, (a.height + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_bilateral, grid, dim3(kTile, kTile), 0, s, a);
    return hipGetLastError();
}

}  // namespace mcrt
