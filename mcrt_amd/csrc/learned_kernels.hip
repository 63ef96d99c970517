// learned_kernels.hip — learned denoiser on BF16 MFMA (gfx950), FP32 accumulate.
// No counterpart in the reference (its only denoiser is the bilateral filter,
// src/shaders/denoise.wgsl); the bilateral pass stays the fallback.
//
// Network: six 3x3 convolutions 10 -> 32 -> 32 -> 32 -> 32 -> 32 -> 3, ReLU
// after the first five, global skip: out_log = log1p(noisy) + conv6(...),
// denoised = expm1(out_log).  Input channels: log1p radiance (3), first-hit
// albedo (3), normal (3), log-depth (1).  81,216 FLOP per pixel (inference).
//
// Activations are NHWC BF16 with the channel dimension padded to 32 (64 B per
// pixel); padded channels hold zeros and padded weights stay zero under Adam
// (their gradients are exactly zero).
//
// Convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16: a block of 4
// waves computes a 16 x 16 pixel output tile; each wave owns 4 rows, and for
// each of the 9 taps multiplies A = (16 pixels of a row) x (32 input
// channels) by B = (32 input channels) x (16 output channels), twice for the
// 32 outputs.  The input tile with its 1-pixel halo (zero padded at the
// image border) is staged in LDS; the 9 x 32 x 32 weights live in registers
// as MFMA B fragments for the whole (grid-stride) block.
// The same kernel computes the data gradient (flipped, transposed weights,
// ReLU mask of the layer output fused into the tile load).
// The weight gradient is a second MFMA GEMM with the pixels as the K axis.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace mcrt {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kC = 32;        // padded channel count
constexpr int kT = 16;        // output tile edge
constexpr int kH = kT + 2;    // tile edge with halo
constexpr int kPS = 40;       // LDS pixel stride (bf16) of NHWC tiles
constexpr int kWSize = 9 * kC * kC;

__device__ __forceinline__ bf16 f2b(float f) { return (bf16)f; }
__device__ __forceinline__ float b2f(bf16 b) { return (float)b; }

// Zero the elements of v whose mask element is not positive.
__device__ __forceinline__ uint4 apply_mask(uint4 v, uint4 m) {
    uint32_t vv[4] = {v.x, v.y, v.z, v.w}, mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t lo = mm[k] & 0xFFFFu, hi = mm[k] >> 16;
        bool plo = !(lo & 0x8000u) && (lo & 0x7FFFu);
        bool phi = !(hi & 0x8000u) && (hi & 0x7FFFu);
        vv[k] = (plo ? (vv[k] & 0xFFFFu) : 0u) | (phi ? (vv[k] & 0xFFFF0000u) : 0u);
    }
    return make_uint4(vv[0], vv[1], vv[2], vv[3]);
}

// ---------------------------------------------------------------------------
// Input packing: renders -> BF16 NHWC32 network input
// ---------------------------------------------------------------------------
// The network's 10 input channels of one pixel (channels 0-7 in lo, 8-9 and
// zeros in hi): log1p radiance, albedo, normal, scaled log depth.
__device__ __forceinline__ void input_channels(float4 L, float4 al, float4 g, bf16x8 &lo, bf16x8 &hi) {
    const float d = fminf(fmaxf(g.w, 0.0f), 1e4f);
    float v[16] = {log1pf(fmaxf(L.x, 0.0f)), log1pf(fmaxf(L.y, 0.0f)), log1pf(fmaxf(L.z, 0.0f)),
                   al.x, al.y, al.z, g.x, g.y, g.z, log1pf(d) * (1.0f / 9.21044f),
                   0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        lo[k] = f2b(v[k]);
        hi[k] = f2b(v[8 + k]);
    }
}

__global__ __launch_bounds__(256) void k_pack(PackArgs a) {
    const size_t n = (size_t)a.B * a.H * a.W;
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    bf16x8 lo, hi;
    input_channels(a.noisy[p], a.albedo[p], a.gbuf[p], lo, hi);
    bf16x8 z = {};
    bf16x8 *o = reinterpret_cast<bf16x8 *>(a.out + p * kC);
    o[0] = lo;
    o[1] = hi;
    o[2] = z;
    o[3] = z;
}

// Inference input with 16 channels (32 B per pixel): the 10 network inputs
// and 6 zeros.  The first layer's 16-channel kernel (k_conv3x3 with
// kCinT = 16) reads it; training keeps the 32-channel layout, whose padded
// channels the weight gradient reads.
__global__ __launch_bounds__(256) void k_pack16(PackArgs a) {
    const size_t n = (size_t)a.B * a.H * a.W;
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    bf16x8 lo, hi;
    input_channels(a.noisy[p], a.albedo[p], a.gbuf[p], lo, hi);
    bf16x8 *o = reinterpret_cast<bf16x8 *>(a.out + p * 16);
    o[0] = lo;
    o[1] = hi;
}

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM on MFMA
// ---------------------------------------------------------------------------
// A block loops over tiles (grid-stride) and keeps the layer's B fragments
// in registers for its lifetime: lane (m, kq) holds
// w[tap][16 h + m][8 kq .. 8 kq + 7] for the 9 taps and both output halves
// h, loaded from global memory once.  Per tile, the 18 x 18 input tile is
// staged in LDS; wave w owns output rows 4w .. 4w+3, which read input rows
// 4w .. 4w+5 of the halo tile.  Each A fragment (16 pixels of one input row
// at one horizontal tap kx, 32 channels) is read from LDS once and feeds
// every output row r it contributes to (vertical tap ky = input row - r).
// kFinal: the last layer, whose 3 outputs need only the first half.
// Block layout: the 16 x 16 output tile is split into 4 groups of 4 rows;
// a wave computes one row group for one 16-channel half of the outputs
// (8 waves per block), or for the 3 output channels of the last layer
// (4 waves).  Holding one half of the weights per wave (36 VGPRs of B
// fragments instead of 72) doubles the resident waves, which hide the
// tile loads of the other blocks.
// kEnd (inference, last layer only): write the denoised image
// (ConvArgs::out_final) instead of the log-space output, so no unpack pass
// follows; a separate instantiation, so the training kernel carries none of
// it.
// kCinT = 16 (inference, first layer): the input has 16 channels (k_pack16),
// and one MFMA's K = 32 spans two horizontal taps of 16 channels each: taps
// (kx 0, 1) and (kx 2, zeros), 6 MFMAs per output row instead of 9, with
// half the input bytes.  Its weights are the layer's 32-channel weights,
// whose channels 16..31 are zero.
template <bool kFinal, bool kEnd = false, int kCinT = 32>
__global__ __launch_bounds__(kFinal ? 256 : 512) void k_conv3x3(ConvArgs a) {
    constexpr int kThreads = kFinal ? 256 : 512;
    constexpr bool kIn16 = kCinT == 16;
    constexpr int kPSin = kIn16 ? 24 : kPS;  // LDS pixel stride of the input tile
    constexpr int kQuads = kCinT / 8;        // 16-B loads per input pixel
    constexpr int kLdsIn = kH * kH * kPSin, kLdsOut = kT * kT * kPS;
    __shared__ __align__(16) bf16 s_in[kLdsIn > kLdsOut ? kLdsIn : kLdsOut];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m = lane & 15, kq = lane >> 4;
    const int rg = wave & 3;                     // row group: output rows 4*rg .. 4*rg+3
    const int half = kFinal ? 0 : (wave >> 2);   // output channels 16*half .. 16*half+15
    // B fragments: per tap (32 channels), or for kIn16 per (ky, tap pair):
    // lanes kq 0-1 hold tap kx = 2p channels 8kq.., lanes kq 2-3 tap 2p+1.
    constexpr int kFrags = kIn16 ? 6 : 9;
    bf16x8 wfrag[kFrags];
#pragma unroll
    for (int f = 0; f < kFrags; ++f) {
        if (kIn16) {
            const int ky = f >> 1, pr = f & 1;
            const int kx = 2 * pr + (kq >> 1), ci = (kq & 1) * 8;
            wfrag[f] = bf16x8{};
            if (kx < 3)
                wfrag[f] = *reinterpret_cast<const bf16x8 *>(a.w + ((ky * 3 + kx) * kC + 16 * half + m) * kC + ci);
        } else {
            wfrag[f] = *reinterpret_cast<const bf16x8 *>(a.w + (f * kC + 16 * half + m) * kC + kq * 8);
        }
    }
    const float bias = a.bias ? a.bias[16 * half + m] : 0.0f;

    const int tiles_x = (a.W + kT - 1) / kT, tiles_y = (a.H + kT - 1) / kT;
    const int per_image = tiles_x * tiles_y, n_tiles = per_image * a.B;
    // Software pipeline: each thread's share of a tile's input (with halo,
    // zero outside the image) is loaded into registers one tile ahead, so
    // the loads are in flight while the current tile computes.
    constexpr int kLoads = (kH * kH * kQuads + kThreads - 1) / kThreads;
    uint4 pre[kLoads], pmask[kLoads];
    auto fetch = [&](int t) {
        const int b = t / per_image, rem = t - b * per_image;
        const int y0 = (rem / tiles_x) * kT, x0 = (rem % tiles_x) * kT;
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
            const int i = tid + k * kThreads;
            const int pix = i / kQuads, q = i % kQuads;
            const int hy = pix / kH, hx = pix - hy * kH;
            const int gy = y0 + hy - 1, gx = x0 + hx - 1;
            pre[k] = make_uint4(0, 0, 0, 0);
            pmask[k] = make_uint4(0, 0, 0, 0);
            if (i < kH * kH * kQuads && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                const size_t px = ((size_t)b * a.H + gy) * a.W + gx;
                const size_t off = px * kCinT + q * 8;
                pre[k] = *reinterpret_cast<const uint4 *>(a.in + off);
                if (!kIn16 && a.mask) pmask[k] = *reinterpret_cast<const uint4 *>(a.mask + off);
            }
        }
    };
    if ((int)blockIdx.x < n_tiles) fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int b = tile / per_image, rem = tile - b * per_image;
        const int y0 = (rem / tiles_x) * kT, x0 = (rem % tiles_x) * kT;
        __syncthreads();  // the previous tile is done with s_in
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
            const int i = tid + k * kThreads;
            if (i < kH * kH * kQuads) {
                const int pix = i / kQuads, q = i % kQuads;
                *reinterpret_cast<uint4 *>(&s_in[pix * kPSin + q * 8]) =
                    (!kIn16 && a.mask) ? apply_mask(pre[k], pmask[k]) : pre[k];
            }
        }
        __syncthreads();
        if (tile + (int)gridDim.x < n_tiles) fetch(tile + gridDim.x);
        floatx4 acc[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ir = 0; ir < 6; ++ir) {
            if constexpr (kIn16) {
#pragma unroll
                for (int pr = 0; pr < 2; ++pr) {
                    const int kx = 2 * pr + (kq >> 1);
                    bf16x8 av = bf16x8{};
                    if (kx < 3)
                        av = *reinterpret_cast<const bf16x8 *>(
                            &s_in[((rg * 4 + ir) * kH + (m + kx)) * kPSin + (kq & 1) * 8]);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int ky = ir - r;
                        if (ky < 0 || ky > 2) continue;
                        acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, wfrag[ky * 2 + pr], acc[r], 0, 0, 0);
                    }
                }
            } else {
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const bf16x8 av =
                        *reinterpret_cast<const bf16x8 *>(&s_in[((rg * 4 + ir) * kH + (m + kx)) * kPS + kq * 8]);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int ky = ir - r;
                        if (ky < 0 || ky > 2) continue;
                        acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, wfrag[ky * 3 + kx], acc[r], 0, 0, 0);
                    }
                }
            }
        }
        // D layout: lane holds rows 4*kq + j (pixel x), column m (output channel).
        if (kFinal) {
            // out_log = log1p(max(noisy, 0)) + conv + bias, channels 0..2:
            // the conv terms go through LDS so that each thread then reads
            // and writes one pixel's float4.
            float *s_f = reinterpret_cast<float *>(s_in);  // [pixel][3]
            __syncthreads();  // every wave is done reading the input tile
            if (m < 3) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int j = 0; j < 4; ++j) s_f[((rg * 4 + r) * kT + 4 * kq + j) * 3 + m] = acc[r][j] + bias;
            }
            __syncthreads();
            for (int px = tid; px < kT * kT; px += kThreads) {
                const int gy = y0 + (px >> 4), gx = x0 + (px & 15);
                if (gy < a.H && gx < a.W) {
                    const size_t p = ((size_t)b * a.H + gy) * a.W + gx;
                    const float4 nz = a.noisy[p];
                    const float4 o = make_float4(log1pf(fmaxf(nz.x, 0.0f)) + s_f[px * 3 + 0],
                                                 log1pf(fmaxf(nz.y, 0.0f)) + s_f[px * 3 + 1],
                                                 log1pf(fmaxf(nz.z, 0.0f)) + s_f[px * 3 + 2], 0.0f);
                    if (kEnd)  // inference: the denoised image, as k_unpack
                        a.out_final[p] = make_float4(expm1f(o.x), expm1f(o.y), expm1f(o.z), 1.0f);
                    else
                        a.out_log[p] = o;
                }
            }
            continue;
        }
        __syncthreads();  // s_in is reused as the output tile
        const int co = 16 * half + m;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int y = rg * 4 + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v = acc[r][j] + bias;
                if (a.mode == 0) v = fmaxf(v, 0.0f);
                s_in[(y * kT + 4 * kq + j) * kPS + co] = f2b(v);
            }
        }
        __syncthreads();
        for (int i = tid; i < kT * kT * 4; i += kThreads) {
            const int pix = i >> 2, q = i & 3;
            const int y = pix >> 4, x = pix & 15;
            const int gy = y0 + y, gx = x0 + x;
            if (gy < a.H && gx < a.W)
                *reinterpret_cast<uint4 *>(a.out + (((size_t)b * a.H + gy) * a.W + gx) * kC + q * 8) =
                    *reinterpret_cast<const uint4 *>(&s_in[pix * kPS + q * 8]);
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[tap][co][ci] = sum_p dPre[p][co] * X[p + tap][ci],
// dB[co] = sum_p dPre[p][co].  Pixels are the K axis of the MFMA GEMM; each
// block loops over tiles and writes one partial sum, reduced in block order.
// ---------------------------------------------------------------------------
constexpr int kDP = kT * kT;  // pixels per tile

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    __shared__ bf16 s_dp[kC * kDP];             // [co][p], p = y*16 + x
    __shared__ bf16 s_x[3 * kC * kH * kT];      // [shift][ci][hy][x]
    __shared__ float s_db[256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m = lane & 15, kq = lane >> 4;
    // Taps of this wave: wave, wave + 4, wave + 8 (< 9).
    floatx4 acc[3][2][2];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[t][i][0] = acc[t][i][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    float db = 0.0f;
    const int tiles_x = (a.W + kT - 1) / kT, tiles_y = (a.H + kT - 1) / kT;
    const int n_tiles = tiles_x * tiles_y * a.B;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int b = tile / (tiles_x * tiles_y);
        const int rem = tile - b * tiles_x * tiles_y;
        const int y0 = (rem / tiles_x) * kT, x0 = (rem % tiles_x) * kT;
        __syncthreads();
        for (int i = tid; i < kDP * 4; i += 256) {
            const int p = i >> 2, q = i & 3;
            const int gy = y0 + (p >> 4), gx = x0 + (p & 15);
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy < a.H && gx < a.W) {
                const size_t off = (((size_t)b * a.H + gy) * a.W + gx) * kC + q * 8;
                v = *reinterpret_cast<const uint4 *>(a.dy + off);
                if (a.mask) v = apply_mask(v, *reinterpret_cast<const uint4 *>(a.mask + off));
            }
            const bf16 *e = reinterpret_cast<const bf16 *>(&v);
#pragma unroll
            for (int k = 0; k < 8; ++k) s_dp[(q * 8 + k) * kDP + p] = e[k];
        }
        for (int i = tid; i < kH * kH * 4; i += 256) {
            const int pix = i >> 2, q = i & 3;
            const int hy = pix / kH, hx = pix - hy * kH;
            const int gy = y0 + hy - 1, gx = x0 + hx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
                v = *reinterpret_cast<const uint4 *>(a.x + (((size_t)b * a.H + gy) * a.W + gx) * kC + q * 8);
            const bf16 *e = reinterpret_cast<const bf16 *>(&v);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int x = hx - s;
                if (x < 0 || x >= kT) continue;
#pragma unroll
                for (int k = 0; k < 8; ++k) s_x[((s * kC + q * 8 + k) * kH + hy) * kT + x] = e[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int tap = wave + 4 * t;
            if (tap < 9) {
                const int ky = tap / 3, kx = tap - ky * 3;
#pragma unroll 2
                for (int kk = 0; kk < kDP / 32; ++kk) {
                    const int p = kk * 32 + kq * 8;  // 8 pixels in one row
                    const int y = p >> 4, x = p & 15;
#pragma unroll
                    for (int mt = 0; mt < 2; ++mt) {
                        const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&s_dp[(16 * mt + m) * kDP + p]);
#pragma unroll
                        for (int nt = 0; nt < 2; ++nt) {
                            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(
                                &s_x[((kx * kC + 16 * nt + m) * kH + (y + ky)) * kT + x]);
                            acc[t][mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t][mt][nt], 0, 0, 0);
                        }
                    }
                }
            }
        }
        {  // bias gradient: thread -> channel tid % 32, 32-pixel slice tid / 32
            const int co = tid & 31, part = tid >> 5;
            float s = 0.0f;
            for (int k = 0; k < 32; ++k) s += b2f(s_dp[co * kDP + part * 32 + k]);
            db += s;
        }
    }
    float *out = a.partial + (size_t)blockIdx.x * (kWSize + kC);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int tap = wave + 4 * t;
        if (tap < 9) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int co = 16 * mt + 4 * kq + j, ci = 16 * nt + m;
                        out[tap * kC * kC + co * kC + ci] = acc[t][mt][nt][j];
                    }
        }
    }
    s_db[tid] = db;
    __syncthreads();
    if (tid < kC) {
        float s = 0.0f;
        for (int part = 0; part < 8; ++part) s += s_db[part * 32 + tid];
        out[kWSize + tid] = s;
    }
}

__global__ void k_sum_partials(const float *partial, int n_blocks, int n, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int b = 0; b < n_blocks; ++b) s += partial[(size_t)b * n + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------
// Loss over RGB, mean over pixels and channels; gradient with respect to the
// final layer's output o = log1p(x) (BF16 NHWC32, channels 0..2) and
// per-block sums.  kind 0: (o - log1p(r))^2.  kind 1: relMSE
// (x - r)^2 / (r^2 + 1e-2) with x = expm1(o), the reference r held constant,
// so d/do = 2 (x - r) / (r^2 + 1e-2) * exp(o).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float loss_term(int kind, float o, float r, float &g) {
    r = fmaxf(r, 0.0f);
    if (kind == 0) {
        const float d = o - log1pf(r);
        g = 2.0f * d;
        return d * d;
    }
    const float x = expm1f(fminf(o, 30.0f));
    const float d = x - r;
    const float w = 1.0f / (r * r + 1e-2f);
    g = 2.0f * d * w * (x + 1.0f);
    return d * d * w;
}

__global__ __launch_bounds__(256) void k_loss(LossArgs a) {
    __shared__ float s[256];
    const size_t n = (size_t)a.B * a.H * a.W;
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < n) {
        const float4 o = a.out_log[p], r = a.ref[p];
        float g0, g1, g2;
        l = loss_term(a.kind, o.x, r.x, g0) + loss_term(a.kind, o.y, r.y, g1) +
            loss_term(a.kind, o.z, r.z, g2);
        const float sc = a.inv_count;
        bf16x8 g = {};
        g[0] = f2b(sc * g0);
        g[1] = f2b(sc * g1);
        g[2] = f2b(sc * g2);
        bf16x8 z = {};
        bf16x8 *o8 = reinterpret_cast<bf16x8 *>(a.grad + p * kC);
        o8[0] = g;
        o8[1] = z;
        o8[2] = z;
        o8[3] = z;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// Adam on FP32 master weights [tap][co][ci] + bias, then the BF16 copies:
// forward [tap][co][ci] and data-gradient [8 - tap][ci][co].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kWSize + kC) return;
    float g = a.grad[i];
    float m = a.m[i] = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
    float v = a.v[i] = a.beta2 * a.v[i] + (1.0f - a.beta2) * g * g;
    float mh = m / a.bc1, vh = v / a.bc2;
    float w = a.w[i] - a.lr * mh / (sqrtf(vh) + a.eps);
    a.w[i] = w;
    if (i < kWSize) {
        const int tap = i / (kC * kC), co = (i / kC) % kC, ci = i % kC;
        a.wf[i] = f2b(w);
        a.wb[(8 - tap) * kC * kC + ci * kC + co] = f2b(w);
    }
}

// Refresh the BF16 copies from the master weights (after loading weights).
__global__ __launch_bounds__(256) void k_weights_to_bf16(const float *w, bf16 *wf, bf16 *wb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kWSize) return;
    const int tap = i / (kC * kC), co = (i / kC) % kC, ci = i % kC;
    wf[i] = f2b(w[i]);
    wb[(8 - tap) * kC * kC + ci * kC + co] = f2b(w[i]);
}

__global__ __launch_bounds__(256) void k_unpack(const float4 *out_log, float4 *out, size_t n) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const float4 o = out_log[p];
    out[p] = make_float4(expm1f(o.x), expm1f(o.y), expm1f(o.z), 1.0f);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + 255) / 256); }

hipError_t launch_pack(const PackArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((size_t)a.B * a.H * a.W)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pack16(const PackArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_pack16, dim3(blocks_for((size_t)a.B * a.H * a.W)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Grid-stride launch sized to the blocks that are resident at once.
template <bool kFinal>
uint32_t conv_grid(size_t tiles) {
    static uint32_t resident = 0;
    if (!resident) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_conv3x3<kFinal, false>),
                                                         kFinal ? 256 : 512, 0) != hipSuccess ||
            cus < 1 || per_cu < 1)
            resident = 1024;
        else
            resident = (uint32_t)(cus * per_cu);
    }
    return (uint32_t)std::min<size_t>(tiles, resident);
}

hipError_t launch_conv3x3(const ConvArgs &a, hipStream_t s) {
    const size_t tiles = (size_t)((a.W + kT - 1) / kT) * ((a.H + kT - 1) / kT) * a.B;
    if (a.mode == 2 && a.out_final)
        hipLaunchKernelGGL((k_conv3x3<true, true>), dim3(conv_grid<true>(tiles)), dim3(256), 0, s, a);
    else if (a.mode == 2)
        hipLaunchKernelGGL((k_conv3x3<true, false>), dim3(conv_grid<true>(tiles)), dim3(256), 0, s, a);
    else if (a.cin16)
        hipLaunchKernelGGL((k_conv3x3<false, false, 16>), dim3(conv_grid<false>(tiles)), dim3(512), 0, s, a);
    else
        hipLaunchKernelGGL((k_conv3x3<false, false>), dim3(conv_grid<false>(tiles)), dim3(512), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_wgrad(const WgradArgs &a, uint32_t blocks, float *out_grad, hipStream_t s) {
    hipLaunchKernelGGL(k_wgrad, dim3(blocks), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_partials, dim3(blocks_for(kWSize + kC)), dim3(256), 0, s, a.partial,
                       (int)blocks, kWSize + kC, out_grad);
    return hipGetLastError();
}

hipError_t launch_loss(const LossArgs &a, float *loss_out, hipStream_t s) {
    const uint32_t nb = blocks_for((size_t)a.B * a.H * a.W);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, a.partial, (int)nb, 1, loss_out);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kWSize + kC)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_weights_to_bf16(const float *w, void *wf, void *wb, hipStream_t s) {
    hipLaunchKernelGGL(k_weights_to_bf16, dim3(blocks_for(kWSize)), dim3(256), 0, s, w, (bf16 *)wf,
                       (bf16 *)wb);
    return hipGetLastError();
}

hipError_t launch_unpack(const float4 *out_log, float4 *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_unpack, dim3(blocks_for(n)), dim3(256), 0, s, out_log, out, n);
    return hipGetLastError();
}

}  // namespace mcrt
