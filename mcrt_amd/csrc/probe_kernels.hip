// probe_kernels.hip — on-device peak probes for the roofline fractions of the
// benchmark (no counterpart in the reference).
//
//  - k_copy_probe: a float4 grid-stride copy between two buffers of the same
//    size; bytes moved = 2 x size (one read, one write per byte).  Its rate
//    is the HBM bandwidth a streaming kernel reaches on this device.
//  - k_mfma_probe: every wave issues a fixed number of
//    v_mfma_f32_16x16x32_bf16 on register operands, eight independent
//    accumulators deep so the matrix core never waits on a result; 16,384
//    FLOP each.  Its rate is the dense BF16 MFMA issue peak.
// Both are timed with HIP events on the caller's stream.
#include <hip/hip_runtime.h>

#include "../../include/mcrt_abi.h"
#include "kernels.h"

namespace mcrt {

typedef __bf16 pbf16x8 __attribute__((ext_vector_type(8)));
typedef float pfloatx4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                    size_t n) {
    const size_t stride = (size_t)gridDim.x * 256u;
    size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

constexpr int kMfmaChains = 8;

__global__ __launch_bounds__(256) void k_mfma_probe(uint32_t iters, float seed, float *out) {
    pbf16x8 a, b;
    for (int k = 0; k < 8; ++k) {
        a[k] = (__bf16)(seed * (float)(threadIdx.x + k));
        b[k] = (__bf16)(seed * (float)(blockIdx.x + k));
    }
    pfloatx4 acc[kMfmaChains];
#pragma unroll
    for (int c = 0; c < kMfmaChains; ++c) acc[c] = pfloatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kMfmaChains; ++c) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kMfmaChains; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    // Keeps the chains live; the host never reads a value it depends on.
    if (s == 12345.678f) out[blockIdx.x * 256u + threadIdx.x] = s;
}

}  // namespace mcrt

using namespace mcrt;

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" int mcrt_measure_peaks(int device, uint64_t copy_bytes, double *copy_gbs, double *mfma_bf16_tflops) {
    if (!copy_gbs || !mfma_bf16_tflops || copy_bytes < (1u << 20)) return MCRT_ERR_BAD_ARG;
    if (hipSetDevice(device) != hipSuccess) return MCRT_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MCRT_ERR_HIP;
    const size_t n = (size_t)(copy_bytes / sizeof(float4));
    float4 *src = nullptr, *dst = nullptr;
    float *sink = nullptr;
    int rc = MCRT_OK;
    EventPair ev;
    const uint32_t cu = (uint32_t)prop.multiProcessorCount;
    if (hipMalloc(&src, n * sizeof(float4)) != hipSuccess || hipMalloc(&dst, n * sizeof(float4)) != hipSuccess ||
        hipMalloc(&sink, (size_t)cu * 8u * 256u * sizeof(float)) != hipSuccess) {
        rc = MCRT_ERR_OOM;
    } else if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) {
        rc = MCRT_ERR_HIP;
    } else {
        // copy: best of 5 after one warm-up
        const uint32_t blocks = cu * 8u;
        float best = 1e30f;
        (void)hipMemsetAsync(src, 0, n * sizeof(float4), nullptr);
        for (int r = 0; r < 6 && rc == MCRT_OK; ++r) {
            (void)hipEventRecord(ev.a, nullptr);
            hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, src, dst, n);
            (void)hipEventRecord(ev.b, nullptr);
            float ms = 0.0f;
            if (hipEventSynchronize(ev.b) != hipSuccess || hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess)
                rc = MCRT_ERR_HIP;
            else if (r > 0 && ms < best)
                best = ms;
        }
        if (rc == MCRT_OK) *copy_gbs = 2.0 * (double)n * sizeof(float4) / (best * 1e-3) / 1e9;
        // MFMA: 8 waves per CU's 4 SIMDs x 2, best of 3 after one warm-up
        const uint32_t iters = 20000;
        best = 1e30f;
        for (int r = 0; r < 4 && rc == MCRT_OK; ++r) {
            (void)hipEventRecord(ev.a, nullptr);
            hipLaunchKernelGGL(k_mfma_probe, dim3(cu * 8u), dim3(256), 0, nullptr, iters, 1e-3f, sink);
            (void)hipEventRecord(ev.b, nullptr);
            float ms = 0.0f;
            if (hipEventSynchronize(ev.b) != hipSuccess || hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess)
                rc = MCRT_ERR_HIP;
            else if (r > 0 && ms < best)
                best = ms;
        }
        if (rc == MCRT_OK) {
            const double waves = (double)cu * 8.0 * 4.0;
            const double flop = waves * iters * kMfmaChains * 16.0 * 16.0 * 32.0 * 2.0;
            *mfma_bf16_tflops = flop / (best * 1e-3) / 1e12;
        }
    }
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (sink) (void)hipFree(sink);
    return rc;
}
