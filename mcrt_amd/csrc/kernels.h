// kernels.h — argument blocks and host launchers of the HIP kernels
// (internal to libmcrt_hip.so).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"

namespace mcrt {

struct GenArgs {
    CameraUniform cam;
    uint32_t width, height, wh;
    uint32_t frame0;   // frame of path 0 (frame_count of the first frame)
    uint32_t n_paths;  // wh * frames in this batch
    float4 *A, *B, *C;
};

struct ExtArgs {
    DevScene scene;
    const float4 *A, *B;
    float4 *H;
    const uint32_t *count;
    uint32_t *work;
    uint32_t *ovf;  // traversal-stack overflow, (depth - short) x ovf_stride
    uint32_t ovf_stride;  // threads in the persistent grid
};

struct ShadeArgs {
    DevScene scene;
    const float4 *A, *B, *C, *H;
    float4 *A_out, *B_out, *C_out;
    const uint32_t *count_in;
    uint32_t *count_out;
    uint32_t *work;
    float4 *samples;  // per path id
    float4 *gbuf;     // may be null
    uint32_t gbuf_frame_index;  // batch-relative frame whose G-buffer is kept
    uint32_t wh;
    uint32_t max_depth;
};

struct AccArgs {
    const float4 *samples;
    const float4 *accum_in;  // may be null (zeros)
    float4 *accum_out;
    uint32_t wh, n_frames, frame0;
};

struct TraceArgs {
    DevScene scene;
    const float *origins, *dirs;
    uint32_t n;
    float tmin, tmax;
    uint32_t *out_id;
    float *out_t;
    uint32_t *work;
    uint32_t *ovf;
    uint32_t ovf_stride;
};

struct BilateralArgs {
    const float4 *accum, *gbuf;
    float4 *out;
    int32_t width, height;
    int32_t radius;
    float inv2ds, normal_pow;
    const float *spatial;  // (2R+1)^2 weights exp(-r^2/(2 sigma_s^2)), row-major
};

constexpr int kMaxParams = 8;

struct DiffArgs {
    DevScene scene;
    CameraUniform cam;
    uint32_t width, height, spp, frame0, max_depth, n_params;
    float4 params[kMaxParams];  // (albedo.rgb, emission)
    int8_t pmap[MCRT_MAX_MATERIALS];  // material slot -> parameter index or -1
    float4 *image;           // primal: W*H mean radiance
    const float4 *d_image;   // adjoint: dLoss/dimage
    float *partial;          // adjoint: per-block sums, kMaxParams*4 per block
    float *grad;             // adjoint: n_params*4
    uint32_t *ovf;
    uint32_t ovf_stride;
};

hipError_t launch_diff(const DiffArgs &a, bool adjoint, hipStream_t s);
uint32_t diff_grid(uint32_t wh);

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s);
hipError_t launch_generate(const GenArgs &a, hipStream_t s);
hipError_t launch_extend(const ExtArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_shade(const ShadeArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_accumulate(const AccArgs &a, hipStream_t s);
hipError_t launch_trace(const TraceArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_bilateral(const BilateralArgs &a, hipStream_t s);
// which: 0 extend, 1 shade, 2 trace
hipError_t occupancy(int which, bool spheres, int *blocks_per_cu);
int trace_block();
uint32_t short_stack();

}  // namespace mcrt
