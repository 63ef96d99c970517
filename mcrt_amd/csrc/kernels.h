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
    uint32_t n_paths;  // pixels * frames generated by this launch
    // This launch's share of the batch: frames fo .. fo + n_paths / pn - 1
    // (batch-relative) of pixels po .. po + pn - 1.  Path id = (fo + frame) * wh
    // + pixel: frame-major over the whole batch whatever the split.
    uint32_t fo, po, pn;
    float4 *A, *B, *C;  // C null: the first k_shade recomputes the throughput
    float4 *samples;   // next-event estimation: zeroed per path (else null)
    float *P;          // next-event estimation: BSDF pdf of the arriving ray, 0 = camera
};

// Camera rays of a first-bounce queue, recomputed from the queue slot instead
// of being written by k_generate and read back (pinhole cameras, renders
// without next-event estimation).  k_shade recomputes the ray, PCG state and
// unit throughput (camera_ray), so k_generate writes no throughput; the
// sphere-free plain-stack k_extend recomputes the ray too (pinhole_ray), and
// k_generate is then skipped.  Slot order as k_generate's: pixel-major, nf
// frames per pixel, of pixels po.. and frames fo.. of the batch; frame0 =
// frame_count of the launch's first frame.  GenArgs::C may be null.
struct CamGen {
    CameraUniform cam;
    uint32_t width, height, wh;
    uint32_t frame0, nf, fo, po;
    uint32_t on;  // nonzero: the queue is this launch's camera rays
};

struct ExtArgs {
    DevScene scene;
    CamGen gen;
    const float4 *A, *B;
    float4 *H;
    // Optional shadow rays traced in the same launch, queued after the *count
    // extension rays (null occ_count: none); layout as OccArgs.
    const float4 *S_o, *S_d, *S_c;
    const uint32_t *occ_count;
    float4 *samples;
    const uint32_t *count;
    uint32_t *work;
    uint32_t *ovf;  // traversal-stack overflow, (depth - short) x ovf_stride
    uint32_t ovf_stride;  // threads in the persistent grid
};

// Threads per k_shade block.  Survivors are enqueued with one atomic per
// block on a single device-scope counter, so larger blocks mean fewer of
// those serialised atomics.
constexpr uint32_t kShadeBlock = 512;

struct ShadeArgs {
    DevScene scene;
    CamGen gen;
    const float4 *A, *B, *C, *H;
    float4 *A_out, *B_out, *C_out;
    const uint32_t *count_in;
    uint32_t *count_out;
    uint32_t *work;
    float4 *samples;  // per path id
    float4 *gbuf;     // may be null
    float4 *albedo;   // may be null: first-hit albedo of the G-buffer frame
    uint32_t gbuf_frame_index;  // batch-relative frame whose G-buffer is kept
    uint32_t wh;
    uint32_t max_depth;
    // Next-event estimation (nee != 0): P / P_out = BSDF solid-angle pdf of
    // the arriving / leaving ray (0 after the camera or a specular event);
    // shadow-ray queue S_o = (origin, path id), S_d = (dir, t_max),
    // S_c = (MIS-weighted contribution, -), appended at *shadow_count.
    uint32_t nee;
    const float *P;
    float *P_out;
    float4 *S_o, *S_d, *S_c;
    uint32_t *shadow_count;
};

struct OccArgs {
    DevScene scene;
    const float4 *S_o, *S_d, *S_c;
    const uint32_t *count;
    uint32_t *work;
    float4 *samples;
    uint32_t *ovf;
    uint32_t ovf_stride;
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
    int32_t radius;        // any value up to kMaxBilateralRadius; negative = no tap
    float inv2ss, inv2ds, normal_pow;
};

// Largest radius the bilateral launcher accepts: a 513 x 513 footprint,
// 263k taps per pixel.
constexpr int32_t kMaxBilateralRadius = 256;

constexpr int kMaxParams = 8;

struct DiffArgs {
    DevScene scene;
    CameraUniform cam;
    uint32_t width, height, spp, frame0, max_depth, n_params;
    uint32_t groups;         // frame groups per pixel (one thread each); diff_groups()
    float4 params[kMaxParams];  // (albedo.rgb, emission)
    int8_t pmap[MCRT_MAX_MATERIALS];  // material slot -> parameter index or -1
    float4 *image;           // primal: W*H mean radiance
    float4 *image_part;      // primal with groups > 1: groups x W*H radiance sums
    const float4 *d_image;   // adjoint: dLoss/dimage
    float *partial;          // adjoint: per-block sums, kMaxParams*4 per block
    float4 *tape;            // tape adjoint: per-thread vertices, 3 float4 x max_depth x tape_stride;
                             // null selects the replay adjoint
    uint32_t tape_stride;    // adjoint: threads in the grid
    float *grad;             // adjoint: n_params*4
    uint32_t *ovf;
    uint32_t ovf_stride;
};

hipError_t launch_diff(const DiffArgs &a, bool adjoint, hipStream_t s);
uint32_t diff_groups(uint32_t wh, uint32_t spp);
uint32_t diff_grid(uint32_t wh, uint32_t groups);

// --- learned denoiser (learned_kernels.hip) ---------------------------------
struct PackArgs {
    const float4 *noisy, *albedo, *gbuf;
    __bf16 *out;  // [B][H][W][32]
    int B, H, W;
};

struct ConvArgs {
    const __bf16 *in;    // [B][H][W][32]
    const __bf16 *mask;  // optional: in *= (mask > 0)
    const __bf16 *w;     // [9][32 co][32 ci]
    const float *bias;   // optional [32]
    __bf16 *out;         // modes 0, 1
    float4 *out_log;     // mode 2
    const float4 *noisy; // mode 2 residual source
    // Inference only (mode 2): write the denoised image expm1(out_log),
    // alpha 1, instead of out_log.
    float4 *out_final;
    int B, H, W;
    int mode;            // 0: ReLU, BF16 out; 1: linear, BF16 out; 2: final layer
    int cin16;           // mode 0 only: the input has 16 channels ([B][H][W][16], k_pack16)
};

struct WgradArgs {
    const __bf16 *x;     // layer input
    const __bf16 *dy;    // gradient w.r.t. the layer output
    const __bf16 *mask;  // optional: dy *= (mask > 0), mask = layer output
    float *partial;      // [blocks][9*32*32 + 32]
    int B, H, W;
};

struct LossArgs {
    const float4 *out_log, *ref;
    __bf16 *grad;        // [B][H][W][32]
    float *partial;      // per block
    float inv_count;     // 1 / (B*H*W*3)
    int B, H, W;
    int kind;            // 0: L2 in log1p space; 1: relMSE, (x - r)^2 / (r^2 + 1e-2)
};

struct AdamArgs {
    float *w, *m, *v;
    const float *grad;
    __bf16 *wf, *wb;
    float lr, beta1, beta2, eps, bc1, bc2;
};

constexpr int kDenoiseLayers = 6;
constexpr int kLayerParams = 9 * 32 * 32 + 32;

hipError_t launch_pack(const PackArgs &a, hipStream_t s);
hipError_t launch_pack16(const PackArgs &a, hipStream_t s);  // out: [B][H][W][16]
hipError_t launch_conv3x3(const ConvArgs &a, hipStream_t s);
hipError_t launch_wgrad(const WgradArgs &a, uint32_t blocks, float *out_grad, hipStream_t s);
hipError_t launch_loss(const LossArgs &a, float *loss_out, hipStream_t s);
hipError_t launch_adam(const AdamArgs &a, hipStream_t s);
hipError_t launch_weights_to_bf16(const float *w, void *wf, void *wb, hipStream_t s);
hipError_t launch_unpack(const float4 *out_log, float4 *out, size_t n, hipStream_t s);

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s);
hipError_t launch_generate(const GenArgs &a, hipStream_t s);
hipError_t launch_extend(const ExtArgs &a, uint32_t grid, hipStream_t s);
// The first bounce of a sphere-free scene, as wave-uniform packets.
hipError_t launch_extend_packet(const ExtArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_shade(const ShadeArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_accumulate(const AccArgs &a, hipStream_t s);
hipError_t launch_trace(const TraceArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_occlude(const OccArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_bilateral(const BilateralArgs &a, hipStream_t s);
// which: 0 extend, 1 shade, 2 trace, 3 occlude (the traversal variant the scene selects)
hipError_t occupancy(int which, const DevScene &sc, int *blocks_per_cu);
int trace_block();
// LDS entries and 32-bit words per entry of the plain (cull = false) or
// culling traversal stack, for sizing the overflow columns.
void stack_layout(bool cull, uint32_t *short_entries, uint32_t *words_per_entry);

}  // namespace mcrt
