/*
 * mcrt_abi.h — C-ABI of libmcrt_hip.so, the MI355X (gfx950) render hot path.
 *
 * Part 1 restates the Pod buffer contract of the reference
 * (src/gpu_layout.rs:78-234 bindings, the repr(C) structs listed per type).
 * The byte layouts are identical, so buffers the reference host builds can be
 * handed over unchanged.
 *
 * Part 2 declares the entry points.  Each one replaces a step of the
 * reference's headless driver (src/headless.rs:112-386) or of its wgpu
 * binding group, cited per function.
 *
 * Conventions
 *  - Every function returns an int status: 0 = MCRT_OK, negative = error.
 *    The per-context message is returned by mcrt_last_error().  Nothing in
 *    the library aborts the process.
 *  - Host pointers are read during the call only and are never retained.
 *  - "_dev" variants take device pointers and a stream (hipStream_t passed as
 *    void*; NULL = the null stream).  They enqueue work and return without
 *    synchronising; the caller owns the memory and keeps it alive until the
 *    stream has drained.
 *  - One context per host thread.  Calls on one context are not thread-safe.
 */
#ifndef MCRT_ABI_H
#define MCRT_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Part 1: Pod contract types (byte-identical to the reference)             */
/* ------------------------------------------------------------------------ */

/* src/scene.rs:22-29 — 32 B.  center_r.xyz = centre, .w = radius;
 * mat_and_pad[0] = material index. */
typedef struct GpuSphere {
    float center_r[4];
    uint32_t mat_and_pad[4];
} GpuSphere;

/* src/bvh.rs:54-68 — 48 B, pre-order flat BVH, root = 0.
 * Internal (prim_count == 0): left child = index + 1, right = right_or_offset.
 * Leaf (prim_count > 0): primitives [right_or_offset, +prim_count). */
typedef struct GpuBvhNode {
    float aabb_min[4];
    float aabb_max[4];
    uint32_t right_or_offset;
    uint32_t prim_count;
    uint32_t _pad[2];
} GpuBvhNode;

/* src/mesh.rs:24-34 — 48 B. */
typedef struct GpuVertex {
    float position[4];
    float normal[4];
    float uv[4];
} GpuVertex;

/* src/mesh.rs:39-46 — 16 B.  Bit 31 of mat_idx = back-face cull flag
 * (src/scene.rs:709-718). */
typedef struct GpuTriangle {
    uint32_t v[3];
    uint32_t mat_idx;
} GpuTriangle;

/* src/material.rs:35-44 — 48 B.  type_pad[0]: 0 lambertian, 1 metal,
 * 2 dielectric, 3 emissive; type_pad[1]: albedo texture layer (0 = solid).
 * Emissive: albedo_fuzz.xyz = colour, ior_pad[0] = strength. */
typedef struct GpuMaterial {
    uint32_t type_pad[4];
    float albedo_fuzz[4];
    float ior_pad[4];
} GpuMaterial;

#define MCRT_MAX_MATERIALS 64u /* src/material.rs:17 */

/* src/material.rs:54-63 — 3088 B. */
typedef struct GpuMaterialData {
    uint32_t mat_count;
    uint32_t n_emissive;
    uint32_t _pad[2];
    GpuMaterial materials[MCRT_MAX_MATERIALS];
} GpuMaterialData;

/* src/camera.rs:35-53 — 112 B. */
typedef struct CameraUniform {
    float origin[4];
    float lower_left[4];
    float horizontal[4];
    float vertical[4];
    float defocus_u[4];
    float defocus_v[4];
    uint32_t frame_count;
    uint32_t _pad[3];
} CameraUniform;

/* src/gpu.rs:164-183 — 16 B.  Defaults: radius 3, sigma_s 2, sigma_d 0.5,
 * normal_pow 8. */
typedef struct DenoiseParams {
    int32_t radius;
    float sigma_s;
    float sigma_d;
    float normal_pow;
} DenoiseParams;

/* Constants of src/shaders/path_trace.wgsl:36-66. */
#define MCRT_T_MIN 1e-4f
#define MCRT_T_MAX 1e9f
#define MCRT_MAX_FRAME_COUNT 65535u
#define MCRT_MAX_BOUNCES 8u
#define MCRT_MAT_LAMBERTIAN 0u
#define MCRT_MAT_METAL 1u
#define MCRT_MAT_DIELECTRIC 2u
#define MCRT_MAT_EMISSIVE 3u
#define MCRT_BACKFACE_CULL_FLAG 0x80000000u
#define MCRT_ENV_MAP_WIDTH 1024u  /* src/texture.rs:33 */
#define MCRT_ENV_MAP_HEIGHT 512u  /* src/texture.rs:36 */

/* Primitive ids reported by the closest-hit query. */
#define MCRT_ID_MISS 0xFFFFFFFFu
#define MCRT_ID_SPHERE_BIT 0x80000000u /* id = bit | original sphere index */

/* ------------------------------------------------------------------------ */
/* Part 2: entry points                                                      */
/* ------------------------------------------------------------------------ */

enum {
    MCRT_OK = 0,
    MCRT_ERR_NO_DEVICE = -1,
    MCRT_ERR_BAD_ARG = -2,
    MCRT_ERR_OOM = -3,
    MCRT_ERR_HIP = -4,
    MCRT_ERR_LIMIT = -5,    /* e.g. more than 64 materials */
    MCRT_ERR_NO_SCENE = -6, /* render/trace before mcrt_upload_scene */
};

typedef struct mcrt_ctx mcrt_ctx;

/* Replaces the adapter/device request of src/headless.rs:135-148: binds a
 * context to HIP device `device`.  Fails with MCRT_ERR_NO_DEVICE when the
 * device does not exist (the reference: "no suitable GPU adapter"). */
int mcrt_create(mcrt_ctx **out_ctx, int device);
void mcrt_destroy(mcrt_ctx *ctx);

/* Message of the last failed call on `ctx` ("" if none).  With ctx == NULL,
 * the message of the last failed mcrt_create or context-free call
 * (mcrt_validate_bvh) on this thread. */
const char *mcrt_last_error(const mcrt_ctx *ctx);

/* Replaces gpu_layout::build_scene_buffers (src/gpu_layout.rs:259-390) after
 * the BVHs are built: uploads the scene in the reference layouts and derives
 * the device-side traversal layout from them.
 *  - spheres/sphere_bvh: output of build_bvh (src/bvh.rs:239), BVH-ordered.
 *  - verts/tris/mesh_bvh: output of build_mesh_bvh (src/mesh.rs:82),
 *    triangles in BVH leaf order.
 *  - tri_orig_index: per ordered triangle, its index before the BVH build
 *    (reported by the closest-hit query); NULL = identity.
 *  - sphere_orig_index: likewise for spheres; NULL = identity.
 *  - Empty sets may be passed as n = 0 or as the reference's stub entries
 *    (STUB_SPHERE, zeroed node, src/gpu_layout.rs:264-312).
 *  - At most 2^30 - 1 triangles and 2^30 - 1 spheres (MCRT_ERR_LIMIT).
 *  - env_rgba32f: env_w x env_h RGBA32F equirectangular map, row 0 = zenith
 *    (src/texture.rs:158-220).  Required.
 *  - albedo_rgba8: n_layers x size x size RGBA8 sRGB texture array
 *    (src/gpu_layout.rs:435-477); NULL / 0 layers = solid colours only, as
 *    the shader's empty-array guard (path_trace.wgsl:509-511). */
int mcrt_upload_scene(mcrt_ctx *ctx,
                      const GpuSphere *spheres, uint32_t n_spheres,
                      const uint32_t *sphere_orig_index,
                      const GpuBvhNode *sphere_bvh, uint32_t n_sphere_nodes,
                      const GpuVertex *verts, uint32_t n_verts,
                      const GpuTriangle *tris, uint32_t n_tris,
                      const uint32_t *tri_orig_index,
                      const GpuBvhNode *mesh_bvh, uint32_t n_mesh_nodes,
                      const GpuMaterialData *materials,
                      const float *env_rgba32f, uint32_t env_w, uint32_t env_h,
                      const uint8_t *albedo_rgba8, uint32_t n_layers,
                      uint32_t layer_size);

/* Replaces the spp loop of src/headless.rs:249-290 (one cs_main dispatch per
 * frame, path_trace.wgsl:626-662).  Frame i of the call uses
 * camera->frame_count = first_frame + i for seeding and the running-mean
 * weight 1/(min(f,65535)+1).  max_depth = bounce limit (the reference's
 * MAX_BOUNCES = 8; 1..64 accepted).
 *  - accum_in: W*H RGBA32F running mean after frames [0, first_frame), or
 *    NULL for zeros (a fresh render, like the zero-initialised texture).
 *  - out_accum: W*H RGBA32F running mean after the call (alpha = 1).
 *  - out_gbuf: W*H RGBA32F G-buffer of the last frame (normal.xyz, depth),
 *    sky = (0,0,0,1e9); may be NULL.
 * Host-pointer variant; blocks until done. */
int mcrt_render(mcrt_ctx *ctx, const CameraUniform *camera,
                uint32_t width, uint32_t height, uint32_t spp,
                uint32_t first_frame, uint32_t max_depth,
                const float *accum_in, float *out_accum, float *out_gbuf);

int mcrt_render_dev(mcrt_ctx *ctx, const CameraUniform *camera,
                    uint32_t width, uint32_t height, uint32_t spp,
                    uint32_t first_frame, uint32_t max_depth,
                    const float *accum_in_dev, float *out_accum_dev,
                    float *out_gbuf_dev, void *stream);

/* Opt-in next-event estimation for the mcrt_render* calls of this context
 * (default off: the reference estimator, BSDF sampling only,
 * path_trace.wgsl:583-620).  When on, every Lambertian vertex also samples
 * one emitter (emissive triangles and the emissive spheres that binding 12
 * lists, src/gpu_layout.rs:51-52, src/scene.rs:912-923) and adds its light
 * through a shadow ray, combined with BSDF sampling by multiple importance
 * sampling (power heuristic).  The expectation is the reference
 * estimator's; the noise is lower.  BSDF draws are unchanged: light samples
 * use a separate PCG stream. */
int mcrt_set_nee(mcrt_ctx *ctx, int enabled);

/* mcrt_render_dev plus a first-hit albedo buffer (W*H RGBA32F, last frame;
 * dielectric = 1, sky = 0).  No counterpart in the reference, whose G-buffer
 * holds normal and depth only: it is the extra input of the learned
 * denoiser. */
int mcrt_render_aux_dev(mcrt_ctx *ctx, const CameraUniform *camera,
                        uint32_t width, uint32_t height, uint32_t spp,
                        uint32_t first_frame, uint32_t max_depth,
                        const float *accum_in_dev, float *out_accum_dev,
                        float *out_gbuf_dev, float *out_albedo_dev, void *stream);

/* Replaces cs_denoise (src/shaders/denoise.wgsl:52-120) as dispatched by
 * src/gpu.rs:1150-1169.  accum/gbuf/out: W*H RGBA32F.  params == NULL uses
 * the reference defaults (src/gpu.rs:177-183).  Any radius up to 256 is
 * filtered; a negative radius copies every pixel (the shader's loops run no
 * tap and its weight sum stays 0).  Radii above 256 ((2R+1)^2 > 263k taps
 * per pixel) return MCRT_ERR_LIMIT.  The _dev call uploads nothing and never
 * synchronises: the spatial weights are formed in the kernel. */
int mcrt_denoise_bilateral(mcrt_ctx *ctx, const DenoiseParams *params,
                           uint32_t width, uint32_t height,
                           const float *accum, const float *gbuf, float *out);

int mcrt_denoise_bilateral_dev(mcrt_ctx *ctx, const DenoiseParams *params,
                               uint32_t width, uint32_t height,
                               const float *accum_dev, const float *gbuf_dev,
                               float *out_dev, void *stream);

/* Closest-hit query over the uploaded scene: world_hit
 * (path_trace.wgsl:471-479) for n rays.  origins/dirs: n*3 floats (xyz
 * interleaved); dirs need not be unit length.  Outputs: primitive id (original
 * triangle index, MCRT_ID_SPHERE_BIT|sphere index, or MCRT_ID_MISS) and hit t
 * (-1 on a miss).  Among primitives whose t is bitwise equal, the lowest
 * original index wins; a sphere wins a tie against a triangle, as in
 * world_hit.  Valid t lies in [tmin, tmax). */
int mcrt_trace_closest(mcrt_ctx *ctx, const float *origins, const float *dirs,
                       uint32_t n, float tmin, float tmax,
                       uint32_t *out_id, float *out_t);

int mcrt_trace_closest_dev(mcrt_ctx *ctx, const float *origins_dev,
                           const float *dirs_dev, uint32_t n, float tmin,
                           float tmax, uint32_t *out_id_dev, float *out_t_dev,
                           void *stream);

/* mcrt_trace_closest_dev that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * scene is still uploaded by mcrt_upload_scene; the call's own scratch (the
 * persistent grid's work counter and its traversal-stack overflow) lives in a
 * static 32 MiB device array that each GPU the library runs on holds until
 * the process ends.  Deep trees shrink the persistent grid to fit it; a tree
 * too deep for one block returns MCRT_ERR_LIMIT.  Every call on a GPU shares
 * that GPU's array, so calls that overlap on one GPU (other streams, other
 * contexts, other host threads) race: serialise them.  Results are those of
 * mcrt_trace_closest_dev. */
int mcrt_trace_closest_noalloc_dev(mcrt_ctx *ctx, const float *origins_dev,
                                   const float *dirs_dev, uint32_t n, float tmin,
                                   float tmax, uint32_t *out_id_dev, float *out_t_dev,
                                   void *stream);

/* Binned-SAH BVH builders, src/mesh.rs:82-275 and src/bvh.rs:239-463 (8 bins,
 * leaf <= 4, pre-order, left = i + 1).  Host-only; no device needed.
 * Outputs are caller-allocated: out_nodes with room for 2*n - 1 nodes,
 * out_ordered / out_orig_index with n entries.  *out_n_nodes receives the
 * node count (0 for n = 0, as the reference).  Centroid ties in the split
 * sort are broken by original index (the reference's sort_unstable_by leaves
 * them unspecified). */
int mcrt_build_mesh_bvh(const GpuVertex *verts, uint32_t n_verts,
                        const GpuTriangle *tris, uint32_t n_tris,
                        GpuBvhNode *out_nodes, uint32_t *out_n_nodes,
                        GpuTriangle *out_ordered, uint32_t *out_orig_index);

/* Checks a reference pre-order BVH (src/bvh.rs:54-68 layout, as produced by
 * build_mesh_bvh / build_bvh) the way mcrt_upload_scene does before it
 * converts one: children inside the pre-order layout, every node reached from
 * the root exactly once, leaves within [0, n_prims) and disjoint, every
 * primitive in some leaf.  The reference's empty forms (n_nodes = 0, or one
 * zeroed node) pass.  Host-only; MCRT_ERR_BAD_ARG with the reason in
 * mcrt_last_error(NULL).  No counterpart call in the reference, whose host
 * builds its trees itself (src/gpu_layout.rs:259-340 uploads them unchecked). */
int mcrt_validate_bvh(const GpuBvhNode *nodes, uint32_t n_nodes, uint32_t n_prims);

int mcrt_build_sphere_bvh(const GpuSphere *spheres, uint32_t n_spheres,
                          GpuBvhNode *out_nodes, uint32_t *out_n_nodes,
                          GpuSphere *out_ordered, uint32_t *out_orig_index);

/* Differentiable render + path-replay adjoint.  No counterpart in the
 * reference: the reference estimator (path_trace.wgsl:583-662) extended with
 * up to 8 "parameter materials" (Lambertian, RGB albedo, white emission added
 * at every vertex on them; with emission 0 exactly the reference Lambertian).
 *  - params: n_params x (albedo.r, albedo.g, albedo.b, emission), host memory.
 *  - param_of_material: 64 entries, material slot -> parameter index or -1
 *    (the material table's own entry is used).
 *  - image: W*H RGBA32F, the mean of frames first_frame..first_frame+spp-1.
 *  - d_image: W*H RGBA32F, dLoss/dimage (alpha ignored).
 *  - grad: n_params x 4 floats, dLoss/dparams, summed in a fixed order.
 * The adjoint re-traces every sample's path from its PCG stream (path
 * replay) and needs no memory beyond its output; see mcrt_diff_set_mode. */
int mcrt_diff_render(mcrt_ctx *ctx, const CameraUniform *camera, uint32_t width,
                     uint32_t height, uint32_t spp, uint32_t first_frame,
                     uint32_t max_depth, const float *params, uint32_t n_params,
                     const int32_t *param_of_material, float *out_image);
int mcrt_diff_render_dev(mcrt_ctx *ctx, const CameraUniform *camera, uint32_t width,
                         uint32_t height, uint32_t spp, uint32_t first_frame,
                         uint32_t max_depth, const float *params, uint32_t n_params,
                         const int32_t *param_of_material, float *out_image_dev,
                         void *stream);
int mcrt_diff_adjoint(mcrt_ctx *ctx, const CameraUniform *camera, uint32_t width,
                      uint32_t height, uint32_t spp, uint32_t first_frame,
                      uint32_t max_depth, const float *params, uint32_t n_params,
                      const int32_t *param_of_material, const float *d_image,
                      float *out_grad);
int mcrt_diff_adjoint_dev(mcrt_ctx *ctx, const CameraUniform *camera, uint32_t width,
                          uint32_t height, uint32_t spp, uint32_t first_frame,
                          uint32_t max_depth, const float *params, uint32_t n_params,
                          const int32_t *param_of_material, const float *d_image_dev,
                          float *out_grad_dev, void *stream);

/* Adjoint form of this context's mcrt_diff_adjoint* calls:
 *  - MCRT_DIFF_REPLAY (default): one re-trace per sample, parameter
 *    derivatives carried along the path; no tape, no division by albedos.
 *  - MCRT_DIFF_TAPE: the trace stores each vertex (48 B x max_depth per
 *    thread of a grid of about 2^22 threads, held by the context), and the
 *    gradient is summed walking the tape backwards.
 * Both are exact for zero albedos and agree to FP32 rounding. */
#define MCRT_DIFF_REPLAY 0
#define MCRT_DIFF_TAPE 1
int mcrt_diff_set_mode(mcrt_ctx *ctx, int mode);

/* Learned denoiser.  No counterpart in the reference (its only denoiser is
 * the bilateral filter, which stays the fallback): six 3x3 convolutions
 * 10 -> 32 x5 -> 3 with ReLU and a global skip in log space, BF16 MFMA
 * kernels with FP32 accumulation, Adam on FP32 master weights.
 * Inputs per pixel (RGBA32F images, batch-major): noisy radiance (from
 * mcrt_render*), first-hit albedo (mcrt_render_aux_dev), G-buffer.
 * Weights in public order: per layer W[co][ci][ky][kx] then b[co];
 * mcrt_denoiser_param_count() = 40771. */
typedef struct mcrt_denoiser mcrt_denoiser;
int mcrt_denoiser_create(mcrt_ctx *ctx, uint32_t seed, mcrt_denoiser **out);
void mcrt_denoiser_destroy(mcrt_denoiser *dn);
uint32_t mcrt_denoiser_param_count(void);
int mcrt_denoiser_get_weights(mcrt_denoiser *dn, float *out);
int mcrt_denoiser_set_weights(mcrt_denoiser *dn, const float *in);
int mcrt_denoiser_infer_dev(mcrt_denoiser *dn, uint32_t batch, uint32_t width, uint32_t height,
                            const float *noisy_dev, const float *albedo_dev,
                            const float *gbuf_dev, float *out_dev, void *stream);
/* Training loss, mean over pixels and RGB (default LOG_L2):
 *   LOG_L2:  (log1p(out) - log1p(ref))^2
 *   REL_MSE: (out - ref)^2 / (ref^2 + 1e-2), the evaluation metric. */
#define MCRT_DENOISE_LOSS_LOG_L2 0
#define MCRT_DENOISE_LOSS_REL_MSE 1
int mcrt_denoiser_set_loss(mcrt_denoiser *dn, int kind);
/* One Adam step on the selected loss between the output and ref.
 * loss_out (host, optional) synchronises the stream. */
int mcrt_denoiser_train_step_dev(mcrt_denoiser *dn, uint32_t batch, uint32_t width,
                                 uint32_t height, const float *noisy_dev,
                                 const float *albedo_dev, const float *gbuf_dev,
                                 const float *ref_dev, float lr, float *loss_out,
                                 void *stream);
/* Loss gradient w.r.t. the public weights for one batch (no update);
 * device inputs, host outputs; blocks. */
int mcrt_denoiser_gradient(mcrt_denoiser *dn, uint32_t batch, uint32_t width, uint32_t height,
                           const float *noisy_dev, const float *albedo_dev,
                           const float *gbuf_dev, const float *ref_dev, float *out_grad,
                           float *loss_out);

/* Greedy voxel meshing (src/world/mesher.rs:97-400 as driven by
 * src/gpu.rs:386-458).  Host-only.  `blocks` holds chunks_x * chunks_z chunks
 * of 16 x 256 x 16 u16 block ids, chunk (ix, iz) at offset
 * (ix * chunks_z + iz) * 65536, each chunk XZY-major
 * (x * 4096 + y * 16 + z, src/world/chunk.rs:28-57).  Chunk (ix, iz) is
 * world chunk (cx0 + ix, cz0 + iz).  opaque[id] and block_material[id]
 * (256 entries each) give opacity and the material slot.  With out_verts and
 * out_tris NULL only the counts are returned; otherwise the buffers receive
 * the mesh (MCRT_ERR_LIMIT if max_verts / max_tris are too small). */
int mcrt_mesh_voxel_chunks(const uint16_t *blocks, uint32_t chunks_x, uint32_t chunks_z,
                           int32_t cx0, int32_t cz0, float voxel_size,
                           const uint8_t *opaque, const uint32_t *block_material,
                           GpuVertex *out_verts, uint64_t max_verts,
                           GpuTriangle *out_tris, uint64_t max_tris,
                           uint64_t *n_verts, uint64_t *n_tris);

/* Kernel timing.  When enabled, every launch of the named kernels is
 * bracketed by HIP events on its own stream; mcrt_kernel_stats returns the
 * summed milliseconds and the launch count since the last reset.
 * Names: "trace_closest", "extend", "shade", "generate", "accumulate",
 * "bilateral". */
int mcrt_set_profiling(mcrt_ctx *ctx, int enabled);
int mcrt_kernel_stats(mcrt_ctx *ctx, const char *kernel, double *total_ms,
                      uint64_t *launches);
int mcrt_reset_stats(mcrt_ctx *ctx);

/* Counters of the last mcrt_render*: path segments traced (rays) and
 * samples produced.  Synchronises the context's last stream. */
int mcrt_render_counters(mcrt_ctx *ctx, uint64_t *segments, uint64_t *samples);

/* Peak probes for roofline fractions, no counterpart in the reference: a
 * one-float4-per-thread copy between two buffers of copy_bytes each (GB/s of
 * bytes read + written), loops of register-operand MFMAs on every SIMD:
 * v_mfma_f32_32x32x16_bf16 (the dense BF16 peak) and v_mfma_f32_16x16x32_bf16
 * (TFLOP/s), and independent chains of non-packed FP32 fma at 8 waves per
 * SIMD (TFLOP/s), timed with HIP events, best of several runs.  Allocates
 * 2 x copy_bytes on `device` for the call; blocks. */
int mcrt_measure_peaks(int device, uint64_t copy_bytes, double *copy_gbs, double *mfma_bf16_tflops,
                       double *mfma_bf16_16x16x32_tflops, double *valu_fp32_tflops);

#ifdef __cplusplus
}
#endif

#endif /* MCRT_ABI_H */
