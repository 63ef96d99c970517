/*
 * mcrt_oracle.h — CPU oracle for the render hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP32 restatement of the
 * reference shaders src/shaders/path_trace.wgsl and src/shaders/denoise.wgsl
 * and of the headless tonemap (src/headless.rs:93-98).  Only tests, the smoke
 * check and the CPU-baseline leg of the benchmark may call it.  The product
 * (libmcrt_hip.so) never links or calls it.
 *
 * Built with -ffp-contract=off so every arithmetic step rounds as written.
 */
#ifndef MCRT_ORACLE_H
#define MCRT_ORACLE_H

#include <stdint.h>
#include "../include/mcrt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleScene {
    const GpuSphere *spheres;
    uint32_t n_spheres;
    const uint32_t *sphere_orig; /* NULL = identity */
    const GpuBvhNode *sphere_bvh;
    uint32_t n_sphere_nodes;
    const GpuVertex *verts;
    uint32_t n_verts;
    const GpuTriangle *tris; /* BVH leaf order */
    uint32_t n_tris;
    const uint32_t *tri_orig; /* NULL = identity */
    const GpuBvhNode *mesh_bvh;
    uint32_t n_mesh_nodes;
    const GpuMaterialData *mats;
    const float *env; /* env_w * env_h * 4 */
    uint32_t env_w, env_h;
} OracleScene;

/* Traversal statistics of the reference-order traversal. */
typedef struct OracleCounts {
    uint64_t nodes; /* BVH nodes fetched (popped and slab-tested) */
    uint64_t tris;  /* Möller–Trumbore tests */
    uint64_t spheres;
} OracleCounts;

/* PCG hash and uniform draw, path_trace.wgsl:194-204. */
uint32_t oracle_pcg_hash(uint32_t v);
float oracle_rand_f32(uint32_t *state);
/* Per-pixel seed, path_trace.wgsl:636-637. */
uint32_t oracle_pixel_seed(uint32_t x, uint32_t y, uint32_t width,
                           uint32_t frame);

/* Möller–Trumbore, path_trace.wgsl:300-317.  out[4] = (t,u,v,1) or
 * (-1,0,0,0). */
void oracle_ray_triangle(const float o[3], const float d[3], const float v0[3],
                         const float v1[3], const float v2[3], float t_min,
                         float t_max, float out[4]);
/* Slab test, path_trace.wgsl:248-256. */
int oracle_ray_aabb(const float o[3], const float d[3], const float bmin[3],
                    const float bmax[3], float t_min, float t_max);

/* Closest hit, literal reference order (world_hit, path_trace.wgsl:343-479):
 * first-visited wins ties.  Returns id as in mcrt_trace_closest, t in *t. */
uint32_t oracle_closest_ref(const OracleScene *s, const float o[3],
                            const float d[3], float t_min, float t_max,
                            float *t_out, OracleCounts *counts);
/* Closest hit with the order-free tie rule of mcrt_trace_closest: lowest
 * original index among bitwise-equal t; sphere beats triangle at equal t. */
uint32_t oracle_closest_exact(const OracleScene *s, const float o[3],
                              const float d[3], float t_min, float t_max,
                              float *t_out);
/* Brute force over every primitive with the same tie rule (no BVH). */
uint32_t oracle_closest_brute(const OracleScene *s, const float o[3],
                              const float d[3], float t_min, float t_max,
                              float *t_out);

/* Batched queries over n rays (origins/dirs n*3), multithreaded.
 * mode 0 = ref, 1 = exact, 2 = brute.  counts may be NULL (mode 0 only). */
void oracle_trace_batch(const OracleScene *s, const float *origins,
                        const float *dirs, uint64_t n, float t_min,
                        float t_max, int mode, uint32_t *out_id, float *out_t,
                        OracleCounts *counts, int n_threads);

/* Render: cs_main (path_trace.wgsl:626-662) for frames
 * first_frame..first_frame+spp-1 with running-mean accumulation.
 * accum_in may be NULL (zeros).  out_accum/out_gbuf W*H*4 (gbuf may be NULL).
 * max_depth replaces MAX_BOUNCES.  counts may be NULL; if given, receives
 * traversal statistics summed over every segment.  segments may be NULL. */
void oracle_render(const OracleScene *s, const CameraUniform *cam,
                   uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t first_frame, uint32_t max_depth,
                   const float *accum_in, float *out_accum, float *out_gbuf,
                   OracleCounts *counts, uint64_t *segments, int n_threads);

/* oracle_render with the build's opt-in next-event estimator (emitters =
 * emissive triangles and spheres, chosen by area x luminance; power-heuristic
 * MIS against the cosine BSDF sampling; light draws from a separate PCG
 * stream).  No reference counterpart; its expectation is the reference
 * estimator's. */
void oracle_render_nee(const OracleScene *s, const CameraUniform *cam,
                       uint32_t width, uint32_t height, uint32_t spp,
                       uint32_t first_frame, uint32_t max_depth,
                       const float *accum_in, float *out_accum, float *out_gbuf,
                       int n_threads);

/* One sample of one pixel (path_color of cs_main) with its G-buffer texel. */
void oracle_sample(const OracleScene *s, const CameraUniform *cam,
                   uint32_t width, uint32_t height, uint32_t x, uint32_t y,
                   uint32_t frame, uint32_t max_depth, float out_rgb[3],
                   float out_gbuf[4]);

/* cs_denoise, denoise.wgsl:52-120.  Images W*H*4. */
void oracle_denoise(const DenoiseParams *p, uint32_t width, uint32_t height,
                    const float *accum, const float *gbuf, float *out,
                    int n_threads);

/* Differentiable-render reference: the reference estimator with up to 8
 * parameter materials (Lambertian albedo a_p + white emission E_p added at
 * each vertex on them), FP64 radiance over the FP32 path geometry of
 * oracle_render.  params: n_params x (r, g, b, E); pmap: 64 entries.
 * Images are W*H*3 doubles (mean over spp).  The gradient of
 * sum(d_img * image) is computed by forward-mode counting. */
void oracle_diff_render64(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                          uint32_t spp, uint32_t first, uint32_t depth, const double *params,
                          uint32_t n_params, const int32_t *pmap, double *out_img, int n_threads);
void oracle_diff_grad64(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                        uint32_t spp, uint32_t first, uint32_t depth, const double *params,
                        uint32_t n_params, const int32_t *pmap, const double *d_img,
                        double *grad, int n_threads);

/* Converged reference ("oracle B") on a pixel window of a W x H frame: the
 * reference estimator with FP64 accumulation of spp frames from `first`.
 * out: ch * cw * 3 doubles. */
void oracle_reference64_window(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                               uint32_t cx0, uint32_t cy0, uint32_t cw, uint32_t ch, uint32_t spp,
                               uint32_t first, uint32_t depth, double *out, int n_threads);

/* One scatter event (path_trace.wgsl:521-573) for tests. */
void oracle_scatter(const GpuMaterial *m, const float rd[3], const float n[3], int front,
                    uint32_t *rng, float out_dir[3], float out_att[3], int *absorbed);

/* Headless tonemap, src/headless.rs:93-98 and :359-362 (non-finite -> 0). */
uint8_t oracle_tonemap_u8(float v);

/* Procedural gradient sky, src/texture.rs:158-199.  out: w*h*4. */
void oracle_gradient_env(uint32_t w, uint32_t h, float *out);
/* sample_env (path_trace.wgsl:233-242) for one direction: out = rgb. */
void oracle_sample_env(const OracleScene *s, const float dir[3], float out[3]);

#ifdef __cplusplus
}
#endif

#endif
