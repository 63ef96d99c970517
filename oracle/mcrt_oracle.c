/*
 * mcrt_oracle.c — CPU oracle for the render hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C FP32 restatement of
 * src/shaders/path_trace.wgsl, src/shaders/denoise.wgsl and the headless
 * tonemap.  Each function cites the reference lines it follows.  Built with
 * -ffp-contract=off: every operation rounds as written, in the order written
 * in the shader.  Vector helpers evaluate component-wise left to right
 * (dot = x*x' + y*y' + z*z'; normalize = v * (1/sqrt(dot(v,v)))).
 *
 * Parity pins (see tests/test_oracle.py): PCG known answers, brute-force vs
 * BVH closest hit, closed-form quad hits, sampling statistics, Schlick closed
 * form, the furnace and direct-emission known answers, bilateral invariants.
 */
#include "mcrt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define TAU 6.283185307179586f /* path_trace.wgsl:36 */
#define PI_F 3.141592653589793f /* :37 */
#define MAX_MAT_IDX 63u         /* :43 */
#define STACK_SIZE 64u          /* :49 */

/* ------------------------------------------------------------------------ */
/* vec3 helpers                                                              */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 normalize(v3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return scl(inv, a);
}
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline float fminv(float a, float b) { return fminf(a, b); }
static inline float fmaxv(float a, float b) { return fmaxf(a, b); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* WGSL f32 -> u32 conversion saturates (negative and NaN -> 0). */
static inline uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

/* ------------------------------------------------------------------------ */
/* PCG, path_trace.wgsl:63-66, :194-204, :636-637                             */
/* ------------------------------------------------------------------------ */
uint32_t oracle_pcg_hash(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t rot = s >> 28u;
    s = ((s >> (rot + 4u)) ^ s) * 277803737u;
    return (s >> 22u) ^ s;
}

float oracle_rand_f32(uint32_t *state) {
    *state = oracle_pcg_hash(*state);
    uint32_t bits = (*state >> 9u) | 0x3f800000u;
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

uint32_t oracle_pixel_seed(uint32_t x, uint32_t y, uint32_t width, uint32_t frame) {
    return oracle_pcg_hash(oracle_pcg_hash(x + y * width) ^ (frame * 2654435761u));
}

/* ------------------------------------------------------------------------ */
/* Sampling helpers, path_trace.wgsl:210-227                                  */
/* ------------------------------------------------------------------------ */
static v3 random_unit_sphere(uint32_t *rng) {
    float cos_phi = 1.0f - 2.0f * oracle_rand_f32(rng);
    float sin_phi = sqrtf(fmaxv(0.0f, 1.0f - cos_phi * cos_phi));
    float theta = TAU * oracle_rand_f32(rng);
    return mk(sin_phi * cosf(theta), sin_phi * sinf(theta), cos_phi);
}

static v3 cosine_scatter(uint32_t *rng, v3 n) {
    v3 s = add(n, random_unit_sphere(rng));
    if (dot(s, s) < 1e-6f) return n;
    return normalize(s);
}

static void random_in_unit_disk(uint32_t *rng, float *ox, float *oy) {
    float r = sqrtf(oracle_rand_f32(rng));
    float theta = TAU * oracle_rand_f32(rng);
    *ox = r * cosf(theta);
    *oy = r * sinf(theta);
}

/* ------------------------------------------------------------------------ */
/* Environment map, path_trace.wgsl:233-242                                   */
/* ------------------------------------------------------------------------ */
static v3 sample_env(const OracleScene *s, v3 dir) {
    float phi = atan2f(dir.z, dir.x);
    float theta = asinf(clampf(dir.y, -1.0f, 1.0f));
    float u = phi / TAU + 0.5f;
    float v = 0.5f - theta / PI_F;
    uint32_t ix = f2u_sat(u * (float)s->env_w) % s->env_w;
    uint32_t iy = f2u_sat(v * (float)s->env_h) % s->env_h;
    const float *p = s->env + 4ull * ((uint64_t)iy * s->env_w + ix);
    return mk(p[0], p[1], p[2]);
}

/* Exported for the known-answer tests of the direction -> texel mapping. */
void oracle_sample_env(const OracleScene *s, const float dir[3], float out[3]) {
    v3 c = sample_env(s, mk(dir[0], dir[1], dir[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* Gradient sky, src/texture.rs:158-199. */
void oracle_gradient_env(uint32_t w, uint32_t h, float *out) {
    for (uint32_t y = 0; y < h; ++y) {
        float v = ((float)y + 0.5f) / (float)h;
        float e = 1.0f - 2.0f * v;
        float r, g, b;
        if (e >= 0.0f) {
            float t = e;
            r = 1.0f - 0.5f * t;
            g = 1.0f - 0.3f * t;
            b = 1.0f;
        } else {
            float t = fminf(-e, 1.0f);
            r = (1.0f - t) * 0.8f;
            g = (1.0f - t) * 0.4f;
            b = (1.0f - t) * 0.1f;
        }
        for (uint32_t x = 0; x < w; ++x) {
            float *p = out + 4ull * ((uint64_t)y * w + x);
            p[0] = r; p[1] = g; p[2] = b; p[3] = 1.0f;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Slab test, path_trace.wgsl:248-256                                         */
/* ------------------------------------------------------------------------ */
static inline void slab(v3 o, v3 d, const float *bmin, const float *bmax,
                        float *tenter, float *texit) {
    v3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    v3 t0 = mul(sub(ld3(bmin), o), inv);
    v3 t1 = mul(sub(ld3(bmax), o), inv);
    *tenter = fmaxv(fmaxv(fminv(t0.x, t1.x), fminv(t0.y, t1.y)), fminv(t0.z, t1.z));
    *texit = fminv(fminv(fmaxv(t0.x, t1.x), fmaxv(t0.y, t1.y)), fmaxv(t0.z, t1.z));
}

int oracle_ray_aabb(const float o[3], const float d[3], const float bmin[3],
                    const float bmax[3], float t_min, float t_max) {
    float te, tx;
    slab(ld3(o), ld3(d), bmin, bmax, &te, &tx);
    return tx > fmaxv(te, t_min) && te < t_max;
}

/* ------------------------------------------------------------------------ */
/* Ray-sphere, path_trace.wgsl:262-293                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    float t;
    v3 p, n;
    float u, v;
    int front;
    uint32_t mat;
    uint32_t id;   /* primitive id as reported by the closest-hit query */
    uint32_t prim; /* leaf-order index; bit 31 marks a sphere */
} Hit;

static Hit sphere_hit(v3 o, v3 d, const GpuSphere *sp, float t_min, float t_max) {
    Hit h;
    h.t = -1.0f;
    v3 c = ld3(sp->center_r);
    float radius = sp->center_r[3];
    v3 oc = sub(o, c);
    float a = dot(d, d);
    float hh = dot(oc, d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = hh * hh - a * cc;
    if (disc < 0.0f) return h;
    float sq = sqrtf(disc);
    float t = (-hh - sq) / a;
    if (t <= t_min || t >= t_max) {
        t = (-hh + sq) / a;
        if (t <= t_min || t >= t_max) return h;
    }
    h.t = t;
    h.p = add(o, scl(t, d));
    v3 outn = mk((h.p.x - c.x) / radius, (h.p.y - c.y) / radius, (h.p.z - c.z) / radius);
    h.front = dot(d, outn) < 0.0f;
    h.n = h.front ? outn : neg(outn);
    h.mat = sp->mat_and_pad[0];
    float phi = atan2f(outn.z, outn.x);
    float th = asinf(clampf(outn.y, -1.0f, 1.0f));
    h.u = phi / TAU + 0.5f;
    h.v = 0.5f - th / PI_F;
    return h;
}

/* ------------------------------------------------------------------------ */
/* Möller–Trumbore, path_trace.wgsl:300-317                                   */
/* ------------------------------------------------------------------------ */
static inline void mt(v3 o, v3 d, v3 v0, v3 v1, v3 v2, float t_min, float t_max,
                      float out[4]) {
    v3 e1 = sub(v1, v0);
    v3 e2 = sub(v2, v0);
    v3 h = cross(d, e2);
    float a = dot(e1, h);
    out[0] = -1.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 0.0f;
    if (fabsf(a) < 1e-8f) return;
    float f = 1.0f / a;
    v3 s = sub(o, v0);
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return;
    v3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return;
    float t = f * dot(e2, q);
    if (t < t_min || t >= t_max) return;
    out[0] = t; out[1] = u; out[2] = v; out[3] = 1.0f;
}

void oracle_ray_triangle(const float o[3], const float d[3], const float v0[3],
                         const float v1[3], const float v2[3], float t_min,
                         float t_max, float out[4]) {
    mt(ld3(o), ld3(d), ld3(v0), ld3(v1), ld3(v2), t_min, t_max, out);
}

/* Triangle candidate: interpolated normal, cull test and hit record,
 * path_trace.wgsl:434-460.  Returns 0 when back-face culled. */
static int tri_candidate(const OracleScene *s, uint32_t ti, v3 o, v3 d,
                         const float res[4], Hit *out) {
    const GpuTriangle *tri = &s->tris[ti];
    const GpuVertex *va = &s->verts[tri->v[0]];
    const GpuVertex *vb = &s->verts[tri->v[1]];
    const GpuVertex *vc = &s->verts[tri->v[2]];
    int cull = (tri->mat_idx >> 31u) != 0u;
    uint32_t clean = tri->mat_idx & 0x7FFFFFFFu;
    float t = res[0], u = res[1], v = res[2];
    float w = 1.0f - u - v;
    v3 n = normalize(add(add(scl(w, ld3(va->normal)), scl(u, ld3(vb->normal))),
                         scl(v, ld3(vc->normal))));
    int front = dot(d, n) < 0.0f;
    if (cull && !front) return 0;
    out->t = t;
    out->p = add(o, scl(t, d));
    out->n = front ? n : neg(n);
    out->front = front;
    out->mat = clean;
    out->u = w * va->uv[0] + u * vb->uv[0] + v * vc->uv[0];
    out->v = w * va->uv[1] + u * vb->uv[1] + v * vc->uv[1];
    out->id = s->tri_orig ? s->tri_orig[ti] : ti;
    out->prim = ti;
    return 1;
}

static inline v3 tri_pos(const OracleScene *s, uint32_t ti, int k) {
    return ld3(s->verts[s->tris[ti].v[k]].position);
}

/* ------------------------------------------------------------------------ */
/* Reference-order traversal, path_trace.wgsl:343-465                         */
/* ------------------------------------------------------------------------ */
static uint32_t next_leaf(const GpuBvhNode *nodes, v3 o, v3 d, float t_min,
                          float t_max, uint32_t *stack, uint32_t *top,
                          uint64_t *node_count) {
    while (*top > 0u) {
        *top -= 1u;
        uint32_t ni = stack[*top];
        const GpuBvhNode *nd = &nodes[ni];
        if (node_count) ++*node_count;
        float te, tx;
        slab(o, d, nd->aabb_min, nd->aabb_max, &te, &tx);
        if (!(tx > fmaxv(te, t_min) && te < t_max)) continue;
        if (nd->prim_count > 0u) return ni;
        if (*top + 2u <= STACK_SIZE) {
            stack[*top] = nd->right_or_offset; *top += 1u;
            stack[*top] = ni + 1u; *top += 1u;
        }
    }
    return 0xFFFFFFFFu;
}

static Hit sphere_bvh_hit_ref(const OracleScene *s, v3 o, v3 d, float t_min,
                              float t_max_in, OracleCounts *c) {
    Hit best;
    best.t = -1.0f;
    if (s->n_sphere_nodes == 0u) return best;
    float t_max = t_max_in;
    uint32_t stack[STACK_SIZE];
    uint32_t top = 0;
    stack[top++] = 0u;
    for (;;) {
        uint32_t li = next_leaf(s->sphere_bvh, o, d, t_min, t_max, stack, &top,
                                c ? &c->nodes : NULL);
        if (li == 0xFFFFFFFFu) break;
        const GpuBvhNode *nd = &s->sphere_bvh[li];
        for (uint32_t j = 0; j < nd->prim_count; ++j) {
            uint32_t si = nd->right_or_offset + j;
            if (c) ++c->spheres;
            Hit h = sphere_hit(o, d, &s->spheres[si], t_min, t_max);
            if (h.t > 0.0f) {
                h.id = MCRT_ID_SPHERE_BIT | (s->sphere_orig ? s->sphere_orig[si] : si);
                h.prim = 0x80000000u | si;
                best = h;
                t_max = h.t;
            }
        }
    }
    return best;
}

static Hit mesh_bvh_hit_ref(const OracleScene *s, v3 o, v3 d, float t_min,
                            float t_max_in, OracleCounts *c) {
    Hit best;
    best.t = -1.0f;
    if (s->n_mesh_nodes == 0u) return best;
    float t_max = t_max_in;
    uint32_t stack[STACK_SIZE];
    uint32_t top = 0;
    stack[top++] = 0u;
    for (;;) {
        uint32_t li = next_leaf(s->mesh_bvh, o, d, t_min, t_max, stack, &top,
                                c ? &c->nodes : NULL);
        if (li == 0xFFFFFFFFu) break;
        const GpuBvhNode *nd = &s->mesh_bvh[li];
        for (uint32_t j = 0; j < nd->prim_count; ++j) {
            uint32_t ti = nd->right_or_offset + j;
            float res[4];
            if (c) ++c->tris;
            mt(o, d, tri_pos(s, ti, 0), tri_pos(s, ti, 1), tri_pos(s, ti, 2), t_min,
               t_max, res);
            if (res[0] > 0.0f) {
                Hit h;
                if (!tri_candidate(s, ti, o, d, res, &h)) continue;
                best = h;
                t_max = h.t;
            }
        }
    }
    return best;
}

/* world_hit, path_trace.wgsl:471-479. */
static Hit world_hit_ref(const OracleScene *s, v3 o, v3 d, float t_min, float t_max,
                         OracleCounts *c) {
    Hit sh = sphere_bvh_hit_ref(s, o, d, t_min, t_max, c);
    float mesh_tmax = sh.t > 0.0f ? sh.t : t_max;
    Hit mh = mesh_bvh_hit_ref(s, o, d, t_min, mesh_tmax, c);
    if (mh.t > 0.0f) return mh;
    return sh;
}

uint32_t oracle_closest_ref(const OracleScene *s, const float o[3], const float d[3],
                            float t_min, float t_max, float *t_out, OracleCounts *c) {
    Hit h = world_hit_ref(s, ld3(o), ld3(d), t_min, t_max, c);
    if (h.t > 0.0f) {
        *t_out = h.t;
        return h.id;
    }
    *t_out = -1.0f;
    return MCRT_ID_MISS;
}

/* ------------------------------------------------------------------------ */
/* Order-free closest hit (the tie rule of mcrt_trace_closest).               */
/* Node visit rule: the reference slab test against the query interval,       */
/* plus a conservative cull `tenter > best_t * CULL_SLACK` that never removes */
/* a node whose entry point precedes the current best hit.                    */
/* ------------------------------------------------------------------------ */
#define CULL_SLACK 1.0009765625f /* 1 + 2^-10 */

static inline int better(float t, uint32_t id, float bt, uint32_t bid) {
    return t < bt || (t == bt && id < bid);
}

static Hit sphere_exact(const OracleScene *s, v3 o, v3 d, float t_min, float t_max) {
    Hit best;
    best.t = -1.0f;
    best.id = MCRT_ID_MISS;
    if (s->n_sphere_nodes == 0u) return best;
    float bt = t_max;
    uint32_t stack[STACK_SIZE], top = 0;
    stack[top++] = 0u;
    while (top > 0u) {
        uint32_t ni = stack[--top];
        const GpuBvhNode *nd = &s->sphere_bvh[ni];
        float te, tx;
        slab(o, d, nd->aabb_min, nd->aabb_max, &te, &tx);
        if (!(tx > fmaxv(te, t_min) && te < t_max)) continue;
        if (te > bt * CULL_SLACK) continue;
        if (nd->prim_count > 0u) {
            for (uint32_t j = 0; j < nd->prim_count; ++j) {
                uint32_t si = nd->right_or_offset + j;
                Hit h = sphere_hit(o, d, &s->spheres[si], t_min, t_max);
                if (h.t > 0.0f) {
                    h.id = MCRT_ID_SPHERE_BIT | (s->sphere_orig ? s->sphere_orig[si] : si);
                    h.prim = 0x80000000u | si;
                    if (best.t < 0.0f || better(h.t, h.id, best.t, best.id)) {
                        best = h;
                        bt = h.t;
                    }
                }
            }
        } else if (top + 2u <= STACK_SIZE) {
            stack[top++] = nd->right_or_offset;
            stack[top++] = ni + 1u;
        }
    }
    return best;
}

static Hit mesh_exact(const OracleScene *s, v3 o, v3 d, float t_min, float t_max) {
    Hit best;
    best.t = -1.0f;
    best.id = MCRT_ID_MISS;
    if (s->n_mesh_nodes == 0u) return best;
    float bt = t_max;
    uint32_t stack[STACK_SIZE], top = 0;
    stack[top++] = 0u;
    while (top > 0u) {
        uint32_t ni = stack[--top];
        const GpuBvhNode *nd = &s->mesh_bvh[ni];
        float te, tx;
        slab(o, d, nd->aabb_min, nd->aabb_max, &te, &tx);
        if (!(tx > fmaxv(te, t_min) && te < t_max)) continue;
        if (te > bt * CULL_SLACK) continue;
        if (nd->prim_count > 0u) {
            for (uint32_t j = 0; j < nd->prim_count; ++j) {
                uint32_t ti = nd->right_or_offset + j;
                float res[4];
                mt(o, d, tri_pos(s, ti, 0), tri_pos(s, ti, 1), tri_pos(s, ti, 2), t_min,
                   t_max, res);
                if (res[0] > 0.0f) {
                    Hit h;
                    if (!tri_candidate(s, ti, o, d, res, &h)) continue;
                    if (best.t < 0.0f || better(h.t, h.id, best.t, best.id)) {
                        best = h;
                        bt = h.t;
                    }
                }
            }
        } else if (top + 2u <= STACK_SIZE) {
            stack[top++] = nd->right_or_offset;
            stack[top++] = ni + 1u;
        }
    }
    return best;
}

static Hit world_exact(const OracleScene *s, v3 o, v3 d, float t_min, float t_max) {
    Hit sh = sphere_exact(s, o, d, t_min, t_max);
    float mesh_tmax = sh.t > 0.0f ? sh.t : t_max;
    Hit mh = mesh_exact(s, o, d, t_min, mesh_tmax);
    if (mh.t > 0.0f) return mh;
    return sh;
}

uint32_t oracle_closest_exact(const OracleScene *s, const float o[3], const float d[3],
                              float t_min, float t_max, float *t_out) {
    Hit h = world_exact(s, ld3(o), ld3(d), t_min, t_max);
    *t_out = h.t > 0.0f ? h.t : -1.0f;
    return h.t > 0.0f ? h.id : MCRT_ID_MISS;
}

uint32_t oracle_closest_brute(const OracleScene *s, const float o3[3], const float d3[3],
                              float t_min, float t_max, float *t_out) {
    v3 o = ld3(o3), d = ld3(d3);
    Hit sb;
    sb.t = -1.0f;
    sb.id = MCRT_ID_MISS;
    /* Spheres: every non-stub sphere (the stub has radius 0 far away). */
    for (uint32_t si = 0; si < s->n_spheres; ++si) {
        Hit h = sphere_hit(o, d, &s->spheres[si], t_min, t_max);
        if (h.t > 0.0f) {
            h.id = MCRT_ID_SPHERE_BIT | (s->sphere_orig ? s->sphere_orig[si] : si);
            if (sb.t < 0.0f || better(h.t, h.id, sb.t, sb.id)) sb = h;
        }
    }
    float mesh_tmax = sb.t > 0.0f ? sb.t : t_max;
    Hit mb;
    mb.t = -1.0f;
    mb.id = MCRT_ID_MISS;
    for (uint32_t ti = 0; ti < s->n_tris; ++ti) {
        float res[4];
        mt(o, d, tri_pos(s, ti, 0), tri_pos(s, ti, 1), tri_pos(s, ti, 2), t_min,
           mesh_tmax, res);
        if (res[0] > 0.0f) {
            Hit h;
            if (!tri_candidate(s, ti, o, d, res, &h)) continue;
            if (mb.t < 0.0f || better(h.t, h.id, mb.t, mb.id)) mb = h;
        }
    }
    Hit w = mb.t > 0.0f ? mb : sb;
    *t_out = w.t > 0.0f ? w.t : -1.0f;
    return w.t > 0.0f ? w.id : MCRT_ID_MISS;
}

/* ------------------------------------------------------------------------ */
/* Scatter, path_trace.wgsl:485-573                                           */
/* ------------------------------------------------------------------------ */
static inline v3 reflect_vec(v3 v, v3 n) { return sub(v, scl(2.0f * dot(v, n), n)); }

static v3 refract_vec(v3 v, v3 n, float eta) {
    float cos_theta = fminf(dot(neg(v), n), 1.0f);
    v3 r_perp = scl(eta, add(v, scl(cos_theta, n)));
    v3 r_para = scl(-sqrtf(fmaxf(0.0f, 1.0f - dot(r_perp, r_perp))), n);
    return add(r_perp, r_para);
}

static float schlick(float cos_theta, float eta) {
    float r0 = (1.0f - eta) / (1.0f + eta);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * powf(1.0f - cos_theta, 5.0f);
}

typedef struct { v3 dir, att; int absorbed; } Scatter;

static Scatter scatter(const GpuMaterial *m, v3 rd, const Hit *h, uint32_t *rng) {
    Scatter r;
    r.absorbed = 0;
    /* material_albedo (:506-515): the oracle carries no texture array, which
     * is the shader's empty-array branch: solid colour. */
    v3 albedo = ld3(m->albedo_fuzz);
    switch (m->type_pad[0]) {
    case 1u: { /* scatter_metal :530-541 */
        float fuzz = m->albedo_fuzz[3];
        v3 refl = reflect_vec(normalize(rd), h->n);
        v3 pert = add(refl, scl(fuzz, random_unit_sphere(rng)));
        r.dir = normalize(dot(pert, pert) >= 1e-6f ? pert : refl);
        r.att = albedo;
        r.absorbed = dot(r.dir, h->n) <= 0.0f;
        break;
    }
    case 2u: { /* scatter_dielectric :543-564 */
        r.att = mk(1.0f, 1.0f, 1.0f);
        float ior = m->ior_pad[0];
        float eta = h->front ? 1.0f / ior : ior;
        float cos_theta = fminf(dot(neg(rd), h->n), 1.0f);
        float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
        int cannot = eta * sin_theta > 1.0f;
        float sc = (!h->front && !cannot)
                       ? sqrtf(fmaxf(0.0f, 1.0f - eta * eta * (1.0f - cos_theta * cos_theta)))
                       : cos_theta;
        float refl = schlick(sc, eta);
        if (cannot || refl > oracle_rand_f32(rng))
            r.dir = reflect_vec(rd, h->n);
        else
            r.dir = normalize(refract_vec(rd, h->n, eta));
        break;
    }
    default: /* scatter_lambertian :521-528 */
        r.dir = cosine_scatter(rng, h->n);
        r.att = albedo;
        break;
    }
    return r;
}

/* Test hook: one scatter event for a given incident direction and normal. */
void oracle_scatter(const GpuMaterial *m, const float rd[3], const float n[3], int front,
                    uint32_t *rng, float out_dir[3], float out_att[3], int *absorbed) {
    Hit h;
    memset(&h, 0, sizeof h);
    h.n = ld3(n);
    h.front = front;
    Scatter s = scatter(m, ld3(rd), &h, rng);
    out_dir[0] = s.dir.x; out_dir[1] = s.dir.y; out_dir[2] = s.dir.z;
    out_att[0] = s.att.x; out_att[1] = s.att.y; out_att[2] = s.att.z;
    *absorbed = s.absorbed;
}

/* ------------------------------------------------------------------------ */
/* path_color + cs_main, path_trace.wgsl:583-662                              */
/* ------------------------------------------------------------------------ */
static v3 path_color(const OracleScene *s, v3 o, v3 d, uint32_t *rng, uint32_t max_depth,
                     float gbuf[4], OracleCounts *c, uint64_t *segments) {
    v3 thr = mk(1.0f, 1.0f, 1.0f);
    int first = 1;
    for (uint32_t b = 0; b < max_depth; ++b) {
        if (segments) ++*segments;
        Hit h = world_hit_ref(s, o, d, MCRT_T_MIN, MCRT_T_MAX, c);
        if (h.t < 0.0f) {
            if (first) { gbuf[0] = 0.0f; gbuf[1] = 0.0f; gbuf[2] = 0.0f; gbuf[3] = MCRT_T_MAX; }
            return mul(thr, sample_env(s, d));
        }
        if (first) {
            gbuf[0] = h.n.x; gbuf[1] = h.n.y; gbuf[2] = h.n.z; gbuf[3] = h.t;
            first = 0;
        }
        uint32_t mi = h.mat < MAX_MAT_IDX ? h.mat : MAX_MAT_IDX;
        const GpuMaterial *m = &s->mats->materials[mi];
        if (m->type_pad[0] == MCRT_MAT_EMISSIVE)
            return scl(m->ior_pad[0], mul(thr, ld3(m->albedo_fuzz)));
        Scatter sr = scatter(m, d, &h, rng);
        if (sr.absorbed) return mk(0.0f, 0.0f, 0.0f);
        thr = mul(thr, sr.att);
        o = h.p;
        d = sr.dir;
    }
    return mk(0.0f, 0.0f, 0.0f);
}

/* ------------------------------------------------------------------------ */
/* Next-event estimation: the build's opt-in estimator (no reference        */
/* counterpart; the shader reserves binding 12, src/gpu_layout.rs:51-52,    */
/* for it).  Emitter chosen with probability proportional to area x         */
/* luminance; triangle points uniform in area; sphere directions uniform in */
/* the subtended cone; power heuristic against the cosine BSDF sampling;    */
/* light draws from their own PCG stream hashed from the path state after   */
/* the scatter.  An independent CPU statement of that estimator, in the     */
/* same operation order as the kernels.                                     */
/* ------------------------------------------------------------------------ */
#define INV_PI 0.31830988618379067f
#define LIGHT_STREAM 0x68E31DA4u
#define EMIT_SPHERE 0x80000000u

typedef struct {
    uint32_t n;
    float *cdf, *prob, *area;
    uint32_t *ref;
    v3 *le;
    uint32_t *tri_emit, *sph_emit;
} Emitters;

static float lum3(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

static int emission_of(const OracleScene *s, uint32_t mat, v3 *le) {
    const GpuMaterial *g = &s->mats->materials[mat < MAX_MAT_IDX ? mat : MAX_MAT_IDX];
    if (g->type_pad[0] != MCRT_MAT_EMISSIVE) return 0;
    *le = mk(g->ior_pad[0] * g->albedo_fuzz[0], g->ior_pad[0] * g->albedo_fuzz[1],
             g->ior_pad[0] * g->albedo_fuzz[2]);
    return 1;
}

static void emitters_free(Emitters *E) {
    free(E->cdf); free(E->prob); free(E->area); free(E->ref); free(E->le);
    free(E->tri_emit); free(E->sph_emit);
    memset(E, 0, sizeof *E);
}

static void emitters_build(const OracleScene *s, Emitters *E) {
    memset(E, 0, sizeof *E);
    uint32_t cap = s->n_tris + s->n_spheres;
    float *w = (float *)malloc(sizeof(float) * (cap ? cap : 1));
    E->ref = (uint32_t *)malloc(sizeof(uint32_t) * (cap ? cap : 1));
    E->area = (float *)malloc(sizeof(float) * (cap ? cap : 1));
    E->le = (v3 *)malloc(sizeof(v3) * (cap ? cap : 1));
    uint32_t n = 0;
    for (uint32_t i = 0; i < s->n_tris; ++i) {
        v3 le;
        if (!emission_of(s, s->tris[i].mat_idx & 0x7FFFFFFFu, &le)) continue;
        v3 p0 = tri_pos(s, i, 0);
        v3 cr = cross(sub(tri_pos(s, i, 1), p0), sub(tri_pos(s, i, 2), p0));
        float area = 0.5f * sqrtf(dot(cr, cr));
        float wt = area * lum3(le);
        if (!(wt > 0.0f) || !isfinite(wt)) continue;
        E->ref[n] = i; E->area[n] = area; E->le[n] = le; w[n] = wt; ++n;
    }
    for (uint32_t i = 0; i < s->n_spheres; ++i) {
        v3 le;
        if (!emission_of(s, s->spheres[i].mat_and_pad[0], &le)) continue;
        float r = s->spheres[i].center_r[3];
        float area = 4.0f * 3.14159265358979f * r * r;
        float wt = area * lum3(le);
        if (!(wt > 0.0f) || !isfinite(wt)) continue;
        E->ref[n] = EMIT_SPHERE | i; E->area[n] = area; E->le[n] = le; w[n] = wt; ++n;
    }
    E->n = n;
    E->cdf = (float *)malloc(sizeof(float) * (n ? n : 1));
    E->prob = (float *)malloc(sizeof(float) * (n ? n : 1));
    E->tri_emit = (uint32_t *)malloc(sizeof(uint32_t) * (s->n_tris ? s->n_tris : 1));
    E->sph_emit = (uint32_t *)malloc(sizeof(uint32_t) * (s->n_spheres ? s->n_spheres : 1));
    for (uint32_t i = 0; i < s->n_tris; ++i) E->tri_emit[i] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < s->n_spheres; ++i) E->sph_emit[i] = 0xFFFFFFFFu;
    double total = 0.0, run = 0.0;
    for (uint32_t k = 0; k < n; ++k) total += w[k];
    for (uint32_t k = 0; k < n; ++k) {
        E->prob[k] = (float)(w[k] / total);
        run += w[k];
        E->cdf[k] = k + 1 == n ? 1.0f : (float)(run / total);
        if (E->ref[k] & EMIT_SPHERE) E->sph_emit[E->ref[k] & ~EMIT_SPHERE] = k;
        else E->tri_emit[E->ref[k]] = k;
    }
    free(w);
}

static inline float power_heuristic(float a, float b) {
    float a2 = a * a;
    return a2 / (a2 + b * b);
}

static inline float cone_one_minus_cos(float r2, float d2) {
    float s2 = r2 / d2;
    return s2 / (1.0f + sqrtf(1.0f - s2));
}

/* Orthonormal basis (Duff et al., JCGT 6(1), 2017). */
static void onb(v3 n, v3 *t1, v3 *t2) {
    float sg = copysignf(1.0f, n.z);
    float a = -1.0f / (sg + n.z);
    float b = n.x * n.y * a;
    *t1 = mk(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x);
    *t2 = mk(b, sg + n.y * n.y * a, -n.y);
}

static float nee_light_pdf(const OracleScene *s, const Emitters *E, uint32_t e, v3 x, v3 d, float t) {
    uint32_t ref = E->ref[e];
    if (ref & EMIT_SPHERE) {
        const GpuSphere *sp = &s->spheres[ref & ~EMIT_SPHERE];
        v3 w = sub(ld3(sp->center_r), x);
        float d2 = dot(w, w), r2 = sp->center_r[3] * sp->center_r[3];
        if (!(d2 > r2)) return 0.0f;
        return E->prob[e] / (TAU * cone_one_minus_cos(r2, d2));
    }
    v3 p0 = tri_pos(s, ref, 0);
    v3 ng = cross(sub(tri_pos(s, ref, 1), p0), sub(tri_pos(s, ref, 2), p0));
    float cosl = fabsf(dot(ng, d)) / sqrtf(dot(ng, ng));
    if (!(cosl > 1e-7f)) return 0.0f;
    return E->prob[e] * (t * t) / (E->area[e] * cosl);
}

static int nee_sample(const OracleScene *s, const Emitters *E, v3 x, uint32_t *lrng, v3 *dir,
                      float *dist, v3 *le, float *pdf) {
    float u0 = oracle_rand_f32(lrng), u1 = oracle_rand_f32(lrng), u2 = oracle_rand_f32(lrng);
    uint32_t lo = 0, hi = E->n - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (u0 < E->cdf[mid]) hi = mid; else lo = mid + 1u;
    }
    uint32_t e = lo, ref = E->ref[e];
    *le = E->le[e];
    if (ref & EMIT_SPHERE) {
        const GpuSphere *sp = &s->spheres[ref & ~EMIT_SPHERE];
        v3 w = sub(ld3(sp->center_r), x);
        float d2 = dot(w, w), r2 = sp->center_r[3] * sp->center_r[3];
        if (!(d2 > r2)) return 0;
        float omc = cone_one_minus_cos(r2, d2);
        float ct = 1.0f - u1 * omc;
        float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float phi = TAU * u2;
        v3 wn = scl(1.0f / sqrtf(d2), w);
        v3 t1, t2;
        onb(wn, &t1, &t2);
        *dir = normalize(add(add(scl(st * cosf(phi), t1), scl(st * sinf(phi), t2)), scl(ct, wn)));
        float b = dot(*dir, w);
        *dist = b - sqrtf(fmaxf(r2 - (d2 - b * b), 0.0f));
        *pdf = E->prob[e] / (TAU * omc);
        return *dist > 0.0f && *pdf > 0.0f && *pdf < 3.0e38f;
    }
    v3 v0 = tri_pos(s, ref, 0);
    v3 e1 = sub(tri_pos(s, ref, 1), v0), e2 = sub(tri_pos(s, ref, 2), v0);
    float su = sqrtf(u1);
    float bu = su * (1.0f - u2), bv = su * u2;
    v3 y = add(add(v0, scl(bu, e1)), scl(bv, e2));
    v3 w = sub(y, x);
    float d2 = dot(w, w);
    if (!(d2 > 0.0f)) return 0;
    *dist = sqrtf(d2);
    *dir = scl(1.0f / *dist, w);
    if (s->tris[ref].mat_idx >> 31u) {
        /* a back-face-culled emitter is invisible from behind (:438-452) */
        const GpuTriangle *tri = &s->tris[ref];
        float bw = 1.0f - bu - bv;
        v3 n = normalize(add(add(scl(bw, ld3(s->verts[tri->v[0]].normal)), scl(bu, ld3(s->verts[tri->v[1]].normal))),
                             scl(bv, ld3(s->verts[tri->v[2]].normal))));
        if (!(dot(*dir, n) < 0.0f)) return 0;
    }
    v3 ng = cross(e1, e2);
    float cosl = fabsf(dot(ng, *dir)) / sqrtf(dot(ng, ng));
    if (!(cosl > 1e-7f)) return 0;
    *pdf = E->prob[e] * d2 / (E->area[e] * cosl);
    return *pdf < 3.0e38f;
}

static v3 path_color_nee(const OracleScene *s, const Emitters *E, v3 o, v3 d, uint32_t *rng,
                         uint32_t max_depth, float gbuf[4]) {
    v3 thr = mk(1.0f, 1.0f, 1.0f), L = mk(0.0f, 0.0f, 0.0f);
    float pdf_prev = 0.0f; /* BSDF pdf of the arriving ray; 0 after the camera or a specular event */
    int first = 1;
    for (uint32_t b = 0; b < max_depth; ++b) {
        Hit h = world_exact(s, o, d, MCRT_T_MIN, MCRT_T_MAX);
        if (h.t < 0.0f) {
            if (first) { gbuf[0] = 0.0f; gbuf[1] = 0.0f; gbuf[2] = 0.0f; gbuf[3] = MCRT_T_MAX; }
            return add(L, mul(thr, sample_env(s, d)));
        }
        if (first) {
            gbuf[0] = h.n.x; gbuf[1] = h.n.y; gbuf[2] = h.n.z; gbuf[3] = h.t;
            first = 0;
        }
        uint32_t mi = h.mat < MAX_MAT_IDX ? h.mat : MAX_MAT_IDX;
        const GpuMaterial *m = &s->mats->materials[mi];
        if (m->type_pad[0] == MCRT_MAT_EMISSIVE) {
            v3 rad = scl(m->ior_pad[0], mul(thr, ld3(m->albedo_fuzz)));
            if (pdf_prev > 0.0f && E->n) {
                uint32_t e = (h.prim & EMIT_SPHERE) ? E->sph_emit[h.prim & ~EMIT_SPHERE] : E->tri_emit[h.prim];
                if (e != 0xFFFFFFFFu) rad = scl(power_heuristic(pdf_prev, nee_light_pdf(s, E, e, o, d, h.t)), rad);
            }
            return add(L, rad);
        }
        Scatter sr = scatter(m, d, &h, rng);
        if (sr.absorbed) return L;
        thr = mul(thr, sr.att);
        if (b + 1u >= max_depth) return L;
        pdf_prev = 0.0f;
        if (m->type_pad[0] != MCRT_MAT_METAL && m->type_pad[0] != MCRT_MAT_DIELECTRIC) {
            pdf_prev = fmaxf(dot(h.n, sr.dir), 0.0f) * INV_PI;
            uint32_t lrng = oracle_pcg_hash(*rng ^ LIGHT_STREAM);
            v3 ldir, le;
            float dist, pdf;
            if (E->n && nee_sample(s, E, h.p, &lrng, &ldir, &dist, &le, &pdf)) {
                float cosx = dot(h.n, ldir);
                if (cosx > 0.0f) {
                    float wgt = power_heuristic(pdf, cosx * INV_PI) * cosx * INV_PI / pdf;
                    v3 c = scl(wgt, mul(thr, le));
                    if (world_exact(s, h.p, ldir, MCRT_T_MIN, dist * (1.0f - 1e-3f)).t < 0.0f) L = add(L, c);
                }
            }
        }
        o = h.p;
        d = sr.dir;
    }
    return L;
}

static v3 camera_sample_ex(const OracleScene *s, const Emitters *E, const CameraUniform *cam,
                           uint32_t W, uint32_t H, uint32_t x, uint32_t y, uint32_t frame,
                           uint32_t max_depth, float gbuf[4], OracleCounts *c, uint64_t *segments);

static v3 camera_sample(const OracleScene *s, const CameraUniform *cam, uint32_t W,
                        uint32_t H, uint32_t x, uint32_t y, uint32_t frame,
                        uint32_t max_depth, float gbuf[4], OracleCounts *c,
                        uint64_t *segments) {
    return camera_sample_ex(s, NULL, cam, W, H, x, y, frame, max_depth, gbuf, c, segments);
}

static v3 camera_sample_ex(const OracleScene *s, const Emitters *E, const CameraUniform *cam,
                           uint32_t W, uint32_t H, uint32_t x, uint32_t y, uint32_t frame,
                           uint32_t max_depth, float gbuf[4], OracleCounts *c, uint64_t *segments) {
    uint32_t rng = oracle_pixel_seed(x, y, W, frame);
    float u = ((float)x + oracle_rand_f32(&rng)) / (float)W;
    float v = 1.0f - ((float)y + oracle_rand_f32(&rng)) / (float)H;
    v3 focal = add(add(ld3(cam->lower_left), scl(u, ld3(cam->horizontal))),
                   scl(v, ld3(cam->vertical)));
    float lx, ly;
    random_in_unit_disk(&rng, &lx, &ly);
    v3 lens = add(scl(lx, ld3(cam->defocus_u)), scl(ly, ld3(cam->defocus_v)));
    v3 ro = add(ld3(cam->origin), lens);
    v3 rd = normalize(sub(focal, ro));
    /* `mix(prev, s, w)` of the caller follows; the pixel's sample only here. */
    if (E) return path_color_nee(s, E, ro, rd, &rng, max_depth, gbuf);
    return path_color(s, ro, rd, &rng, max_depth, gbuf, c, segments);
}

void oracle_sample(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                   uint32_t x, uint32_t y, uint32_t frame, uint32_t max_depth,
                   float out_rgb[3], float out_gbuf[4]) {
    float g[4] = {0, 0, 0, 0};
    v3 c = camera_sample(s, cam, W, H, x, y, frame, max_depth, g, NULL, NULL);
    out_rgb[0] = c.x; out_rgb[1] = c.y; out_rgb[2] = c.z;
    if (out_gbuf) memcpy(out_gbuf, g, sizeof g);
}

/* GLSL / SPIR-V FMix: x*(1-a) + y*a. */
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

typedef struct {
    const OracleScene *s;
    const CameraUniform *cam;
    uint32_t W, H, spp, first, depth;
    const float *accum_in;
    float *accum, *gbuf;
    int tid, nthreads;
    OracleCounts counts;
    uint64_t segments;
    int want_counts;
    const Emitters *E; /* next-event estimation, or NULL */
} RenderJob;

static void *render_worker(void *arg) {
    RenderJob *j = (RenderJob *)arg;
    memset(&j->counts, 0, sizeof j->counts);
    j->segments = 0;
    for (uint32_t y = (uint32_t)j->tid; y < j->H; y += (uint32_t)j->nthreads) {
        for (uint32_t x = 0; x < j->W; ++x) {
            uint64_t p = (uint64_t)y * j->W + x;
            float acc[3] = {0.0f, 0.0f, 0.0f};
            if (j->accum_in) {
                acc[0] = j->accum_in[4 * p];
                acc[1] = j->accum_in[4 * p + 1];
                acc[2] = j->accum_in[4 * p + 2];
            }
            float g[4] = {0, 0, 0, 0};
            for (uint32_t i = 0; i < j->spp; ++i) {
                uint32_t f = j->first + i;
                v3 smp = camera_sample_ex(j->s, j->E, j->cam, j->W, j->H, x, y, f, j->depth, g,
                                          j->want_counts ? &j->counts : NULL,
                                          j->want_counts ? &j->segments : NULL);
                uint32_t fc = f < MCRT_MAX_FRAME_COUNT ? f : MCRT_MAX_FRAME_COUNT;
                float w = 1.0f / (float)(fc + 1u);
                acc[0] = mixf(acc[0], smp.x, w);
                acc[1] = mixf(acc[1], smp.y, w);
                acc[2] = mixf(acc[2], smp.z, w);
            }
            j->accum[4 * p] = acc[0];
            j->accum[4 * p + 1] = acc[1];
            j->accum[4 * p + 2] = acc[2];
            j->accum[4 * p + 3] = 1.0f;
            if (j->gbuf) memcpy(&j->gbuf[4 * p], g, sizeof g);
        }
    }
    return NULL;
}

static void render_run(const OracleScene *s, const Emitters *E, const CameraUniform *cam, uint32_t W,
                       uint32_t H, uint32_t spp, uint32_t first_frame, uint32_t max_depth,
                       const float *accum_in, float *out_accum, float *out_gbuf,
                       OracleCounts *counts, uint64_t *segments, int n_threads);

void oracle_render(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                   uint32_t spp, uint32_t first_frame, uint32_t max_depth,
                   const float *accum_in, float *out_accum, float *out_gbuf,
                   OracleCounts *counts, uint64_t *segments, int n_threads) {
    render_run(s, NULL, cam, W, H, spp, first_frame, max_depth, accum_in, out_accum, out_gbuf, counts,
               segments, n_threads);
}

void oracle_render_nee(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                       uint32_t spp, uint32_t first_frame, uint32_t max_depth,
                       const float *accum_in, float *out_accum, float *out_gbuf, int n_threads) {
    Emitters E;
    emitters_build(s, &E);
    render_run(s, &E, cam, W, H, spp, first_frame, max_depth, accum_in, out_accum, out_gbuf, NULL, NULL,
               n_threads);
    emitters_free(&E);
}

static void render_run(const OracleScene *s, const Emitters *E, const CameraUniform *cam, uint32_t W,
                       uint32_t H, uint32_t spp, uint32_t first_frame, uint32_t max_depth,
                       const float *accum_in, float *out_accum, float *out_gbuf,
                       OracleCounts *counts, uint64_t *segments, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    RenderJob *jobs = (RenderJob *)calloc((size_t)n_threads, sizeof(RenderJob));
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int t = 0; t < n_threads; ++t) {
        RenderJob *j = &jobs[t];
        j->s = s; j->cam = cam; j->W = W; j->H = H; j->spp = spp;
        j->first = first_frame; j->depth = max_depth;
        j->accum_in = accum_in; j->accum = out_accum; j->gbuf = out_gbuf;
        j->tid = t; j->nthreads = n_threads;
        j->want_counts = counts != NULL || segments != NULL;
        j->E = E;
        pthread_create(&th[t], NULL, render_worker, j);
    }
    if (counts) memset(counts, 0, sizeof *counts);
    if (segments) *segments = 0;
    for (int t = 0; t < n_threads; ++t) {
        pthread_join(th[t], NULL);
        if (counts) {
            counts->nodes += jobs[t].counts.nodes;
            counts->tris += jobs[t].counts.tris;
            counts->spheres += jobs[t].counts.spheres;
        }
        if (segments) *segments += jobs[t].segments;
    }
    free(jobs);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* Differentiable-render reference (extended estimator, see the header).      */
/* Path geometry in FP32 exactly as oracle_render; radiance and gradients in  */
/* FP64.  The gradient is the power rule per radiance term: t = T_k * Le_k    */
/* with T_k = a_m^{n_m} * O_m, where n_m counts the earlier vertices on       */
/* parameter material m and O_m multiplies every other earlier attenuation,   */
/* so dt/da_m[c] = n_m * a_m[c]^{n_m - 1} * O_m[c] * Le[c] (no division, so   */
/* a zero albedo is exact); dL/dE_p sums T_k over vertices on p.  This is an  */
/* independent derivation from the GPU kernels' product-rule accumulation.    */
/* ------------------------------------------------------------------------ */
#define ORACLE_MAX_PARAMS 8

typedef struct {
    const OracleScene *s;
    const CameraUniform *cam;
    uint32_t W, H, spp, first, depth, n_params;
    uint32_t cx0, cy0, cw, ch; /* pixel window (the whole image by default) */
    const double *params;
    const int32_t *pmap;
    const double *d_img; /* NULL: primal only */
    double *img;
    int tid, nthreads;
    double grad[ORACLE_MAX_PARAMS * 4];
} DiffJob;

static void diff_term(DiffJob *j, const double T[3], const double Le[3], const int counts[],
                      double other[][3], const double delta[3], double L[3]) {
    double t[3] = {T[0] * Le[0], T[1] * Le[1], T[2] * Le[2]};
    L[0] += t[0]; L[1] += t[1]; L[2] += t[2];
    if (!j->d_img) return;
    for (uint32_t m = 0; m < j->n_params; ++m) {
        if (!counts[m]) continue;
        for (int c = 0; c < 3; ++c)
            j->grad[4 * m + c] += delta[c] * Le[c] * (double)counts[m] *
                                  pow(j->params[4 * m + c], (double)(counts[m] - 1)) * other[m][c];
    }
}

/* O_m *= att for every parameter m except `skip` (-1: all). */
static void diff_other(double other[][3], uint32_t n_params, int skip, const double att[3]) {
    for (uint32_t m = 0; m < n_params; ++m) {
        if ((int)m == skip) continue;
        other[m][0] *= att[0]; other[m][1] *= att[1]; other[m][2] *= att[2];
    }
}

static void *diff_worker(void *arg) {
    DiffJob *j = (DiffJob *)arg;
    const OracleScene *s = j->s;
    const CameraUniform *cam = j->cam;
    memset(j->grad, 0, sizeof j->grad);
    GpuMaterial lam;
    memset(&lam, 0, sizeof lam);
    lam.albedo_fuzz[0] = lam.albedo_fuzz[1] = lam.albedo_fuzz[2] = 0.5f;
    for (uint32_t y = j->cy0 + (uint32_t)j->tid; y < j->cy0 + j->ch; y += (uint32_t)j->nthreads) {
        for (uint32_t x = j->cx0; x < j->cx0 + j->cw; ++x) {
            uint64_t pix = (uint64_t)(y - j->cy0) * j->cw + (x - j->cx0);
            double delta[3] = {0, 0, 0};
            if (j->d_img)
                for (int c = 0; c < 3; ++c) delta[c] = j->d_img[3 * pix + c] / (double)j->spp;
            double sum[3] = {0, 0, 0};
            for (uint32_t i = 0; i < j->spp; ++i) {
                uint32_t rng = oracle_pixel_seed(x, y, j->W, j->first + i);
                float u = ((float)x + oracle_rand_f32(&rng)) / (float)j->W;
                float v = 1.0f - ((float)y + oracle_rand_f32(&rng)) / (float)j->H;
                v3 focal = add(add(ld3(cam->lower_left), scl(u, ld3(cam->horizontal))),
                               scl(v, ld3(cam->vertical)));
                float lx, ly;
                random_in_unit_disk(&rng, &lx, &ly);
                v3 lens = add(scl(lx, ld3(cam->defocus_u)), scl(ly, ld3(cam->defocus_v)));
                v3 o = add(ld3(cam->origin), lens);
                v3 d = normalize(sub(focal, o));
                double T[3] = {1, 1, 1}, L[3] = {0, 0, 0};
                int counts[ORACLE_MAX_PARAMS] = {0};
                double other[ORACLE_MAX_PARAMS][3];
                for (int m = 0; m < ORACLE_MAX_PARAMS; ++m) other[m][0] = other[m][1] = other[m][2] = 1.0;
                for (uint32_t b = 0; b < j->depth; ++b) {
                    Hit h = world_hit_ref(s, o, d, MCRT_T_MIN, MCRT_T_MAX, NULL);
                    if (h.t < 0.0f) {
                        v3 e = sample_env(s, d);
                        double Le[3] = {e.x, e.y, e.z};
                        diff_term(j, T, Le, counts, other, delta, L);
                        break;
                    }
                    uint32_t mi = h.mat < MAX_MAT_IDX ? h.mat : MAX_MAT_IDX;
                    int p = j->pmap[mi];
                    const GpuMaterial *m = &s->mats->materials[mi];
                    double att[3];
                    if (p >= 0) {
                        const double *q = &j->params[4 * p];
                        double Le[3] = {q[3], q[3], q[3]};
                        diff_term(j, T, Le, counts, other, delta, L);
                        if (j->d_img) j->grad[4 * p + 3] += delta[0] * T[0] + delta[1] * T[1] + delta[2] * T[2];
                        Scatter sr = scatter(&lam, d, &h, &rng);
                        att[0] = q[0]; att[1] = q[1]; att[2] = q[2];
                        diff_other(other, j->n_params, p, att);
                        counts[p] += 1;
                        d = sr.dir;
                    } else if (m->type_pad[0] == MCRT_MAT_EMISSIVE) {
                        double Le[3] = {(double)(m->ior_pad[0] * m->albedo_fuzz[0]),
                                        (double)(m->ior_pad[0] * m->albedo_fuzz[1]),
                                        (double)(m->ior_pad[0] * m->albedo_fuzz[2])};
                        diff_term(j, T, Le, counts, other, delta, L);
                        break;
                    } else {
                        Scatter sr = scatter(m, d, &h, &rng);
                        if (sr.absorbed) break;
                        att[0] = sr.att.x; att[1] = sr.att.y; att[2] = sr.att.z;
                        diff_other(other, j->n_params, -1, att);
                        d = sr.dir;
                    }
                    T[0] *= att[0]; T[1] *= att[1]; T[2] *= att[2];
                    o = h.p;
                }
                sum[0] += L[0]; sum[1] += L[1]; sum[2] += L[2];
            }
            if (j->img)
                for (int c = 0; c < 3; ++c) j->img[3 * pix + c] = sum[c] / (double)j->spp;
        }
    }
    return NULL;
}

static void diff_run_window(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                            uint32_t cx0, uint32_t cy0, uint32_t cw, uint32_t ch, uint32_t spp,
                            uint32_t first, uint32_t depth, const double *params,
                            uint32_t n_params, const int32_t *pmap, const double *d_img,
                            double *img, double *grad, int n_threads);

static void diff_run(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                     uint32_t spp, uint32_t first, uint32_t depth, const double *params,
                     uint32_t n_params, const int32_t *pmap, const double *d_img, double *img,
                     double *grad, int n_threads) {
    diff_run_window(s, cam, W, H, 0, 0, W, H, spp, first, depth, params, n_params, pmap, d_img, img,
                    grad, n_threads);
}

/* Converged reference on a pixel window: the reference estimator (no
 * parameter materials) with FP64 accumulation, mean over spp frames starting
 * at `first`.  out: ch x cw x 3 doubles. */
void oracle_reference64_window(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                               uint32_t cx0, uint32_t cy0, uint32_t cw, uint32_t ch, uint32_t spp,
                               uint32_t first, uint32_t depth, double *out, int n_threads) {
    int32_t pmap[MCRT_MAX_MATERIALS];
    for (uint32_t m = 0; m < MCRT_MAX_MATERIALS; ++m) pmap[m] = -1;
    double dummy[4] = {0, 0, 0, 0};
    diff_run_window(s, cam, W, H, cx0, cy0, cw, ch, spp, first, depth, dummy, 0, pmap, NULL, out, NULL,
                    n_threads);
}

static void diff_run_window(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                            uint32_t cx0, uint32_t cy0, uint32_t cw, uint32_t ch, uint32_t spp,
                            uint32_t first, uint32_t depth, const double *params,
                            uint32_t n_params, const int32_t *pmap, const double *d_img,
                            double *img, double *grad, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    if (n_params > ORACLE_MAX_PARAMS) n_params = ORACLE_MAX_PARAMS;
    DiffJob *jobs = (DiffJob *)calloc((size_t)n_threads, sizeof(DiffJob));
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int t = 0; t < n_threads; ++t) {
        DiffJob *j = &jobs[t];
        j->s = s; j->cam = cam; j->W = W; j->H = H; j->spp = spp; j->first = first;
        j->cx0 = cx0; j->cy0 = cy0; j->cw = cw; j->ch = ch;
        j->depth = depth; j->n_params = n_params; j->params = params; j->pmap = pmap;
        j->d_img = d_img; j->img = img; j->tid = t; j->nthreads = n_threads;
        pthread_create(&th[t], NULL, diff_worker, j);
    }
    if (grad) memset(grad, 0, sizeof(double) * 4 * n_params);
    for (int t = 0; t < n_threads; ++t) {
        pthread_join(th[t], NULL);
        if (grad)
            for (uint32_t k = 0; k < 4 * n_params; ++k) grad[k] += jobs[t].grad[k];
    }
    free(jobs);
    free(th);
}

void oracle_diff_render64(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                          uint32_t spp, uint32_t first, uint32_t depth, const double *params,
                          uint32_t n_params, const int32_t *pmap, double *out_img, int n_threads) {
    diff_run(s, cam, W, H, spp, first, depth, params, n_params, pmap, NULL, out_img, NULL, n_threads);
}

void oracle_diff_grad64(const OracleScene *s, const CameraUniform *cam, uint32_t W, uint32_t H,
                        uint32_t spp, uint32_t first, uint32_t depth, const double *params,
                        uint32_t n_params, const int32_t *pmap, const double *d_img,
                        double *grad, int n_threads) {
    diff_run(s, cam, W, H, spp, first, depth, params, n_params, pmap, d_img, NULL, grad, n_threads);
}

/* ------------------------------------------------------------------------ */
/* Batched closest-hit                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    const OracleScene *s;
    const float *o, *d;
    uint64_t n;
    float tmin, tmax;
    int mode, tid, nthreads;
    uint32_t *id;
    float *t;
    OracleCounts counts;
} TraceJob;

static void *trace_worker(void *arg) {
    TraceJob *j = (TraceJob *)arg;
    memset(&j->counts, 0, sizeof j->counts);
    uint64_t lo = j->n * (uint64_t)j->tid / (uint64_t)j->nthreads;
    uint64_t hi = j->n * (uint64_t)(j->tid + 1) / (uint64_t)j->nthreads;
    for (uint64_t i = lo; i < hi; ++i) {
        const float *o = j->o + 3 * i, *d = j->d + 3 * i;
        if (j->mode == 0)
            j->id[i] = oracle_closest_ref(j->s, o, d, j->tmin, j->tmax, &j->t[i], &j->counts);
        else if (j->mode == 1)
            j->id[i] = oracle_closest_exact(j->s, o, d, j->tmin, j->tmax, &j->t[i]);
        else
            j->id[i] = oracle_closest_brute(j->s, o, d, j->tmin, j->tmax, &j->t[i]);
    }
    return NULL;
}

void oracle_trace_batch(const OracleScene *s, const float *origins, const float *dirs,
                        uint64_t n, float t_min, float t_max, int mode, uint32_t *out_id,
                        float *out_t, OracleCounts *counts, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    TraceJob *jobs = (TraceJob *)calloc((size_t)n_threads, sizeof(TraceJob));
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int t = 0; t < n_threads; ++t) {
        TraceJob *j = &jobs[t];
        j->s = s; j->o = origins; j->d = dirs; j->n = n; j->tmin = t_min; j->tmax = t_max;
        j->mode = mode; j->tid = t; j->nthreads = n_threads; j->id = out_id; j->t = out_t;
        pthread_create(&th[t], NULL, trace_worker, j);
    }
    if (counts) memset(counts, 0, sizeof *counts);
    for (int t = 0; t < n_threads; ++t) {
        pthread_join(th[t], NULL);
        if (counts) {
            counts->nodes += jobs[t].counts.nodes;
            counts->tris += jobs[t].counts.tris;
            counts->spheres += jobs[t].counts.spheres;
        }
    }
    free(jobs);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* Bilateral denoiser, denoise.wgsl:52-120                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    const DenoiseParams *p;
    uint32_t W, H;
    const float *accum, *gbuf;
    float *out;
    int tid, nthreads;
} DenoiseJob;

static void *denoise_worker(void *arg) {
    DenoiseJob *j = (DenoiseJob *)arg;
    const DenoiseParams *dp = j->p;
    int W = (int)j->W, H = (int)j->H;
    float inv2ss = 1.0f / (2.0f * dp->sigma_s * dp->sigma_s);
    float inv2ds = 1.0f / (2.0f * dp->sigma_d * dp->sigma_d);
    for (int py = j->tid; py < H; py += j->nthreads) {
        for (int px = 0; px < W; ++px) {
            const float *gp = &j->gbuf[4 * ((size_t)py * W + px)];
            v3 np = ld3(gp);
            float dpth = gp[3];
            int on_sky = dpth < 0.0f;
            float ca[3] = {0.0f, 0.0f, 0.0f};
            float wa = 0.0f;
            for (int dy = -dp->radius; dy <= dp->radius; ++dy) {
                for (int dx = -dp->radius; dx <= dp->radius; ++dx) {
                    int qx = px + dx, qy = py + dy;
                    qx = qx < 0 ? 0 : (qx > W - 1 ? W - 1 : qx);
                    qy = qy < 0 ? 0 : (qy > H - 1 ? H - 1 : qy);
                    const float *gq = &j->gbuf[4 * ((size_t)qy * W + qx)];
                    v3 nq = ld3(gq);
                    float dq = gq[3];
                    int q_sky = dq < 0.0f;
                    if (on_sky != q_sky) continue;
                    float dist_sq = (float)(dx * dx + dy * dy);
                    float w_s = expf(-dist_sq * inv2ss);
                    float w_n = on_sky ? 1.0f : powf(fmaxf(0.0f, dot(np, nq)), dp->normal_pow);
                    float drel = (dpth - dq) / fmaxf(dpth, 1e-4f);
                    float w_d = on_sky ? 1.0f : expf(-(drel * drel) * inv2ds);
                    float w = w_s * w_n * w_d;
                    const float *a = &j->accum[4 * ((size_t)qy * W + qx)];
                    ca[0] = ca[0] + w * a[0];
                    ca[1] = ca[1] + w * a[1];
                    ca[2] = ca[2] + w * a[2];
                    wa = wa + w;
                }
            }
            float *o = &j->out[4 * ((size_t)py * W + px)];
            const float *a = &j->accum[4 * ((size_t)py * W + px)];
            if (wa > 0.0f) {
                o[0] = ca[0] / wa; o[1] = ca[1] / wa; o[2] = ca[2] / wa;
            } else {
                o[0] = a[0]; o[1] = a[1]; o[2] = a[2];
            }
            o[3] = 1.0f;
        }
    }
    return NULL;
}

void oracle_denoise(const DenoiseParams *p, uint32_t W, uint32_t H, const float *accum,
                    const float *gbuf, float *out, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    DenoiseJob *jobs = (DenoiseJob *)calloc((size_t)n_threads, sizeof(DenoiseJob));
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int t = 0; t < n_threads; ++t) {
        DenoiseJob *j = &jobs[t];
        j->p = p; j->W = W; j->H = H; j->accum = accum; j->gbuf = gbuf; j->out = out;
        j->tid = t; j->nthreads = n_threads;
        pthread_create(&th[t], NULL, denoise_worker, j);
    }
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* Tonemap, src/headless.rs:93-98 with the NaN/Inf sanitising of :359-362      */
/* ------------------------------------------------------------------------ */
uint8_t oracle_tonemap_u8(float v) {
    if (!isfinite(v)) v = 0.0f;
    v = fmaxf(v, 0.0f);
    float mapped = v / (v + 1.0f);
    float g = clampf(powf(mapped, 1.0f / 2.2f), 0.0f, 1.0f);
    return (uint8_t)(g * 255.0f + 0.5f);
}
