// capi.hip — the extern "C" entry points of libmcrt_hip.so.
//
// Host logic of the drop-in boundary: context lifetime, validation and
// conversion of the reference's buffer contract (src/gpu_layout.rs:259-390)
// into the device layout of dev_common.h, and the stream-ordered launch
// sequence that replaces the headless spp loop (src/headless.rs:249-290).
// Nothing here computes radiance on the host: every render, trace and
// denoise result comes from the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/mcrt_abi.h"
#include "bvh_refine.h"
#include "bvh_wide.h"
#include "kernels.h"

using namespace mcrt;

namespace {

thread_local std::string g_create_error;

// Scratch of mcrt_trace_closest_noalloc_dev: the first 256 B hold the
// persistent grid's work counter, the rest the traversal-stack overflow
// columns.  One copy per device, for the life of the process; host code uses
// the device address from hipGetSymbolAddress (g_scratch), never this name.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr size_t kNoAllocCounterBytes = 256;
static __device__ __attribute__((used)) __align__(256) uint32_t scratch[kNoAllocScratchBytes / sizeof(uint32_t)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};          // device address of `scratch`, per device
hipError_t g_scratch_err[kMaxDevices] = {};  // result of its lookup

// Looks up `scratch` on `device` (which must be current) once per process.
hipError_t scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch[device]) {
        void *p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_scratch_err[device] == hipSuccess) g_scratch[device] = p;
    }
    return g_scratch[device] ? hipSuccess : g_scratch_err[device];
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Launch {
    std::string name;
    hipEvent_t start, stop;
};

}  // namespace

struct mcrt_ctx {
    int device = 0;
    std::string err;
    int num_cus = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t last_stream = nullptr;
    // scene
    bool has_scene = false;
    DevBuf mnodes, mnodes_full, mwide, mtris, tshade, snodes, sph, mats, env;
    DevBuf emit, emit_cdf, tri_emit, sph_emit;  // next-event estimation emitters
    DevBuf ovf, ovf2;  // traversal-stack overflow for trees deeper than the LDS short stack (per stream)
    hipStream_t side_stream = nullptr;  // second stream of a split batch (default priority)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DevScene scene{};
    uint32_t n_tris = 0, n_spheres = 0;
    bool nee = false;  // next-event estimation (mcrt_set_nee)
    int diff_mode = MCRT_DIFF_REPLAY;  // mcrt_diff_set_mode
    // work buffers
    DevBuf A[2], B[2], C[2], H, samples;
    DevBuf P[2], S_o, S_d, S_c;  // next-event estimation path pdf and shadow queue
    size_t cap_paths = 0, cap_nee = 0;
    DevBuf counters;
    size_t cap_counters = 0;
    DevBuf host_accum, host_gbuf, host_out, host_accum_in;
    DevBuf trace_o, trace_d, trace_id, trace_t, trace_work;
    DevBuf diff_partial, diff_grad, diff_image, diff_tape;
    // last render bookkeeping
    uint32_t last_batches = 0, last_depth = 0, last_halves = 1;
    uint64_t last_samples = 0;
    // profiling
    bool prof = false;
    std::vector<Launch> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, double> total_ms;
    std::map<std::string, uint64_t> launches;
};

namespace {

int fail(mcrt_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_create_error = buf;  // context-free calls (mcrt_validate_bvh)
    return code;
}

#define HIP_TRY(ctx, expr)                                                                \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(ctx, e_ == hipErrorOutOfMemory ? MCRT_ERR_OOM : MCRT_ERR_HIP,      \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                   \
    } while (0)

int ensure(mcrt_ctx *c, DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return MCRT_OK;
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    HIP_TRY(c, hipMalloc(&b.p, want));
    b.bytes = want;
    return MCRT_OK;
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int upload(mcrt_ctx *c, DevBuf &b, const void *src, size_t bytes) {
    int rc = ensure(c, b, bytes);
    if (rc) return rc;
    if (bytes) HIP_TRY(c, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return MCRT_OK;
}

int bind(mcrt_ctx *c) {
    if (!c) return MCRT_ERR_BAD_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    return MCRT_OK;
}

// --- profiling -------------------------------------------------------------
hipEvent_t take_event(mcrt_ctx *c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct Timed {
    mcrt_ctx *c;
    const char *name;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(mcrt_ctx *c_, const char *n, hipStream_t s_) : c(c_), name(n), s(s_) {
        if (c->prof) {
            a = take_event(c);
            b = take_event(c);
            if (a) (void)hipEventRecord(a, s);
        }
    }
    ~Timed() {
        if (a && b) {
            (void)hipEventRecord(b, s);
            c->pending.push_back(Launch{name, a, b});
        }
    }
};

int drain_profile(mcrt_ctx *c) {
    for (Launch &l : c->pending) {
        HIP_TRY(c, hipEventSynchronize(l.stop));
        float ms = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&ms, l.start, l.stop));
        c->total_ms[l.name] += ms;
        c->launches[l.name] += 1;
        c->pool.push_back(l.start);
        c->pool.push_back(l.stop);
    }
    c->pending.clear();
    return MCRT_OK;
}

// --- scene conversion -------------------------------------------------------
constexpr uint32_t kTopNodes = 64;  // device slots of the breadth-first top levels

struct TreeOut {
    std::vector<float4> nodes;  // 4 per internal node
    uint32_t top = 0;           // slots [0, top) hold the breadth-first top levels
    uint32_t root = kNone;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    uint32_t depth = 1;
    std::vector<uint8_t> last;  // per primitive: 1 if last in its leaf
};

bool is_stub_tree(const GpuBvhNode *nodes, uint32_t n) {
    return n == 0 || (n == 1 && nodes[0].prim_count == 0);
}

// Validates a reference pre-order tree and converts it to child pairs.
int convert_tree(mcrt_ctx *c, const char *what, const GpuBvhNode *nodes, uint32_t n,
                 uint32_t n_prims, TreeOut &out, bool pair_leaves = false) {
    out.last.assign(n_prims, 0);
    if (is_stub_tree(nodes, n)) return MCRT_OK;
    if (!nodes) return fail(c, MCRT_ERR_BAD_ARG, "%s BVH: null node array", what);
    std::vector<uint32_t> dev_index(n, kNone);
    std::vector<uint8_t> covered(n_prims, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const GpuBvhNode &nd = nodes[i];
        if (nd.prim_count == 0) {
            if (i + 1 >= n || nd.right_or_offset <= i + 1 || nd.right_or_offset >= n)
                return fail(c, MCRT_ERR_BAD_ARG,
                            "%s BVH: internal node %u has children (%u, %u) outside the pre-order "
                            "layout of %u nodes",
                            what, i, i + 1, nd.right_or_offset, n);
        } else {
            uint64_t end = (uint64_t)nd.right_or_offset + nd.prim_count;
            if (end > n_prims)
                return fail(c, MCRT_ERR_BAD_ARG, "%s BVH: leaf %u covers [%u, %llu) beyond %u primitives",
                            what, i, nd.right_or_offset, (unsigned long long)end, n_prims);
            for (uint64_t k = nd.right_or_offset; k < end; ++k) {
                if (covered[k])
                    return fail(c, MCRT_ERR_BAD_ARG, "%s BVH: primitive %llu is in two leaves", what,
                                (unsigned long long)k);
                covered[k] = 1;
            }
            out.last[end - 1] = 1;
        }
    }
    // Tree shape: every node is reached from the root exactly once (a node
    // with two parents, or one no parent reaches, is rejected before any
    // index is derived from it), and every primitive lies in a reached leaf.
    {
        std::vector<uint8_t> seen(n, 0);
        std::vector<uint32_t> todo{0u};
        while (!todo.empty()) {
            const uint32_t i = todo.back();
            todo.pop_back();
            if (seen[i])
                return fail(c, MCRT_ERR_BAD_ARG, "%s BVH: node %u is reached from two parents", what, i);
            seen[i] = 1;
            if (nodes[i].prim_count == 0) {
                todo.push_back(nodes[i].right_or_offset);
                todo.push_back(i + 1);
            }
        }
        for (uint32_t i = 0; i < n; ++i)
            if (!seen[i])
                return fail(c, MCRT_ERR_BAD_ARG, "%s BVH: node %u is not reachable from the root", what, i);
        for (uint32_t k = 0; k < n_prims; ++k)
            if (!covered[k])
                return fail(c, MCRT_ERR_BAD_ARG, "%s BVH: primitive %u is in no leaf", what, k);
    }
    // Device order: the internal children of a node take consecutive indices,
    // a pair starting at an even index, so siblings share one 128-B line (one
    // 64-B node each) and a traversal that visits both fetches the line once.
    // Subtrees are laid out depth-first.  Unused slots stay zero and are never
    // referenced.
    uint32_t n_dev = 0;
    out.top = 0;
    if (nodes[0].prim_count == 0) {
        dev_index[0] = 0;
        n_dev = 2;  // the root alone in the first line
        // The top levels breadth-first, as far as kTopNodes device slots go
        // (the traversal keeps them in LDS), then each remaining subtree
        // depth-first.
        std::vector<uint32_t> frontier, q{0u};
        for (size_t h = 0; h < q.size(); ++h) {
            const uint32_t i = q[h];
            const uint32_t a = i + 1, b = nodes[i].right_or_offset;
            const bool ia = nodes[a].prim_count == 0, ib = nodes[b].prim_count == 0;
            const uint32_t need = (ia && ib) ? 2u + (n_dev & 1u) : (ia || ib) ? 1u : 0u;
            if (!frontier.empty() || n_dev + need > kTopNodes) {
                frontier.push_back(i);
                continue;
            }
            if (ia && ib) {
                n_dev += n_dev & 1u;
                dev_index[a] = n_dev;
                dev_index[b] = n_dev + 1;
                n_dev += 2;
            } else if (ia) {
                dev_index[a] = n_dev++;
            } else if (ib) {
                dev_index[b] = n_dev++;
            }
            if (ia) q.push_back(a);
            if (ib) q.push_back(b);
        }
        out.top = n_dev;
        std::vector<uint32_t> todo(frontier.rbegin(), frontier.rend());
        while (!todo.empty()) {
            const uint32_t i = todo.back();
            todo.pop_back();
            const uint32_t a = i + 1, b = nodes[i].right_or_offset;
            const bool ia = nodes[a].prim_count == 0, ib = nodes[b].prim_count == 0;
            if (ia && ib) {
                n_dev += n_dev & 1u;
                dev_index[a] = n_dev;
                dev_index[b] = n_dev + 1;
                n_dev += 2;
            } else if (ia) {
                dev_index[a] = n_dev++;
            } else if (ib) {
                dev_index[b] = n_dev++;
            }
            if (ib) todo.push_back(b);
            if (ia) todo.push_back(a);
        }
    }
    auto ref = [&](uint32_t i) -> uint32_t {
        if (nodes[i].prim_count == 0) return dev_index[i];
        const bool pair = pair_leaves && nodes[i].prim_count == 2 && n_prims < kLeafPair;
        return kLeafBit | (pair ? kLeafPair : 0u) | nodes[i].right_or_offset;
    };
    out.nodes.assign((size_t)std::max<uint32_t>(n_dev, 1) * 4, make_float4(0, 0, 0, 0));
    for (uint32_t i = 0; i < n; ++i) {
        if (nodes[i].prim_count != 0) continue;
        const GpuBvhNode &a = nodes[i + 1];
        const GpuBvhNode &b = nodes[nodes[i].right_or_offset];
        float4 *q = &out.nodes[4 * (size_t)dev_index[i]];
        q[0] = make_float4(a.aabb_min[0], a.aabb_max[0], a.aabb_min[1], a.aabb_max[1]);
        q[1] = make_float4(b.aabb_min[0], b.aabb_max[0], b.aabb_min[1], b.aabb_max[1]);
        q[2] = make_float4(a.aabb_min[2], a.aabb_max[2], b.aabb_min[2], b.aabb_max[2]);
        uint32_t r0 = ref(i + 1), r1 = ref(nodes[i].right_or_offset);
        float f0, f1;
        std::memcpy(&f0, &r0, 4);
        std::memcpy(&f1, &r1, 4);
        q[3] = make_float4(f0, f1, 0.0f, 0.0f);
    }
    out.root = ref(0);
    for (int k = 0; k < 3; ++k) {
        out.lo[k] = nodes[0].aabb_min[k];
        out.hi[k] = nodes[0].aabb_max[k];
    }
    // Largest number of internal nodes on a root-to-leaf path: the traversal
    // pushes at most one far child per internal node on its current path.
    uint32_t maxd = 0;
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.push_back({0u, 0u});
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        if (nodes[i].prim_count == 0) {
            st.push_back({i + 1, d + 1});
            st.push_back({nodes[i].right_or_offset, d + 1});
        } else {
            maxd = std::max(maxd, d);
        }
    }
    out.depth = std::max<uint32_t>(maxd, 1);
    return MCRT_OK;
}

// The child pairs of a converted tree as quantized 32-B pairs
// (node_quant.h); unused slots stay zero.  False where a box cannot be
// encoded.
bool quantize_tree(const std::vector<float4> &nodes, std::vector<uint32_t> &out) {
    const size_t n = nodes.size() / 4;
    out.assign(std::max<size_t>(n, 1) * kQuantWords, 0u);
    for (size_t i = 0; i < n; ++i) {
        const float4 *q = &nodes[4 * i];
        uint32_t r0, r1;
        std::memcpy(&r0, &q[3].x, 4);
        std::memcpy(&r1, &q[3].y, 4);
        if (r0 == 0u && r1 == 0u) continue;  // an unused slot (child 0 is never the root)
        const float lo0[3] = {q[0].x, q[0].z, q[2].x}, hi0[3] = {q[0].y, q[0].w, q[2].y};
        const float lo1[3] = {q[1].x, q[1].z, q[2].z}, hi1[3] = {q[1].y, q[1].w, q[2].w};
        if (!qnode_encode(lo0, hi0, lo1, hi1, r0, r1, &out[kQuantWords * i])) return false;
    }
    return true;
}

}  // namespace

extern "C" int mcrt_validate_bvh(const GpuBvhNode *nodes, uint32_t n_nodes, uint32_t n_prims) {
    TreeOut out;
    return convert_tree(nullptr, "given", nodes, n_nodes, n_prims, out);
}

namespace {

inline float ubits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Device counters per batch: path counts [max_depth+1], extend and shade work
// counters [max_depth+1] each, then shadow-ray counts and occlusion work
// counters [max_depth] each.
uint32_t counters_per_batch(uint32_t max_depth) {
    return 3u * (max_depth + 1u) + 2u * max_depth;
}

// Frames per batch: a batch holds at most 2^26 paths, 8.6 GB of path state
// (11.8 GB with next-event estimation).  Larger batches give a wave more
// frames of the same pixels and drain fewer persistent launches: the 1080p
// voxel terrain at 16 spp goes from 2 batches to 1 (+7%); the icosphere
// Cornell scene at 1024^2, 64 spp, from 4 to 1 (+12%).
constexpr uint64_t kMaxBatchPaths = 1ull << 26;

uint32_t frames_per_batch(uint32_t wh, uint32_t spp) {
    uint64_t cap = kMaxBatchPaths / std::max<uint32_t>(wh, 1);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, spp));
}

int grid_for(mcrt_ctx *c, int which, uint32_t *grid) {
    int per_cu = 0;
    hipError_t e = occupancy(which, c->scene, &per_cu);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    *grid = (uint32_t)(c->num_cus * per_cu);
    return MCRT_OK;
}

// Overflow columns for traversal stacks deeper than the LDS short stack:
// (depth - short) entries of the stack layout's words (node reference, and
// the entry distance for the culling stack) for every thread of a grid.
int ensure_overflow(mcrt_ctx *c, uint32_t grid_threads, bool cull, uint32_t which, uint32_t **ptr) {
    *ptr = nullptr;
    uint32_t sh = 0, words = 0;
    stack_layout(cull, &sh, &words);
    const uint32_t depth = c->scene.stack_depth;
    if (depth <= sh) return MCRT_OK;
    DevBuf &buf = which ? c->ovf2 : c->ovf;
    int rc = ensure(c, buf, (size_t)(depth - sh) * words * grid_threads * sizeof(uint32_t));
    if (rc) return rc;
    *ptr = (uint32_t *)buf.p;
    return MCRT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int mcrt_create(mcrt_ctx **out_ctx, int device) {
    if (!out_ctx) return MCRT_ERR_BAD_ARG;
    *out_ctx = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_create_error = "no suitable GPU adapter: no HIP device";
        return MCRT_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_error = "no suitable GPU adapter: device index out of range";
        return MCRT_ERR_NO_DEVICE;
    }
    mcrt_ctx *c = new mcrt_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreate(&c->own_stream) != hipSuccess) {
        g_create_error = "no suitable GPU adapter: HIP initialisation failed";
        delete c;
        return MCRT_ERR_NO_DEVICE;
    }
    c->num_cus = prop.multiProcessorCount;
    // The no-allocation trace's scratch: its device address, per device.  A
    // failed lookup is reported by that entry point, not here.
    (void)scratch_setup(device);
    *out_ctx = c;
    return MCRT_OK;
}

extern "C" void mcrt_destroy(mcrt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->last_stream);
    (void)hipStreamSynchronize(c->own_stream);
    if (c->side_stream) (void)hipStreamSynchronize(c->side_stream);
    for (DevBuf *b : {&c->mnodes, &c->mnodes_full, &c->mwide, &c->mtris, &c->tshade, &c->ovf, &c->ovf2, &c->snodes, &c->sph, &c->mats,
                      &c->env, &c->A[0], &c->A[1], &c->B[0], &c->B[1], &c->C[0], &c->C[1], &c->H,
                      &c->samples, &c->counters, &c->host_accum, &c->host_gbuf, &c->host_out,
                      &c->host_accum_in, &c->trace_o, &c->trace_d, &c->trace_id, &c->trace_t,
                      &c->trace_work, &c->diff_partial, &c->diff_grad, &c->diff_image, &c->diff_tape, &c->emit,
                      &c->emit_cdf, &c->tri_emit, &c->sph_emit, &c->P[0], &c->P[1], &c->S_o, &c->S_d,
                      &c->S_c})
        release(*b);
    for (Launch &l : c->pending) {
        (void)hipEventDestroy(l.start);
        (void)hipEventDestroy(l.stop);
    }
    for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    if (c->side_stream) (void)hipStreamDestroy(c->side_stream);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    delete c;
}

extern "C" const char *mcrt_last_error(const mcrt_ctx *c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

extern "C" int mcrt_upload_scene(mcrt_ctx *c, const GpuSphere *spheres, uint32_t n_spheres,
                                 const uint32_t *sphere_orig_index, const GpuBvhNode *sphere_bvh,
                                 uint32_t n_sphere_nodes, const GpuVertex *verts, uint32_t n_verts,
                                 const GpuTriangle *tris, uint32_t n_tris,
                                 const uint32_t *tri_orig_index, const GpuBvhNode *mesh_bvh,
                                 uint32_t n_mesh_nodes, const GpuMaterialData *materials,
                                 const float *env_rgba32f, uint32_t env_w, uint32_t env_h,
                                 const uint8_t *albedo_rgba8, uint32_t n_layers,
                                 uint32_t layer_size) {
    int rc = bind(c);
    if (rc) return rc;
    c->has_scene = false;
    if (!materials) return fail(c, MCRT_ERR_BAD_ARG, "materials must not be null");
    if (materials->mat_count > MCRT_MAX_MATERIALS)
        return fail(c, MCRT_ERR_LIMIT, "scene uses %u materials but limit is %u", materials->mat_count,
                    MCRT_MAX_MATERIALS);
    if (!env_rgba32f || env_w == 0 || env_h == 0)
        return fail(c, MCRT_ERR_BAD_ARG, "environment map is required (got %ux%u)", env_w, env_h);
    if ((n_spheres && !spheres) || (n_tris && (!tris || !verts)))
        return fail(c, MCRT_ERR_BAD_ARG, "null primitive array with a non-zero count");
    if (n_tris >= kLeafPair || n_spheres >= kLeafPair)  // primitive indices share a word with two flag bits
        return fail(c, MCRT_ERR_LIMIT, "at most %u triangles and %u spheres (got %u and %u)", kLeafPair - 1u,
                    kLeafPair - 1u, n_tris, n_spheres);
    if (n_layers > 0 && albedo_rgba8) {
        for (uint32_t m = 0; m < MCRT_MAX_MATERIALS; ++m) {
            const GpuMaterial &g = materials->materials[m];
            if (g.type_pad[1] != 0 && g.type_pad[0] != MCRT_MAT_EMISSIVE && m < materials->mat_count)
                return fail(c, MCRT_ERR_BAD_ARG,
                            "material %u samples albedo texture layer %u; texture sampling is not "
                            "implemented by this library",
                            m, g.type_pad[1]);
        }
    }
    (void)layer_size;
    for (uint32_t i = 0; i < n_tris; ++i)
        for (int k = 0; k < 3; ++k)
            if (tris[i].v[k] >= n_verts)
                return fail(c, MCRT_ERR_BAD_ARG, "triangle %u references vertex %u of %u", i,
                            tris[i].v[k], n_verts);

    TreeOut mt, st;
    rc = convert_tree(c, "mesh", mesh_bvh, n_mesh_nodes, n_tris, mt, true);
    if (rc) return rc;
    // Leaves of at most 2 triangles: fewer triangle tests and shorter
    // dependent leaf loops for about 3% more node steps.
    std::vector<GpuBvhNode> rnodes;
    std::vector<uint32_t> perm, rorig;
    std::vector<GpuTriangle> rtris;
    // given[j]: the index in the given triangle order of leaf-order record j
    // (the emitter list keeps the given order whatever the records' order).
    std::vector<uint32_t> given(n_tris);
    for (uint32_t j = 0; j < n_tris; ++j) given[j] = j;
    if (refine_mesh_tree(mesh_bvh, n_mesh_nodes, verts, tris, n_tris, rnodes, perm) &&
        rnodes.size() < 0xFFFFFFFFu) {
        TreeOut rt;
        rc = convert_tree(c, "mesh", rnodes.data(), (uint32_t)rnodes.size(), n_tris, rt, true);
        if (rc) return rc;
        if (rt.depth <= 96u || rt.depth <= mt.depth) {
            mt = std::move(rt);
            rtris.resize(n_tris);
            rorig.resize(n_tris);
            for (uint32_t j = 0; j < n_tris; ++j) {
                rtris[j] = tris[perm[j]];
                rorig[j] = tri_orig_index ? tri_orig_index[perm[j]] : perm[j];
            }
            tris = rtris.data();
            tri_orig_index = rorig.data();
            given = perm;
        }
    }
    rc = convert_tree(c, "sphere", sphere_bvh, n_sphere_nodes, n_spheres, st);
    if (rc) return rc;
    // Bytes a traversal can touch (nodes + triangles + spheres); the per-scene
    // policies below compare it with the 32 MB L2.
    const size_t traversal_bytes = (mt.nodes.size() + st.nodes.size()) * sizeof(float4) +
                                   (size_t)n_tris * 48 + (size_t)n_spheres * 32;
    const bool fits_l2 = traversal_bytes <= (32u << 20);

    // Scenes that spill the L2: the culling-stack kernels traverse the mesh
    // tree as 4-wide 48-B nodes (bvh_wide.h).  The leaf records are
    // reordered here, before any per-record array is built, so that the
    // leaves of each wide node follow one another; the binary tree's leaf
    // references follow the new order.
    WidePlan wplan;
    bool wide = false;
    std::vector<GpuTriangle> wtris;
    std::vector<uint32_t> worig;
    if (!fits_l2 && n_tris && mt.root != kNone && !(mt.root & kLeafBit)) {
        const WidePairs bp{mt.nodes.data(), mt.nodes.size() / 4, kLeafBit, kLeafPair, kNone};
        std::vector<uint32_t> wperm, wstart;
        if (wide_plan(bp, mt.root, wplan) && wplan.stack_need <= 96u &&
            wide_leaf_order(bp, wplan, mt.last.data(), n_tris, wperm, wstart)) {
            wtris.resize(n_tris);
            worig.resize(n_tris);
            std::vector<uint8_t> wlast(n_tris);
            std::vector<uint32_t> wgiven(n_tris);
            for (uint32_t j = 0; j < n_tris; ++j) {
                wtris[j] = tris[wperm[j]];
                worig[j] = tri_orig_index ? tri_orig_index[wperm[j]] : wperm[j];
                wlast[j] = mt.last[wperm[j]];
                wgiven[j] = given[wperm[j]];
            }
            given = std::move(wgiven);
            for (size_t i = 0; i < mt.nodes.size() / 4; ++i) {
                float4 &q = mt.nodes[4 * i + 3];
                for (float *f : {&q.x, &q.y}) {
                    uint32_t r;
                    std::memcpy(&r, f, 4);
                    if (r == kNone || !(r & kLeafBit)) continue;
                    r = (r & (kLeafBit | kLeafPair)) | wstart[r & ~(kLeafBit | kLeafPair)];
                    std::memcpy(f, &r, 4);
                }
            }
            tris = wtris.data();
            tri_orig_index = worig.data();
            mt.last = std::move(wlast);
            wide = true;
        }
    }

    // Triangles in leaf order: v0, e1 = v1 - v0, e2 = v2 - v0.
    // Scenes that spill the L2 pad each triangle to 64 B, so a record gathered
    // from memory does not straddle two 128-B lines (the 1080p voxel terrain:
    // +0.5%); cache-resident scenes keep the denser 48 B (padding cost the
    // 70k-triangle icosphere scene 2%).
    const uint32_t tsd = fits_l2 ? 3u : 4u;
    std::vector<float4> dt((size_t)std::max<uint32_t>(n_tris, 1) * tsd, make_float4(0, 0, 0, 0));
    for (uint32_t i = 0; i < n_tris; ++i) {
        const float *p0 = verts[tris[i].v[0]].position;
        const float *p1 = verts[tris[i].v[1]].position;
        const float *p2 = verts[tris[i].v[2]].position;
        uint32_t flags = (mt.last[i] ? kPrimLast : 0u) | ((tris[i].mat_idx >> 31) ? kPrimCull : 0u);
        uint32_t orig = tri_orig_index ? tri_orig_index[i] : i;
        dt[tsd * (size_t)i + 0] = make_float4(p0[0], p0[1], p0[2], ubits(orig));
        dt[tsd * (size_t)i + 1] =
            make_float4(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], ubits(flags));
        dt[tsd * (size_t)i + 2] = make_float4(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.0f);
    }
    // Shading records in leaf order: vertex normals + mat_idx (with flag).
    // Scenes whose traversal set spills the L2 gather them from memory, so
    // each is padded to one aligned 64-B sector (the 1080p voxel terrain's
    // shade pass: -2.7%); cache-resident scenes keep the denser 48 B.
    const uint32_t ts = fits_l2 ? 3u : 4u;
    std::vector<float4> dsh((size_t)std::max<uint32_t>(n_tris, 1) * ts, make_float4(0, 0, 0, 0));
    for (uint32_t i = 0; i < n_tris; ++i) {
        for (int k = 0; k < 3; ++k) {
            const float *nn = verts[tris[i].v[k]].normal;
            dsh[ts * (size_t)i + k] = make_float4(nn[0], nn[1], nn[2], k == 0 ? ubits(tris[i].mat_idx) : 0.0f);
        }
    }
    std::vector<float4> ds((size_t)std::max<uint32_t>(n_spheres, 1) * 2, make_float4(0, 0, 0, 0));
    for (uint32_t i = 0; i < n_spheres; ++i) {
        const GpuSphere &s = spheres[i];
        uint32_t orig = sphere_orig_index ? sphere_orig_index[i] : i;
        ds[2 * (size_t)i] = make_float4(s.center_r[0], s.center_r[1], s.center_r[2], s.center_r[3]);
        ds[2 * (size_t)i + 1] = make_float4(ubits(s.mat_and_pad[0]), ubits(orig),
                                            ubits(st.last[i] ? kPrimLast : 0u), 0.0f);
    }
    std::vector<float4> denv((size_t)env_w * env_h);
    std::memcpy(denv.data(), env_rgba32f, denv.size() * sizeof(float4));

    // Emitters for next-event estimation: triangles and spheres whose material
    // is emissive (emission = colour x strength, :608-610), chosen with
    // probability proportional to area x luminance.  Triangle area from the
    // same FP32 edges the kernels use.  Triangles are listed in the given
    // order (the order of the caller's triangle array), so a light sample
    // picks the same triangle whatever order the records take here.
    std::vector<float4> demit;
    std::vector<float> dcdf;
    std::vector<uint32_t> dtri_emit, dsph_emit;
    {
        struct Em {
            uint32_t ref;
            float area, le[3], w;
        };
        std::vector<Em> em;
        auto emission = [&](uint32_t mat, float le[3]) {
            const GpuMaterial &g = materials->materials[std::min<uint32_t>(mat, MCRT_MAX_MATERIALS - 1)];
            if (g.type_pad[0] != MCRT_MAT_EMISSIVE) return false;
            for (int k = 0; k < 3; ++k) le[k] = g.ior_pad[0] * g.albedo_fuzz[k];
            return true;
        };
        auto lum = [](const float le[3]) { return 0.2126f * le[0] + 0.7152f * le[1] + 0.0722f * le[2]; };
        std::vector<uint32_t> at(n_tris);  // leaf-order record of each given index
        for (uint32_t j = 0; j < n_tris; ++j) at[given[j]] = j;
        for (uint32_t g = 0; g < n_tris; ++g) {
            const uint32_t i = at[g];
            Em e{};
            if (!emission(tris[i].mat_idx & 0x7FFFFFFFu, e.le)) continue;
            const float4 e1 = dt[tsd * (size_t)i + 1], e2 = dt[tsd * (size_t)i + 2];
            const float cx = e1.y * e2.z - e1.z * e2.y, cy = e1.z * e2.x - e1.x * e2.z,
                        cz = e1.x * e2.y - e1.y * e2.x;
            e.area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
            e.w = e.area * lum(e.le);
            e.ref = i;
            if (e.w > 0.0f && std::isfinite(e.w)) em.push_back(e);
        }
        for (uint32_t i = 0; i < n_spheres; ++i) {
            Em e{};
            if (!emission(spheres[i].mat_and_pad[0], e.le)) continue;
            const float r = spheres[i].center_r[3];
            e.area = 4.0f * 3.14159265358979f * r * r;
            e.w = e.area * lum(e.le);
            e.ref = 0x80000000u | i;
            if (e.w > 0.0f && std::isfinite(e.w)) em.push_back(e);
        }
        double total = 0.0;
        for (const Em &e : em) total += e.w;
        if (!em.empty()) {
            dtri_emit.assign(std::max<uint32_t>(n_tris, 1), kNone);
            dsph_emit.assign(std::max<uint32_t>(n_spheres, 1), kNone);
            double run = 0.0;
            for (size_t k = 0; k < em.size(); ++k) {
                const Em &e = em[k];
                const float prob = (float)(e.w / total);
                run += e.w;
                dcdf.push_back(k + 1 == em.size() ? 1.0f : (float)(run / total));
                demit.push_back(make_float4(prob, ubits(e.ref), e.area, 0.0f));
                demit.push_back(make_float4(e.le[0], e.le[1], e.le[2], 0.0f));
                if (e.ref & 0x80000000u)
                    dsph_emit[e.ref & 0x7FFFFFFFu] = (uint32_t)k;
                else
                    dtri_emit[e.ref] = (uint32_t)k;
            }
        }
    }

    // Scenes that spill the L2 traverse quantized 32-B child pairs with the
    // culling stack (voxel terrain 1186 -> 1237 Msamples/s: half the node
    // bytes per step, and the address unit is the busiest unit of the
    // traversal); cache-resident scenes keep the 64-B pairs, which decode for
    // free (icosphere Cornell: quantized pairs cost 4% of the render and 9% of
    // the 16M-ray closest-hit launch).  A tree with a box the grid cannot
    // encode keeps the 64-B pairs and the plain stack.
    std::vector<uint32_t> qm, qs;
    // Quad leaves (scenes that spill the L2, 64-B triangle records): a leaf of
    // two triangles (a, b, c) + (a, c, d), in either order, keeps both in its
    // first record (dev_common.h tri_leaf); other 2-triangle leaves lose the
    // kLeafPair flag and take the per-record loop.  Equality is bitwise.
    if (!fits_l2 && tsd == 4u) {
        auto same3 = [](const float4 &x, const float4 &y) {
            return std::memcmp(&x, &y, 3 * sizeof(float)) == 0;
        };
        auto quad = [&](uint32_t r) -> uint32_t {
            if (r == kNone || !(r & kLeafBit) || !(r & kLeafPair)) return r;
            const uint32_t i = r & ~(kLeafBit | kLeafPair);
            float4 *x = &dt[4 * (size_t)i], *y = &dt[4 * (size_t)i + 4];
            if (!same3(x[0], y[0])) return r & ~kLeafPair;
            uint32_t fl;
            if (same3(y[1], x[2])) {  // x = (a, b, c), y = (a, c, d)
                std::memcpy(&fl, &y[1].w, 4);
                x[3] = make_float4(y[2].x, y[2].y, y[2].z, ubits(fl));
            } else if (same3(x[1], y[2])) {  // x = (a, c, d), y = (a, b, c)
                std::memcpy(&fl, &y[1].w, 4);
                x[3] = make_float4(y[1].x, y[1].y, y[1].z, ubits(fl | kQuadSwap));
            } else {
                return r & ~kLeafPair;
            }
            x[2].w = y[0].w;  // the second triangle's id
            return r;
        };
        for (size_t n = 0; n < mt.nodes.size() / 4; ++n) {
            float4 &q = mt.nodes[4 * n + 3];
            uint32_t r0, r1;
            std::memcpy(&r0, &q.x, 4);
            std::memcpy(&r1, &q.y, 4);
            r0 = quad(r0);
            r1 = quad(r1);
            std::memcpy(&q.x, &r0, 4);
            std::memcpy(&q.y, &r1, 4);
        }
        mt.root = quad(mt.root);
    }
    bool quant = !fits_l2 && quantize_tree(mt.nodes, qm) && quantize_tree(st.nodes, qs);
    // The culling-stack kernels read the mesh tree as 4-wide nodes; a mesh
    // tree with internal nodes but no 4-wide form keeps the plain stack.
    std::vector<uint32_t> wwords;
    if (quant && wide) {
        const WidePairs bp{mt.nodes.data(), mt.nodes.size() / 4, kLeafBit, kLeafPair, kNone};
        wide = wide_emit(bp, wplan, wwords);
    }
    if (quant && n_tris && mt.root != kNone && !(mt.root & kLeafBit) && !wide) quant = false;
    // Quad leaves serve the culling-stack kernels only (the 1080p voxel
    // terrain: +1.7% for fetching a leaf's two records at once, +1.7% more
    // for one 64-B quad record); the plain-stack kernels of cache-resident
    // scenes compile no quad path, so their trees drop the kLeafPair flag.
    if (!quant) {
        auto strip = [](uint32_t r) { return (r != kNone && (r & kLeafBit)) ? (r & ~kLeafPair) : r; };
        for (size_t i = 0; i < mt.nodes.size() / 4; ++i) {
            float4 &q = mt.nodes[4 * i + 3];
            uint32_t r0, r1;
            std::memcpy(&r0, &q.x, 4);
            std::memcpy(&r1, &q.y, 4);
            r0 = strip(r0);
            r1 = strip(r1);
            std::memcpy(&q.x, &r0, 4);
            std::memcpy(&q.y, &r1, 4);
        }
        mt.root = strip(mt.root);
    }
    if (quant) {
        if ((rc = upload(c, c->mnodes, qm.data(), qm.size() * sizeof(uint32_t)))) return rc;
        // the exact 64-B pairs too, for the first bounce's packet traversal
        // (scalar loads of whole pairs, no plane decode: +2% on the voxel terrain)
        if ((rc = upload(c, c->mnodes_full, mt.nodes.data(), mt.nodes.size() * sizeof(float4)))) return rc;
        if (wide && (rc = upload(c, c->mwide, wwords.data(), wwords.size() * sizeof(uint32_t)))) return rc;
    } else {
        if ((rc = upload(c, c->mnodes, mt.nodes.data(), mt.nodes.size() * sizeof(float4)))) return rc;
    }
    if ((rc = upload(c, c->mtris, dt.data(), dt.size() * sizeof(float4)))) return rc;
    if ((rc = upload(c, c->tshade, dsh.data(), dsh.size() * sizeof(float4)))) return rc;
    if (quant) {
        if ((rc = upload(c, c->snodes, qs.data(), qs.size() * sizeof(uint32_t)))) return rc;
    } else {
        if ((rc = upload(c, c->snodes, st.nodes.data(), st.nodes.size() * sizeof(float4)))) return rc;
    }
    if ((rc = upload(c, c->sph, ds.data(), ds.size() * sizeof(float4)))) return rc;
    if ((rc = upload(c, c->mats, materials->materials, sizeof(materials->materials)))) return rc;
    if ((rc = upload(c, c->env, denv.data(), denv.size() * sizeof(float4)))) return rc;
    if (!demit.empty()) {
        if ((rc = upload(c, c->emit, demit.data(), demit.size() * sizeof(float4))) ||
            (rc = upload(c, c->emit_cdf, dcdf.data(), dcdf.size() * sizeof(float))) ||
            (rc = upload(c, c->tri_emit, dtri_emit.data(), dtri_emit.size() * sizeof(uint32_t))) ||
            (rc = upload(c, c->sph_emit, dsph_emit.data(), dsph_emit.size() * sizeof(uint32_t))))
            return rc;
    }

    DevScene &d = c->scene;
    d = DevScene{};
    d.n_emit = (uint32_t)dcdf.size();
    d.emit = d.n_emit ? (const float4 *)c->emit.p : nullptr;
    d.emit_cdf = d.n_emit ? (const float *)c->emit_cdf.p : nullptr;
    d.tri_emit = d.n_emit ? (const uint32_t *)c->tri_emit.p : nullptr;
    d.sph_emit = d.n_emit ? (const uint32_t *)c->sph_emit.p : nullptr;
    d.mnodes = (const float4 *)c->mnodes.p;
    d.mnodes_full = (const float4 *)(quant ? c->mnodes_full.p : c->mnodes.p);
    d.mtris = (const float4 *)c->mtris.p;
    d.tshade = (const float4 *)c->tshade.p;
    d.tshade_stride = ts;
    d.tri_stride = tsd;
    d.mroot = n_tris ? mt.root : kNone;
    d.mwide = (quant && wide) ? (const uint4 *)c->mwide.p : nullptr;
    d.mwroot = !n_tris ? kNone : (quant && wide) ? 0u : mt.root;
    std::memcpy(d.mroot_lo, mt.lo, sizeof mt.lo);
    std::memcpy(d.mroot_hi, mt.hi, sizeof mt.hi);
    d.snodes = (const float4 *)c->snodes.p;
    d.sph = (const float4 *)c->sph.p;
    d.sroot = n_spheres ? st.root : kNone;
    std::memcpy(d.sroot_lo, st.lo, sizeof st.lo);
    std::memcpy(d.sroot_hi, st.hi, sizeof st.hi);
    d.mats = (const GpuMaterial *)c->mats.p;
    d.env = (const float4 *)c->env.p;
    d.env_w = env_w;
    d.env_h = env_h;
    d.stack_depth = std::max(mt.depth, st.depth);
    if (quant && wide) d.stack_depth = std::max(d.stack_depth, wplan.stack_need);
    // A refill stops the whole wave to load new rays.  Where the traversal
    // set (nodes + triangles + spheres) outgrows the 32 MB of L2 that load's
    // latency is long, and refilling more lanes per stop pays: at 40 idle
    // lanes the 2.16M-triangle voxel terrain renders 4.5% faster than at 16.
    // A cache-resident scene keeps lanes fuller instead (the 70k-triangle
    // icosphere Cornell scene: 3% slower at 32, 8% at 48).
    d.refill = fits_l2 ? 16u : 40u;
    // Culling pops skip node fetches that miss the L2 on such scenes (voxel
    // terrain: +2.8%); on a cache-resident scene the bookkeeping and the
    // shorter LDS stack cost more than they save (icosphere Cornell: -3.5%).
    // The culling-stack kernels read the quantized pairs.
    d.stack_cull = quant ? 1u : 0u;
    d.mtop = quant ? 0u : std::min<uint32_t>(mt.top, kNodeCache);
    if (d.stack_depth > 96)
        return fail(c, MCRT_ERR_LIMIT, "BVH depth %u exceeds the traversal stack limit of 96",
                    d.stack_depth);
    c->n_tris = n_tris;
    c->n_spheres = n_spheres;
    c->has_scene = true;
    return MCRT_OK;
}

extern "C" int mcrt_render_aux_dev(mcrt_ctx *c, const CameraUniform *cam, uint32_t width,
                                   uint32_t height, uint32_t spp, uint32_t first_frame,
                                   uint32_t max_depth, const float *accum_in_dev,
                                   float *out_accum_dev, float *out_gbuf_dev,
                                   float *out_albedo_dev, void *stream);

extern "C" int mcrt_set_nee(mcrt_ctx *c, int enabled) {
    if (!c) return MCRT_ERR_BAD_ARG;
    c->nee = enabled != 0;
    return MCRT_OK;
}

extern "C" int mcrt_render_dev(mcrt_ctx *c, const CameraUniform *cam, uint32_t width,
                               uint32_t height, uint32_t spp, uint32_t first_frame,
                               uint32_t max_depth, const float *accum_in_dev, float *out_accum_dev,
                               float *out_gbuf_dev, void *stream) {
    return mcrt_render_aux_dev(c, cam, width, height, spp, first_frame, max_depth, accum_in_dev,
                               out_accum_dev, out_gbuf_dev, nullptr, stream);
}

extern "C" int mcrt_render_aux_dev(mcrt_ctx *c, const CameraUniform *cam, uint32_t width,
                                   uint32_t height, uint32_t spp, uint32_t first_frame,
                                   uint32_t max_depth, const float *accum_in_dev,
                                   float *out_accum_dev, float *out_gbuf_dev,
                                   float *out_albedo_dev, void *stream) {
    int rc = bind(c);
    if (rc) return rc;
    if (!c->has_scene) return fail(c, MCRT_ERR_NO_SCENE, "render called before mcrt_upload_scene");
    if (!cam || !out_accum_dev) return fail(c, MCRT_ERR_BAD_ARG, "camera and out_accum are required");
    if (width == 0 || height == 0 || spp == 0)
        return fail(c, MCRT_ERR_BAD_ARG, "width, height and spp must be positive");
    if (max_depth < 1 || max_depth > 64)
        return fail(c, MCRT_ERR_BAD_ARG, "max_depth must be in [1, 64] (got %u)", max_depth);
    uint64_t wh64 = (uint64_t)width * height;
    if (wh64 > (1ull << 26)) return fail(c, MCRT_ERR_LIMIT, "image of %llu pixels is too large",
                                         (unsigned long long)wh64);
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    const uint32_t wh = (uint32_t)wh64;
    const uint32_t F = frames_per_batch(wh, spp);
    const size_t cap = (size_t)wh * F;
    if (cap > c->cap_paths) {
        for (DevBuf *b : {&c->A[0], &c->A[1], &c->B[0], &c->B[1], &c->C[0], &c->C[1], &c->H, &c->samples})
            release(*b);
        for (DevBuf *b : {&c->A[0], &c->A[1], &c->B[0], &c->B[1], &c->C[0], &c->C[1], &c->H, &c->samples})
            if ((rc = ensure(c, *b, cap * sizeof(float4)))) return rc;
        c->cap_paths = cap;
    }
    const bool nee = c->nee;
    if (nee && cap > c->cap_nee) {
        for (DevBuf *b : {&c->P[0], &c->P[1], &c->S_o, &c->S_d, &c->S_c}) release(*b);
        for (DevBuf *b : {&c->S_o, &c->S_d, &c->S_c})
            if ((rc = ensure(c, *b, cap * sizeof(float4)))) return rc;
        for (DevBuf *b : {&c->P[0], &c->P[1]})
            if ((rc = ensure(c, *b, cap * sizeof(float)))) return rc;
        c->cap_nee = cap;
    }
    const uint32_t n_batches = (spp + F - 1) / F;
    // Cache-resident scenes trace a batch as two halves on two streams: a
    // bounce's persistent launch on one half drains its slowest rays while the
    // other half's launches fill the device.  A batch of at least 4 frames is
    // split into frame halves (the icosphere Cornell frame at 1080p: 6 spp 5%
    // faster, 28 spp 2%), a batch of 1-3 frames into pixel halves (1 spp 11% faster, with
    // next-event estimation 11-13%).  Scenes that spill the L2 stay on one
    // stream, and so do batches of 4+ frames with next-event estimation
    // (measured flat to slower).
    const bool split = c->scene.stack_cull == 0 && (!nee || F < 4);
    const uint32_t halves = split ? 2u : 1u;
    const uint32_t per_batch = counters_per_batch(max_depth);
    const size_t n_counters = (size_t)per_batch * n_batches * halves;
    if ((rc = ensure(c, c->counters, n_counters * sizeof(uint32_t)))) return rc;
    uint32_t *counters = (uint32_t *)c->counters.p;
    HIP_TRY(c, launch_fill_u32(counters, (uint32_t)n_counters, 0u, s));

    uint32_t g_ext = 0, g_occ = 0;
    grid_for(c, 0, &g_ext);
    if (nee) grid_for(c, 3, &g_occ);
    const uint32_t ovf_threads = std::max(g_ext, g_occ) * (uint32_t)trace_block();
    uint32_t *ovf[2] = {nullptr, nullptr};
    for (uint32_t h = 0; h < halves; ++h)
        if ((rc = ensure_overflow(c, ovf_threads, c->scene.stack_cull != 0, h, &ovf[h]))) return rc;
    hipStream_t hs[2] = {s, s};
    if (split) {
        if (!c->side_stream) {
            HIP_TRY(c, hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking));
            HIP_TRY(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
            HIP_TRY(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        }
        hs[1] = c->side_stream;
    }

    // Pinhole camera (zero defocus, no -0 origin component) without next-event
    // estimation: the first k_shade recomputes each path's camera ray instead
    // of reading it, and for cache-resident sphere-free scenes k_extend does
    // too, so k_generate is skipped (CamGen).
    bool pinhole = true;
    for (int k = 0; k < 3; ++k)
        pinhole = pinhole && cam->defocus_u[k] == 0.0f && cam->defocus_v[k] == 0.0f &&
                  !(cam->origin[k] == 0.0f && std::signbit(cam->origin[k]));
    const bool shade_gen = pinhole && !nee;
    const bool ext_gen = shade_gen && c->scene.stack_cull == 0 && c->scene.sroot == kNone;
    const uint32_t last_frame = first_frame + spp - 1u;
    for (uint32_t bi = 0; bi < n_batches; ++bi) {
        const uint32_t f0 = first_frame + bi * F;
        const uint32_t nf = std::min<uint32_t>(F, spp - bi * F);
        const bool by_pixels = split && nf < 4 && wh >= 2u;  // both halves hold pixels
        const bool two = split && (nf >= 4 || by_pixels);
        const uint32_t nh = two ? 2u : 1u;
        // Half h: frames [fo[h], fo[h] + fn[h]) of pixels [po[h], po[h] + pn[h]);
        // its queue slots start at base[h] (wh * fo[h] for frame halves,
        // nf * po[h] for pixel halves).
        uint32_t fo[2] = {0u, 0u}, fn[2] = {nf, nf};
        uint32_t po[2] = {0u, 0u}, pn[2] = {wh, wh};
        size_t base[2] = {0u, 0u};
        if (by_pixels) {
            pn[0] = wh / 2u;
            po[1] = pn[0];
            pn[1] = wh - pn[0];
            base[1] = (size_t)nf * pn[0];
        } else if (two) {
            fn[0] = nf / 2u;
            fo[1] = fn[0];
            fn[1] = nf - fn[0];
            base[1] = (size_t)wh * fo[1];
        }
        const bool has_last = last_frame >= f0 && last_frame < f0 + nf;
        if (two) {
            HIP_TRY(c, hipEventRecord(c->ev_fork, s));
            HIP_TRY(c, hipStreamWaitEvent(hs[1], c->ev_fork, 0));
        }
        uint32_t *cnt[2], *wext[2], *wsh[2], *scnt[2], *wocc[2];
        CamGen cg[2] = {};
        for (uint32_t h = 0; h < nh; ++h) {
            const uint32_t np = pn[h] * fn[h];
            cnt[h] = counters + ((size_t)bi * halves + h) * per_batch;  // [max_depth+1] path counts
            wext[h] = cnt[h] + (max_depth + 1u);
            wsh[h] = wext[h] + (max_depth + 1u);
            scnt[h] = wsh[h] + (max_depth + 1u);  // [max_depth] shadow rays per bounce
            wocc[h] = scnt[h] + max_depth;
            HIP_TRY(c, launch_fill_u32(cnt[h], 1u, np, hs[h]));
            GenArgs g{};
            g.cam = *cam;
            g.width = width;
            g.height = height;
            g.wh = wh;
            g.frame0 = f0 + fo[h];
            g.n_paths = np;
            g.fo = fo[h];
            g.po = po[h];
            g.pn = pn[h];
            g.A = (float4 *)c->A[0].p + base[h];
            g.B = (float4 *)c->B[0].p + base[h];
            g.C = (float4 *)c->C[0].p + base[h];
            g.samples = nee ? (float4 *)c->samples.p : nullptr;
            g.P = nee ? (float *)c->P[0].p + base[h] : nullptr;
            if (shade_gen) {
                cg[h].cam = *cam;
                cg[h].width = width;
                cg[h].height = height;
                cg[h].wh = wh;
                cg[h].frame0 = g.frame0;
                cg[h].nf = fn[h];
                cg[h].fo = fo[h];
                cg[h].po = po[h];
                cg[h].on = 1u;
                if (ext_gen) continue;
                g.C = nullptr;
            }
            Timed t(c, "generate", hs[h]);
            HIP_TRY(c, launch_generate(g, hs[h]));
        }
        for (uint32_t b = 0; b < max_depth; ++b) {
            const int cur = b & 1, nxt = cur ^ 1;
            for (uint32_t h = 0; h < nh; ++h) {
                const size_t bs = base[h];
                ExtArgs e{};
                e.scene = c->scene;
                if (b == 0 && ext_gen) e.gen = cg[h];
                e.A = (const float4 *)c->A[cur].p + bs;
                e.B = (const float4 *)c->B[cur].p + bs;
                e.H = (float4 *)c->H.p + bs;
                e.count = cnt[h] + b;
                e.work = wext[h] + b;
                e.ovf = ovf[h];
                e.ovf_stride = g_ext * (uint32_t)trace_block();
                if (nee && b > 0) {  // the previous bounce's shadow rays ride along
                    e.S_o = (const float4 *)c->S_o.p + bs;
                    e.S_d = (const float4 *)c->S_d.p + bs;
                    e.S_c = (const float4 *)c->S_c.p + bs;
                    e.occ_count = scnt[h] + (b - 1);
                    e.samples = (float4 *)c->samples.p;
                }
                Timed t(c, "extend", hs[h]);
                if (b == 0 && c->scene.sroot == kNone && !e.occ_count)
                    HIP_TRY(c, launch_extend_packet(e, g_ext, hs[h]));
                else
                    HIP_TRY(c, launch_extend(e, g_ext, hs[h]));
            }
            for (uint32_t h = 0; h < nh; ++h) {
                const uint32_t np = pn[h] * fn[h];
                const size_t bs = base[h];
                ShadeArgs sh{};
                sh.scene = c->scene;
                if (b == 0) sh.gen = cg[h];
                sh.A = (const float4 *)c->A[cur].p + bs;
                sh.B = (const float4 *)c->B[cur].p + bs;
                sh.C = (const float4 *)c->C[cur].p + bs;
                sh.H = (const float4 *)c->H.p + bs;
                sh.A_out = (float4 *)c->A[nxt].p + bs;
                sh.B_out = (float4 *)c->B[nxt].p + bs;
                sh.C_out = (float4 *)c->C[nxt].p + bs;
                sh.count_in = cnt[h] + b;
                sh.count_out = cnt[h] + b + 1;
                sh.work = wsh[h] + b;
                sh.samples = (float4 *)c->samples.p;
                sh.gbuf = (has_last && b == 0) ? (float4 *)out_gbuf_dev : nullptr;
                sh.albedo = (has_last && b == 0) ? (float4 *)out_albedo_dev : nullptr;
                sh.gbuf_frame_index = last_frame - f0;
                sh.wh = wh;
                sh.max_depth = max_depth;
                if (nee) {
                    sh.nee = 1u;
                    sh.P = (const float *)c->P[cur].p + bs;
                    sh.P_out = (float *)c->P[nxt].p + bs;
                    sh.S_o = (float4 *)c->S_o.p + bs;
                    sh.S_d = (float4 *)c->S_d.p + bs;
                    sh.S_c = (float4 *)c->S_c.p + bs;
                    sh.shadow_count = scnt[h] + b;
                }
                {
                    Timed t(c, "shade", hs[h]);
                    HIP_TRY(c, launch_shade(sh, (np + kShadeBlock - 1u) / kShadeBlock, hs[h]));
                }
                if (nee && b + 1 == max_depth) {  // the last bounce's shadow rays
                    OccArgs oa{};
                    oa.scene = c->scene;
                    oa.S_o = (const float4 *)c->S_o.p + bs;
                    oa.S_d = (const float4 *)c->S_d.p + bs;
                    oa.S_c = (const float4 *)c->S_c.p + bs;
                    oa.count = scnt[h] + b;
                    oa.work = wocc[h] + b;
                    oa.samples = (float4 *)c->samples.p;
                    oa.ovf = ovf[h];
                    oa.ovf_stride = g_occ * (uint32_t)trace_block();
                    Timed t(c, "occlude", hs[h]);
                    HIP_TRY(c, launch_occlude(oa, g_occ, hs[h]));
                }
            }
        }
        if (two) {
            HIP_TRY(c, hipEventRecord(c->ev_join, hs[1]));
            HIP_TRY(c, hipStreamWaitEvent(s, c->ev_join, 0));
        }
        AccArgs acc{};
        acc.samples = (const float4 *)c->samples.p;
        acc.accum_in = bi == 0 ? (const float4 *)accum_in_dev : (const float4 *)out_accum_dev;
        acc.accum_out = (float4 *)out_accum_dev;
        acc.wh = wh;
        acc.n_frames = nf;
        acc.frame0 = f0;
        {
            Timed t(c, "accumulate", s);
            HIP_TRY(c, launch_accumulate(acc, s));
        }
    }
    c->last_halves = halves;
    c->last_batches = n_batches;
    c->last_depth = max_depth;
    c->last_samples = (uint64_t)wh * spp;
    return MCRT_OK;
}

extern "C" int mcrt_render(mcrt_ctx *c, const CameraUniform *cam, uint32_t width, uint32_t height,
                           uint32_t spp, uint32_t first_frame, uint32_t max_depth,
                           const float *accum_in, float *out_accum, float *out_gbuf) {
    int rc = bind(c);
    if (rc) return rc;
    if (!out_accum) return fail(c, MCRT_ERR_BAD_ARG, "out_accum is required");
    if (width == 0 || height == 0) return fail(c, MCRT_ERR_BAD_ARG, "width and height must be positive");
    size_t bytes = (size_t)width * height * 16;
    if ((rc = ensure(c, c->host_accum, bytes))) return rc;
    if ((rc = ensure(c, c->host_gbuf, bytes))) return rc;
    hipStream_t s = c->own_stream;
    const float *din = nullptr;
    if (accum_in) {
        if ((rc = ensure(c, c->host_accum_in, bytes))) return rc;
        HIP_TRY(c, hipMemcpyAsync(c->host_accum_in.p, accum_in, bytes, hipMemcpyHostToDevice, s));
        din = (const float *)c->host_accum_in.p;
    }
    rc = mcrt_render_dev(c, cam, width, height, spp, first_frame, max_depth, din,
                         (float *)c->host_accum.p, out_gbuf ? (float *)c->host_gbuf.p : nullptr, s);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(out_accum, c->host_accum.p, bytes, hipMemcpyDeviceToHost, s));
    if (out_gbuf)
        HIP_TRY(c, hipMemcpyAsync(out_gbuf, c->host_gbuf.p, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return MCRT_OK;
}

extern "C" int mcrt_render_counters(mcrt_ctx *c, uint64_t *segments, uint64_t *samples) {
    int rc = bind(c);
    if (rc) return rc;
    if (c->last_batches == 0) return fail(c, MCRT_ERR_NO_SCENE, "no render has run on this context");
    HIP_TRY(c, hipStreamSynchronize(c->last_stream));
    const uint32_t per_batch = counters_per_batch(c->last_depth);
    const uint32_t blocks = c->last_batches * c->last_halves;  // counter blocks, one per batch half
    std::vector<uint32_t> h((size_t)per_batch * blocks);
    HIP_TRY(c, hipMemcpy(h.data(), c->counters.p, h.size() * 4, hipMemcpyDeviceToHost));
    uint64_t seg = 0;
    for (uint32_t bi = 0; bi < blocks; ++bi)
        for (uint32_t b = 0; b < c->last_depth; ++b) seg += h[(size_t)bi * per_batch + b];
    if (segments) *segments = seg;
    if (samples) *samples = c->last_samples;
    return MCRT_OK;
}

extern "C" int mcrt_trace_closest_dev(mcrt_ctx *c, const float *origins_dev, const float *dirs_dev,
                                      uint32_t n, float tmin, float tmax, uint32_t *out_id_dev,
                                      float *out_t_dev, void *stream) {
    int rc = bind(c);
    if (rc) return rc;
    if (!c->has_scene) return fail(c, MCRT_ERR_NO_SCENE, "trace called before mcrt_upload_scene");
    if (n == 0) return MCRT_OK;
    if (!origins_dev || !dirs_dev || !out_id_dev || !out_t_dev)
        return fail(c, MCRT_ERR_BAD_ARG, "null ray or output pointer");
    if (n > 0xFFFFFFFFu - 64u) return fail(c, MCRT_ERR_LIMIT, "too many rays");
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    if ((rc = ensure(c, c->trace_work, 4))) return rc;
    HIP_TRY(c, launch_fill_u32((uint32_t *)c->trace_work.p, 1u, 0u, s));
    TraceArgs a{};
    a.scene = c->scene;
    a.origins = origins_dev;
    a.dirs = dirs_dev;
    a.n = n;
    a.tmin = tmin;
    a.tmax = tmax;
    a.out_id = out_id_dev;
    a.out_t = out_t_dev;
    a.work = (uint32_t *)c->trace_work.p;
    uint32_t grid = 0;
    grid_for(c, 2, &grid);
    uint32_t need = (n + 63u) / 64u;  // waves of work
    uint32_t waves_per_block = (uint32_t)trace_block() / 64u;
    grid = std::max<uint32_t>(1, std::min(grid, (need + waves_per_block - 1) / waves_per_block));
    if ((rc = ensure_overflow(c, grid * (uint32_t)trace_block(), c->scene.stack_cull != 0, 0, &a.ovf))) return rc;
    a.ovf_stride = grid * (uint32_t)trace_block();
    {
        Timed t(c, "trace_closest", s);
        HIP_TRY(c, launch_trace(a, grid, s));
    }
    return MCRT_OK;
}

extern "C" int mcrt_trace_closest_noalloc_dev(mcrt_ctx *c, const float *origins_dev,
                                              const float *dirs_dev, uint32_t n, float tmin, float tmax,
                                              uint32_t *out_id_dev, float *out_t_dev, void *stream) {
    int rc = bind(c);
    if (rc) return rc;
    if (!c->has_scene) return fail(c, MCRT_ERR_NO_SCENE, "trace called before mcrt_upload_scene");
    if (n == 0) return MCRT_OK;
    if (!origins_dev || !dirs_dev || !out_id_dev || !out_t_dev)
        return fail(c, MCRT_ERR_BAD_ARG, "null ray or output pointer");
    if (n > 0xFFFFFFFFu - 64u) return fail(c, MCRT_ERR_LIMIT, "too many rays");
    hipError_t se = scratch_setup(c->device);
    if (se != hipSuccess)
        return fail(c, MCRT_ERR_HIP, "no-allocation scratch unavailable on device %d: %s", c->device,
                    hipGetErrorString(se));
    uint8_t *base = (uint8_t *)g_scratch[c->device];
    // The persistent grid of mcrt_trace_closest_dev, shrunk until every
    // thread's overflow column fits the scratch.
    uint32_t grid = 0;
    grid_for(c, 2, &grid);
    const uint32_t block = (uint32_t)trace_block();
    const uint32_t need = (n + 63u) / 64u;
    grid = std::max<uint32_t>(1, std::min(grid, (need + block / 64u - 1u) / (block / 64u)));
    uint32_t sh = 0, words = 0;
    stack_layout(c->scene.stack_cull != 0, &sh, &words);
    const uint32_t depth = c->scene.stack_depth;
    const size_t per_thread = depth > sh ? (size_t)(depth - sh) * words * sizeof(uint32_t) : 0;
    if (per_thread) {
        const size_t fit = (kNoAllocScratchBytes - kNoAllocCounterBytes) / (per_thread * block);
        if (fit == 0)
            return fail(c, MCRT_ERR_LIMIT,
                        "BVH depth %u needs %zu B of traversal-stack scratch per block of %u threads; the "
                        "no-allocation trace holds %zu B",
                        depth, per_thread * block, block, kNoAllocScratchBytes - kNoAllocCounterBytes);
        grid = (uint32_t)std::min<size_t>(grid, fit);
    }
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    TraceArgs a{};
    a.scene = c->scene;
    a.origins = origins_dev;
    a.dirs = dirs_dev;
    a.n = n;
    a.tmin = tmin;
    a.tmax = tmax;
    a.out_id = out_id_dev;
    a.out_t = out_t_dev;
    a.work = (uint32_t *)base;
    a.ovf = per_thread ? (uint32_t *)(base + kNoAllocCounterBytes) : nullptr;
    a.ovf_stride = grid * block;
    HIP_TRY(c, launch_fill_u32(a.work, 1u, 0u, s));  // a fill kernel, not a memset
    HIP_TRY(c, launch_trace(a, grid, s));
    return MCRT_OK;
}

extern "C" int mcrt_trace_closest(mcrt_ctx *c, const float *origins, const float *dirs, uint32_t n,
                                  float tmin, float tmax, uint32_t *out_id, float *out_t) {
    int rc = bind(c);
    if (rc) return rc;
    if (n == 0) return MCRT_OK;
    if (!origins || !dirs || !out_id || !out_t)
        return fail(c, MCRT_ERR_BAD_ARG, "null ray or output pointer");
    size_t b3 = (size_t)n * 12, b1 = (size_t)n * 4;
    if ((rc = ensure(c, c->trace_o, b3)) || (rc = ensure(c, c->trace_d, b3)) ||
        (rc = ensure(c, c->trace_id, b1)) || (rc = ensure(c, c->trace_t, b1)))
        return rc;
    hipStream_t s = c->own_stream;
    HIP_TRY(c, hipMemcpyAsync(c->trace_o.p, origins, b3, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->trace_d.p, dirs, b3, hipMemcpyHostToDevice, s));
    rc = mcrt_trace_closest_dev(c, (const float *)c->trace_o.p, (const float *)c->trace_d.p, n, tmin,
                                tmax, (uint32_t *)c->trace_id.p, (float *)c->trace_t.p, s);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(out_id, c->trace_id.p, b1, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(out_t, c->trace_t.p, b1, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return MCRT_OK;
}

extern "C" int mcrt_denoise_bilateral_dev(mcrt_ctx *c, const DenoiseParams *params, uint32_t width,
                                          uint32_t height, const float *accum_dev,
                                          const float *gbuf_dev, float *out_dev, void *stream) {
    int rc = bind(c);
    if (rc) return rc;
    DenoiseParams p = params ? *params : DenoiseParams{3, 2.0f, 0.5f, 8.0f};
    if (width == 0 || height == 0) return MCRT_OK;
    if (!accum_dev || !gbuf_dev || !out_dev) return fail(c, MCRT_ERR_BAD_ARG, "null image pointer");
    if (p.radius > kMaxBilateralRadius)
        return fail(c, MCRT_ERR_LIMIT,
                    "denoise radius %d exceeds %d (a (2R+1)^2 = %lld-tap footprint per pixel)", p.radius,
                    kMaxBilateralRadius, (long long)(2 * kMaxBilateralRadius + 1) * (2 * kMaxBilateralRadius + 1));
    if (width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return fail(c, MCRT_ERR_LIMIT, "image too large");
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    // inv2ss / inv2ds as denoise.wgsl:64-65; the spatial weights are formed
    // in the kernel, so nothing is uploaded and the call is stream-ordered.
    BilateralArgs a{};
    a.accum = (const float4 *)accum_dev;
    a.gbuf = (const float4 *)gbuf_dev;
    a.out = (float4 *)out_dev;
    a.width = (int32_t)width;
    a.height = (int32_t)height;
    a.radius = p.radius;
    a.inv2ss = 1.0f / (2.0f * p.sigma_s * p.sigma_s);
    a.inv2ds = 1.0f / (2.0f * p.sigma_d * p.sigma_d);
    a.normal_pow = p.normal_pow;
    {
        Timed t(c, "bilateral", s);
        HIP_TRY(c, launch_bilateral(a, s));
    }
    return MCRT_OK;
}

extern "C" int mcrt_denoise_bilateral(mcrt_ctx *c, const DenoiseParams *params, uint32_t width,
                                      uint32_t height, const float *accum, const float *gbuf,
                                      float *out) {
    int rc = bind(c);
    if (rc) return rc;
    if (width == 0 || height == 0) return MCRT_OK;
    if (!accum || !gbuf || !out) return fail(c, MCRT_ERR_BAD_ARG, "null image pointer");
    size_t bytes = (size_t)width * height * 16;
    if ((rc = ensure(c, c->host_accum, bytes)) || (rc = ensure(c, c->host_gbuf, bytes)) ||
        (rc = ensure(c, c->host_out, bytes)))
        return rc;
    hipStream_t s = c->own_stream;
    HIP_TRY(c, hipMemcpyAsync(c->host_accum.p, accum, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->host_gbuf.p, gbuf, bytes, hipMemcpyHostToDevice, s));
    rc = mcrt_denoise_bilateral_dev(c, params, width, height, (const float *)c->host_accum.p,
                                    (const float *)c->host_gbuf.p, (float *)c->host_out.p, s);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(out, c->host_out.p, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return MCRT_OK;
}

// ---------------------------------------------------------------------------
// Differentiable render and path-replay adjoint
// ---------------------------------------------------------------------------
namespace {

int diff_setup(mcrt_ctx *c, const CameraUniform *cam, uint32_t width, uint32_t height, uint32_t spp,
               uint32_t first_frame, uint32_t max_depth, const float *params, uint32_t n_params,
               const int32_t *param_of_material, DiffArgs &a) {
    if (!c->has_scene) return fail(c, MCRT_ERR_NO_SCENE, "diff render called before mcrt_upload_scene");
    if (!cam || !params || !param_of_material) return fail(c, MCRT_ERR_BAD_ARG, "null argument");
    if (width == 0 || height == 0 || spp == 0) return fail(c, MCRT_ERR_BAD_ARG, "empty render");
    if (max_depth < 1 || max_depth > 64) return fail(c, MCRT_ERR_BAD_ARG, "max_depth must be in [1, 64]");
    if (n_params > (uint32_t)kMaxParams)
        return fail(c, MCRT_ERR_LIMIT, "at most %d parameter materials (got %u)", kMaxParams, n_params);
    if ((uint64_t)width * height > (1ull << 26)) return fail(c, MCRT_ERR_LIMIT, "image too large");
    a = DiffArgs{};
    a.scene = c->scene;
    a.cam = *cam;
    a.width = width;
    a.height = height;
    a.spp = spp;
    a.frame0 = first_frame;
    a.max_depth = max_depth;
    a.n_params = n_params;
    for (uint32_t p = 0; p < n_params; ++p) {
        for (int k = 0; k < 4; ++k)
            if (!std::isfinite(params[4 * p + k])) return fail(c, MCRT_ERR_BAD_ARG, "non-finite parameter");
        a.params[p] = make_float4(params[4 * p], params[4 * p + 1], params[4 * p + 2], params[4 * p + 3]);
    }
    for (uint32_t m = 0; m < MCRT_MAX_MATERIALS; ++m) {
        int32_t q = param_of_material[m];
        if (q < -1 || q >= (int32_t)n_params)
            return fail(c, MCRT_ERR_BAD_ARG, "param_of_material[%u] = %d out of range", m, q);
        a.pmap[m] = (int8_t)q;
    }
    a.groups = diff_groups(width * height, spp);
    const uint32_t grid = diff_grid(width * height, a.groups);
    a.ovf_stride = grid * 256u;
    int rc = ensure_overflow(c, a.ovf_stride, false, 0, &a.ovf);
    if (rc) return rc;
    return MCRT_OK;
}

}  // namespace

extern "C" int mcrt_diff_render_dev(mcrt_ctx *c, const CameraUniform *cam, uint32_t width,
                                    uint32_t height, uint32_t spp, uint32_t first_frame,
                                    uint32_t max_depth, const float *params, uint32_t n_params,
                                    const int32_t *param_of_material, float *out_image_dev,
                                    void *stream) {
    int rc = bind(c);
    if (rc) return rc;
    DiffArgs a;
    if ((rc = diff_setup(c, cam, width, height, spp, first_frame, max_depth, params, n_params,
                         param_of_material, a)))
        return rc;
    if (!out_image_dev) return fail(c, MCRT_ERR_BAD_ARG, "null image pointer");
    a.image = (float4 *)out_image_dev;
    if (a.groups > 1u) {
        if ((rc = ensure(c, c->diff_image, (size_t)a.groups * width * height * sizeof(float4)))) return rc;
        a.image_part = (float4 *)c->diff_image.p;
    }
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    Timed t(c, "diff_primal", s);
    HIP_TRY(c, launch_diff(a, false, s));
    return MCRT_OK;
}

extern "C" int mcrt_diff_adjoint_dev(mcrt_ctx *c, const CameraUniform *cam, uint32_t width,
                                     uint32_t height, uint32_t spp, uint32_t first_frame,
                                     uint32_t max_depth, const float *params, uint32_t n_params,
                                     const int32_t *param_of_material, const float *d_image_dev,
                                     float *out_grad_dev, void *stream) {
    int rc = bind(c);
    if (rc) return rc;
    DiffArgs a;
    if ((rc = diff_setup(c, cam, width, height, spp, first_frame, max_depth, params, n_params,
                         param_of_material, a)))
        return rc;
    if (!d_image_dev || !out_grad_dev) return fail(c, MCRT_ERR_BAD_ARG, "null pointer");
    const uint32_t grid = diff_grid(width * height, a.groups);
    if ((rc = ensure(c, c->diff_partial, (size_t)grid * kMaxParams * 4 * sizeof(float)))) return rc;
    if (c->diff_mode == MCRT_DIFF_TAPE) {
        a.tape_stride = grid * 256u;
        if ((rc = ensure(c, c->diff_tape, (size_t)a.tape_stride * max_depth * 3 * sizeof(float4)))) return rc;
        a.tape = (float4 *)c->diff_tape.p;
    } else {
        a.tape = nullptr;  // replay: no tape
    }
    a.d_image = (const float4 *)d_image_dev;
    a.partial = (float *)c->diff_partial.p;
    a.grad = out_grad_dev;
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    Timed t(c, "diff_adjoint", s);
    HIP_TRY(c, launch_diff(a, true, s));
    return MCRT_OK;
}

extern "C" int mcrt_diff_set_mode(mcrt_ctx *c, int mode) {
    if (!c) return MCRT_ERR_BAD_ARG;
    if (mode != MCRT_DIFF_REPLAY && mode != MCRT_DIFF_TAPE)
        return fail(c, MCRT_ERR_BAD_ARG, "unknown adjoint mode %d", mode);
    c->diff_mode = mode;
    if (mode == MCRT_DIFF_REPLAY) release(c->diff_tape);
    return MCRT_OK;
}

extern "C" int mcrt_diff_render(mcrt_ctx *c, const CameraUniform *cam, uint32_t width, uint32_t height,
                                uint32_t spp, uint32_t first_frame, uint32_t max_depth,
                                const float *params, uint32_t n_params,
                                const int32_t *param_of_material, float *out_image) {
    int rc = bind(c);
    if (rc) return rc;
    if (!out_image || width == 0 || height == 0) return fail(c, MCRT_ERR_BAD_ARG, "bad image");
    size_t bytes = (size_t)width * height * 16;
    if ((rc = ensure(c, c->host_out, bytes))) return rc;
    hipStream_t s = c->own_stream;
    rc = mcrt_diff_render_dev(c, cam, width, height, spp, first_frame, max_depth, params, n_params,
                              param_of_material, (float *)c->host_out.p, s);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(out_image, c->host_out.p, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return MCRT_OK;
}

extern "C" int mcrt_diff_adjoint(mcrt_ctx *c, const CameraUniform *cam, uint32_t width,
                                 uint32_t height, uint32_t spp, uint32_t first_frame,
                                 uint32_t max_depth, const float *params, uint32_t n_params,
                                 const int32_t *param_of_material, const float *d_image,
                                 float *out_grad) {
    int rc = bind(c);
    if (rc) return rc;
    if (!d_image || !out_grad || width == 0 || height == 0) return fail(c, MCRT_ERR_BAD_ARG, "bad args");
    size_t bytes = (size_t)width * height * 16;
    if ((rc = ensure(c, c->host_accum_in, bytes)) || (rc = ensure(c, c->diff_grad, 4 * kMaxParams * 4)))
        return rc;
    hipStream_t s = c->own_stream;
    HIP_TRY(c, hipMemcpyAsync(c->host_accum_in.p, d_image, bytes, hipMemcpyHostToDevice, s));
    rc = mcrt_diff_adjoint_dev(c, cam, width, height, spp, first_frame, max_depth, params, n_params,
                               param_of_material, (const float *)c->host_accum_in.p,
                               (float *)c->diff_grad.p, s);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(out_grad, c->diff_grad.p, (size_t)n_params * 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return MCRT_OK;
}

// ---------------------------------------------------------------------------
// Learned denoiser
// ---------------------------------------------------------------------------
namespace {
constexpr int kCin[kDenoiseLayers] = {10, 32, 32, 32, 32, 32};
constexpr int kCout[kDenoiseLayers] = {32, 32, 32, 32, 32, 3};
constexpr uint32_t kWgradBlocks = 512;
}  // namespace

struct mcrt_denoiser {
    mcrt_ctx *c = nullptr;
    DevBuf w[kDenoiseLayers], m[kDenoiseLayers], v[kDenoiseLayers], g[kDenoiseLayers];
    DevBuf wf[kDenoiseLayers], wb[kDenoiseLayers];
    DevBuf act[kDenoiseLayers], dA, dB, out_log, partial, loss_partial, loss;
    size_t cap_px = 0;
    uint32_t step = 0;
    int loss_kind = MCRT_DENOISE_LOSS_LOG_L2;
};

namespace {

uint32_t public_param_count() {
    uint32_t n = 0;
    for (int l = 0; l < kDenoiseLayers; ++l) n += kCout[l] * kCin[l] * 9 + kCout[l];
    return n;
}

int dn_workspace(mcrt_denoiser *d, size_t px) {
    mcrt_ctx *c = d->c;
    if (px <= d->cap_px) return MCRT_OK;
    int rc;
    const size_t act_bytes = px * 32 * sizeof(uint16_t);
    for (int l = 0; l < kDenoiseLayers; ++l)
        if ((rc = ensure(c, d->act[l], act_bytes))) return rc;
    if ((rc = ensure(c, d->dA, act_bytes)) || (rc = ensure(c, d->dB, act_bytes)) ||
        (rc = ensure(c, d->out_log, px * 16)) ||
        (rc = ensure(c, d->loss_partial, ((px + 255) / 256) * sizeof(float))))
        return rc;
    d->cap_px = px;
    return MCRT_OK;
}

int dn_check(mcrt_denoiser *d, uint32_t batch, uint32_t width, uint32_t height) {
    if (!d) return MCRT_ERR_BAD_ARG;
    int rc = bind(d->c);
    if (rc) return rc;
    if (batch == 0 || width == 0 || height == 0) return fail(d->c, MCRT_ERR_BAD_ARG, "empty batch");
    const uint64_t px = (uint64_t)batch * width * height;
    if (px > (1ull << 27) || height > 65535 * 16 || batch > 65535)
        return fail(d->c, MCRT_ERR_LIMIT, "denoiser batch of %llu pixels is too large",
                    (unsigned long long)px);
    return dn_workspace(d, (size_t)px);
}

// Forward pass into d->out_log; activations kept for training.
int dn_forward(mcrt_denoiser *d, int B, int W, int H, const float *noisy, const float *albedo,
               const float *gbuf, hipStream_t s) {
    mcrt_ctx *c = d->c;
    PackArgs pa{(const float4 *)noisy, (const float4 *)albedo, (const float4 *)gbuf,
                (__bf16 *)d->act[0].p, B, H, W};
    HIP_TRY(c, launch_pack(pa, s));
    for (int l = 0; l < kDenoiseLayers; ++l) {
        ConvArgs ca{};
        ca.in = (const __bf16 *)d->act[l].p;
        ca.w = (const __bf16 *)d->wf[l].p;
        ca.bias = (const float *)d->w[l].p + 9 * 32 * 32;
        ca.B = B;
        ca.H = H;
        ca.W = W;
        if (l + 1 < kDenoiseLayers) {
            ca.out = (__bf16 *)d->act[l + 1].p;
            ca.mode = 0;
        } else {
            ca.out_log = (float4 *)d->out_log.p;
            ca.noisy = (const float4 *)noisy;
            ca.mode = 2;
        }
        Timed t(c, "denoise_conv", s);
        HIP_TRY(c, launch_conv3x3(ca, s));
    }
    return MCRT_OK;
}

}  // namespace

extern "C" int mcrt_denoiser_create(mcrt_ctx *c, uint32_t seed, mcrt_denoiser **out) {
    if (!out) return MCRT_ERR_BAD_ARG;
    *out = nullptr;
    int rc = bind(c);
    if (rc) return rc;
    mcrt_denoiser *d = new mcrt_denoiser();
    d->c = c;
    std::mt19937 gen(seed);
    for (int l = 0; l < kDenoiseLayers; ++l) {
        std::vector<float> host(kLayerParams, 0.0f);
        // He initialisation for the ReLU layers; the last layer starts small
        // so the network begins close to the identity (the global skip).
        const float std = l + 1 < kDenoiseLayers ? std::sqrt(2.0f / (9.0f * kCin[l])) : 1e-3f;
        std::normal_distribution<float> nd(0.0f, std);
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < kCout[l]; ++co)
                for (int ci = 0; ci < kCin[l]; ++ci) host[tap * 1024 + co * 32 + ci] = nd(gen);
        std::vector<float> zeros(kLayerParams, 0.0f);
        if ((rc = upload(c, d->w[l], host.data(), kLayerParams * 4)) ||
            (rc = upload(c, d->m[l], zeros.data(), kLayerParams * 4)) ||
            (rc = upload(c, d->v[l], zeros.data(), kLayerParams * 4)) ||
            (rc = ensure(c, d->g[l], kLayerParams * 4)) || (rc = ensure(c, d->wf[l], 9 * 1024 * 2)) ||
            (rc = ensure(c, d->wb[l], 9 * 1024 * 2))) {
            mcrt_denoiser_destroy(d);
            return rc;
        }
        if (launch_weights_to_bf16((const float *)d->w[l].p, d->wf[l].p, d->wb[l].p, nullptr) != hipSuccess) {
            mcrt_denoiser_destroy(d);
            return fail(c, MCRT_ERR_HIP, "weight conversion launch failed");
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        mcrt_denoiser_destroy(d);
        return fail(c, MCRT_ERR_HIP, "denoiser initialisation failed");
    }
    *out = d;
    return MCRT_OK;
}

extern "C" void mcrt_denoiser_destroy(mcrt_denoiser *d) {
    if (!d) return;
    (void)hipSetDevice(d->c->device);
    (void)hipDeviceSynchronize();
    for (int l = 0; l < kDenoiseLayers; ++l)
        for (DevBuf *b : {&d->w[l], &d->m[l], &d->v[l], &d->g[l], &d->wf[l], &d->wb[l], &d->act[l]})
            release(*b);
    for (DevBuf *b : {&d->dA, &d->dB, &d->out_log, &d->partial, &d->loss_partial, &d->loss}) release(*b);
    delete d;
}

extern "C" uint32_t mcrt_denoiser_param_count(void) { return public_param_count(); }

extern "C" int mcrt_denoiser_get_weights(mcrt_denoiser *d, float *out) {
    if (!d || !out) return MCRT_ERR_BAD_ARG;
    int rc = bind(d->c);
    if (rc) return rc;
    HIP_TRY(d->c, hipDeviceSynchronize());
    std::vector<float> host(kLayerParams);
    size_t k = 0;
    for (int l = 0; l < kDenoiseLayers; ++l) {
        HIP_TRY(d->c, hipMemcpy(host.data(), d->w[l].p, kLayerParams * 4, hipMemcpyDeviceToHost));
        for (int co = 0; co < kCout[l]; ++co)
            for (int ci = 0; ci < kCin[l]; ++ci)
                for (int tap = 0; tap < 9; ++tap) out[k++] = host[tap * 1024 + co * 32 + ci];
        for (int co = 0; co < kCout[l]; ++co) out[k++] = host[9 * 1024 + co];
    }
    return MCRT_OK;
}

extern "C" int mcrt_denoiser_set_loss(mcrt_denoiser *d, int kind) {
    if (!d) return MCRT_ERR_BAD_ARG;
    if (kind != MCRT_DENOISE_LOSS_LOG_L2 && kind != MCRT_DENOISE_LOSS_REL_MSE)
        return fail(d->c, MCRT_ERR_BAD_ARG, "unknown denoiser loss kind");
    d->loss_kind = kind;
    return MCRT_OK;
}

extern "C" int mcrt_denoiser_set_weights(mcrt_denoiser *d, const float *in) {
    if (!d || !in) return MCRT_ERR_BAD_ARG;
    int rc = bind(d->c);
    if (rc) return rc;
    HIP_TRY(d->c, hipDeviceSynchronize());
    size_t k = 0;
    for (int l = 0; l < kDenoiseLayers; ++l) {
        std::vector<float> host(kLayerParams, 0.0f), zeros(kLayerParams, 0.0f);
        for (int co = 0; co < kCout[l]; ++co)
            for (int ci = 0; ci < kCin[l]; ++ci)
                for (int tap = 0; tap < 9; ++tap) host[tap * 1024 + co * 32 + ci] = in[k++];
        for (int co = 0; co < kCout[l]; ++co) host[9 * 1024 + co] = in[k++];
        for (size_t i = 0; i < host.size(); ++i)
            if (!std::isfinite(host[i])) return fail(d->c, MCRT_ERR_BAD_ARG, "non-finite weight");
        if ((rc = upload(d->c, d->w[l], host.data(), kLayerParams * 4)) ||
            (rc = upload(d->c, d->m[l], zeros.data(), kLayerParams * 4)) ||
            (rc = upload(d->c, d->v[l], zeros.data(), kLayerParams * 4)))
            return rc;
        HIP_TRY(d->c, launch_weights_to_bf16((const float *)d->w[l].p, d->wf[l].p, d->wb[l].p, nullptr));
    }
    d->step = 0;
    HIP_TRY(d->c, hipDeviceSynchronize());
    return MCRT_OK;
}

extern "C" int mcrt_denoiser_infer_dev(mcrt_denoiser *d, uint32_t batch, uint32_t width,
                                       uint32_t height, const float *noisy, const float *albedo,
                                       const float *gbuf, float *out, void *stream) {
    int rc = dn_check(d, batch, width, height);
    if (rc) return rc;
    if (!noisy || !albedo || !gbuf || !out) return fail(d->c, MCRT_ERR_BAD_ARG, "null image pointer");
    hipStream_t s = (hipStream_t)stream;
    d->c->last_stream = s;
    // The training forward pass without its log-space output: the last
    // layer writes the denoised image, and the first reads a 16-channel
    // input (32 B per pixel instead of 64).
    PackArgs pa{(const float4 *)noisy, (const float4 *)albedo, (const float4 *)gbuf, (__bf16 *)d->act[0].p,
                (int)batch, (int)height, (int)width};
    HIP_TRY(d->c, launch_pack16(pa, s));
    for (int l = 0; l < kDenoiseLayers; ++l) {
        ConvArgs ca{};
        ca.cin16 = l == 0;
        ca.in = (const __bf16 *)d->act[l].p;
        ca.noisy = (const float4 *)noisy;
        ca.w = (const __bf16 *)d->wf[l].p;
        ca.bias = (const float *)d->w[l].p + 9 * 32 * 32;
        ca.B = (int)batch;
        ca.H = (int)height;
        ca.W = (int)width;
        if (l + 1 < kDenoiseLayers) {
            ca.out = (__bf16 *)d->act[l + 1].p;
            ca.mode = 0;
        } else {
            ca.out_final = (float4 *)out;
            ca.mode = 2;
        }
        Timed t(d->c, "denoise_conv", s);
        HIP_TRY(d->c, launch_conv3x3(ca, s));
    }
    return MCRT_OK;
}

extern "C" int mcrt_denoiser_train_step_dev(mcrt_denoiser *d, uint32_t batch, uint32_t width,
                                            uint32_t height, const float *noisy, const float *albedo,
                                            const float *gbuf, const float *ref, float lr,
                                            float *loss_out, void *stream) {
    int rc = dn_check(d, batch, width, height);
    if (rc) return rc;
    mcrt_ctx *c = d->c;
    if (!noisy || !albedo || !gbuf || !ref) return fail(c, MCRT_ERR_BAD_ARG, "null image pointer");
    if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(c, MCRT_ERR_BAD_ARG, "learning rate must be positive");
    hipStream_t s = (hipStream_t)stream;
    c->last_stream = s;
    const int B = (int)batch, W = (int)width, H = (int)height;
    if ((rc = dn_forward(d, B, W, H, noisy, albedo, gbuf, s))) return rc;
    if ((rc = ensure(c, d->partial, (size_t)kWgradBlocks * kLayerParams * 4)) ||
        (rc = ensure(c, d->loss, 4)))
        return rc;
    LossArgs la{(const float4 *)d->out_log.p, (const float4 *)ref, (__bf16 *)d->dA.p,
                (float *)d->loss_partial.p, 1.0f / (3.0f * (float)B * H * W), B, H, W,
                d->loss_kind};
    HIP_TRY(c, launch_loss(la, (float *)d->loss.p, s));
    // Backward: layer 5 (no ReLU), then 4..0 with the ReLU mask of their output.
    __bf16 *cur = (__bf16 *)d->dA.p, *nxt = (__bf16 *)d->dB.p;
    for (int l = kDenoiseLayers - 1; l >= 0; --l) {
        const __bf16 *mask = l + 1 < kDenoiseLayers ? (const __bf16 *)d->act[l + 1].p : nullptr;
        WgradArgs wa{(const __bf16 *)d->act[l].p, cur, mask, (float *)d->partial.p, B, H, W};
        {
            Timed t(c, "denoise_wgrad", s);
            HIP_TRY(c, launch_wgrad(wa, kWgradBlocks, (float *)d->g[l].p, s));
        }
        if (l > 0) {
            ConvArgs ca{};
            ca.in = cur;
            ca.mask = mask;
            ca.w = (const __bf16 *)d->wb[l].p;
            ca.out = nxt;
            ca.B = B;
            ca.H = H;
            ca.W = W;
            ca.mode = 1;
            {
                Timed t(c, "denoise_dgrad", s);
                HIP_TRY(c, launch_conv3x3(ca, s));
            }
            std::swap(cur, nxt);
        }
    }
    d->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int l = 0; l < kDenoiseLayers; ++l) {
        AdamArgs aa{(float *)d->w[l].p, (float *)d->m[l].p, (float *)d->v[l].p, (const float *)d->g[l].p,
                    (__bf16 *)d->wf[l].p, (__bf16 *)d->wb[l].p, lr, b1, b2, 1e-8f,
                    1.0f - std::pow(b1, (float)d->step), 1.0f - std::pow(b2, (float)d->step)};
        HIP_TRY(c, launch_adam(aa, s));
    }
    if (loss_out) {
        HIP_TRY(c, hipMemcpyAsync(loss_out, d->loss.p, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        *loss_out *= la.inv_count;
    }
    return MCRT_OK;
}

// Gradient of the loss w.r.t. the public weights for one batch, without an
// Adam step (for checking against an autograd reference).
extern "C" int mcrt_denoiser_gradient(mcrt_denoiser *d, uint32_t batch, uint32_t width,
                                      uint32_t height, const float *noisy, const float *albedo,
                                      const float *gbuf, const float *ref, float *out_grad,
                                      float *loss_out) {
    int rc = dn_check(d, batch, width, height);
    if (rc) return rc;
    mcrt_ctx *c = d->c;
    if (!noisy || !albedo || !gbuf || !ref || !out_grad) return fail(c, MCRT_ERR_BAD_ARG, "null pointer");
    hipStream_t s = c->own_stream;
    const int B = (int)batch, W = (int)width, H = (int)height;
    if ((rc = dn_forward(d, B, W, H, noisy, albedo, gbuf, s))) return rc;
    if ((rc = ensure(c, d->partial, (size_t)kWgradBlocks * kLayerParams * 4)) ||
        (rc = ensure(c, d->loss, 4)))
        return rc;
    LossArgs la{(const float4 *)d->out_log.p, (const float4 *)ref, (__bf16 *)d->dA.p,
                (float *)d->loss_partial.p, 1.0f / (3.0f * (float)B * H * W), B, H, W,
                d->loss_kind};
    HIP_TRY(c, launch_loss(la, (float *)d->loss.p, s));
    __bf16 *cur = (__bf16 *)d->dA.p, *nxt = (__bf16 *)d->dB.p;
    for (int l = kDenoiseLayers - 1; l >= 0; --l) {
        const __bf16 *mask = l + 1 < kDenoiseLayers ? (const __bf16 *)d->act[l + 1].p : nullptr;
        WgradArgs wa{(const __bf16 *)d->act[l].p, cur, mask, (float *)d->partial.p, B, H, W};
        HIP_TRY(c, launch_wgrad(wa, kWgradBlocks, (float *)d->g[l].p, s));
        if (l > 0) {
            ConvArgs ca{};
            ca.in = cur;
            ca.mask = mask;
            ca.w = (const __bf16 *)d->wb[l].p;
            ca.out = nxt;
            ca.B = B;
            ca.H = H;
            ca.W = W;
            ca.mode = 1;
            HIP_TRY(c, launch_conv3x3(ca, s));
            std::swap(cur, nxt);
        }
    }
    HIP_TRY(c, hipStreamSynchronize(s));
    std::vector<float> host(kLayerParams);
    size_t k = 0;
    for (int l = 0; l < kDenoiseLayers; ++l) {
        HIP_TRY(c, hipMemcpy(host.data(), d->g[l].p, kLayerParams * 4, hipMemcpyDeviceToHost));
        for (int co = 0; co < kCout[l]; ++co)
            for (int ci = 0; ci < kCin[l]; ++ci)
                for (int tap = 0; tap < 9; ++tap) out_grad[k++] = host[tap * 1024 + co * 32 + ci];
        for (int co = 0; co < kCout[l]; ++co) out_grad[k++] = host[9 * 1024 + co];
    }
    if (loss_out) {
        HIP_TRY(c, hipMemcpy(loss_out, d->loss.p, 4, hipMemcpyDeviceToHost));
        *loss_out *= la.inv_count;
    }
    return MCRT_OK;
}

extern "C" int mcrt_set_profiling(mcrt_ctx *c, int enabled) {
    if (!c) return MCRT_ERR_BAD_ARG;
    c->prof = enabled != 0;
    return MCRT_OK;
}

extern "C" int mcrt_kernel_stats(mcrt_ctx *c, const char *kernel, double *total_ms,
                                 uint64_t *launches) {
    int rc = bind(c);
    if (rc) return rc;
    if (!kernel) return fail(c, MCRT_ERR_BAD_ARG, "kernel name is required");
    if ((rc = drain_profile(c))) return rc;
    if (total_ms) *total_ms = c->total_ms.count(kernel) ? c->total_ms[kernel] : 0.0;
    if (launches) *launches = c->launches.count(kernel) ? c->launches[kernel] : 0;
    return MCRT_OK;
}

extern "C" int mcrt_reset_stats(mcrt_ctx *c) {
    int rc = bind(c);
    if (rc) return rc;
    if ((rc = drain_profile(c))) return rc;
    c->total_ms.clear();
    c->launches.clear();
    return MCRT_OK;
}
