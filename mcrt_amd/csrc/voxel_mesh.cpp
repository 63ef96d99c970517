// voxel_mesh.cpp — greedy voxel meshing of a grid of 16 x 256 x 16 chunks
// (host-side C-ABI helper for the voxel-terrain workload).
//
// Semantics follow the reference mesher src/world/mesher.rs:97-400 and its
// caller src/gpu.rs:386-458:
//  - chunks are meshed one by one in grid order (cx outer, cz inner, as
//    visible_chunk_coords, src/world/mod.rs:63-74); neighbours inside the
//    grid are used for face culling, outside the grid count as air;
//  - per chunk, the six directions +X, -X, +Y, -Y, +Z, -Z; per direction,
//    layers along the axis; a mask over (u, v) = (Y, Z) / (X, Z) / (X, Y)
//    holding the block id of every opaque voxel whose neighbour in the face
//    direction is not opaque;
//  - greedy merge: scan u outer, v inner; grow along v, then along u while
//    whole rows match and are unvisited;
//  - one quad per rectangle, corners (u0,v0) (u0+h,v0) (u0+h,v0+w) (u0,v0+w)
//    on the plane layer + (sign > 0), uv (0,0) (1,0) (1,1) (0,1), flat normal,
//    winding (0,1,2)(0,2,3) for +sign and (0,2,1)(0,3,2) for -sign;
//  - position = chunk origin (cx*16*s, 0, cz*16*s) + voxel offset * s.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mcrt_abi.h"

namespace {

constexpr int kXZ = 16, kY = 256;
constexpr size_t kChunk = (size_t)kXZ * kY * kXZ;

inline size_t cidx(int x, int y, int z) { return (size_t)x * kY * kXZ + (size_t)y * kXZ + z; }

struct Out {
    GpuVertex *v;
    GpuTriangle *t;
    uint64_t nv = 0, nt = 0, max_v, max_t;
    bool overflow = false;
};

void emit(Out &o, const float origin[3], float s, int axis, int ua, int va, int layer, int u0, int v0,
          int h, int w, int sign, uint32_t mat) {
    uint64_t base = o.nv;
    if (o.nv + 4 > o.max_v || o.nt + 2 > o.max_t) {
        o.overflow = true;
        o.nv += 4;
        o.nt += 2;
        return;
    }
    const int face = sign > 0 ? 1 : 0;
    auto pos = [&](int u, int v, float *p) {
        float q[3] = {0.0f, 0.0f, 0.0f};
        q[axis] = (float)(layer + face) * s;
        q[ua] = (float)u * s;
        q[va] = (float)v * s;
        p[0] = origin[0] + q[0];
        p[1] = origin[1] + q[1];
        p[2] = origin[2] + q[2];
    };
    const int cu[4] = {u0, u0 + h, u0 + h, u0};
    const int cv[4] = {v0, v0, v0 + w, v0 + w};
    const float uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int k = 0; k < 4; ++k) {
        GpuVertex &gv = o.v[o.nv + k];
        std::memset(&gv, 0, sizeof gv);
        pos(cu[k], cv[k], gv.position);
        gv.position[3] = 1.0f;
        gv.normal[axis] = (float)sign;
        gv.uv[0] = uv[k][0];
        gv.uv[1] = uv[k][1];
    }
    uint32_t b = (uint32_t)base;
    GpuTriangle &t0 = o.t[o.nt], &t1 = o.t[o.nt + 1];
    if (sign > 0) {
        t0 = GpuTriangle{{b, b + 1, b + 2}, mat};
        t1 = GpuTriangle{{b, b + 2, b + 3}, mat};
    } else {
        t0 = GpuTriangle{{b, b + 2, b + 1}, mat};
        t1 = GpuTriangle{{b, b + 3, b + 2}, mat};
    }
    o.nv += 4;
    o.nt += 2;
}

}  // namespace

extern "C" int mcrt_mesh_voxel_chunks(const uint16_t *blocks, uint32_t chunks_x, uint32_t chunks_z,
                                      int32_t cx0, int32_t cz0, float voxel_size,
                                      const uint8_t *opaque, const uint32_t *block_material,
                                      GpuVertex *out_verts, uint64_t max_verts,
                                      GpuTriangle *out_tris, uint64_t max_tris,
                                      uint64_t *n_verts, uint64_t *n_tris) {
    if (!blocks || !opaque || !block_material || !n_verts || !n_tris) return MCRT_ERR_BAD_ARG;
    if ((out_verts == nullptr) != (out_tris == nullptr)) return MCRT_ERR_BAD_ARG;
    Out o;
    o.v = out_verts;
    o.t = out_tris;
    o.max_v = out_verts ? max_verts : 0;
    o.max_t = out_tris ? max_tris : 0;
    const int dim[3] = {kXZ, kY, kXZ};
    std::vector<uint16_t> mask;
    std::vector<uint8_t> visited;
    for (uint32_t ix = 0; ix < chunks_x; ++ix) {
        for (uint32_t iz = 0; iz < chunks_z; ++iz) {
            const uint16_t *ch = blocks + ((size_t)ix * chunks_z + iz) * kChunk;
            // neighbours [+X, -X, +Y, -Y, +Z, -Z]; Y never has one.
            const uint16_t *nb[6] = {
                ix + 1 < chunks_x ? blocks + ((size_t)(ix + 1) * chunks_z + iz) * kChunk : nullptr,
                ix > 0 ? blocks + ((size_t)(ix - 1) * chunks_z + iz) * kChunk : nullptr,
                nullptr,
                nullptr,
                iz + 1 < chunks_z ? blocks + ((size_t)ix * chunks_z + iz + 1) * kChunk : nullptr,
                iz > 0 ? blocks + ((size_t)ix * chunks_z + iz - 1) * kChunk : nullptr,
            };
            const int32_t cx = cx0 + (int32_t)ix, cz = cz0 + (int32_t)iz;
            // chunk_world_origin, src/world/chunk.rs:131-137
            const float origin[3] = {(float)cx * (float)kXZ * voxel_size, 0.0f,
                                     (float)cz * (float)kXZ * voxel_size};
            for (int f = 0; f < 6; ++f) {
                const int axis = f / 2, sign = (f % 2 == 0) ? 1 : -1;
                const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
                const int ld = dim[axis], ud = dim[ua], vd = dim[va];
                mask.assign((size_t)ud * vd, 0);
                visited.assign((size_t)ud * vd, 0);
                for (int layer = 0; layer < ld; ++layer) {
                    std::fill(mask.begin(), mask.end(), 0);
                    std::fill(visited.begin(), visited.end(), 0);
                    for (int u = 0; u < ud; ++u) {
                        for (int v = 0; v < vd; ++v) {
                            int p[3];
                            p[axis] = layer;
                            p[ua] = u;
                            p[va] = v;
                            uint16_t id = ch[cidx(p[0], p[1], p[2])];
                            if (id == 0 || !opaque[id & 0xFF] || id > 0xFF) continue;
                            int c = p[axis] + sign;
                            uint16_t nid;
                            if (c < 0 || c >= dim[axis]) {
                                const uint16_t *n = nb[axis * 2 + (c < 0 ? 1 : 0)];
                                if (n) {
                                    int q[3] = {p[0], p[1], p[2]};
                                    q[axis] = c < 0 ? dim[axis] - 1 : 0;
                                    nid = n[cidx(q[0], q[1], q[2])];
                                } else {
                                    nid = 0;
                                }
                            } else {
                                int q[3] = {p[0], p[1], p[2]};
                                q[axis] = c;
                                nid = ch[cidx(q[0], q[1], q[2])];
                            }
                            if (!(nid <= 0xFF && opaque[nid])) mask[(size_t)u * vd + v] = id;
                        }
                    }
                    for (int u0 = 0; u0 < ud; ++u0) {
                        int v0 = 0;
                        while (v0 < vd) {
                            uint16_t sid = mask[(size_t)u0 * vd + v0];
                            if (sid == 0 || visited[(size_t)u0 * vd + v0]) {
                                ++v0;
                                continue;
                            }
                            int w = 1;
                            while (v0 + w < vd && mask[(size_t)u0 * vd + v0 + w] == sid &&
                                   !visited[(size_t)u0 * vd + v0 + w])
                                ++w;
                            int h = 1;
                            for (;;) {
                                if (u0 + h >= ud) break;
                                bool ok = true;
                                for (int dv = 0; dv < w; ++dv) {
                                    size_t k = (size_t)(u0 + h) * vd + v0 + dv;
                                    if (mask[k] != sid || visited[k]) {
                                        ok = false;
                                        break;
                                    }
                                }
                                if (!ok) break;
                                ++h;
                            }
                            for (int du = 0; du < h; ++du)
                                for (int dv = 0; dv < w; ++dv) visited[(size_t)(u0 + du) * vd + v0 + dv] = 1;
                            emit(o, origin, voxel_size, axis, ua, va, layer, u0, v0, h, w, sign,
                                 block_material[sid]);
                            v0 += w;
                        }
                    }
                }
            }
        }
    }
    *n_verts = o.nv;
    *n_tris = o.nt;
    if (o.overflow && out_verts) return MCRT_ERR_LIMIT;
    return MCRT_OK;
}
