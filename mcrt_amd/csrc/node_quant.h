// node_quant.h — quantized BVH child pairs (32 B per internal node) for the
// traversal of scenes whose traversal set spills the L2, shared by the host
// encoder (capi.hip) and the device decoder (dev_common.h node_step).
//
// Layout, 8 words:
//   w0..w2  per axis x, y, z: the grid origin as FP32 bits whose low byte is
//           replaced by the biased exponent e of the grid step 2^(e - 127)
//           (the origin keeps its low byte zero);
//   w3..w5  12 plane bytes q: child 0 (lo.x, hi.x, lo.y, hi.y, lo.z, hi.z),
//           then child 1 in the same order, 4 to a word, low byte first;
//   w6, w7  the two child references.
// A plane decodes as fma(q, step, origin) in FP32 (one rounding).  The
// encoder evaluates that same expression: a lower plane takes the largest q
// whose decoded value is <= the reference plane, an upper plane the smallest
// q whose decoded value is >= it.  Every decoded box therefore contains its
// reference box, and the reference slab test (path_trace.wgsl:248-256), being
// monotone in each plane, accepts on the decoded box every ray interval it
// accepts on the reference box: the traversal visits a superset of the
// reference's nodes, and closest-hit results, decided by Moller-Trumbore in
// the leaves, are unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mcrt {

constexpr uint32_t kQuantWords = 8;  // 32 B per child pair

__host__ __device__ inline float qnode_origin(uint32_t w) {
    const uint32_t b = w & 0xFFFFFF00u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

__host__ __device__ inline float qnode_step(uint32_t w) {
    const uint32_t b = (w & 0xFFu) << 23;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

__host__ __device__ inline float qnode_plane(uint32_t w, int byte, float step, float origin) {
    return fmaf((float)((w >> (8 * byte)) & 0xFFu), step, origin);
}

// Grid of one axis over [mn, mx]: the word (origin bits with the step's
// biased exponent in the low byte), or false where a plane is not finite or
// 255 steps overflow FP32.
inline bool qgrid_axis(float mn, float mx, uint32_t &word) {
    if (!std::isfinite(mn) || !std::isfinite(mx)) return false;
    // Origin: mn with its low byte cleared, moved down one unit of 256 ulps
    // where clearing raised it (a negative mn).
    uint32_t bits;
    memcpy(&bits, &mn, 4);
    uint32_t ob = bits & 0xFFFFFF00u;
    float org;
    memcpy(&org, &ob, 4);
    if (org > mn) {
        ob = ((bits | 0xFFu) + 1u) & 0xFFFFFF00u;
        memcpy(&org, &ob, 4);
    }
    if (!std::isfinite(org) || org > mn) return false;
    // Step: the smallest power of two whose 255th grid plane reaches mx.
    uint32_t e = 1;
    for (; e <= 254u; ++e) {
        const float top = qnode_plane(255u, 0, qnode_step(e), org);
        if (std::isfinite(top) && top >= mx) break;
    }
    if (e > 254u) return false;
    word = ob | e;
    // Where the word format allows it, a multiple of the step as origin
    // instead (the same or the next step): then every plane on the step grid
    // (voxel coordinates, for one) decodes exactly and adds no thickness.
    for (uint32_t ea = e; ea <= std::min<uint32_t>(e + 1u, 254u); ++ea) {
        const double st = (double)qnode_step(ea);
        const float a = (float)(std::floor((double)mn / st) * st);
        uint32_t ab;
        memcpy(&ab, &a, 4);
        if ((ab & 0xFFu) != 0u || !std::isfinite(a) || a > mn) continue;
        const float top = qnode_plane(255u, 0, qnode_step(ea), a);
        if (!std::isfinite(top) || top < mx) continue;
        word = ab | ea;
        break;
    }
    return true;
}

// The plane byte of a lower (upper) plane v on an axis grid: the largest q
// whose decoded value is <= v (the smallest q whose value is >= v).  The
// caller checks that the decoded plane lies on the outer side of v.
inline uint32_t qgrid_lower(uint32_t word, float v) {
    const float st = qnode_step(word), o = qnode_origin(word);
    auto dec = [&](uint32_t q) { return qnode_plane(q, 0, st, o); };
    double g = std::floor(((double)v - (double)o) / (double)st);
    uint32_t q = (uint32_t)std::min(255.0, std::max(0.0, g));
    while (q > 0u && dec(q) > v) --q;
    while (q < 255u && dec(q + 1u) <= v) ++q;
    return q;
}

inline uint32_t qgrid_upper(uint32_t word, float v) {
    const float st = qnode_step(word), o = qnode_origin(word);
    auto dec = [&](uint32_t q) { return qnode_plane(q, 0, st, o); };
    double g = std::ceil(((double)v - (double)o) / (double)st);
    uint32_t q = (uint32_t)std::min(255.0, std::max(0.0, g));
    while (q < 255u && dec(q) < v) ++q;
    while (q > 0u && dec(q - 1u) >= v) --q;
    return q;
}

// Plane bytes of n boxes on one shared grid, 6 per box in the order
// (lo.x, hi.x, lo.y, hi.y, lo.z, hi.z), and the three grid words.  False
// where the grid cannot express a box.
inline bool qgrid_encode(const float (*lo)[3], const float (*hi)[3], int n, uint32_t grid[3], uint32_t *bytes) {
    for (int k = 0; k < 3; ++k) {
        float mn = lo[0][k], mx = hi[0][k];
        for (int c = 1; c < n; ++c) {
            mn = std::min(mn, lo[c][k]);
            mx = std::max(mx, hi[c][k]);
        }
        for (int c = 0; c < n; ++c)
            if (!std::isfinite(lo[c][k]) || !std::isfinite(hi[c][k])) return false;
        if (!qgrid_axis(mn, mx, grid[k])) return false;
        const float st = qnode_step(grid[k]), o = qnode_origin(grid[k]);
        for (int c = 0; c < n; ++c) {
            const uint32_t ql = qgrid_lower(grid[k], lo[c][k]), qh = qgrid_upper(grid[k], hi[c][k]);
            if (qnode_plane(ql, 0, st, o) > lo[c][k] || qnode_plane(qh, 0, st, o) < hi[c][k]) return false;
            bytes[6 * c + 2 * k] = ql;
            bytes[6 * c + 2 * k + 1] = qh;
        }
    }
    return true;
}

// Encodes the child pair (lo0, hi0), (lo1, hi1) with references ref0, ref1.
// Returns false where the grid cannot express a box (a non-finite plane, or
// a range whose 255 steps overflow FP32); w is then unspecified.
inline bool qnode_encode(const float lo0[3], const float hi0[3], const float lo1[3], const float hi1[3],
                         uint32_t ref0, uint32_t ref1, uint32_t w[kQuantWords]) {
    const float lo[2][3] = {{lo0[0], lo0[1], lo0[2]}, {lo1[0], lo1[1], lo1[2]}};
    const float hi[2][3] = {{hi0[0], hi0[1], hi0[2]}, {hi1[0], hi1[1], hi1[2]}};
    uint32_t bytes[12];
    if (!qgrid_encode(lo, hi, 2, w, bytes)) return false;
    for (int j = 0; j < 3; ++j)
        w[3 + j] = bytes[4 * j] | (bytes[4 * j + 1] << 8) | (bytes[4 * j + 2] << 16) | (bytes[4 * j + 3] << 24);
    w[6] = ref0;
    w[7] = ref1;
    return true;
}

// Decodes child c (0 or 1) of an encoded pair: lo[3], hi[3].
__host__ __device__ inline void qnode_decode(const uint32_t w[kQuantWords], int c, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        const float st = qnode_step(w[k]), o = qnode_origin(w[k]);
        const int b = 6 * c + 2 * k;  // byte index of lo; hi follows
        lo[k] = qnode_plane(w[3 + b / 4], b % 4, st, o);
        hi[k] = qnode_plane(w[3 + (b + 1) / 4], (b + 1) % 4, st, o);
    }
}

}  // namespace mcrt
