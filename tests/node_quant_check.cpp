// CPU check of the quantized BVH child pairs of node_quant.h, built with
// hipcc as a host program (it calls no HIP API).  For adversarial pairs of
// boxes (flat, tiny, large, negative, straddling zero, sharing planes,
// denormal, far from the origin) it encodes, decodes with the device's
// formula, and counts:
//   fail     finite pairs the encoder rejected;
//   outside  decoded planes on the inner side of a reference plane;
//   loose    decoded planes further than one grid step from the reference;
//   nonfin   pairs with a non-finite plane the encoder accepted.
// With `plant` it moves one lower plane of every pair one step inward after
// encoding, which the containment count must catch.
//   node_quant_check <cases> <seed> [plant]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../mcrt_amd/csrc/node_quant.h"

using namespace mcrt;

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const long cases = atol(argv[1]);
    const bool plant = argc > 3;
    std::mt19937_64 rng(strtoull(argv[2], nullptr, 10));
    std::uniform_real_distribution<double> u(0.0, 1.0);
    long fail = 0, outside = 0, loose = 0, nonfin = 0;
    for (long c = 0; c < cases; ++c) {
        // scale of the scene and of the boxes
        const double scale = std::ldexp(1.0, (int)(u(rng) * 40.0) - 20);
        const double ext = scale * std::ldexp(1.0, -(int)(u(rng) * 30.0));
        float lo[2][3], hi[2][3];
        const double centre = (u(rng) * 2.0 - 1.0) * scale * (u(rng) < 0.2 ? 1000.0 : 1.0);
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 3; ++k) {
                double a = centre + (u(rng) * 2.0 - 1.0) * ext, d = (u(rng) < 0.2) ? 0.0 : u(rng) * ext;
                if (u(rng) < 0.1) a = std::round(a);                     // grid-aligned planes
                if (u(rng) < 0.05) a = (u(rng) < 0.5 ? 1 : -1) * 1e-40;  // denormal
                if (b == 1 && u(rng) < 0.2) a = lo[0][k];  // shared plane
                lo[b][k] = (float)a;
                hi[b][k] = (float)(a + d);
                if (hi[b][k] < lo[b][k]) hi[b][k] = lo[b][k];
            }
        uint32_t w[kQuantWords];
        if (!qnode_encode(lo[0], hi[0], lo[1], hi[1], 7u, 0x80000009u, w)) {
            ++fail;
            continue;
        }
        if (plant) w[3] = (w[3] & ~0xFFu) | ((w[3] & 0xFFu) + 1u > 255u ? 255u : (w[3] & 0xFFu) + 1u);
        for (int b = 0; b < 2; ++b) {
            float dl[3], dh[3];
            qnode_decode(w, b, dl, dh);
            for (int k = 0; k < 3; ++k) {
                const float st = qnode_step(w[k]);
                if (dl[k] > lo[b][k] || dh[k] < hi[b][k]) ++outside;
                // one grid step, or two ulps where the step is finer than the
                // FP32 spacing at the plane (decoded values then repeat)
                const double tl = std::max((double)st, 2.0 * (std::nextafter(std::fabs(lo[b][k]), INFINITY) - std::fabs(lo[b][k])));
                const double th = std::max((double)st, 2.0 * (std::nextafter(std::fabs(hi[b][k]), INFINITY) - std::fabs(hi[b][k])));
                if ((double)lo[b][k] - dl[k] > tl * 1.0000001 || (double)dh[k] - hi[b][k] > th * 1.0000001) ++loose;
            }
        }
        if (w[6] != 7u || w[7] != 0x80000009u) ++outside;
        // a non-finite plane must be rejected
        float bad[3] = {lo[0][0], lo[0][1], lo[0][2]};
        bad[c % 3] = (c & 1) ? INFINITY : NAN;
        if (qnode_encode(bad, hi[0], lo[1], hi[1], 1u, 2u, w)) ++nonfin;
    }
    printf("cases %ld fail %ld outside %ld loose %ld nonfin %ld\n", cases, fail, outside, loose, nonfin);
    return 0;
}
