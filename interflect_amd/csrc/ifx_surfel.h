// Surfel placement shared by the host and device paths (src/surfel.rs:147-255):
// Halton seeds in the scene box, Newton projection onto the zero level set,
// and the Poisson-disc grid of the acceptance test.
#pragma once
#include "ifx_shade.h"

namespace ifx {

// src/surfel.rs (radical inverse in f32, `f *= 1/base`).
IFX_HD float halton(uint32_t i, uint32_t base) {
    float f = 1.0f, r = 0.0f;
    float inv = 1.0f / (float)base;
    while (i > 0) {
        f *= inv;
        r += f * (float)(i % base);
        i /= base;
    }
    return r;
}

struct SurfelParams {
    f3 lo, hi;
    float diag, d_min, inv_cell, eps, d_min_sq, area;
    int32_t dim[3];
    uint32_t seeds;
};

IFX_HD SurfelParams surfel_params(f3 lo, f3 hi, uint64_t target) {
    SurfelParams P;
    P.lo = lo;
    P.hi = hi;
    P.diag = len3(hi - lo);
    P.d_min = fmx(P.diag / sqrtf((float)target), 1e-4f);
    P.inv_cell = 1.0f / P.d_min;
    P.dim[0] = clampi(sat_i32(ceilf((hi.x - lo.x) * P.inv_cell)), 1, 512);
    P.dim[1] = clampi(sat_i32(ceilf((hi.y - lo.y) * P.inv_cell)), 1, 512);
    P.dim[2] = clampi(sat_i32(ceilf((hi.z - lo.z) * P.inv_cell)), 1, 512);
    P.eps = P.d_min * 0.02f;
    P.d_min_sq = P.d_min * P.d_min;
    P.area = P.d_min * P.d_min / 0.65f;
    const uint64_t s24 = (uint64_t)(uint32_t)target * 24ull;
    P.seeds = (uint32_t)(s24 < 6000000ull ? s24 : 6000000ull);
    return P;
}

IFX_HD void surfel_cell(const SurfelParams& P, f3 p, int32_t c[3]) {
    c[0] = clampi(sat_i32((p.x - P.lo.x) * P.inv_cell), 0, P.dim[0] - 1);
    c[1] = clampi(sat_i32((p.y - P.lo.y) * P.inv_cell), 0, P.dim[1] - 1);
    c[2] = clampi(sat_i32((p.z - P.lo.z) * P.inv_cell), 0, P.dim[2] - 1);
}

// Seed i (0-based; the reference's Halton index is i + 1): Newton steps onto
// the surface, then the acceptance conditions that do not involve other
// surfels.  Returns whether the seed survives; p is its projected position.
template <class Stk>
IFX_HD bool project_seed(const SceneView& S, const SurfelParams& P, uint32_t i, f3& p, Stk& stk) {
    f3 s = mk(halton(i + 1, 2), halton(i + 1, 3), halton(i + 1, 5));
    p = P.lo + (P.hi - P.lo) * s;
    bool ok = false;
    for (int it = 0; it < 10; ++it) {
        float d = dist(S, p, stk);
        if (fabsf(d) < P.eps) {
            ok = true;
            break;
        }
        if (!(fabsf(d) <= kFltMax) || fabsf(d) > P.diag) break;  // not finite, or too far
        p = p - scene_normal(S, p, stk) * d;
    }
    return ok && fabsf(dist(S, p, stk)) < P.eps &&
           !(p.x < P.lo.x || p.y < P.lo.y || p.z < P.lo.z || p.x > P.hi.x || p.y > P.hi.y || p.z > P.hi.z);
}

}  // namespace ifx
