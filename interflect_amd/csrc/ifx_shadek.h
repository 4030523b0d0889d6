// Camera-ray shading kernels (render_frame, src/main.rs:229-367): included by
// the shade_*.hip translation units, each of which instantiates one group of
// them behind a launcher of ifx_launch.h.
#pragma once
#include "ifx_kern.h"
#include "ifx_launch.h"

namespace ifx {

// 16x16 pixel blocks, one 8x8 tile per wavefront.
template <bool COUNT, bool LDS, bool WAVE>
__global__ __launch_bounds__(kBlock) void k_shade(SceneView Sg, LightsView Lg, LdsPlan P, ShadeArgs A, float4* film,
                                                  unsigned long long* evals, unsigned long long* ops) {
    IFX_SETUP_W(LDS, true, COUNT, WAVE);
    uint32_t wv = threadIdx.x >> 6, l = threadIdx.x & 63u;
    uint32_t x = blockIdx.x * 16u + (wv & 1u) * 8u + (l & 7u);
    uint32_t y = blockIdx.y * 16u + (wv >> 1) * 8u + (l >> 3);
    uint32_t local = 0;
    const bool on = x < A.W && y < A.H;  // every lane calls (pixels off the film follow along)
    f3 c = shade_pixel(S, L, A, on ? x : 0u, on ? y : 0u, local, stk, on);
    if (on) film[(size_t)y * A.W + x] = st4(c);
    count_steps(evals, local);
    stk_flush(stk, ops);
}

// Many cameras over one solve (the turntable, src/main.rs:492-520): grid z
// is the view; each view is k_shade with its own camera.
template <bool COUNT, bool LDS, bool WAVE>
__global__ __launch_bounds__(kBlock) void k_shade_views(SceneView Sg, LightsView Lg, LdsPlan P, ShadeArgs A0,
                                                        const CamV* __restrict__ cams, float4* film,
                                                        unsigned long long* evals, unsigned long long* ops) {
    IFX_SETUP_W(LDS, true, COUNT, WAVE);
    ShadeArgs A = A0;
    A.cam = cams[blockIdx.z];
    uint32_t wv = threadIdx.x >> 6, l = threadIdx.x & 63u;
    uint32_t x = blockIdx.x * 16u + (wv & 1u) * 8u + (l & 7u);
    uint32_t y = blockIdx.y * 16u + (wv >> 1) * 8u + (l >> 3);
    uint32_t local = 0;
    const bool on = x < A.W && y < A.H;
    f3 c = shade_pixel(S, L, A, on ? x : 0u, on ? y : 0u, local, stk, on);
    if (on) film[(size_t)blockIdx.z * A.W * A.H + (size_t)y * A.W + x] = st4(c);
    count_steps(evals, local);
    stk_flush(stk, ops);
}

template <bool COUNT, bool LDS, bool WAVE>
inline void launch_shade_t(const ShadeLaunch& a) {
    if (a.cams)
        k_shade_views<COUNT, LDS, WAVE><<<a.grid, kBlock, a.lds_bytes, a.stream>>>(a.S, a.L, a.P, a.A, a.cams, a.film,
                                                                                   a.evals, a.ops);
    else
        k_shade<COUNT, LDS, WAVE><<<a.grid, kBlock, a.lds_bytes, a.stream>>>(a.S, a.L, a.P, a.A, a.film, a.evals,
                                                                             a.ops);
}

}  // namespace ifx
