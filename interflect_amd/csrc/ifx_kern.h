// Device helpers, launch macros and host utilities shared by the kernel
// translation units (engine.hip: transfer / solve / shade; adjoint.hip:
// backward pass).
#pragma once
#include "engine.h"

namespace ifx {

// ====================================================================== device helpers
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// Adds a per-lane step count to a 64-bit counter, one atomic per wave.
__device__ __forceinline__ void count_steps(unsigned long long* ctr, uint32_t steps) {
    uint32_t t = wave_sum(steps);
    if (lane_id() == 0 && t) atomicAdd(ctr, (unsigned long long)t);
}

// Number of set bits of `m` at lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Dynamic LDS of every traversal kernel: [scene records][lights, materials]
// [traversal stack].  Staging the scene puts every node / object / light
// read of the divergent traversal on the LDS path instead of the vector
// cache; scenes too large for the budget are read from global memory and
// only the stack lives in LDS.
constexpr uint32_t kLdsBudget = 64u * 1024u;

__device__ __forceinline__ void copy_f4(float4* dst, const float4* __restrict__ src, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Copies the scene (and lights) into LDS and repoints the views.  Every
// thread of the block must call it (it ends in a barrier).
template <bool LDS, bool LIGHTS>
__device__ __forceinline__ void stage_in(float4* lds, SceneView& S, LightsView& L) {
    if constexpr (LDS) {
        float4* q = lds;
        uint32_t n = 4u * S.n_obj;
        copy_f4(q, S.obj, n);
        S.obj = q;
        q += n;
        n = 4u * S.n_planes;
        copy_f4(q, S.planes, n);
        S.planes = q;
        q += n;
        n = 4u * S.n_subs;
        copy_f4(q, S.subs, n);
        S.subs = q;
        q += n;
        n = 2u * S.n_nodes;
        copy_f4(q, S.nodes, n);
        S.nodes = q;
        q += n;
        n = (S.n_obj + 3u) / 4u;
        copy_f4(q, (const float4*)S.lin, n);
        S.lin = (const uint32_t*)q;
        q += n;
        if constexpr (LIGHTS) {
            n = 7u * L.n_quad;
            copy_f4(q, L.quad, n);
            L.quad = q;
            q += n;
            n = 2u * L.n_sph;
            copy_f4(q, L.sph, n);
            L.sph = q;
            q += n;
            n = 2u * L.n_pnt;
            copy_f4(q, L.pnt, n);
            L.pnt = q;
            q += n;
            n = 4u * L.n_mat;
            copy_f4(q, L.mat, n);
            L.mat = q;
        }
        __syncthreads();
    }
}

// Timed kernels traverse with LdsStack; the instrumented instantiation
// (COUNT=true) uses CountStack and adds its operation tally to `ops`.
template <bool COUNT>
struct StackOf {
    using T = LdsStack;
};
template <>
struct StackOf<true> {
    using T = CountStack;
};
__device__ __forceinline__ void stk_init(LdsStack&) {}
__device__ __forceinline__ void stk_init(CountStack& s) { s.ops = 0ull; }
__device__ __forceinline__ void stk_flush(LdsStack&, unsigned long long*) {}
__device__ __forceinline__ void stk_flush(CountStack& s, unsigned long long* ops) {
    if (s.ops) atomicAdd(ops, s.ops);
}

// Kernel prologue: stage, then place this lane's stack column after the staged records.
#define IFX_SETUP(LDS, LIGHTS, COUNT)                                            \
    extern __shared__ float4 ifx_lds[];                                         \
    SceneView S = Sg;                                                           \
    LightsView L = Lg;                                                          \
    stage_in<LDS, LIGHTS>(ifx_lds, S, L);                                       \
    typename StackOf<COUNT>::T stk;                                             \
    stk.dbase = (float*)(ifx_lds + P.scene_f4 + P.light_f4) + threadIdx.x;      \
    stk.base = (uint16_t*)(stk.dbase - threadIdx.x + P.stack_depth * kBlock) +  \
               threadIdx.x;                                                     \
    stk.stride = (uint32_t)kBlock;                                              \
    stk_init(stk);                                                              \
    (void)L

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// Items a wavefront takes per queue atomic: about an eighth of its even
// share of the work, between one and four wavefront widths.
inline uint32_t wave_chunk(uint64_t items, uint32_t blocks) {
    uint64_t waves = (uint64_t)blocks * (kBlock / 64);
    uint64_t c = items / (waves * 8 + 1);
    return (uint32_t)(c < 64 ? 64 : (c > 256 ? 256 : c));
}

inline uint32_t bits_for(uint64_t maxv) {
    uint32_t b = 1;
    while (b < 64 && (maxv >> b) != 0) ++b;
    return b;
}

template <class F>
void Engine::launch(const char* name, F&& f) {
    if (!prof_) {
        f();
        return;
    }
    hipEvent_t a, b;
    if (event_pool_.size() >= 2) {
        a = event_pool_.back();
        event_pool_.pop_back();
        b = event_pool_.back();
        event_pool_.pop_back();
    } else {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    (void)hipEventRecord(a, stream);
    f();
    (void)hipEventRecord(b, stream);
    pending_.push_back(Pending{name, a, b});
}

}  // namespace ifx

#define LAUNCH_CL(K, COUNT, LDS, GRID, DYN, ...)                                          \
    do {                                                                                 \
        if (COUNT) {                                                                     \
            if (LDS)                                                                     \
                K<true, true><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);               \
            else                                                                         \
                K<true, false><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);              \
        } else {                                                                         \
            if (LDS)                                                                     \
                K<false, true><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);              \
            else                                                                         \
                K<false, false><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);             \
        }                                                                                \
    } while (0)

#define LAUNCH_L(K, LDS, GRID, DYN, ...)                                                 \
    do {                                                                                 \
        if (LDS)                                                                         \
            K<true><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                         \
        else                                                                             \
            K<false><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                        \
    } while (0)

#define CK(call)                                       \
    do {                                               \
        int rc_ = check((call), #call, err);           \
        if (rc_) return rc_;                           \
    } while (0)
