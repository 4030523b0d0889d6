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

// Order-preserving integer keys of finite floats: key order is float order.
__device__ __forceinline__ int32_t f2key(float x) {
    const int32_t b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float key2f(int32_t k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }

// One DPP step of wave_min_keys: each key of the rows in RM takes the
// minimum with the key CTRL moves to it; other rows see the largest key
// (the identity of min) and keep theirs.
template <int CTRL, int RM, int N>
__device__ __forceinline__ void dpp_min_step(int32_t (&k)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int32_t t = __builtin_amdgcn_update_dpp(0x7fffffff, k[i], CTRL, RM, 0xF, false);
        k[i] = t < k[i] ? t : k[i];
    }
}

// Minimum of each of N keys over all 64 lanes, which must all be active;
// the results are wave-uniform.  DPP within each row of 16 lanes (xor 1,
// xor 2, half-row mirror, row mirror), then row 0 into row 1 and row 2
// into row 3 (row_bcast:15), then row 1 into rows 2 and 3 (row_bcast:31):
// lane 63 holds the minimum.  The N chains are interleaved, so no step
// waits on the one before it.
template <int N>
__device__ __forceinline__ void wave_min_keys(int32_t (&k)[N]) {
    dpp_min_step<0xB1, 0xF>(k);   // quad_perm [1,0,3,2]
    dpp_min_step<0x4E, 0xF>(k);   // quad_perm [2,3,0,1]
    dpp_min_step<0x141, 0xF>(k);  // row_half_mirror
    dpp_min_step<0x140, 0xF>(k);  // row_mirror
    dpp_min_step<0x142, 0xA>(k);  // row_bcast:15
    dpp_min_step<0x143, 0xC>(k);  // row_bcast:31
#pragma unroll
    for (int i = 0; i < N; ++i) k[i] = __builtin_amdgcn_readlane(k[i], 63);
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
        n = (2u * S.n_obj + 3u) / 4u;  // slot table + leaf table
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

// Timed kernels traverse with LdsStack; the instrumented instantiations use
// CountStack (MODE 1: the reference traversal's operation tally) or
// ExecStack (MODE 2: primitive evaluations of the timed traversal) and add
// their tallies to `ops`.
template <int MODE>
struct StackOf {
    using T = LdsStack;
};
template <>
struct StackOf<1> {
    using T = CountStack;
};
template <>
struct StackOf<2> {
    using T = ExecStack;
};

// Stacks of the camera-ray shading kernels: LdsStack (wavefront-shared
// queries) or LaneStack, an LdsStack whose march loops keep the per-lane
// traversal (kShared false).
struct LaneStack : LdsStack {
    static constexpr bool kShared = false;
};

template <bool COUNT, bool WAVE>
struct ShadeStackOf {
    using T = LdsStack;
};
template <>
struct ShadeStackOf<false, false> {
    using T = LaneStack;
};
template <bool WAVE>
struct ShadeStackOf<true, WAVE> {
    using T = CountStack;
};

// Camera-ray shading of a scene small enough for the uniform scan keeps the
// per-lane traversal from this many pixels per launch up: with the GPU full,
// the scan's extra objects per lane cost more than its coherence saves
// (measured on MI355X, 7-object scene: one 640x480 view 10 % faster shared,
// twelve views in one launch 13 % slower).
constexpr uint64_t kWaveShadeMaxPixels = 1ull << 21;

// Counter slots of a tracing kernel, from its `ops` pointer: ops[0] counted
// FP32 operations and ops[4] primitive evaluations of the reference
// traversal (CountStack), ops[8] primitive evaluations executed by the
// timed traversal (ExecStack), ops[12] lane slots of the persistent
// kernels' march steps, ops[16] / ops[20] wavefront queries answered by the
// flat list / the BVH walk, ops[24] objects the flat lists held.  Every lane
// of the wavefront calls stk_flush.
__device__ __forceinline__ void stk_init(LdsStack&) {}
__device__ __forceinline__ void stk_init(ExecStack& s) { s.xe = s.xw = s.xf = s.xb = s.xl = 0u; }
__device__ __forceinline__ void stk_init(CountStack& s) {
    s.ops = 0ull;
    s.prims = 0u;
}
__device__ __forceinline__ void stk_flush(LdsStack&, unsigned long long*) {}
__device__ __forceinline__ void stk_flush(ExecStack& s, unsigned long long* ops) {
    const uint32_t t = wave_sum(s.xe);
    if (lane_id() == (uint32_t)__builtin_ctzll(__ballot(true))) {
        if (t + s.xw) atomicAdd(ops + 8, (unsigned long long)t + s.xw);
        if (s.xf) atomicAdd(ops + 16, (unsigned long long)s.xf);
        if (s.xb) atomicAdd(ops + 20, (unsigned long long)s.xb);
        if (s.xl) atomicAdd(ops + 24, (unsigned long long)s.xl);
    }
}
__device__ __forceinline__ void stk_flush(CountStack& s, unsigned long long* ops) {
    if (s.ops) atomicAdd(ops, s.ops);
    if (s.prims) atomicAdd(ops + 4, (unsigned long long)s.prims);
}

// LDS bytes of a block's lane columns: per lane `depth` f32 distances, then
// `depth` u16 node indices.
inline size_t stack_bytes(uint32_t depth) { return ((size_t)kBlock * depth * 6u + 15u) & ~(size_t)15u; }

// This lane's columns, after `f4` staged float4s.
template <class Stk>
__device__ __forceinline__ void place_stack(Stk& stk, float4* lds, uint32_t f4, uint32_t depth) {
    stk.dbase = (float*)(lds + f4) + threadIdx.x;
    stk.base = (uint16_t*)((float*)(lds + f4) + depth * kBlock) + threadIdx.x;
    stk.stride = (uint32_t)kBlock;
}

// ---------------------------------------------------------------- wavefront distance query
// The distance of src/scene.rs:191-208 (the value dist() returns) at one
// point per lane, answered by one BVH traversal shared by the wavefront:
// the node sequence is uniform, a node is opened when any lane that needs a
// value could find a relevant object under it, and every lane evaluates every
// object of an opened leaf, so the loop never diverges and each object record
// is one kind for the whole wavefront.  Lanes with `need` false only follow
// along.  All 64 lanes must call it together (uniform control flow).
//
// Equality with the per-lane traversal:
//  * SDF values are exact distances or larger (src/sdf.rs primitives;
//    subtractions only raise them), so an object is no nearer than its leaf
//    box, up to f32 rounding; delta = 1e-4 (1 + max|p|) bounds that rounding.
//    A node is opened while its box distance < the running bound B + 2 delta,
//    with B = max(hard, 0) without blending and min(hard, smin) + 1.25 K with
//    it; so every object left unevaluated is at least B_final + delta away,
//    where it can change neither the nearest hard object nor (farther than
//    min(hard, smin) + 1.25 K) the blend pair's result.  Each lane thus holds
//    the minima over all objects that matter.
//  * The per-lane traversal returns the minimum over the objects it visits,
//    which can miss an object whose value rounds below its leaf box distance
//    when its leaf is pruned.  It visits a leaf whose box distance is below
//    its final prune threshold (running thresholds only fall, ancestors' box
//    distances are no larger in f32).  Each lane checks that the leaves of
//    the objects deciding its result pass that test (without blending: box
//    distance <= result, or below every other object's value and the bound
//    of the unevaluated ones); a lane whose check fails takes the per-lane
//    traversal, so every lane returns dist() bit for bit.
//  Scenes of at most kWaveScanObjs objects skip the interior nodes: their
//  objects (or, without object boxes, their leaves) are tested in slot
//  order, each under the same opening test.
//  Flat list (S.g_obox set, all 64 lanes active, every lane that needs a
//  value has a bound ub): instead of walking the BVH, each lane tests the boxes of objects lane,
//  lane + 64, ... against the box of the wavefront's points, grown by the
//  largest opening bound; the objects that pass are taken in slot order, each
//  after its own box passes some lane's opening test.  An object left out is
//  then beyond every lane's bound, as under an unopened node, so the same
//  checks apply (a leaf's objects keep their slot order).
//  Records are read from their global copies (S.g_*) through
//  constant-address-space pointers: with the node and object indices uniform,
//  those reads are scalar loads.  The per-lane fallback reads S as given
//  (LDS when the scene is staged).
constexpr uint32_t kWaveScanObjs = 16;

typedef float __attribute__((ext_vector_type(4))) NativeF4;
struct ConstF4 {
    const __attribute__((address_space(4))) NativeF4* q;
    __device__ __forceinline__ float4 operator[](uint32_t i) const {
        const NativeF4 v = q[i];
        return make_float4(v.x, v.y, v.z, v.w);
    }
    __device__ __forceinline__ ConstF4 operator+(uint32_t i) const { return ConstF4{q + i}; }
};
__device__ __forceinline__ ConstF4 as_const(const float4* p) {
    return ConstF4{(const __attribute__((address_space(4))) NativeF4*)p};
}

//  MAT: also the material dist_mat reports (the first object reaching the
//  minimum in the per-lane order); a lane whose minimum is tied across leaves
//  (where the orders differ) takes the per-lane traversal.
//  ub: an upper bound on the lane's scene distance at p, when the caller has
//  one (a march's previous sample h at step s away: h + s, SDFs being
//  1-Lipschitz), else kFltMax.  The opening bound starts from it, so nodes
//  farther than ub (plus the same margins) are never opened: every object
//  under them is then at least ub + delta >= the result + delta away.
//  hint (optional, per lane): in, the slot of the object that decided the
//  lane's previous sample (or kNone); out, the object deciding this one.
//  See the flat list.
template <bool MAT, class Stk>
__device__ __forceinline__ DistHit wave_dist_mat(const SceneView& S, f3 p, bool need, Stk& stk,
                                                 float ub = kFltMax, uint32_t* hint = nullptr) {
    if (!S.use_bvh || S.n_nodes == 0) return need ? dist_mat(S, p, stk) : DistHit{0.0f, 0u};
    const ConstF4 nodes = as_const(S.g_nodes), objs = as_const(S.g_obj), subs = as_const(S.g_subs);
    const ConstF4 planes = as_const(S.g_planes);
    const __attribute__((address_space(4))) uint32_t* lin =
        (const __attribute__((address_space(4))) uint32_t*)S.g_lin;  // uniform reads: scalar loads
    if (__ballot(need) == 0ull) return DistHit{0.0f, 0u};
    constexpr uint32_t kNone = 0xffffffffu;
    const bool blend = S.blend_k > 0.0f;
    const float K = S.blend_k;
    const float delta = 1e-4f * (1.0f + maxe(abs3(p)));
    const uint32_t lanes = (uint32_t)__popcll(__ballot(true));  // lanes evaluating each opened object
    float hard = kFltMax, b1 = kFltMax, b2 = kFltMax, h2 = kFltMax;
    uint32_t a = kNone, la = kNone, s1 = kNone, s2 = kNone, hmat = 0u, b1mat = 0u;
    for (uint32_t i = 0; i < S.n_planes; ++i) {
        uint32_t mat;
        bool bl;
        float d = obj_dist_at(planes + 4 * i, subs, p, mat, bl, stk);
        if (d < hard) hard = d, hmat = mat;
    }
    // Squared box distances: the opening test only compares them with the
    // bound, and the margins above absorb the rounding.
    auto box2 = [&](uint32_t n) {
        const f3 e = max3(max3(ld3(nodes[2 * n]) - p, p - ld3(nodes[2 * n + 1])), spl(0.0f));
        return dot3(e, e);
    };
    // the opening bound (blin; open2 its square, -1 when not positive),
    // refreshed when a leaf has been evaluated
    const float ub0 = fmx(ub, 0.0f);  // objects containing p (box distance 0) stay openable
    const float ucap = ub < kFltMax ? (blend ? (ub0 + 1.25f * K) + 2.0f * delta : ub0 + 2.0f * delta) : kFltMax;
    float blin = kFltMax, open2 = kFltMax;
    float sm = blend ? smin_poly(b1, b2, K) : kFltMax;  // the blend pair's smooth minimum, kept current
    auto refresh = [&]() {
        float b = blend ? (fmn(hard, sm) + 1.25f * K) + 2.0f * delta : fmx(hard, 0.0f) + 2.0f * delta;
        blin = fmn(b, ucap);
        open2 = blin > 0.0f ? blin * blin : -1.0f;
    };
    refresh();
    // object k of leaf ni at p: the lane's running minima.  A CSG object is
    // skipped when its base primitive (a lower bound: subtraction only raises
    // the value) is beyond every lane's bound, as an unopened node is.
    auto take = [&](uint32_t k, uint32_t ni) {
        const ConstF4 rec = objs + 4 * k;
        const float4 g0 = rec[0], g1 = rec[1], g2 = rec[2], g3 = rec[3];
        const uint32_t mat = fbits(g3.x), sb = fbits(g3.y), sc = fbits(g3.z);
        const bool bl = fbits(g3.w) != 0u;
        float d = prim_dist(g0, g1, g2, p);
        stk.count_wave_prims(lanes);
        if (sc > 0) {
            if (__ballot(need && d < blin) == 0ull) return;
            for (uint32_t q = sb; q < sb + sc; ++q) {
                const ConstF4 r = subs + 4 * q;
                d = fmx(d, -prim_dist(r[0], r[1], r[2], p));
                stk.count_wave_prims(lanes);
            }
        }
        if (blend && bl) {
            if (d < b1) {
                b2 = b1, s2 = s1;
                b1 = d, s1 = k, b1mat = mat;
            } else if (d < b2) {
                b2 = d, s2 = k;
            }
            sm = smin_poly(b1, b2, K);
        } else if (d < hard) {
            if (ni != la) h2 = hard;  // everything evaluated so far lies outside this leaf
            hard = d, a = k, la = ni, hmat = mat;
        } else if (ni != la) {
            h2 = fmn(h2, d);
        } else if (MAT && d == hard && k < a) {
            // a tie within the leaf: the per-lane traversal takes the leaf's
            // objects in slot order, so the lower slot's material wins (hints
            // can be taken before the rest of their leaf)
            a = k, hmat = mat;
        }
    };
    // The flat list needs the whole wavefront: lane l tests objects l, l + 64, ...
    // and the reductions read every lane.
    const bool flat = S.g_obox != nullptr && S.n_obj > kWaveScanObjs && __ballot(true) == ~0ull &&
                      __ballot(need && !(ub < kFltMax)) == 0ull;
    if (S.n_obj > kWaveScanObjs) stk.count_query(flat);
    if (flat) {
        const ConstF4 obox = as_const(S.g_obox);
        // Hints first: any one object's value bounds the scene distance from
        // above, so once the objects that decided the lanes' previous samples
        // are taken, most lanes' opening bounds sit near their final values
        // and the list below stays short.  At most two distinct objects, each
        // the hint of the first lane not yet covered; the list skips them (a
        // blend object must not be taken twice).
        uint32_t t0 = kNone, t1 = kNone;
        if (hint) {
            const uint32_t hv = *hint;
            uint64_t c = __ballot(need && hv < S.n_obj);
            if (c) {
                t0 = (uint32_t)__builtin_amdgcn_readlane((int)hv, (int)__builtin_ctzll(c));
                take(t0, fbits(obox[2 * t0].w));
                refresh();
                c = __ballot(need && hv < S.n_obj && hv != t0);
                if (c) {
                    t1 = (uint32_t)__builtin_amdgcn_readlane((int)hv, (int)__builtin_ctzll(c));
                    take(t1, fbits(obox[2 * t1].w));
                    refresh();
                }
            }
        }
        // the needing lanes' point box and largest opening bound (maxima as
        // minima of complemented keys)
        int32_t kv[7] = {f2key(need ? p.x : kFltMax),   f2key(need ? p.y : kFltMax),   f2key(need ? p.z : kFltMax),
                         ~f2key(need ? p.x : -kFltMax), ~f2key(need ? p.y : -kFltMax), ~f2key(need ? p.z : -kFltMax),
                         ~f2key(need ? blin : 0.0f)};
        wave_min_keys(kv);
        const f3 wlo = mk(key2f(kv[0]), key2f(kv[1]), key2f(kv[2]));
        const f3 whi = mk(key2f(~kv[3]), key2f(~kv[4]), key2f(~kv[5]));
        const float wb = key2f(~kv[6]);
        const float wb2 = wb * wb;
        const uint32_t lane = lane_id();
        for (uint32_t g = 0; g < S.n_obj; g += 64u) {
            bool pass = false;
            if (g + lane < S.n_obj) {
                const float4 lo = S.g_obox[2 * (g + lane)], hi = S.g_obox[2 * (g + lane) + 1];
                const f3 e = max3(max3(ld3(lo) - whi, wlo - ld3(hi)), spl(0.0f));
                pass = dot3(e, e) < wb2;
            }
            uint64_t m = __ballot(pass);
            while (m) {
                const uint32_t k = g + (uint32_t)__builtin_ctzll(m);
                m &= m - 1ull;
                if (k == t0 || k == t1) continue;
                stk.count_listed();
                const float4 lo = obox[2 * k];
                const f3 e = max3(max3(ld3(lo) - p, p - ld3(obox[2 * k + 1])), spl(0.0f));
                if (__ballot(need && dot3(e, e) < open2) == 0ull) continue;
                take(k, fbits(lo.w));
                refresh();
            }
        }
    } else if (S.n_obj <= kWaveScanObjs && S.g_obox) {
        // few objects: every object in slot order, each taken when its own
        // box passes some lane's opening test (tighter than its leaf's box)
        const ConstF4 obox = as_const(S.g_obox);
        for (uint32_t k = 0; k < S.n_obj; ++k) {
            const float4 lo = obox[2 * k];
            const f3 e = max3(max3(ld3(lo) - p, p - ld3(obox[2 * k + 1])), spl(0.0f));
            if (__ballot(need && dot3(e, e) < open2) == 0ull) continue;
            take(k, fbits(lo.w));
            refresh();
        }
    } else if (S.n_obj <= kWaveScanObjs) {
        // few objects: the leaves in slot order (no interior nodes), each
        // opened under the same test
        for (uint32_t k = 0; k < S.n_obj;) {
            const uint32_t ni = lin[S.n_obj + k];
            const uint32_t cnt = fbits(nodes[2 * ni + 1].w);
            if (__ballot(need && box2(ni) < open2) != 0ull) {
                for (uint32_t j = k; j < k + cnt; ++j) take(j, ni);
                refresh();
            }
            k += cnt ? cnt : 1u;  // leaves are non-empty (checked when the scene is loaded)
        }
    } else {
        int sp = 0;
        stk.put(sp++, 0u, box2(0u));
        while (sp > 0) {
            float dn2;
            const uint32_t ni = __builtin_amdgcn_readfirstlane(stk.get(--sp, dn2));
            if (__ballot(need && dn2 < open2) == 0ull) continue;
            const uint32_t w0 = fbits(nodes[2 * ni].w), cnt = fbits(nodes[2 * ni + 1].w);
            if (cnt > 0) {
                if (S.g_obox) {
                    // each object of the leaf when its own box passes some lane's opening test
                    const ConstF4 obox = as_const(S.g_obox);
                    for (uint32_t k = w0; k < w0 + cnt; ++k) {
                        const f3 e = max3(max3(ld3(obox[2 * k]) - p, p - ld3(obox[2 * k + 1])), spl(0.0f));
                        if (__ballot(need && dot3(e, e) < open2) == 0ull) continue;
                        take(k, ni);
                        refresh();
                    }
                } else {
                    for (uint32_t k = w0; k < w0 + cnt; ++k) take(k, ni);
                    refresh();
                }
            } else {
                const uint32_t l = ni + 1, r = w0;
                const float dl = box2(l), dr = box2(r);
                // open first the child nearer for most lanes
                const uint32_t nl = (uint32_t)__popcll(__ballot(need && dl < dr));
                const uint32_t nn = (uint32_t)__popcll(__ballot(need));
                if (2u * nl >= nn) {
                    stk.put(sp++, r, dr);
                    stk.put(sp++, l, dl);
                } else {
                    stk.put(sp++, l, dl);
                    stk.put(sp++, r, dr);
                }
            }
        }
    }
    if (!need) return DistHit{0.0f, 0u};
    if (hint) *hint = blend && sm < hard ? s1 : a;
    // The checks compare aabb_dist of a leaf's box at p (the correctly
    // rounded length of the box offset) with thresholds.  The squared
    // offset d2 decides unless it lies within a relative 1e-5 of the
    // threshold's square, which bounds the rounding of d2, t * t and the
    // factor; there the length itself is taken.  la is the leaf of a.
    auto leaf_d2 = [&](uint32_t n) {
        const f3 e = max3(max3(ld3(S.nodes[2 * n]) - p, p - ld3(S.nodes[2 * n + 1])), spl(0.0f));
        return dot3(e, e);
    };
    auto len_below = [](float d2, float t, bool strict) {
        if (t > 1e-15f && t < 1e15f) {
            const float t2 = t * t;
            if (d2 < t2 * 0.99999f) return true;
            if (d2 > t2 * 1.00001f) return false;
        }
        const float l = sqrtf(d2);
        return strict ? l < t : l <= t;
    };
    auto slot_leaf = [&](uint32_t slot) { return S.lin[S.n_obj + slot]; };
    bool ok = true;
    float res;
    uint32_t rmat = hmat;
    if (!blend) {
        res = hard;
        if (a != kNone) {
            const float d2 = leaf_d2(la);
            ok = len_below(d2, hard, false) || (len_below(d2, h2, true) && len_below(d2, fmx(hard, 0.0f) + delta, true));
            if (MAT && !(h2 > hard)) ok = false;
        }
    } else {
        res = sm < hard ? sm : hard;
        if (MAT) {
            if (sm < hard) {
                rmat = b1mat;
                if (!(b2 > b1)) ok = false;
            } else if (a != kNone && !(h2 > hard)) {
                ok = false;
            }
        }
        // the per-lane traversal's final prune threshold, lowered by a few ulps of its terms
        float thr = fmn(hard, sm) + K;
        thr = thr - 1e-6f * (fabsf(thr) + K + fabsf(res));
        if (!(sm < hard) && a != kNone && !len_below(leaf_d2(la), thr, true)) ok = false;
        if (sm < hard + 1e-5f * (fabsf(hard) + K)) {  // the blend pair decides (or nearly)
            if (s1 != kNone && !len_below(leaf_d2(slot_leaf(s1)), thr, true)) ok = false;
            if (b2 - b1 < K && s2 != kNone && !len_below(leaf_d2(slot_leaf(s2)), thr, true)) ok = false;
        }
    }
    return ok ? DistHit{res, rmat} : dist_mat(S, p, stk);
}

template <class Stk>
__device__ __forceinline__ float wave_dist(const SceneView& S, f3 p, bool need, Stk& stk, float ub = kFltMax) {
    return wave_dist_mat<false>(S, p, need, stk, ub).d;
}

// Wavefront-shared overloads of the march hooks (the per-lane versions and
// their contracts sit beside the loops): all active lanes call together.
// The bodies are templates over the shared stack types (LdsStack,
// ExecStack); the non-template overloads below select them, since an exact
// match beats the per-lane templates.
template <class Stk>
__device__ __forceinline__ bool shadow_step_shared(const SceneView& S, ShadowMarch& m, bool live, float& vis,
                                                   Stk& stk) {
    f3 p = spl(0.0f);
    bool ended = false, need = false;
    if (live) {
        ended = m.pre(vis, p);
        need = !ended;
    }
    const float h = wave_dist_mat<false>(S, p, need, stk, m.dist_bound(), &m.hint).d;
    if (need) ended = m.post(h, vis);
    return live && ended;
}
// A sphere-trace step when only hit / miss matters (no material): 0
// marching, 1 hit, 2 miss.
template <class Stk>
__device__ __forceinline__ int miss_step_shared(const SceneView& S, TraceMarch& m, bool live, Stk& stk) {
    f3 p = spl(0.0f);
    int st = 0;
    bool need = false;
    if (live) {
        if (m.pre(p)) st = 2;
        else need = true;
    }
    const float h = wave_dist_mat<false>(S, p, need, stk, m.dist_bound(), &m.hint).d;
    if (need) st = m.post(h);
    return st;
}
template <class Stk>
__device__ __forceinline__ int trace_step_shared(const SceneView& S, TraceMarch& m, bool live, uint32_t& mat,
                                                 bool& have_mat, Stk& stk) {
    f3 p = spl(0.0f);
    int st = 0;
    bool need = false;
    if (live) {
        if (m.pre(p)) st = 2;
        else need = true;
    }
    const DistHit h = wave_dist_mat<true>(S, p, need, stk, m.dist_bound(), &m.hint);
    if (need) st = m.post(h.d);
    mat = h.mat;
    have_mat = true;
    return st;
}

__device__ __forceinline__ float query_dist(const SceneView& S, f3 p, LdsStack& stk, float ub = kFltMax,
                                            bool need = true) {
    return wave_dist(S, p, need, stk, ub);
}
__device__ __forceinline__ float query_dist(const SceneView& S, f3 p, ExecStack& stk, float ub = kFltMax,
                                            bool need = true) {
    return wave_dist(S, p, need, stk, ub);
}
__device__ __forceinline__ bool shadow_step(const SceneView& S, ShadowMarch& m, bool live, float& vis,
                                            LdsStack& stk) {
    return shadow_step_shared(S, m, live, vis, stk);
}
__device__ __forceinline__ bool shadow_step(const SceneView& S, ShadowMarch& m, bool live, float& vis,
                                            ExecStack& stk) {
    return shadow_step_shared(S, m, live, vis, stk);
}
__device__ __forceinline__ int miss_step(const SceneView& S, TraceMarch& m, bool live, LdsStack& stk) {
    return miss_step_shared(S, m, live, stk);
}
__device__ __forceinline__ int miss_step(const SceneView& S, TraceMarch& m, bool live, ExecStack& stk) {
    return miss_step_shared(S, m, live, stk);
}
__device__ __forceinline__ int miss_step(const SceneView& S, TraceMarch& m, bool live, CountStack& stk) {
    DistHit h;
    return live ? m.step(S, h, stk) : 0;
}
__device__ __forceinline__ int trace_step(const SceneView& S, TraceMarch& m, bool live, uint32_t& mat,
                                          bool& have_mat, LdsStack& stk) {
    return trace_step_shared(S, m, live, mat, have_mat, stk);
}
__device__ __forceinline__ int trace_step(const SceneView& S, TraceMarch& m, bool live, uint32_t& mat,
                                          bool& have_mat, ExecStack& stk) {
    return trace_step_shared(S, m, live, mat, have_mat, stk);
}

// Links per batch of a Jacobi row: the batch's index, weight and cluster
// loads are issued together, then summed in link order (the order, and so
// every bit, of the serial sum).
constexpr uint32_t kRowBatch = 8;
constexpr uint32_t kSliceBatch = 16;  // the same for the sliced rows

// Kernel prologue: stage, then place this lane's stack column after the staged records.
#define IFX_SETUP(LDS, LIGHTS, COUNT) IFX_SETUP_T(LDS, LIGHTS, typename StackOf<COUNT>::T)
// Camera-ray shading kernels (see ShadeStackOf).
#define IFX_SETUP_W(LDS, LIGHTS, COUNT, WAVE) \
    IFX_SETUP_T(LDS, LIGHTS, typename ShadeStackOf<COUNT, WAVE>::T)
#define IFX_SETUP_T(LDS, LIGHTS, ...)                                           \
    extern __shared__ float4 ifx_lds[];                                         \
    SceneView S = Sg;                                                           \
    LightsView L = Lg;                                                          \
    stage_in<LDS, LIGHTS>(ifx_lds, S, L);                                       \
    __VA_ARGS__ stk;                                                            \
    place_stack(stk, ifx_lds, P.scene_f4 + P.light_f4, P.stack_depth);          \
    stk_init(stk);                                                              \
    (void)L

// src/formfactor.rs:70-89
__device__ __forceinline__ uint32_t normal_bucket(f3 n) {
    f3 a = abs3(n);
    if (a.x >= a.y && a.x >= a.z) return n.x >= 0.0f ? 0u : 1u;
    if (a.y >= a.z) return n.y >= 0.0f ? 2u : 3u;
    return n.z >= 0.0f ? 4u : 5u;
}

// Surfel -> cluster candidate geometry and culls (src/formfactor.rs:169-192).
__device__ __forceinline__ bool ff_geom(f3 sp, f3 sn, float4 cp, float4 cn, float& r, f3& dir, float& ff) {
    f3 d = ld3(cp) - sp;
    float r2 = dot3(d, d);
    if (r2 < 1e-9f) return false;
    r = sqrtf(r2);
    dir = d / r;
    float cos_i = dot3(sn, dir);
    float cos_c = dot3(ld3(cn), -dir);
    if (cos_i <= 1e-3f || cos_c <= 1e-3f) return false;
    float area = cp.w;
    ff = cos_i * cos_c * area / (kPi * r2 + area);
    return !(ff < 2e-5f);
}

// ff_geom's decision alone, for the candidate culls.  The cosines and the
// form factor are first estimated with the hardware square root and
// reciprocal (a few ulps) instead of the correctly rounded operations:
// with unit normals (cluster normals are normalised; a surfel normal whose
// squared length is off 1 by more than 1e-4 takes ff_geom directly) each
// estimated cosine lies within 1e-6 of ff_geom's
// (rounding of the direction, the dot product and 1/r, each a few 1e-8
// of |d| / r), so a cosine more than 1e-5 from the 1e-3 threshold, and a
// form factor further from 2e-5 than its relative error bound (4e-6 over
// each cosine plus 1e-5), decide as ff_geom does.  Near a threshold the
// decision is ff_geom's own.
__device__ __forceinline__ bool ff_pass(f3 sp, f3 sn, float4 cp, float4 cn) {
    const f3 d = ld3(cp) - sp;
    const float r2 = dot3(d, d);
    if (r2 < 1e-9f) return false;
    const float sn2 = dot3(sn, sn);
    const bool unit = sn2 >= 0.9999f && sn2 <= 1.0001f;
    const float ir = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(r2));
    const float ci = dot3(sn, d) * ir, cc = -dot3(ld3(cn), d) * ir;
    constexpr float kCosMargin = 1e-5f, kCosErr = 4e-6f;
    if (unit && (ci <= 1e-3f - kCosMargin || cc <= 1e-3f - kCosMargin)) return false;
    if (unit && ci > 1e-3f + kCosMargin && cc > 1e-3f + kCosMargin) {
        const float ff = ((ci * cc) * cp.w) * __builtin_amdgcn_rcpf(kPi * r2 + cp.w);
        const float rel = kCosErr / ci + kCosErr / cc + 1e-5f;
        if (ff > 2e-5f * (1.0f + rel)) return true;
        if (ff < 2e-5f * (1.0f - rel)) return false;
    }
    float r, ff;
    f3 dir;
    return ff_geom(sp, sn, cp, cn, r, dir, ff);
}

// 15-bit Morton code (5 bits per axis) of each surfel's cell in the scene
// box; it only orders work (coherent wavefronts), never enters a result.
__device__ __forceinline__ uint32_t spread5(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 5; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

// 30-bit Morton code (10 bits per axis); its high bits are the coarser codes.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}
constexpr uint32_t kMortonBits = 30;

// Wavefront-private slice of a global work queue.  A wave takes kChunk
// consecutive items per atomic and hands them to its idle lanes in order;
// all fields are wave-uniform.
// Idle lanes wait until at least min_idle (1..64) of them are idle, so the
// lanes refilled together start their rays together.  The link visibility
// and closure kernels refill a whole wavefront at once (kMarchMinIdle): 64
// consecutive items (one cluster and Morton-adjacent surfels, or one
// direction from Morton-adjacent surfels) march side by side, so the shared
// query's union of nodes or listed objects stays near one lane's (stress
// scene link visibility: 311 -> 222 ms before the flat list, although 27 %
// of the lane slots then idle at the batch tails).
struct WaveQueue {
    uint32_t next, end, chunk;
    bool drained;
    uint32_t min_idle = 1;
};

__device__ __forceinline__ bool refill(bool live, WaveQueue& wq, uint32_t* queue, uint32_t total, uint32_t& item) {
    uint64_t idle = __ballot(!live);
    if (!idle || (uint32_t)__popcll(idle) < wq.min_idle) return false;
    if (wq.next >= wq.end) {
        if (wq.drained) return false;
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(queue, wq.chunk);
        base = __shfl(base, 0, 64);
        if (base >= total) {
            wq.drained = true;
            return false;
        }
        wq.next = base;
        wq.end = total - base < wq.chunk ? total : base + wq.chunk;
    }
    uint32_t nidle = (uint32_t)__popcll(idle);
    uint32_t avail = wq.end - wq.next;
    uint32_t take = nidle < avail ? nidle : avail;
    uint32_t rank = lanes_below(idle);
    item = wq.next + rank;
    wq.next += take;
    return !live && rank < take;
}

// One pixel of render_frame (src/main.rs:247-361).  The reference's early
// returns become a per-lane outcome: every march (camera ray, its normal
// taps, shadows, mirror ray) is entered by all calling lanes, lanes done or
// off (`on` false: no pixel; the result is then meaningless) only following
// along, so a wavefront that calls together keeps whole-wavefront queries.
template <class Stk>
__device__ __forceinline__ f3 shade_pixel(const SceneView& S, const LightsView& L, const ShadeArgs& A, uint32_t x,
                                          uint32_t y, uint32_t& local, Stk& stk, bool on = true) {
    const f3 bg = mk(0.02f, 0.03f, 0.05f);
    const CamV& cam = A.cam;
    float px = (2.0f * ((float)x + 0.5f) / (float)A.W - 1.0f) * cam.aspect * cam.tan_half;
    float py = (1.0f - 2.0f * ((float)y + 0.5f) / (float)A.H) * cam.tan_half;
    f3 rd = nrm3((cam.fwd + cam.right * px) + cam.up * py);
    if (A.mode == M_REFERENCE) {  // src/main.rs:249-258: spp paths, seeded per pixel
        if (!on) return bg;
        PathRng rng(((uint64_t)y << 32) | (uint64_t)x);
        f3 acc = spl(0.0f);
        for (uint32_t k = 0; k < A.spp; ++k) acc = acc + ref_path(S, L, cam.origin, rd, A.tmax, 12u, rng, stk);
        return acc / (float)A.spp;
    }
    uint32_t steps = 0;
    Hit hh = trace<true>(S, cam.origin, rd, 1e-3f, A.tmax, steps, stk, on);
    local += steps;
    bool alive = on;
    f3 res = bg;
    if (alive) {
        float surf_t = hh.hit ? hh.t : kFltMax;
        bool tonemapped = A.mode <= M_INDIRECT;
        f3 e;
        if (emitter_hit(L, cam.origin, rd, fmn(surf_t, A.tmax), e)) {
            res = tonemapped ? e : spl(0.9f);
            alive = false;
        } else if (!hh.hit) {
            res = bg;
            alive = false;
        } else if (A.mode == M_NORMALS) {
            res = hh.n * 0.5f + spl(0.5f);
            alive = false;
        } else if (A.mode == M_DEPTH) {
            float d = 1.0f - clmp(hh.t / A.tmax, 0.0f, 1.0f);
            res = spl(d * d);
            alive = false;
        } else if (A.mode == M_STEPS) {
            float f = (float)steps / (float)kMaxSteps;
            res = mk(f, 1.0f - f, 0.2f);
            alive = false;
        }
    }
    MatV m = load_mat(L, alive ? hh.mat : 0u);
    f3 c = spl(0.0f);
    if (A.mode != M_INDIRECT) {
        uint32_t s = 0;
        f3 d = direct(S, L, hh.p, hh.n, -rd, m, s, stk, alive);
        if (alive) {
            local += s;
            c = c + (d + m.emissive);
        }
    }
    if (A.use_gi) {
        if (alive) {
            f3 irr = gather(A.grid, A.spos, A.snrm, A.irr, hh.p, hh.n, A.radius);
            c = c + m.kd * irr;
        }
        const bool glossy = alive && (m.roughness < 0.45f || m.metallic > 0.05f);
        if (any_lane(glossy)) {
            f3 refl = rd - hh.n * (2.0f * dot3(hh.n, rd));
            uint32_t s = 0;
            Hit rh = trace<true>(S, hh.p + hh.n * 1e-3f, refl, 1e-3f, A.tmax, s, stk, glossy);
            f3 refl_col = bg, e2;
            bool rlit = false;  // the mirror ray hit a surface: its direct light and gather
            if (glossy) {
                local += s;
                float rt = rh.hit ? rh.t : kFltMax;
                if (emitter_hit(L, hh.p, refl, fmn(rt, A.tmax), e2)) refl_col = e2;
                else if (rh.hit) rlit = true;
            }
            MatV m2 = load_mat(L, rlit ? rh.mat : 0u);
            uint32_t s2 = 0;
            f3 d2 = direct(S, L, rh.p, rh.n, -refl, m2, s2, stk, rlit);
            if (rlit) {
                local += s2;
                f3 i2 = gather(A.grid, A.spos, A.snrm, A.irr, rh.p, rh.n, A.radius);
                refl_col = (d2 + m2.kd * i2) + m2.emissive;
            }
            if (glossy) {
                f3 f = fresnel_schlick(fmx(dot3(hh.n, -rd), 0.0f), m.f0);
                float gloss = clmp(1.0f - m.roughness / 0.45f, 0.0f, 1.0f);
                gloss = gloss * gloss;
                c = c + (f * refl_col) * gloss;
            }
        }
    }
    return alive ? c : res;
}


// Film::encode: exposure, ACES fit, sRGB through host-built thresholds
// (thr[k-1] = least f32 that the sRGB transfer maps to >= k).
__device__ __forceinline__ uint8_t srgb_lut(float c, const float* __restrict__ thr) {
    c = clmp(c, 0.0f, 1.0f);
    uint32_t lo = 0, hi = 255;  // count of thresholds <= c
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (c >= thr[mid])
            lo = mid + 1;
        else
            hi = mid;
    }
    return (uint8_t)lo;
}

__device__ __forceinline__ void encode_px(float4 px, float exposure, int32_t tonemap, const float* __restrict__ thr,
                                          uint8_t* out) {
    f3 x = ld3(px) * exposure;
    if (tonemap) {
        f3 num = x * (x * 2.51f + spl(0.03f));
        f3 den = x * (x * 2.43f + spl(0.59f)) + spl(0.14f);
        x = min3(max3(num / den, spl(0.0f)), spl(1.0f));
    }
    out[0] = srgb_lut(x.x, thr);
    out[1] = srgb_lut(x.y, thr);
    out[2] = srgb_lut(x.z, thr);
}

// Glue kernels defined in engine.hip and launched by the batch engine too.
__global__ void k_seg_flags(const uint64_t* keys, uint32_t n, uint32_t* flags);
__global__ void k_seg_begin(const uint32_t* flags, const uint32_t* seg, uint32_t n, uint32_t* begin);
__global__ void k_cluster_build(const uint32_t* begin, uint32_t nseg, const uint32_t* members,
                                const float4* __restrict__ spos, const float4* __restrict__ snrm, float4* cpos,
                                float4* cnrm, uint32_t* valid);
__global__ void k_cluster_compact(uint32_t nseg, const uint32_t* valid, const uint32_t* vidx, const uint32_t* begin,
                                  const float4* cpos_raw, const float4* cnrm_raw, float4* cpos, float4* cnrm,
                                  uint2* range);
__global__ void k_row_sum(uint32_t N, const uint32_t* __restrict__ cstart, const float* __restrict__ cand_v,
                          const uint32_t* __restrict__ open, float* sky_out, uint32_t* lcount, float* scale,
                          float* row_sum);
__global__ void k_compact(uint32_t N, const uint32_t* __restrict__ cstart, const float* __restrict__ cand_v,
                          const uint32_t* __restrict__ cand_ci, const uint32_t* __restrict__ lstart,
                          const float* __restrict__ scale, uint32_t* idx, float* w);
__global__ void k_pack3(const float4* __restrict__ src, uint64_t n, float* dst);
__global__ void k_unpack3(const float* __restrict__ src, uint64_t n, float4* dst);

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
void Engine::launch_on(hipStream_t stream, const char* name, F&& f) {
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

// Three-way instrumentation of the persistent tracing kernels: MODE 0 timed,
// 1 counting (CountStack), 2 executed-evaluation tally (ExecStack).
#define LAUNCH_CM(K, MODE, LDS, GRID, DYN, ...)                                          \
    do {                                                                                 \
        if ((MODE) == 1) {                                                               \
            if (LDS)                                                                     \
                K<1, true><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                  \
            else                                                                         \
                K<1, false><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                 \
        } else if ((MODE) == 2) {                                                        \
            if (LDS)                                                                     \
                K<2, true><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                  \
            else                                                                         \
                K<2, false><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                 \
        } else {                                                                         \
            if (LDS)                                                                     \
                K<0, true><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                  \
            else                                                                         \
                K<0, false><<<GRID, kBlock, DYN, stream>>>(__VA_ARGS__);                 \
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
