// Device scene layout and the scene distance field: primitive and CSG object
// distances, BVH nearest-distance traversal, scene distance with material
// attribution and the tetrahedral normal.  Shared by the gfx950 kernels and
// the host-side surfel generator (one implementation, two targets).
//
// Follows the reference's src/sdf.rs:50-173, src/bvh.rs:45-163 and
// src/scene.rs:187-236: the traversal order, the `>=` prune and the strict
// `<` material update are kept exactly, so the same point yields the same
// (distance, material) bits as the reference.
#pragma once
#include "ifx_math.h"

namespace ifx {

constexpr float kFltMax = 3.40282346638528859812e+38f;

enum PrimKind : uint32_t { K_SPHERE = 0, K_ROUNDBOX = 1, K_PLANE = 2, K_CAPSULE = 3, K_TORUS = 4, K_CYLINDER = 5, K_CONE = 6 };

IFX_HD uint32_t fbits(float f) {
    union {
        float f;
        uint32_t u;
    } x;
    x.f = f;
    return x.u;
}
IFX_HD float bitsf(uint32_t u) {
    union {
        float f;
        uint32_t u;
    } x;
    x.u = u;
    return x.f;
}

// One primitive or CSG object: 4 x float4 (64 B).
//   g0 = (anchor.xyz, s0)   anchor = centre / endpoint a / plane normal
//   g1 = (vec.xyz,    s1)   vec    = box half-size minus round / b - a
//   g2 = (s2, s3, kind, 0)
//   g3 = (mat, sub_begin, sub_count, blended) as u32 bits (objects only)
// Per kind: sphere (c, r); roundbox (c, r | b - r); plane (n, h);
// capsule / cylinder (a, r | b - a, |b - a|^2); torus (c, major | -, minor);
// cone (a, ra | b - a, |b - a|^2 | rb, rb - ra).  The precomputed terms are
// the very f32 expressions the reference evaluates per call.
struct ObjRec {
    float4 g0, g1, g2, g3;
};

// BVH node: 2 x float4.  lo.w = leaf start or interior right-child index,
// hi.w = leaf object count (0 = interior); the left child is node + 1.
struct SceneView {
    const float4* obj;     // bounded objects, stored in BVH leaf order
    const float4* planes;  // unbounded objects, always evaluated first
    const float4* subs;    // CSG subtraction primitives
    const float4* nodes;
    const uint32_t* lin;   // lin[i] = slot of object i (linear scan); lin[n_obj + k] = leaf node of slot k
    uint32_t n_obj, n_planes, n_nodes, use_bvh;
    uint32_t n_subs;
    float blend_k;         // EXT blend-group radius, 0 = reference behaviour
    // The same records in global memory (the pointers above may be moved to
    // an LDS copy); the wavefront-shared traversal reads these.
    const float4 *g_obj, *g_planes, *g_subs, *g_nodes;
    const uint32_t* g_lin;
    // Each object's box (lo with the leaf node index's bits in lo.w, hi;
    // 2 x float4 per slot) in global memory, for the flat candidate list of
    // the wavefront-shared query; null: that query always walks the BVH.
    const float4* g_obox;
    IFX_HD void keep_global() {
        g_obj = obj;
        g_planes = planes;
        g_subs = subs;
        g_nodes = nodes;
        g_lin = lin;
    }
};

struct DistHit {
    float d;
    uint32_t mat;
};

IFX_HD float prim_dist(float4 g0, float4 g1, float4 g2, f3 p) {
    uint32_t kind = fbits(g2.z);
    f3 a = ld3(g0);
    switch (kind) {
        case K_SPHERE:
            return len3(p - a) - g0.w;
        case K_ROUNDBOX: {
            f3 q = abs3(p - a) - ld3(g1);
            return (len3(max3(q, spl(0.0f))) + fmn(fmx(q.x, fmx(q.y, q.z)), 0.0f)) - g0.w;
        }
        case K_PLANE:
            return dot3(p, a) - g0.w;
        case K_CAPSULE: {
            f3 pa = p - a;
            f3 ba = ld3(g1);
            float h = clmp(dot3(pa, ba) / g1.w, 0.0f, 1.0f);
            return len3(pa - ba * h) - g0.w;
        }
        case K_TORUS: {
            f3 q = p - a;
            float xz = sqrtf(q.x * q.x + q.z * q.z) - g0.w;
            return sqrtf(xz * xz + q.y * q.y) - g2.x;
        }
        case K_CYLINDER: {
            f3 ba = ld3(g1);
            float baba = g1.w;
            float r = g0.w;
            if (baba < 1e-12f) return len3(p - a) - r;
            f3 pa = p - a;
            float paba = dot3(pa, ba);
            float x = len3(pa * baba - ba * paba) - r * baba;
            float y = fabsf(paba - baba * 0.5f) - baba * 0.5f;
            float x2 = x * x;
            float y2 = y * y * baba;
            float d = fmx(x, y) < 0.0f ? -fmn(x2, y2) : (x > 0.0f ? x2 : 0.0f) + (y > 0.0f ? y2 : 0.0f);
            float sg = (fbits(d) >> 31) ? -1.0f : 1.0f;  // f32::signum, d is never NaN
            return sg * sqrtf(fabsf(d)) / baba;
        }
        default: {  // K_CONE
            f3 ba = ld3(g1);
            float baba = g1.w;
            float ra = g0.w, rb = g2.x, rba = g2.y;
            if (baba < 1e-12f) return len3(p - a) - fmx(ra, rb);
            f3 pa = p - a;
            float papa = dot3(pa, pa);
            float paba = dot3(pa, ba) / baba;
            float x = sqrtf(fmx(papa - paba * paba * baba, 0.0f));
            float cax = fmx(0.0f, x - (paba < 0.5f ? ra : rb));
            float cay = fabsf(paba - 0.5f) - 0.5f;
            float k = rba * rba + baba;
            float f = clmp((rba * (x - ra) + paba * baba) / k, 0.0f, 1.0f);
            float cbx = x - ra - f * rba;
            float cby = paba - f;
            float s = (cbx < 0.0f && cay < 0.0f) ? -1.0f : 1.0f;
            return s * sqrtf(fmn(cax * cax + cay * cay * baba, cbx * cbx + cby * cby * baba));
        }
    }
}

// src/sdf.rs:164-173: base primitive, then each subtraction in order.  P is
// the record pointer type (a constant-address-space pointer lets uniform
// reads go through the scalar cache).
template <class P, class Stk>
IFX_HD float obj_dist_at(P rec, P subs, f3 p, uint32_t& mat, bool& blended, Stk& stk) {
    float4 g0 = rec[0], g1 = rec[1], g2 = rec[2], g3 = rec[3];
    mat = fbits(g3.x);
    uint32_t sb = fbits(g3.y), sc = fbits(g3.z);
    blended = fbits(g3.w) != 0u;
    float d = prim_dist(g0, g1, g2, p);
    stk.count_prim(fbits(g2.z), 0u);
    for (uint32_t s = sb; s < sb + sc; ++s) {
        P q = subs + 4 * s;
        d = fmx(d, -prim_dist(q[0], q[1], q[2], p));
        stk.count_prim(fbits(q[2].z), 2u);
    }
    return d;
}

template <class Stk>
IFX_HD float obj_dist(const float4* __restrict__ rec, const float4* __restrict__ subs, f3 p, uint32_t& mat,
                      bool& blended, Stk& stk) {
    return obj_dist_at(rec, subs, p, mat, blended, stk);
}

// obj_dist for a caller that only compares the value against a running
// minimum: `cut` for a hard object, `cut_bl` for a blended one.  When the
// base primitive is already >= that cut, the subtractions (which only raise
// the value) cannot bring it below, so they are skipped and the base value
// returned: the caller's comparisons come out the same, bit for bit.  The
// counting stack evaluates everything (the reference algorithm's count).
template <class Stk>
IFX_HD float obj_dist_cut(const float4* __restrict__ rec, const float4* __restrict__ subs, f3 p, float cut,
                          float cut_bl, uint32_t& mat, bool& blended, Stk& stk) {
    if (!Stk::kSkipCsg) return obj_dist_at(rec, subs, p, mat, blended, stk);
    float4 g0 = rec[0], g1 = rec[1], g2 = rec[2], g3 = rec[3];
    mat = fbits(g3.x);
    uint32_t sb = fbits(g3.y), sc = fbits(g3.z);
    blended = fbits(g3.w) != 0u;
    float d = prim_dist(g0, g1, g2, p);
    stk.count_prim(fbits(g2.z), 0u);
    if (d >= (blended ? cut_bl : cut)) return d;
    for (uint32_t s = sb; s < sb + sc; ++s) {
        const float4* q = subs + 4 * s;
        d = fmx(d, -prim_dist(q[0], q[1], q[2], p));
        stk.count_prim(fbits(q[2].z), 2u);
    }
    return d;
}

// src/bvh.rs:45-50
IFX_HD float aabb_dist(f3 p, float4 lo, float4 hi) {
    f3 d = max3(max3(ld3(lo) - p, p - ld3(hi)), spl(0.0f));
    return len3(d);
}

// EXT: polynomial smooth minimum (Quilez), k > 0.
IFX_HD float smin_poly(float a, float b, float k) {
    float h = fmx(k - fabsf(a - b), 0.0f) / k;
    return fmn(a, b) - h * h * k * 0.25f;
}

// Algorithmic FP32 operations (add, sub, mul, div, sqrt, min, max, abs;
// compares and selects not counted) of one primitive distance, by kind,
// counted from prim_dist above; an AABB distance test; and the per-sample
// overhead of a march step (sample point, epsilon / cone update, advance).
constexpr uint32_t kPrimOps[8] = {10u, 23u, 6u, 24u, 13u, 40u, 57u, 0u};
constexpr uint32_t kAabbOps = 18u;
constexpr uint32_t kEvalOps = 15u;

// Per-lane traversal stacks.  Host: a plain array.  Device: a column of a
// block-wide LDS array (entry i of lane t at base[i * stride]), so a
// divergent push / pop never spills to scratch memory.  The count_* hooks
// are empty except in CountStack, the instrumented variant that tallies
// algorithmic operations for the roofline report.
// Each entry is a node index and that node's AABB distance, computed by its
// parent when it was pushed (the reference recomputes the same value on pop,
// src/bvh.rs:116; keeping it is bit-identical and skips the node's loads).
// kShared: the device march loops over this stack answer the active lanes'
// queries with one wavefront-shared traversal (the hooks below).
struct ArrStack {
    static constexpr bool kShared = false;
    static constexpr bool kSkipCsg = true;
    uint32_t v[64];
    float dv[64];
    IFX_HD void put(int i, uint32_t x, float d) {
        v[i] = x;
        dv[i] = d;
    }
    IFX_HD uint32_t get(int i, float& d) const {
        d = dv[i];
        return v[i];
    }
    IFX_HD void count_prim(uint32_t, uint32_t) {}
    IFX_HD void count_aabb(uint32_t) {}
    IFX_HD void count_eval() {}
};

struct LdsStack {
    static constexpr bool kShared = true;
    static constexpr bool kSkipCsg = true;
    uint16_t* base;
    float* dbase;
    uint32_t stride;
    __device__ __forceinline__ void put(int i, uint32_t x, float d) {
        base[i * stride] = (uint16_t)x;
        dbase[i * stride] = d;
    }
    __device__ __forceinline__ uint32_t get(int i, float& d) const {
        d = dbase[i * stride];
        return base[i * stride];
    }
    __device__ __forceinline__ void count_prim(uint32_t, uint32_t) {}
    __device__ __forceinline__ void count_aabb(uint32_t) {}
    __device__ __forceinline__ void count_eval() {}
    __device__ __forceinline__ void count_wave_prims(uint32_t) {}
    __device__ __forceinline__ void count_query(bool) {}
    __device__ __forceinline__ void count_listed() {}
};

// LdsStack of the instrumented instantiation that tallies the primitive
// evaluations the timed traversal executes, for comparison with the
// reference traversal's (CountStack): `xe` per lane on the per-lane paths,
// `xw` per wavefront for the shared traversal (each object it opens is
// evaluated by every active lane).  A separate instantiation: the counters
// cost the timed kernels registers.
// It also counts, per wavefront, the shared queries answered by the flat
// list (xf) or by the BVH walk (xb) and the objects the flat lists held (xl).
struct ExecStack : LdsStack {
    uint32_t xe, xw, xf, xb, xl;
    __device__ __forceinline__ void count_prim(uint32_t, uint32_t) { ++xe; }
    __device__ __forceinline__ void count_wave_prims(uint32_t lanes) { xw += lanes; }
    __device__ __forceinline__ void count_query(bool flat) { flat ? ++xf : ++xb; }
    __device__ __forceinline__ void count_listed() { ++xl; }
};

struct CountStack {
    static constexpr bool kShared = false;
    static constexpr bool kSkipCsg = false;
    uint16_t* base;
    float* dbase;
    uint32_t stride;
    unsigned long long ops;
    uint32_t prims;  // primitive evaluations of the reference traversal
    __device__ __forceinline__ void put(int i, uint32_t x, float d) {
        base[i * stride] = (uint16_t)x;
        dbase[i * stride] = d;
    }
    __device__ __forceinline__ uint32_t get(int i, float& d) const {
        d = dbase[i * stride];
        return base[i * stride];
    }
    __device__ __forceinline__ void count_prim(uint32_t kind, uint32_t extra) {
        ops += kPrimOps[kind & 7u] + extra;
        ++prims;
    }
    __device__ __forceinline__ void count_aabb(uint32_t n) { ops += kAabbOps * n; }
    __device__ __forceinline__ void count_eval() { ops += kEvalOps; }
};

// Blend-group state (EXT): nearest hard hit, nearest two blended objects.
struct BlendState {
    DistHit hard, b1;
    float b2;
};

IFX_HD void blend_take(BlendState& bs, float d, uint32_t mat, bool blended) {
    if (blended) {
        if (d < bs.b1.d) {
            bs.b2 = bs.b1.d;
            bs.b1.d = d;
            bs.b1.mat = mat;
        } else if (d < bs.b2) {
            bs.b2 = d;
        }
    } else if (d < bs.hard.d) {
        bs.hard.d = d;
        bs.hard.mat = mat;
    }
}

// src/bvh.rs:97-150 (BLEND=false) and the EXT blend-group variant.
template <bool BLEND, class Stk>
IFX_HD void bvh_query(const SceneView& S, f3 p, BlendState& bs, Stk& stk) {
    if (S.n_nodes == 0) return;
    int sp = 0;
    stk.count_aabb(1u);
    stk.put(sp++, 0u, aabb_dist(p, S.nodes[0], S.nodes[1]));
    while (sp > 0) {
        float dn;
        uint32_t ni = stk.get(--sp, dn);
        float thr = BLEND ? fmn(bs.hard.d, smin_poly(bs.b1.d, bs.b2, S.blend_k)) + S.blend_k : bs.hard.d;
        if (dn >= thr) continue;
        float4 lo = S.nodes[2 * ni], hi = S.nodes[2 * ni + 1];
        uint32_t w0 = fbits(lo.w), cnt = fbits(hi.w);
        if (cnt > 0) {
            for (uint32_t k = w0; k < w0 + cnt; ++k) {
                uint32_t mat;
                bool bl;
                float d = obj_dist_cut(S.obj + 4 * k, S.subs, p, bs.hard.d, BLEND ? bs.b2 : bs.hard.d, mat, bl, stk);
                if (BLEND) {
                    blend_take(bs, d, mat, bl);
                } else if (d < bs.hard.d) {
                    bs.hard.d = d;
                    bs.hard.mat = mat;
                }
            }
        } else {
            uint32_t l = ni + 1, r = w0;
            float dl = aabb_dist(p, S.nodes[2 * l], S.nodes[2 * l + 1]);
            float dr = aabb_dist(p, S.nodes[2 * r], S.nodes[2 * r + 1]);
            stk.count_aabb(2u);
            if (dl < dr) {
                stk.put(sp++, r, dr);
                stk.put(sp++, l, dl);
            } else {
                stk.put(sp++, l, dl);
                stk.put(sp++, r, dr);
            }
        }
    }
}

// src/scene.rs:191-208
template <class Stk>
IFX_HD DistHit dist_mat(const SceneView& S, f3 p, Stk& stk) {
    BlendState bs;
    bs.hard = DistHit{kFltMax, 0u};
    bs.b1 = DistHit{kFltMax, 0u};
    bs.b2 = kFltMax;
    stk.count_eval();
    for (uint32_t i = 0; i < S.n_planes; ++i) {
        uint32_t mat;
        bool bl;
        float d = obj_dist_cut(S.planes + 4 * i, S.subs, p, bs.hard.d, bs.hard.d, mat, bl, stk);
        if (d < bs.hard.d) {
            bs.hard.d = d;
            bs.hard.mat = mat;
        }
    }
    if (!(S.blend_k > 0.0f)) {
        if (S.use_bvh) {
            bvh_query<false>(S, p, bs, stk);
        } else {
            for (uint32_t i = 0; i < S.n_obj; ++i) {
                uint32_t mat;
                bool bl;
                float d = obj_dist_cut(S.obj + 4 * S.lin[i], S.subs, p, bs.hard.d, bs.hard.d, mat, bl, stk);
                if (d < bs.hard.d) {
                    bs.hard.d = d;
                    bs.hard.mat = mat;
                }
            }
        }
        return bs.hard;
    }
    if (S.use_bvh) {
        bvh_query<true>(S, p, bs, stk);
    } else {
        for (uint32_t i = 0; i < S.n_obj; ++i) {
            uint32_t mat;
            bool bl;
            float d = obj_dist_cut(S.obj + 4 * S.lin[i], S.subs, p, bs.hard.d, bs.b2, mat, bl, stk);
            blend_take(bs, d, mat, bl);
        }
    }
    float s = smin_poly(bs.b1.d, bs.b2, S.blend_k);
    if (s < bs.hard.d) return DistHit{s, bs.b1.mat};
    return bs.hard;
}

template <class Stk>
IFX_HD float dist(const SceneView& S, f3 p, Stk& stk) {
    return dist_mat(S, p, stk).d;
}

// Hooks of the march loops: the distance at one point per active lane, and
// whether any active lane is still marching.  These per-lane versions serve
// the host and the instrumented device stack; device code with an LDS stack
// gets wavefront-shared overloads (found by argument-dependent lookup where
// the loops are instantiated) that answer all active lanes with one
// traversal and return the same values.
// (ub: an upper bound on the distance at p when the caller has one; the
// per-lane version does not need it.  need false: the lane wants no value,
// it only follows its wavefront's shared query, and gets 0.)
template <class Stk>
IFX_HD float query_dist(const SceneView& S, f3 p, Stk& stk, float ub = kFltMax, bool need = true) {
    (void)ub;
    return need ? dist(S, p, stk) : 0.0f;
}

// The material at a point (a trace's hit when its step did not report it).
template <class Stk>
IFX_HD uint32_t query_mat(const SceneView& S, f3 p, Stk& stk) {
    return dist_mat(S, p, stk).mat;
}

IFX_HD bool any_lane(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(v) != 0ull;
#else
    return v;
#endif
}

// src/scene.rs:218-236: tetrahedral 4-tap gradient.
template <class Stk>
IFX_HD f3 scene_normal(const SceneView& S, f3 p, Stk& stk, float d0 = kFltMax, bool need = true) {
    const float H = 5e-4f;
    // d0: the distance at p, when known; each tap lies H sqrt(3) < 8.7e-4 away
    // (need false: the lane only follows along, its result is meaningless)
    const float ub = d0 < kFltMax ? d0 + 8.7e-4f : kFltMax;
    f3 n = spl(0.0f);
    n = n + mk(1.0f, -1.0f, -1.0f) * query_dist(S, p + mk(1.0f, -1.0f, -1.0f) * H, stk, ub, need);
    n = n + mk(-1.0f, -1.0f, 1.0f) * query_dist(S, p + mk(-1.0f, -1.0f, 1.0f) * H, stk, ub, need);
    n = n + mk(-1.0f, 1.0f, -1.0f) * query_dist(S, p + mk(-1.0f, 1.0f, -1.0f) * H, stk, ub, need);
    n = n + mk(1.0f, 1.0f, 1.0f) * query_dist(S, p + mk(1.0f, 1.0f, 1.0f) * H, stk, ub, need);
    float l = len3(n);
    if (l > 1e-12f) return n / l;
    return mk(0.0f, 1.0f, 0.0f);
}

}  // namespace ifx
