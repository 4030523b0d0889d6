// Ray primitives and lighting shared by the gfx950 kernels (and the host
// surfel generator for `trace`-free parts): sphere tracing, cone soft
// shadows, quad form factors, direct light, emitter visibility and the
// surfel irradiance gather.
//
// Follows src/trace.rs:27-148, src/ltc.rs:25-127, src/shade.rs:10-181 and
// src/surfel.rs:127-140,262-293 of the reference, operation for operation.
// EXT terms (sphere / point emitters) are marked and vanish on scenes that
// have none.
#pragma once
#include "ifx_sdf.h"

namespace ifx {

constexpr uint32_t kMaxSteps = 128;    // src/trace.rs:10
constexpr uint32_t kShadowSteps = 96;  // src/trace.rs:11

// Quad light: 7 x float4 = verts[0..3], (emit, area), (centre, radius), (normal, -).
// radius = max(sqrt(area / pi), 1e-4) and normal = normalize((v1-v0) x (v3-v0))
// are the reference's per-call expressions, evaluated once on the host.
// Sphere light (EXT): (centre, r), (emit, -).  Point light (EXT): (pos, -), (emit, -).
// Material: (albedo, roughness), (emissive, metallic), (f0, -), (kd, -).
struct LightsView {
    const float4* quad;
    const float4* sph;
    const float4* pnt;
    const float4* mat;
    uint32_t n_quad, n_sph, n_pnt, n_mat;
};

struct Hit {
    float t;
    f3 p, n;
    uint32_t mat;
    bool hit;
};

// Sphere-tracing march of src/trace.rs:27-58 as a resumable state machine:
// begin() then step() until it reports the end.  trace() below is the plain
// loop over it; persistent kernels interleave marches of many rays per lane.
struct TraceMarch {
    f3 ro, rd;
    float t, tmax, pd, ld;  // pd: previous step's distance; ld: the last sample's
    uint32_t steps;
    uint32_t hint;  // wavefront-shared queries: the object deciding the last sample
    IFX_HD void begin(f3 ro_, f3 rd_, float tmin, float tmax_) {
        ro = ro_;
        rd = rd_;
        t = tmin;
        tmax = tmax_;
        pd = kFltMax;
        ld = kFltMax;
        steps = 0;
        hint = 0xffffffffu;
    }
    // An upper bound on the scene distance at the next sample: the previous
    // sample's distance plus the step taken from it (the step is that
    // distance), the scene SDF being 1-Lipschitz; kFltMax before the first.
    IFX_HD float dist_bound() const { return pd < kFltMax ? pd + pd : kFltMax; }
    // One sample.  Returns 0 while marching, 1 on a hit (h.mat, point at t),
    // 2 on a miss.  On a hit `steps` already includes the normal's 4 taps.
    template <class Stk>
    IFX_HD int step(const SceneView& S, DistHit& h, Stk& stk) {
        f3 p;
        if (pre(p)) return 2;
        h = dist_mat(S, p, stk);
        return post(h.d);
    }
    // The step in two halves around the distance query (for callers that
    // need only the distance): pre is true when the step budget is spent
    // (a miss), else gives the sample point; post takes the distance there.
    IFX_HD bool pre(f3& p) const {
        if (steps >= kMaxSteps) return true;
        p = ro + rd * t;
        return false;
    }
    IFX_HD int post(float d) {
        steps += 1;
        ld = d;
        float eps = 1e-4f * fmx(t, 1.0f);
        if (d < eps) {
            steps += 4;
            return 1;
        }
        pd = d;
        t += d;
        return t > tmax ? 2 : 0;
    }
};

// src/trace.rs:27-58.  NORMAL=false skips the 4-tap normal (closure rays
// discard it); the step count still includes its 4 evaluations.
// One sample of a live lane's trace: 0 marching, 1 hit, 2 miss (0 for a
// lane that is not live).  The per-lane version also reports the hit's
// material (have_mat); see query_dist for the wavefront-shared overload.
template <class Stk>
IFX_HD int trace_step(const SceneView& S, TraceMarch& m, bool live, uint32_t& mat, bool& have_mat, Stk& stk) {
    if (!live) return 0;
    DistHit h;
    const int st = m.step(S, h, stk);
    mat = h.mat;
    have_mat = true;
    return st;
}

// go false: the lane marches nothing (a miss with 0 steps) but follows its
// wavefront's shared queries, the hit normals' included, so lanes that call
// together keep whole-wavefront queries.
template <bool NORMAL, class Stk>
IFX_HD Hit trace(const SceneView& S, f3 ro, f3 rd, float tmin, float tmax, uint32_t& steps_out, Stk& stk,
                 bool go = true) {
    TraceMarch m;
    m.begin(ro, rd, tmin, tmax);
    int st = 0;
    bool have_mat = false;
    uint32_t mat = 0;
    if constexpr (Stk::kShared) {
        bool live = go;
        while (any_lane(live)) {
            uint32_t mm = 0;
            bool hm = false;
            const int s = trace_step(S, m, live, mm, hm, stk);
            if (live && s) {
                st = s;
                mat = mm;
                have_mat = hm;
                live = false;
            }
        }
    } else if (go) {
        DistHit h;
        while ((st = m.step(S, h, stk)) == 0) {
        }
        mat = h.mat;
        have_mat = true;
    }
    steps_out = m.steps;
    Hit r;
    const f3 hp = ro + rd * m.t;
    f3 hn = spl(0.0f);
    if constexpr (NORMAL) hn = scene_normal(S, hp, stk, m.ld, st == 1);  // every calling lane enters
    if (st == 1) {
        r.t = m.t;
        r.p = hp;
        // (the sample's material: the same query as the step's)
        if (!have_mat) mat = query_mat(S, r.p, stk);
        r.n = hn;
        r.mat = mat;
        r.hit = true;
    } else {
        r.t = 0.0f;
        r.p = spl(0.0f);
        r.n = spl(0.0f);
        r.mat = 0;
        r.hit = false;
    }
    return r;
}

// Cone soft shadow of src/trace.rs:84-124 as a resumable state machine.
struct ShadowMarch {
    f3 ro, rd;
    float t, tmax, k, ph, res;
    uint32_t steps;
    uint32_t hint;  // wavefront-shared queries: the object deciding the last sample
    IFX_HD void begin(f3 ro_, f3 rd_, float tmin, float tmax_, float k_) {
        ro = ro_;
        rd = rd_;
        t = tmin;
        tmax = tmax_;
        k = k_;
        ph = 1e20f;
        res = 1.0f;
        steps = 0;
        hint = 0xffffffffu;
    }
    // An upper bound on the scene distance at the next sample: the previous
    // sample's distance plus the step taken from it (1-Lipschitz SDF);
    // kFltMax before the first sample.
    IFX_HD float dist_bound() const { return ph < 1e20f ? ph + clmp(ph, 0.006f, 0.12f) : kFltMax; }
    // A step in two halves around the distance query, so a caller can
    // answer the query for a whole wavefront at once.  pre: true when the
    // march has ended (vis set), else the sample point.
    IFX_HD bool pre(float& vis, f3& p) const {
        if (!(steps < kShadowSteps && t < tmax)) {
            vis = clmp(res, 0.0f, 1.0f);
            return true;
        }
        p = ro + rd * t;
        return false;
    }
    // post: consumes the distance h at the sample; true when the march has ended.
    IFX_HD bool post(float h, float& vis) {
        steps += 1;
        if (h < 1e-4f) {
            vis = 0.0f;
            return true;
        }
        float y, d;
        if (h < ph) {
            y = h * h / (2.0f * ph);
            d = sqrtf(fmx(h * h - y * y, 0.0f));
        } else {
            y = 0.0f;
            d = h;
        }
        res = fmn(res, k * d / fmx(t - y, 1e-4f));
        if (res < 1e-3f) {
            vis = 0.0f;
            return true;
        }
        ph = h;
        t += clmp(h, 0.006f, 0.12f);
        return false;
    }
    // One sample.  Returns true when the march has ended, with `vis` set.
    template <class Stk>
    IFX_HD bool step(const SceneView& S, float& vis, Stk& stk) {
        f3 p;
        if (pre(vis, p)) return true;
        return post(dist(S, p, stk), vis);
    }
};

// src/trace.rs:84-124
// One sample of a live lane's shadow march; true when it has ended (vis
// set).  Per-lane version; see query_dist for the wavefront-shared overload.
template <class Stk>
IFX_HD bool shadow_step(const SceneView& S, ShadowMarch& m, bool live, float& vis, Stk& stk) {
    return live && m.step(S, vis, stk);
}

// `go` false: the lane only follows along (shared stacks: its wavefront's
// other lanes can then keep whole-wavefront queries) and gets vis 0, steps 0.
template <class Stk>
IFX_HD float soft_shadow(const SceneView& S, f3 ro, f3 rd, float tmin, float tmax, float k, uint32_t& steps_out,
                         Stk& stk, bool go = true) {
    ShadowMarch m;
    m.begin(ro, rd, tmin, tmax, k);
    float vis = 0.0f;
    if constexpr (Stk::kShared) {
        bool live = go;
        while (any_lane(live)) {
            float v = 0.0f;
            if (shadow_step(S, m, live, v, stk)) {
                vis = v;
                live = false;
            }
        }
    } else if (go) {
        while (!m.step(S, vis, stk)) {
        }
    }
    steps_out = m.steps;
    return vis;
}

// src/trace.rs:128-148
IFX_HD bool intersect_quad(const float4* L, f3 ro, f3 rd, float& t_out) {
    f3 v0 = ld3(L[0]);
    f3 e1 = ld3(L[1]) - v0;
    f3 e2 = ld3(L[3]) - v0;
    f3 n = cross3(e1, e2);
    float denom = dot3(n, rd);
    if (fabsf(denom) < 1e-9f) return false;
    float t = dot3(v0 - ro, n) / denom;
    if (t < 1e-4f) return false;
    f3 p = (ro + rd * t) - v0;
    float u = dot3(p, e1) / dot3(e1, e1);
    float v = dot3(p, e2) / dot3(e2, e2);
    if (0.0f <= u && u <= 1.0f && 0.0f <= v && v <= 1.0f) {
        t_out = t;
        return true;
    }
    return false;
}

// src/ltc.rs:88-96 (Duff et al. branchless frame)
IFX_HD void frame(f3 n, f3& t, f3& b) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float bb = n.x * n.y * a;
    t = mk(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
    b = mk(bb, s + n.y * n.y * a, -n.y);
}

// One edge term of the Lambert polygon sum (src/ltc.rs:72-80).
IFX_HD float edge_term(f3 a, f3 b) {
    float cos_t = clmp(dot3(a, b), -1.0f, 1.0f);
    float theta = acosf(cos_t);
    f3 cr = cross3(a, b);
    float len = len3(cr);
    return len > 1e-9f ? theta * (cr.z / len) : 0.0f;
}

IFX_HD f3 unit_or_keep(f3 v) {
    float len = len3(v);
    return len > 1e-9f ? v / len : v;
}

// src/ltc.rs:25-83 on the quad given in world space (to_local, :108-115,
// reverses the winding on the way in).  The horizon-clipped polygon is
// streamed vertex by vertex instead of stored: each emitted vertex is
// normalised and closes the edge from its predecessor, and the closing edge
// back to the first vertex is added last, which is the reference's edge
// order (edges 0..n-1, wrapping).  No per-lane array, so nothing spills.
IFX_HD float quad_ff(const float4* L, f3 p, f3 t, f3 b, f3 n) {
    f3 l0, l1, l2, l3;
    {
        f3 d = ld3(L[3]) - p;
        l0 = mk(dot3(d, t), dot3(d, b), dot3(d, n));
        d = ld3(L[2]) - p;
        l1 = mk(dot3(d, t), dot3(d, b), dot3(d, n));
        d = ld3(L[1]) - p;
        l2 = mk(dot3(d, t), dot3(d, b), dot3(d, n));
        d = ld3(L[0]) - p;
        l3 = mk(dot3(d, t), dot3(d, b), dot3(d, n));
    }
    f3 first = spl(0.0f), prev = spl(0.0f);
    int m = 0;
    float sum = 0.0f;
    auto emit = [&](f3 v) {
        v = unit_or_keep(v);
        if (m == 0)
            first = v;
        else
            sum += edge_term(prev, v);
        prev = v;
        ++m;
    };
    auto clip_edge = [&](f3 a, f3 c) {
        float da = a.z, dc = c.z;
        if (da >= 0.0f) emit(a);
        if ((da >= 0.0f) != (dc >= 0.0f)) {
            float tt = da / (da - dc);
            emit(a + (c - a) * tt);
        }
    };
    clip_edge(l0, l1);
    clip_edge(l1, l2);
    clip_edge(l2, l3);
    clip_edge(l3, l0);
    if (m < 3) return 0.0f;
    sum += edge_term(prev, first);
    return fmx(sum * 0.15915494f, 0.0f);
}

// src/ltc.rs:119-127
IFX_HD f3 closest_on_quad(const float4* L, f3 q) {
    f3 o = ld3(L[0]);
    f3 e1 = ld3(L[1]) - o;
    f3 e2 = ld3(L[3]) - o;
    f3 d = q - o;
    float u = clmp(dot3(d, e1) / dot3(e1, e1), 0.0f, 1.0f);
    float v = clmp(dot3(d, e2) / dot3(e2, e2), 0.0f, 1.0f);
    return (o + e1 * u) + e2 * v;
}

// src/shade.rs:10-31
IFX_HD f3 fresnel_schlick(float cos_theta, f3 f0) {
    float m = clmp(1.0f - cos_theta, 0.0f, 1.0f);
    float m2 = m * m;
    return f0 + (spl(1.0f) - f0) * (m2 * m2 * m);
}
IFX_HD float d_ggx(float n_dot_h, float alpha) {
    float a2 = alpha * alpha;
    float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
    return a2 / fmx(kPi * d * d, 1e-9f);
}
IFX_HD float v_smith(float n_dot_v, float n_dot_l, float alpha) {
    float a2 = alpha * alpha;
    float gv = n_dot_l * sqrtf(n_dot_v * n_dot_v * (1.0f - a2) + a2);
    float gl = n_dot_v * sqrtf(n_dot_l * n_dot_l * (1.0f - a2) + a2);
    return 0.5f / fmx(gv + gl, 1e-9f);
}

// EXT: horizon-clipped sphere form factor (Snyder 1996; Lagarde & de
// Rousiers 2014, "Moving Frostbite to PBR", sec. 4.7.2), divided by pi.
IFX_HD float sphere_ff(float cos_t, float s2) {
    float ill;
    if (cos_t * cos_t > s2) {
        ill = kPi * s2 * clmp(cos_t, 0.0f, 1.0f);
    } else {
        float sin_t = sqrtf(fmx(1.0f - cos_t * cos_t, 0.0f));
        float x = sqrtf(1.0f / s2 - 1.0f);
        float y = -x * (cos_t / sin_t);
        float sty = sin_t * sqrtf(fmx(1.0f - y * y, 0.0f));
        ill = (cos_t * acosf(clmp(y, -1.0f, 1.0f)) - x * sty) * s2 + atanf(sty / x);
    }
    return fmx(ill, 0.0f) / kPi;
}

struct ExtGeo {
    f3 ldir;
    float dist, ff, radius, inv_d2;
};

IFX_HD bool sphere_geo(const float4* L, f3 p, f3 n, ExtGeo& g) {
    f3 c = ld3(L[0]);
    float r = L[0].w;
    f3 to = c - p;
    float dist = len3(to);
    if (!(dist > r)) return false;
    g.ldir = to / dist;
    g.dist = dist;
    float cos_t = dot3(n, g.ldir);
    float s2 = (r * r) / (dist * dist);
    g.ff = sphere_ff(cos_t, s2);
    g.radius = fmx(r, 1e-4f);
    g.inv_d2 = 0.0f;
    return g.ff > 1e-6f;
}

IFX_HD bool point_geo(const float4* L, f3 p, f3 n, ExtGeo& g) {
    f3 to = ld3(L[0]) - p;
    float d2 = dot3(to, to);
    float dist = sqrtf(d2);
    if (!(dist > 1e-6f)) return false;
    g.ldir = to / dist;
    g.dist = dist;
    float cos_t = dot3(n, g.ldir);
    if (cos_t <= 0.0f) return false;
    g.ff = cos_t / (kPi * d2);
    g.radius = 1e-4f;
    g.inv_d2 = 1.0f / d2;
    return true;
}

// go false: the lane only follows along (vis 0, no steps; g is not read).
template <class Stk>
IFX_HD float ext_shadow(const SceneView& S, f3 p, f3 n, const ExtGeo& g, uint32_t& steps, Stk& stk,
                        bool go = true) {
    float k = go ? clmp(g.dist / g.radius, 1.5f, 96.0f) : 1.5f;
    float end = go ? fmx(g.dist - g.radius * 1.25f, 0.05f) : 0.05f;
    uint32_t s = 0;
    float vis = soft_shadow(S, p + n * 3e-3f, go ? g.ldir : spl(0.0f), 0.015f, end, k, s, stk, go);
    steps += s;
    return vis;
}

// Light `li` (quad lights first, then EXT sphere and point lights) of the
// diffuse direct term (src/shade.rs:137-162).  Returns false when the
// reference's loop would `continue` past this light without adding.  The
// culls come first and the shadow march after them is entered by every
// calling lane (culled lanes, and lanes with `on` false, only follow along),
// so a wavefront that calls together keeps whole-wavefront queries.
template <class Stk>
IFX_HD bool diffuse_term(const SceneView& S, const LightsView& Lv, uint32_t li, f3 p, f3 n, f3 tan, f3 bit, f3 albedo,
                         f3& term, float& g, uint32_t& steps, Stk& stk, bool on = true) {
    bool go = on;
    f3 ldir = spl(0.0f), emit = spl(0.0f);
    float ff = 0.0f, k = 1.5f, end = 0.05f;
    if (li < Lv.n_quad) {
        const float4* L = Lv.quad + 7 * li;
        if (go) {
            f3 lc = ld3(L[5]);
            f3 to_light = lc - p;
            go = !(dot3(ld3(L[6]), -nrm3(to_light)) <= 0.0f);
            if (go) {
                ff = quad_ff(L, p, tan, bit, n);
                go = !(ff <= 1e-6f);
            }
            if (go) {
                float dist = len3(to_light);
                float radius = L[5].w;
                k = clmp(dist / radius, 1.5f, 96.0f);
                end = fmx(dist - radius * 1.25f, 0.05f);
                ldir = to_light / dist;
                emit = ld3(L[4]);
            }
        }
    } else {
        const uint32_t le = li - Lv.n_quad;
        const bool sph = le < Lv.n_sph;  // EXT
        const float4* L = sph ? Lv.sph + 2 * le : Lv.pnt + 2 * (le - Lv.n_sph);
        if (go) {
            ExtGeo geo;
            go = sph ? sphere_geo(L, p, n, geo) : point_geo(L, p, n, geo);
            if (go) {
                k = clmp(geo.dist / geo.radius, 1.5f, 96.0f);
                end = fmx(geo.dist - geo.radius * 1.25f, 0.05f);
                ldir = geo.ldir;
                ff = geo.ff;
                emit = ld3(L[1]);
            }
        }
    }
    uint32_t s = 0;
    float vis = soft_shadow(S, p + n * 3e-3f, ldir, 0.015f, end, k, s, stk, go);
    if (!go) return false;
    steps += s;
    if (vis <= 1e-4f) return false;
    g = ff * vis;
    term = (albedo * emit) * g;
    return true;
}

IFX_HD uint32_t n_lights(const LightsView& Lv) { return Lv.n_quad + Lv.n_sph + Lv.n_pnt; }

// src/shade.rs:133-164: diffuse direct light, the radiosity E term.
template <class Stk>
IFX_HD f3 direct_diffuse(const SceneView& S, const LightsView& Lv, f3 p, f3 n, f3 albedo, uint32_t& steps_out,
                         Stk& stk) {
    f3 out = spl(0.0f);
    uint32_t steps = 0;
    f3 tan, bit;
    frame(n, tan, bit);
    const uint32_t nl = n_lights(Lv);
    for (uint32_t li = 0; li < nl; ++li) {
        f3 t;
        float g;
        if (diffuse_term(S, Lv, li, p, n, tan, bit, albedo, t, g, steps, stk)) out = out + t;
    }
    steps_out = steps;
    return out;
}

struct MatV {
    f3 albedo, emissive, f0, kd;
    float roughness, metallic;
};

IFX_HD MatV load_mat(const LightsView& Lv, uint32_t id) {
    const float4* m = Lv.mat + 4 * (id < Lv.n_mat ? id : 0u);
    MatV r;
    r.albedo = ld3(m[0]);
    r.roughness = m[0].w;
    r.emissive = ld3(m[1]);
    r.metallic = m[1].w;
    r.f0 = ld3(m[2]);
    r.kd = ld3(m[3]);
    return r;
}

IFX_HD f3 ggx_lobe(f3 n, f3 v, f3 l, float n_dot_v, float n_dot_l, float wide, float energy, f3 f0) {
    f3 h = nrm3(l + v);
    float d = d_ggx(fmx(dot3(n, h), 0.0f), wide);
    float vis_s = v_smith(n_dot_v, n_dot_l, wide);
    f3 f = fresnel_schlick(fmx(dot3(l, h), 0.0f), f0);
    return f * (d * vis_s * energy * n_dot_l);
}

// src/shade.rs:50-125: diffuse + representative-point GGX specular.
// on false: the lane only follows its wavefront's shadow marches (the culls
// of each light come first, then every calling lane enters the march).
template <class Stk>
IFX_HD f3 direct(const SceneView& S, const LightsView& Lv, f3 p, f3 n, f3 v, const MatV& m, uint32_t& steps_out,
                 Stk& stk, bool on = true) {
    f3 out = spl(0.0f);
    uint32_t steps = 0;
    f3 tan, bit;
    frame(n, tan, bit);
    float n_dot_v = fmx(dot3(n, v), 1e-4f);
    float alpha = fmx(m.roughness * m.roughness, 1e-3f);
    for (uint32_t li = 0; li < Lv.n_quad; ++li) {
        const float4* L = Lv.quad + 7 * li;
        f3 lc = ld3(L[5]);
        f3 to_light = lc - p;
        f3 plane_n = ld3(L[6]);
        bool go = on && !(dot3(plane_n, -nrm3(to_light)) <= 0.0f);
        float ff = 0.0f;
        if (go) {
            ff = quad_ff(L, p, tan, bit, n);
            go = !(ff <= 1e-6f);
        }
        float radius = L[5].w, k = 1.5f, end = 0.05f;
        f3 ldir = spl(0.0f);
        if (go) {
            float dist = len3(to_light);
            k = clmp(dist / radius, 1.5f, 96.0f);
            ldir = to_light / dist;
            end = fmx(dist - radius * 1.25f, 0.05f);
        }
        uint32_t s = 0;
        float vis = soft_shadow(S, p + n * 3e-3f, ldir, 0.015f, end, k, s, stk, go);
        if (!go) continue;
        steps += s;
        if (vis <= 1e-4f) continue;
        f3 radiance = ld3(L[4]) * vis;
        out = out + (m.kd * radiance) * ff;
        f3 r = (-v) - n * (2.0f * dot3(n, -v));
        float denom = dot3(r, plane_n);
        f3 rep = lc;
        if (fabsf(denom) > 1e-6f) {
            float t = dot3(ld3(L[0]) - p, plane_n) / denom;
            if (t > 0.0f) rep = closest_on_quad(L, p + r * t);
        }
        f3 l = nrm3(rep - p);
        float n_dot_l = dot3(n, l);
        if (n_dot_l <= 0.0f) continue;
        float rep_dist = fmx(len3(rep - p), 1e-4f);
        float wide = clmp(alpha + radius / (2.0f * rep_dist), alpha, 1.0f);
        float energy = (alpha / wide) * (alpha / wide);
        out = out + ggx_lobe(n, v, l, n_dot_v, n_dot_l, wide, energy, m.f0) * radiance;
    }
    for (uint32_t li = 0; li < Lv.n_sph; ++li) {  // EXT
        const float4* L = Lv.sph + 2 * li;
        ExtGeo g;
        const bool go = on && sphere_geo(L, p, n, g);
        float vis = ext_shadow(S, p, n, g, steps, stk, go);
        if (!go) continue;
        if (vis <= 1e-4f) continue;
        f3 radiance = ld3(L[1]) * vis;
        out = out + (m.kd * radiance) * g.ff;
        f3 r = (-v) - n * (2.0f * dot3(n, -v));
        f3 to = ld3(L[0]) - p;
        f3 ctr = r * dot3(to, r) - to;
        float lc = len3(ctr);
        f3 rep_v = lc > 1e-9f ? to + ctr * clmp(L[0].w / lc, 0.0f, 1.0f) : to;
        f3 l = nrm3(rep_v);
        float n_dot_l = dot3(n, l);
        if (n_dot_l <= 0.0f) continue;
        float rep_dist = fmx(len3(rep_v), 1e-4f);
        float wide = clmp(alpha + g.radius / (2.0f * rep_dist), alpha, 1.0f);
        float energy = (alpha / wide) * (alpha / wide);
        out = out + ggx_lobe(n, v, l, n_dot_v, n_dot_l, wide, energy, m.f0) * radiance;
    }
    for (uint32_t li = 0; li < Lv.n_pnt; ++li) {  // EXT
        const float4* L = Lv.pnt + 2 * li;
        ExtGeo g;
        const bool go = on && point_geo(L, p, n, g);
        float vis = ext_shadow(S, p, n, g, steps, stk, go);
        if (!go) continue;
        if (vis <= 1e-4f) continue;
        f3 emit = ld3(L[1]);
        out = out + (m.kd * emit) * (g.ff * vis);
        f3 l = g.ldir;
        float n_dot_l = dot3(n, l);
        f3 h = nrm3(l + v);
        float d = d_ggx(fmx(dot3(n, h), 0.0f), alpha);
        float vis_s = v_smith(n_dot_v, n_dot_l, alpha);
        f3 f = fresnel_schlick(fmx(dot3(l, h), 0.0f), m.f0);
        out = out + (f * (d * vis_s * n_dot_l)) * (emit * (vis * g.inv_d2));
    }
    steps_out = steps;
    return out;
}

// One light's share of direct() split into the factors the backward pass
// needs: direct() adds (kd * emit) * a + sp * emit for this light (equal in
// exact arithmetic; the forward keeps the reference's rounding order).
// Returns false where direct() skips the light.
template <class Stk>
IFX_HD bool light_coeff(const SceneView& S, const LightsView& Lv, uint32_t li, f3 p, f3 n, f3 tan, f3 bit, f3 v,
                        float n_dot_v, float alpha, f3 f0, float& a, f3& sp, f3& emit, Stk& stk) {
    uint32_t steps = 0;
    a = 0.0f;
    sp = spl(0.0f);
    if (li < Lv.n_quad) {
        const float4* L = Lv.quad + 7 * li;
        f3 lc = ld3(L[5]);
        f3 to_light = lc - p;
        f3 plane_n = ld3(L[6]);
        if (dot3(plane_n, -nrm3(to_light)) <= 0.0f) return false;
        float ff = quad_ff(L, p, tan, bit, n);
        if (ff <= 1e-6f) return false;
        float dist = len3(to_light);
        float radius = L[5].w;
        float k = clmp(dist / radius, 1.5f, 96.0f);
        float end = fmx(dist - radius * 1.25f, 0.05f);
        float vis = soft_shadow(S, p + n * 3e-3f, to_light / dist, 0.015f, end, k, steps, stk);
        if (vis <= 1e-4f) return false;
        a = ff * vis;
        emit = ld3(L[4]);
        f3 r = (-v) - n * (2.0f * dot3(n, -v));
        float denom = dot3(r, plane_n);
        f3 rep = lc;
        if (fabsf(denom) > 1e-6f) {
            float t = dot3(ld3(L[0]) - p, plane_n) / denom;
            if (t > 0.0f) rep = closest_on_quad(L, p + r * t);
        }
        f3 l = nrm3(rep - p);
        float n_dot_l = dot3(n, l);
        if (n_dot_l <= 0.0f) return true;
        float rep_dist = fmx(len3(rep - p), 1e-4f);
        float wide = clmp(alpha + radius / (2.0f * rep_dist), alpha, 1.0f);
        float energy = (alpha / wide) * (alpha / wide);
        sp = ggx_lobe(n, v, l, n_dot_v, n_dot_l, wide, energy, f0) * vis;
        return true;
    }
    li -= Lv.n_quad;
    ExtGeo g;
    if (li < Lv.n_sph) {  // EXT
        const float4* L = Lv.sph + 2 * li;
        if (!sphere_geo(L, p, n, g)) return false;
        float vis = ext_shadow(S, p, n, g, steps, stk);
        if (vis <= 1e-4f) return false;
        a = g.ff * vis;
        emit = ld3(L[1]);
        f3 r = (-v) - n * (2.0f * dot3(n, -v));
        f3 to = ld3(L[0]) - p;
        f3 ctr = r * dot3(to, r) - to;
        float lc = len3(ctr);
        f3 rep_v = lc > 1e-9f ? to + ctr * clmp(L[0].w / lc, 0.0f, 1.0f) : to;
        f3 l = nrm3(rep_v);
        float n_dot_l = dot3(n, l);
        if (n_dot_l <= 0.0f) return true;
        float rep_dist = fmx(len3(rep_v), 1e-4f);
        float wide = clmp(alpha + g.radius / (2.0f * rep_dist), alpha, 1.0f);
        float energy = (alpha / wide) * (alpha / wide);
        sp = ggx_lobe(n, v, l, n_dot_v, n_dot_l, wide, energy, f0) * vis;
        return true;
    }
    const float4* L = Lv.pnt + 2 * (li - Lv.n_sph);  // EXT
    if (!point_geo(L, p, n, g)) return false;
    float vis = ext_shadow(S, p, n, g, steps, stk);
    if (vis <= 1e-4f) return false;
    a = g.ff * vis;
    emit = ld3(L[1]);
    f3 l = g.ldir;
    float n_dot_l = dot3(n, l);
    f3 h = nrm3(l + v);
    float d = d_ggx(fmx(dot3(n, h), 0.0f), alpha);
    float vis_s = v_smith(n_dot_v, n_dot_l, alpha);
    f3 f = fresnel_schlick(fmx(dot3(l, h), 0.0f), f0);
    sp = f * ((d * vis_s * n_dot_l) * (vis * g.inv_d2));
    return true;
}

// Emission of light li (quads, then sphere lights, then point lights).
IFX_HD f3 light_emit(const LightsView& Lv, uint32_t li) {
    if (li < Lv.n_quad) return ld3(Lv.quad[7 * li + 4]);
    li -= Lv.n_quad;
    if (li < Lv.n_sph) return ld3(Lv.sph[2 * li + 1]);
    return ld3(Lv.pnt[2 * (li - Lv.n_sph) + 1]);
}

// src/shade.rs:168-181 (+ EXT visible sphere emitters).
IFX_HD bool emitter_hit(const LightsView& Lv, f3 ro, f3 rd, float tmax, f3& emit, uint32_t* which = nullptr) {
    bool found = false;
    float bt = 0.0f;
    for (uint32_t li = 0; li < Lv.n_quad; ++li) {
        const float4* L = Lv.quad + 7 * li;
        float t;
        if (intersect_quad(L, ro, rd, t)) {
            if (t < tmax && dot3(ld3(L[6]), rd) < 0.0f && (!found || t < bt)) {
                found = true;
                bt = t;
                emit = ld3(L[4]);
                if (which) *which = li;
            }
        }
    }
    for (uint32_t li = 0; li < Lv.n_sph; ++li) {
        const float4* L = Lv.sph + 2 * li;
        f3 oc = ro - ld3(L[0]);
        float b = dot3(oc, rd);
        float c = dot3(oc, oc) - L[0].w * L[0].w;
        float disc = b * b - c;
        if (c <= 0.0f || disc < 0.0f) continue;
        float t = -b - sqrtf(disc);
        if (t < 1e-4f) continue;
        if (t < tmax && (!found || t < bt)) {
            found = true;
            bt = t;
            emit = ld3(L[1]);
            if (which) *which = Lv.n_quad + li;
        }
    }
    return found;
}

// ---------------------------------------------------------------- path tracer
// src/reference.rs:24-46: PCG32, output of the advanced state (XSH-RR), 24-bit float.
struct PathRng {
    uint64_t s;
    IFX_HD explicit PathRng(uint64_t seed) : s(seed * 6364136223846793005ull + 1ull) {}
    IFX_HD float next_f32() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        uint64_t x = ((s >> 18) ^ s) >> 27;
        uint32_t rot = (uint32_t)(s >> 59);
        uint32_t xv = (uint32_t)x;
        uint32_t v = (xv >> rot) | (xv << ((32u - rot) & 31u));
        return (float)(v >> 8) * (1.0f / 16777216.0f);
    }
};

// src/reference.rs:48-58 (sin / cos from the device math library).
IFX_HD f3 cosine_hemisphere(f3 n, PathRng& rng) {
    f3 t, b;
    frame(n, t, b);
    float u1 = rng.next_f32();
    float u2 = rng.next_f32();
    float r = sqrtf(u1);
    float phi = kTau * u2;
    return nrm3((t * (r * cosf(phi)) + b * (r * sinf(phi))) + n * sqrtf(fmx(1.0f - u1, 0.0f)));
}

// src/reference.rs:61-137: one path over the quad lights (the reference's
// light list): next-event estimation with hard shadows, cosine-weighted
// bounces, Russian roulette after depth 3.
template <class Stk>
IFX_HD f3 ref_path(const SceneView& S, const LightsView& Lv, f3 ro, f3 rd, float tmax, uint32_t max_depth,
                   PathRng& rng, Stk& stk) {
    f3 radiance = spl(0.0f), throughput = spl(1.0f);
    f3 origin = ro, dir = rd;
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        uint32_t steps = 0;
        Hit hit = trace<true>(S, origin, dir, 1e-3f, tmax, steps, stk);
        float surf_t = hit.hit ? hit.t : kFltMax;
        bool seen = false;
        f3 e = spl(0.0f);
        float bt = 0.0f;
        for (uint32_t li = 0; li < Lv.n_quad; ++li) {  // shade::emitter_hit over the quad list
            const float4* L = Lv.quad + 7 * li;
            float t;
            if (intersect_quad(L, origin, dir, t) && t < fmn(surf_t, tmax) && dot3(ld3(L[6]), dir) < 0.0f &&
                (!seen || t < bt)) {
                seen = true;
                bt = t;
                e = ld3(L[4]);
            }
        }
        if (seen) {
            if (depth == 0) radiance = radiance + throughput * e;
            break;
        }
        if (!hit.hit) break;
        const MatV m = load_mat(Lv, hit.mat);
        const f3 albedo = m.kd;
        radiance = radiance + throughput * m.emissive;
        for (uint32_t li = 0; li < Lv.n_quad; ++li) {
            const float4* L = Lv.quad + 7 * li;
            const f3 v0 = ld3(L[0]);
            const f3 e1 = ld3(L[1]) - v0, e2 = ld3(L[3]) - v0;
            const float r1 = rng.next_f32();
            const f3 a = v0 + e1 * r1;
            const float r2 = rng.next_f32();
            const f3 lp = a + e2 * r2;
            const f3 to = lp - hit.p;
            const float dist2 = dot3(to, to);
            if (dist2 < 1e-9f) continue;
            const float dist = sqrtf(dist2);
            const f3 l = to / dist;
            const float cos_s = dot3(hit.n, l);
            const float cos_l = dot3(ld3(L[6]), -l);
            if (cos_s <= 0.0f || cos_l <= 0.0f) continue;
            uint32_t s2 = 0;
            Hit occ = trace<false>(S, hit.p + hit.n * 1e-3f, l, 1e-3f, dist - 1e-2f, s2, stk);
            if (occ.hit) continue;
            const float pdf = dist2 / (cos_l * L[4].w);
            radiance = radiance + ((throughput * albedo) * ld3(L[4])) * (cos_s / (kPi * pdf));
        }
        throughput = throughput * albedo;
        dir = cosine_hemisphere(hit.n, rng);
        origin = hit.p + hit.n * 1e-3f;
        if (depth >= 3) {
            const float q = clmp(maxe(throughput), 0.05f, 0.95f);
            if (rng.next_f32() > q) break;
            throughput = throughput / q;
        }
    }
    return radiance;
}

// Uniform grid over surfel positions, CSR by cell, items ascending per cell
// (src/surfel.rs:61-114).
struct GridView {
    f3 lo;
    float inv_cell;
    int32_t dim[3];
    const uint32_t* starts;
    const uint32_t* items;
};

IFX_HD void grid_cell(const GridView& g, f3 p, int32_t c[3]) {
    c[0] = clampi(sat_i32((p.x - g.lo.x) * g.inv_cell), 0, g.dim[0] - 1);
    c[1] = clampi(sat_i32((p.y - g.lo.y) * g.inv_cell), 0, g.dim[1] - 1);
    c[2] = clampi(sat_i32((p.z - g.lo.z) * g.inv_cell), 0, g.dim[2] - 1);
}

// src/surfel.rs:262-293: visit cells z, y, x then items in order.
IFX_HD f3 gather(const GridView& g, const float4* __restrict__ sp, const float4* __restrict__ sn,
                 const float4* __restrict__ irr, f3 p, f3 n, float radius) {
    f3 acc = spl(0.0f);
    float wsum = 0.0f;
    float r2 = radius * radius;
    int32_t c[3];
    grid_cell(g, p, c);
    int32_t z1 = c[2] + 1 < g.dim[2] - 1 ? c[2] + 1 : g.dim[2] - 1;
    int32_t y1 = c[1] + 1 < g.dim[1] - 1 ? c[1] + 1 : g.dim[1] - 1;
    int32_t x1 = c[0] + 1 < g.dim[0] - 1 ? c[0] + 1 : g.dim[0] - 1;
    for (int32_t z = c[2] - 1 > 0 ? c[2] - 1 : 0; z <= z1; ++z)
        for (int32_t y = c[1] - 1 > 0 ? c[1] - 1 : 0; y <= y1; ++y)
            for (int32_t x = c[0] - 1 > 0 ? c[0] - 1 : 0; x <= x1; ++x) {
                uint32_t ci = (uint32_t)(x + g.dim[0] * (y + g.dim[1] * z));
                uint32_t e = g.starts[ci + 1];
                for (uint32_t k = g.starts[ci]; k < e; ++k) {
                    uint32_t it = g.items[k];
                    f3 q = ld3(sp[it]);
                    f3 dq = q - p;
                    float d2 = dot3(dq, dq);
                    if (d2 > r2) continue;
                    float align = dot3(ld3(sn[it]), n);
                    if (align <= 0.15f) continue;
                    float t = 1.0f - d2 / r2;
                    float w = t * t * align;
                    acc = acc + ld3(irr[it]) * w;
                    wsum += w;
                }
            }
    if (wsum > 1e-6f) return acc / wsum;
    return spl(0.0f);
}

}  // namespace ifx
