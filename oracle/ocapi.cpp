// ORACLE — test infrastructure only (see ovec.h).
// C ABI over the CPU restatement, loaded with ctypes by tests/, the smoke
// check and bench.py's CPU-baseline leg.  Vectors cross as packed xyz f32.
#include <cstring>
#include <string>

#include "ogi.h"

using namespace orc;

namespace {
struct Ctx {
    Scene sc;
    std::vector<Surfel> surfels;
    Transfer tr;
    bool have_tr = false;
    uint64_t ff_evals = 0;
    uint32_t spp = 512;
    OpTally vis, direct, shade;  // operation tallies of the last transfer, solve and render
};
V3 ld(const float* p) { return v3(p[0], p[1], p[2]); }
void st(float* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
}  // namespace

extern "C" {

void* orc_create(const char* src, char* err, int errlen) {
    Ctx* c = new Ctx();
    std::string e;
    if (!parse_scene(src ? std::string(src) : std::string(), c->sc, e)) {
        if (err && errlen > 0) {
            std::strncpy(err, e.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        delete c;
        return nullptr;
    }
    return c;
}

void orc_destroy(void* h) { delete (Ctx*)h; }

// objects, planes, quad lights, materials (excluding the implicit default),
// bvh nodes, sphere lights, point lights, blend radius * 1e6
void orc_scene_counts(void* h, int64_t out[8]) {
    Ctx* c = (Ctx*)h;
    out[0] = (int64_t)c->sc.objects.size();
    out[1] = (int64_t)c->sc.planes.size();
    out[2] = (int64_t)c->sc.lights.size();
    out[3] = (int64_t)c->sc.materials.size() - 1;
    out[4] = (int64_t)c->sc.bvh.nodes.size();
    out[5] = (int64_t)c->sc.sphere_lights.size();
    out[6] = (int64_t)c->sc.point_lights.size();
    out[7] = (int64_t)(c->sc.blend_k * 1e6f);
}

// cfg: width, height, surfels, bounces; f: exposure, lo[3], hi[3], cam pos[3], look[3], up[3], fov
void orc_scene_cfg(void* h, uint64_t cfg[4], float f[17]) {
    Ctx* c = (Ctx*)h;
    cfg[0] = c->sc.cfg.width;
    cfg[1] = c->sc.cfg.height;
    cfg[2] = c->sc.cfg.surfels;
    cfg[3] = c->sc.cfg.bounces;
    f[0] = c->sc.cfg.exposure;
    st(f + 1, c->sc.lo);
    st(f + 4, c->sc.hi);
    st(f + 7, c->sc.camera.pos);
    st(f + 10, c->sc.camera.look);
    st(f + 13, c->sc.camera.up);
    f[16] = c->sc.camera.fov;
}

void orc_set_use_bvh(void* h, int on) { ((Ctx*)h)->sc.use_bvh = on != 0; }

// BVH nodes as lo[3],hi[3] and start,count,right; plus the object order.
void orc_bvh(void* h, float* lohi, uint32_t* scr, uint32_t* order) {
    Ctx* c = (Ctx*)h;
    for (size_t i = 0; i < c->sc.bvh.nodes.size(); ++i) {
        const BvhNode& n = c->sc.bvh.nodes[i];
        st(lohi + 6 * i, n.lo);
        st(lohi + 6 * i + 3, n.hi);
        scr[3 * i] = n.start;
        scr[3 * i + 1] = n.count;
        scr[3 * i + 2] = n.right;
    }
    for (size_t i = 0; i < c->sc.bvh.order.size(); ++i) order[i] = c->sc.bvh.order[i];
}

void orc_dist_mat(void* h, const float* pts, int64_t n, float* d, uint32_t* mat, int threads) {
    Ctx* c = (Ctx*)h;
    par_for((size_t)n, threads, [&](size_t i) {
        DistHit r = c->sc.dist_mat(ld(pts + 3 * i));
        d[i] = r.d;
        if (mat) mat[i] = r.mat;
    });
}

void orc_normal(void* h, const float* pts, int64_t n, float* out, int threads) {
    Ctx* c = (Ctx*)h;
    par_for((size_t)n, threads, [&](size_t i) { st(out + 3 * i, c->sc.normal(ld(pts + 3 * i))); });
}

void orc_trace(void* h, const float* ro, const float* rd, int64_t n, float tmin, float tmax, float* t_out,
               float* n_out, uint32_t* mat, int32_t* hit, uint32_t* steps, int threads) {
    Ctx* c = (Ctx*)h;
    par_for((size_t)n, threads, [&](size_t i) {
        uint32_t s = 0;
        Hit hh = trace(c->sc, ld(ro + 3 * i), ld(rd + 3 * i), tmin, tmax, s);
        hit[i] = hh.hit ? 1 : 0;
        t_out[i] = hh.t;
        st(n_out + 3 * i, hh.n);
        mat[i] = hh.mat;
        steps[i] = s;
    });
}

void orc_soft_shadow(void* h, const float* ro, const float* rd, int64_t n, float tmin, const float* tmax,
                     const float* k, float* vis, uint32_t* steps, int threads) {
    Ctx* c = (Ctx*)h;
    par_for((size_t)n, threads, [&](size_t i) {
        uint32_t s = 0;
        vis[i] = soft_shadow(c->sc, ld(ro + 3 * i), ld(rd + 3 * i), tmin, tmax[i], k[i], s);
        steps[i] = s;
    });
}

void orc_closure(void* h, const float* p, const float* nrm, int64_t n, float tmax, uint32_t dirs, float* sky,
                 uint32_t* steps, int threads) {
    Ctx* c = (Ctx*)h;
    par_for((size_t)n, threads, [&](size_t i) {
        uint32_t s = 0;
        sky[i] = hemisphere_closure(c->sc, ld(p + 3 * i), ld(nrm + 3 * i), tmax, dirs, s);
        steps[i] = s;
    });
}

void orc_direct_diffuse(void* h, const float* p, const float* nrm, const float* alb, int64_t n, float* out,
                        uint32_t* steps, int threads) {
    Ctx* c = (Ctx*)h;
    par_for((size_t)n, threads, [&](size_t i) {
        uint32_t s = 0;
        st(out + 3 * i, direct_diffuse(c->sc, ld(p + 3 * i), ld(nrm + 3 * i), ld(alb + 3 * i), s));
        steps[i] = s;
    });
}

int64_t orc_generate(void* h, uint64_t target) {
    Ctx* c = (Ctx*)h;
    c->surfels = generate(c->sc, (size_t)target);
    c->have_tr = false;
    return (int64_t)c->surfels.size();
}

int64_t orc_num_surfels(void* h) { return (int64_t)((Ctx*)h)->surfels.size(); }

void orc_get_surfels(void* h, float* p, float* n, float* albedo, float* area) {
    Ctx* c = (Ctx*)h;
    for (size_t i = 0; i < c->surfels.size(); ++i) {
        const Surfel& s = c->surfels[i];
        if (p) st(p + 3 * i, s.p);
        if (n) st(n + 3 * i, s.n);
        if (albedo) st(albedo + 3 * i, s.albedo);
        if (area) area[i] = s.area;
    }
}

void orc_set_surfels(void* h, const float* p, const float* n, const float* albedo, const float* area, int64_t N) {
    Ctx* c = (Ctx*)h;
    c->surfels.assign((size_t)N, Surfel{});
    for (size_t i = 0; i < (size_t)N; ++i) {
        Surfel& s = c->surfels[i];
        s.p = ld(p + 3 * i);
        s.n = ld(n + 3 * i);
        s.albedo = ld(albedo + 3 * i);
        s.area = area[i];
        s.direct = s.irr = s.rad = splat(0.0f);
    }
    c->have_tr = false;
}

// stats: clusters, links, mean_row_sum, evals
int orc_transfer(void* h, int32_t res, int threads, int64_t row_begin, int64_t row_end, double stats[4]) {
    Ctx* c = (Ctx*)h;
    if (res < 1) return -1;
    std::vector<Cluster> cl = build_clusters(c->sc, c->surfels, res);
    size_t rb = row_begin < 0 ? 0 : (size_t)row_begin;
    size_t re = row_end < 0 ? SIZE_MAX : (size_t)row_end;
    c->ff_evals = build_transfer(c->sc, c->surfels, std::move(cl), c->tr, threads, rb, re);
    c->vis = c->tr.vis_tally;
    c->have_tr = true;
    stats[0] = (double)c->tr.nc;
    stats[1] = (double)c->tr.links;
    stats[2] = (double)c->tr.mean_row_sum;
    stats[3] = (double)c->ff_evals;
    return 0;
}

void orc_get_transfer(void* h, uint32_t* starts, uint32_t* idx, float* w, float* sky, float* row_sum) {
    Ctx* c = (Ctx*)h;
    std::memcpy(starts, c->tr.starts.data(), c->tr.starts.size() * 4);
    if (!c->tr.idx.empty()) {
        std::memcpy(idx, c->tr.idx.data(), c->tr.idx.size() * 4);
        std::memcpy(w, c->tr.w.data(), c->tr.w.size() * 4);
    }
    if (sky) std::memcpy(sky, c->tr.sky.data(), c->tr.sky.size() * 4);
    if (row_sum) std::memcpy(row_sum, c->tr.row_sum.data(), c->tr.row_sum.size() * 4);
}

void orc_get_clusters(void* h, float* p, float* n, float* area, float* radius, uint32_t* mstarts, uint32_t* members) {
    Ctx* c = (Ctx*)h;
    uint32_t off = 0;
    for (size_t i = 0; i < c->tr.clusters.size(); ++i) {
        const Cluster& cl = c->tr.clusters[i];
        st(p + 3 * i, cl.p);
        st(n + 3 * i, cl.n);
        area[i] = cl.area;
        radius[i] = cl.radius;
        mstarts[i] = off;
        for (uint32_t m : cl.members) members[off++] = m;
    }
    mstarts[c->tr.clusters.size()] = off;
}

// stats: direct-light evals, iterations, residual
int orc_solve(void* h, uint32_t bounces, int threads, double stats[3]) {
    Ctx* c = (Ctx*)h;
    if (!c->have_tr) return -1;
    c->direct = OpTally();
    uint64_t ev = light_surfels(c->sc, c->surfels, threads, &c->direct);
    SolveStats s = solve(c->surfels, c->tr, bounces, threads);
    stats[0] = (double)ev;
    stats[1] = (double)s.iterations;
    stats[2] = (double)s.residual;
    return 0;
}

// Light the surfels only (E term).
uint64_t orc_light_surfels(void* h, int threads) {
    Ctx* c = (Ctx*)h;
    return light_surfels(c->sc, c->surfels, threads);
}

void orc_get_radiosity(void* h, float* direct, float* irr, float* rad) {
    Ctx* c = (Ctx*)h;
    for (size_t i = 0; i < c->surfels.size(); ++i) {
        if (direct) st(direct + 3 * i, c->surfels[i].direct);
        if (irr) st(irr + 3 * i, c->surfels[i].irr);
        if (rad) st(rad + 3 * i, c->surfels[i].rad);
    }
}

void orc_set_radiosity(void* h, const float* direct, const float* irr, const float* rad) {
    Ctx* c = (Ctx*)h;
    for (size_t i = 0; i < c->surfels.size(); ++i) {
        if (direct) c->surfels[i].direct = ld(direct + 3 * i);
        if (irr) c->surfels[i].irr = ld(irr + 3 * i);
        if (rad) c->surfels[i].rad = ld(rad + 3 * i);
    }
}

// cam: pos[3], look[3], up[3], fov (NULL = the scene camera).  Rows
// [y0, y1) only; gi_surfels = use the solved surfels for indirect light.
uint64_t orc_render(void* h, int mode, uint32_t w, uint32_t hh, const float* cam, int use_gi, int threads,
                    uint32_t y0, uint32_t y1, float* film_out) {
    Ctx* c = (Ctx*)h;
    const Scene& sc = c->sc;
    V3 pos = sc.camera.pos, look = sc.camera.look, up = sc.camera.up;
    float fov = sc.camera.fov;
    if (cam) {
        pos = ld(cam);
        look = ld(cam + 3);
        up = ld(cam + 6);
        fov = cam[9];
    }
    Cam cm = Cam::make(pos, look, up, fov, w, hh);
    float diag = length(sc.hi - sc.lo);
    float tmax = diag * 4.0f;
    Gi gi;
    const Gi* gip = nullptr;
    if (use_gi && (mode == BEAUTY || mode == INDIRECT)) {
        gi.surfels = &c->surfels;
        gi.radius = (diag / std::sqrt((float)sc.cfg.surfels)) * 2.6f;
        std::vector<V3> pts(c->surfels.size());
        for (size_t i = 0; i < pts.size(); ++i) pts[i] = c->surfels[i].p;
        gi.grid.build(pts, sc.lo, sc.hi, gi.radius);
        gip = &gi;
    }
    std::vector<V3> film;
    c->shade = OpTally();
    uint64_t ev = render_frame(sc, cm, (Mode)mode, gip, tmax, w, hh, film, threads, y0, y1, c->spp, &c->shade);
    for (size_t i = 0; i < film.size(); ++i) st(film_out + 3 * i, film[i]);
    return ev;
}

// Operation tallies (the roofline's counting rule) of the next transfers,
// solves and renders: on = 1 starts tallying, 0 stops.  out = link shadow
// marches, direct light, render: (ops, primitive evaluations) each.
void orc_set_tally(int on) { g_tally_on = on != 0; }
void orc_tallies(void* h, uint64_t out[6]) {
    Ctx* c = (Ctx*)h;
    out[0] = c->vis.ops, out[1] = c->vis.prims;
    out[2] = c->direct.ops, out[3] = c->direct.prims;
    out[4] = c->shade.ops, out[5] = c->shade.prims;
}

// Paths per pixel of the REFERENCE mode (--spp, src/main.rs:81).
void orc_set_spp(void* h, uint32_t spp) { ((Ctx*)h)->spp = spp; }

// One path tracer sample sequence for tests: n floats of the PCG32 seeded with `seed`.
void orc_rng(uint64_t seed, uint32_t n, float* out) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_f32();
}

void orc_gather(void* h, const float* p, const float* n, int64_t np, float radius, float* out) {
    Ctx* c = (Ctx*)h;
    std::vector<V3> pts(c->surfels.size());
    for (size_t i = 0; i < pts.size(); ++i) pts[i] = c->surfels[i].p;
    Grid g;
    g.build(pts, c->sc.lo, c->sc.hi, radius);
    for (int64_t i = 0; i < np; ++i) st(out + 3 * i, gather(c->surfels, g, ld(p + 3 * i), ld(n + 3 * i), radius));
}

void orc_encode(const float* film, int64_t npx, float exposure, int tonemap, uint8_t* out) {
    std::vector<V3> f((size_t)npx);
    for (size_t i = 0; i < f.size(); ++i) f[i] = ld(film + 3 * i);
    std::vector<uint8_t> o;
    encode(f, exposure, tonemap != 0, o);
    std::memcpy(out, o.data(), o.size());
}

// ---- the parsed scene in the layout of the product library's ifx_scene_desc
// arrays (a test fixture for its scene-descriptor entry point): object 13
// words (kind, a[3], b[3], r, r2, sub_first, sub_count, mat, blend), primitive
// 9, BVH node 9 (lo[3], hi[3], start, count, right), material 8, quad light
// 15, sphere light 7, point light 6; integers as their 32-bit patterns.
// counts: objects, planes, subtractions, BVH nodes, materials, quad, sphere
// and point lights.
namespace {
float float_of(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
void prim_words(const Prim& p, float* w) {
    V3 a = p.a, b = p.b;
    float r = p.r, r2 = 0.0f;
    switch (p.kind) {
        case SPHERE: a = p.c; break;
        case ROUNDBOX: a = p.c; break;
        case PLANE: a = p.n, r = p.h; break;
        case TORUS: a = p.c, r = p.major, r2 = p.minor; break;
        case CONE: r = p.ra, r2 = p.rb; break;
        default: break;
    }
    w[0] = float_of((uint32_t)p.kind);
    st(w + 1, a);
    st(w + 4, b);
    w[7] = r;
    w[8] = r2;
}
}  // namespace

void orc_scene_export_counts(void* h, uint32_t counts[8]) {
    const Scene& sc = ((Ctx*)h)->sc;
    uint32_t subs = 0;
    for (const Object& o : sc.objects) subs += (uint32_t)o.subtract.size();
    for (const Object& o : sc.planes) subs += (uint32_t)o.subtract.size();
    counts[0] = (uint32_t)sc.objects.size(), counts[1] = (uint32_t)sc.planes.size(), counts[2] = subs;
    counts[3] = (uint32_t)sc.bvh.nodes.size(), counts[4] = (uint32_t)sc.materials.size();
    counts[5] = (uint32_t)sc.lights.size(), counts[6] = (uint32_t)sc.sphere_lights.size();
    counts[7] = (uint32_t)sc.point_lights.size();
}

void orc_scene_export(void* h, float* objects, float* planes, float* subs, float* nodes, uint32_t* order,
                      float* materials, float* quads, float* spheres, float* points, float* bounds) {
    const Scene& sc = ((Ctx*)h)->sc;
    uint32_t ns = 0;
    auto obj = [&](const Object& o, float* w) {
        prim_words(o.prim, w);
        w[9] = float_of(ns);
        w[10] = float_of((uint32_t)o.subtract.size());
        w[11] = float_of(o.mat);
        w[12] = o.blend;
        for (const Prim& p : o.subtract) prim_words(p, subs + 9 * ns++);
    };
    for (size_t i = 0; i < sc.objects.size(); ++i) obj(sc.objects[i], objects + 13 * i);
    for (size_t i = 0; i < sc.planes.size(); ++i) obj(sc.planes[i], planes + 13 * i);
    for (size_t i = 0; i < sc.bvh.nodes.size(); ++i) {
        const BvhNode& n = sc.bvh.nodes[i];
        float* w = nodes + 9 * i;
        st(w, n.lo);
        st(w + 3, n.hi);
        w[6] = float_of(n.start), w[7] = float_of(n.count), w[8] = float_of(n.right);
    }
    for (size_t i = 0; i < sc.bvh.order.size(); ++i) order[i] = sc.bvh.order[i];
    for (size_t i = 0; i < sc.materials.size(); ++i) {
        const Material& m = sc.materials[i];
        float* w = materials + 8 * i;
        st(w, m.albedo);
        w[3] = m.roughness, w[4] = m.metallic;
        st(w + 5, m.emissive);
    }
    for (size_t i = 0; i < sc.lights.size(); ++i) {
        for (int k = 0; k < 4; ++k) st(quads + 15 * i + 3 * k, sc.lights[i].verts[k]);
        st(quads + 15 * i + 12, sc.lights[i].emit);
    }
    for (size_t i = 0; i < sc.sphere_lights.size(); ++i) {
        st(spheres + 7 * i, sc.sphere_lights[i].c);
        spheres[7 * i + 3] = sc.sphere_lights[i].r;
        st(spheres + 7 * i + 4, sc.sphere_lights[i].emit);
    }
    for (size_t i = 0; i < sc.point_lights.size(); ++i) {
        st(points + 6 * i, sc.point_lights[i].pos);
        st(points + 6 * i + 3, sc.point_lights[i].emit);
    }
    st(bounds, sc.lo);
    st(bounds + 3, sc.hi);
}

// ---- known-answer helpers
// params: kind-specific, packed as c|a[3], b[3], r|h|major|ra, minor|rb, n[3]
float orc_prim_dist(int kind, const float* q, const float* p) {
    Prim pr{};
    pr.kind = (PrimKind)kind;
    pr.c = pr.a = ld(q);
    pr.b = ld(q + 3);
    pr.r = pr.h = pr.major = pr.ra = q[6];
    pr.minor = pr.rb = q[7];
    pr.n = ld(q + 8);
    return pr.dist(ld(p));
}

int orc_prim_bounds(int kind, const float* q, float* lohi) {
    Prim pr{};
    pr.kind = (PrimKind)kind;
    pr.c = pr.a = ld(q);
    pr.b = ld(q + 3);
    pr.r = pr.h = pr.major = pr.ra = q[6];
    pr.minor = pr.rb = q[7];
    pr.n = ld(q + 8);
    V3 lo, hi;
    if (!pr.bounds(lo, hi)) return 0;
    st(lohi, lo);
    st(lohi + 3, hi);
    return 1;
}

float orc_quad_ff(const float* verts, const float* p, const float* n) {
    V3 vv[4] = {ld(verts), ld(verts + 3), ld(verts + 6), ld(verts + 9)};
    V3 t, b, loc[4];
    frame(ld(n), t, b);
    to_local(vv, ld(p), t, b, ld(n), loc);
    return quad_form_factor(loc);
}

float orc_sphere_ff(float cos_t, float s2) { return sphere_form_factor(cos_t, s2); }

// BSDF helpers and direct light (src/shade.rs:8-125), emitters (src/shade.rs:168-181,
// src/trace.rs:128-148)
void orc_fresnel(float cos_theta, const float* f0, float* out) { st(out, fresnel_schlick(cos_theta, ld(f0))); }
float orc_d_ggx(float n_dot_h, float alpha) { return d_ggx(n_dot_h, alpha); }
float orc_v_smith(float n_dot_v, float n_dot_l, float alpha) { return v_smith(n_dot_v, n_dot_l, alpha); }
void orc_f0_of(const float* albedo, float metallic, float* out) {
    Material m;
    m.albedo = ld(albedo);
    m.metallic = metallic;
    st(out, f0_of(m));
}
void orc_direct(void* h, const float* p, const float* n, const float* v, const float* albedo, float roughness,
                float metallic, float* out, uint32_t* steps) {
    Material m;
    m.albedo = ld(albedo);
    m.roughness = roughness;
    m.metallic = metallic;
    uint32_t s = 0;
    st(out, direct(((Ctx*)h)->sc, ld(p), ld(n), ld(v), m, s));
    *steps = s;
}
int orc_emitter_hit(void* h, const float* ro, const float* rd, float tmax, float* emit) {
    V3 e = splat(0.0f);
    const bool hit = emitter_hit(((Ctx*)h)->sc, ld(ro), ld(rd), tmax, e);
    st(emit, e);
    return hit ? 1 : 0;
}
int orc_intersect_quad(const float* verts, const float* ro, const float* rd, float* t) {
    V3 vv[4] = {ld(verts), ld(verts + 3), ld(verts + 6), ld(verts + 9)};
    float tt = 0.0f;
    const bool hit = intersect_quad(vv, ld(ro), ld(rd), tt);
    *t = tt;
    return hit ? 1 : 0;
}
uint32_t orc_srgb8(float c) { return srgb8(c); }
void orc_aces(const float* in, float* out) { st(out, aces(ld(in))); }
float orc_radical_inverse(uint32_t i, uint32_t base) { return radical_inverse(i, base); }
void orc_frame(const float* n, float* t, float* b) {
    V3 tt, bb;
    frame(ld(n), tt, bb);
    st(t, tt);
    st(b, bb);
}

}  // extern "C"
