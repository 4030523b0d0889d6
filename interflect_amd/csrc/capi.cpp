// extern "C" boundary of libinterflect_hip.so (declared in include/interflect.h).
// No C++ exception crosses it: every entry point catches and maps to a status.
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <thread>

#include "../../include/interflect.h"
#include "batch.h"
#include "engine.h"
#include "host_surfels.h"

struct ifx_ctx {
    ifx::Engine eng;
    std::string err;
    bool have_scene = false;
};

struct ifx_batch {
    ifx::Batch b;
    std::string err;
};

namespace {

int host_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return (int)(n == 0 ? 1 : (n > 16 ? 16 : n));
}

template <class F>
ifx_status guarded(ifx_ctx* ctx, F&& f) {
    if (!ctx) return IFX_E_INVALID;
    try {
        ctx->err.clear();
        return f();
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return IFX_E_NOMEM;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return IFX_E_INVALID;
    } catch (...) {
        ctx->err = "unknown error";
        return IFX_E_INVALID;
    }
}

template <class F>
ifx_status bguarded(ifx_batch* b, F&& f) {
    if (!b) return IFX_E_INVALID;
    try {
        b->err.clear();
        return f();
    } catch (const std::bad_alloc&) {
        b->err = "host allocation failed";
        return IFX_E_NOMEM;
    } catch (const std::exception& e) {
        b->err = e.what();
        return IFX_E_INVALID;
    } catch (...) {
        b->err = "unknown error";
        return IFX_E_INVALID;
    }
}

ifx_status need_scene(ifx_ctx* ctx) {
    if (!ctx->have_scene) {
        ctx->err = "no scene loaded";
        return IFX_E_STATE;
    }
    return IFX_OK;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

ifx_status ifx_create(ifx_ctx** out, int32_t device) {
    if (!out) return IFX_E_INVALID;
    *out = nullptr;
    ifx_ctx* c = new (std::nothrow) ifx_ctx();
    if (!c) return IFX_E_NOMEM;
    int rc = c->eng.init(device, c->err);
    if (rc) {
        // Keep the context so the caller can read the message, then destroy it.
        *out = c;
        return rc;
    }
    *out = c;
    return IFX_OK;
}

ifx_status ifx_destroy(ifx_ctx* ctx) {
    delete ctx;
    return IFX_OK;
}

const char* ifx_last_error(const ifx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

ifx_status ifx_load_scene(ifx_ctx* ctx, const char* rad_text) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!rad_text) {
            ctx->err = "null scene text";
            return IFX_E_INVALID;
        }
        int rc = ctx->eng.load_scene(rad_text, ctx->err);
        ctx->have_scene = rc == 0;
        return rc;
    });
}

namespace {
bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// ifx_prim -> the loader's primitive (fields as Prim, src/sdf.rs:9-48).
bool prim_from(const ifx_prim& p, ifx::HPrim& h, std::string& err) {
    if (p.kind > IFX_PRIM_CONE) {
        err = "unknown primitive kind " + std::to_string(p.kind);
        return false;
    }
    if (!finite3(p.a) || !finite3(p.b) || !std::isfinite(p.r) || !std::isfinite(p.r2)) {
        err = "primitive parameters must be finite";
        return false;
    }
    const ifx::f3 a = ifx::mk(p.a[0], p.a[1], p.a[2]), b = ifx::mk(p.b[0], p.b[1], p.b[2]);
    h.kind = p.kind;
    switch (p.kind) {
        case IFX_PRIM_SPHERE: h.c = a, h.r = p.r; break;
        case IFX_PRIM_BOX: h.c = a, h.b = b, h.r = p.r; break;
        case IFX_PRIM_PLANE: h.n = a, h.h = p.r; break;
        case IFX_PRIM_CAPSULE:
        case IFX_PRIM_CYLINDER: h.a = a, h.b = b, h.r = p.r; break;
        case IFX_PRIM_TORUS: h.c = a, h.major = p.r, h.minor = p.r2; break;
        default: h.a = a, h.b = b, h.ra = p.r, h.rb = p.r2;
    }
    return true;
}

bool object_from(const ifx_object& o, const ifx_scene_desc& d, bool plane, ifx::HObject& h, std::string& err) {
    if (!prim_from(o.prim, h.prim, err)) return false;
    if ((o.prim.kind == IFX_PRIM_PLANE) != plane) {
        err = plane ? "the planes list holds planes only" : "planes belong in the planes list (they are unbounded)";
        return false;
    }
    if ((uint64_t)o.sub_first + o.sub_count > d.n_subtractions || (o.sub_count && !d.subtractions)) {
        err = "subtraction range beyond the subtractions array";
        return false;
    }
    if (o.mat >= d.n_materials) {
        err = "material index " + std::to_string(o.mat) + " beyond the materials array";
        return false;
    }
    if (!(o.blend >= 0.0f) || !std::isfinite(o.blend)) {
        err = "blend must be finite and >= 0";
        return false;
    }
    h.subs.resize(o.sub_count);
    for (uint32_t k = 0; k < o.sub_count; ++k)
        if (!prim_from(d.subtractions[o.sub_first + k], h.subs[k], err)) return false;
    h.mat = o.mat;
    h.blend = o.blend;
    return true;
}

// ifx_scene_desc -> a finalized host scene (the loader's output).
ifx_status scene_from(const ifx_scene_desc& d, ifx::HostScene& s, std::string& err) {
    if ((d.n_objects && !d.objects) || (d.n_planes && !d.planes) || (d.n_bvh_nodes && (!d.bvh_nodes || !d.bvh_order)) ||
        (d.n_quad_lights && !d.quad_lights) || (d.n_sphere_lights && !d.sphere_lights) ||
        (d.n_point_lights && !d.point_lights) || !d.materials || d.n_materials == 0) {
        err = "scene descriptor: null array (materials must hold at least the default material)";
        return IFX_E_INVALID;
    }
    if (d.n_quad_lights + d.n_sphere_lights + d.n_point_lights == 0) {
        err = "scene has no `light` block";
        return IFX_E_INVALID;
    }
    s.objects.resize(d.n_objects);
    for (uint32_t i = 0; i < d.n_objects; ++i)
        if (!object_from(d.objects[i], d, false, s.objects[i], err)) return IFX_E_INVALID;
    s.planes.resize(d.n_planes);
    for (uint32_t i = 0; i < d.n_planes; ++i)
        if (!object_from(d.planes[i], d, true, s.planes[i], err)) return IFX_E_INVALID;
    s.mats.resize(d.n_materials);
    for (uint32_t i = 0; i < d.n_materials; ++i) {
        const ifx_material& m = d.materials[i];
        s.mats[i].albedo = ifx::mk(m.albedo[0], m.albedo[1], m.albedo[2]);
        s.mats[i].roughness = m.roughness;
        s.mats[i].metallic = m.metallic;
        s.mats[i].emissive = ifx::mk(m.emissive[0], m.emissive[1], m.emissive[2]);
    }
    for (uint32_t i = 0; i < d.n_quad_lights; ++i) {
        ifx::HQuad q;
        for (int k = 0; k < 4; ++k) q.v[k] = ifx::mk(d.quad_lights[i].verts[k][0], d.quad_lights[i].verts[k][1],
                                                    d.quad_lights[i].verts[k][2]);
        q.emit = ifx::mk(d.quad_lights[i].emit[0], d.quad_lights[i].emit[1], d.quad_lights[i].emit[2]);
        s.quads.push_back(q);
    }
    for (uint32_t i = 0; i < d.n_sphere_lights; ++i) {
        const ifx_sphere_light& l = d.sphere_lights[i];
        s.spheres.push_back(ifx::HSphereLight{ifx::mk(l.center[0], l.center[1], l.center[2]), l.radius,
                                              ifx::mk(l.emit[0], l.emit[1], l.emit[2])});
    }
    for (uint32_t i = 0; i < d.n_point_lights; ++i) {
        const ifx_point_light& l = d.point_lights[i];
        s.points.push_back(ifx::HPointLight{ifx::mk(l.pos[0], l.pos[1], l.pos[2]), ifx::mk(l.emit[0], l.emit[1],
                                                                                          l.emit[2])});
    }
    s.cam.pos = ifx::mk(d.camera.pos[0], d.camera.pos[1], d.camera.pos[2]);
    s.cam.look = ifx::mk(d.camera.look[0], d.camera.look[1], d.camera.look[2]);
    s.cam.up = ifx::mk(d.camera.up[0], d.camera.up[1], d.camera.up[2]);
    s.cam.fov = d.camera.fov;
    s.cfg.width = d.width;
    s.cfg.height = d.height;
    s.cfg.surfels = d.surfels;
    s.cfg.bounces = d.bounces;
    s.cfg.exposure = d.exposure;
    if (!ifx::validate_cfg(s.cfg, err)) return IFX_E_INVALID;
    if (!finite3(d.lo) || !finite3(d.hi) || !(d.lo[0] <= d.hi[0] && d.lo[1] <= d.hi[1] && d.lo[2] <= d.hi[2])) {
        err = "scene bounds must be finite with lo <= hi";
        return IFX_E_INVALID;
    }
    const ifx::f3 bounds[2] = {ifx::mk(d.lo[0], d.lo[1], d.lo[2]), ifx::mk(d.hi[0], d.hi[1], d.hi[2])};
    try {
        if (d.n_bvh_nodes) {
            std::vector<ifx::HBvhNode> nodes(d.n_bvh_nodes);
            for (uint32_t i = 0; i < d.n_bvh_nodes; ++i) {
                const ifx_bvh_node& n = d.bvh_nodes[i];
                nodes[i] = ifx::HBvhNode{ifx::mk(n.lo[0], n.lo[1], n.lo[2]), ifx::mk(n.hi[0], n.hi[1], n.hi[2]),
                                         n.start, n.count, n.right};
            }
            std::vector<uint32_t> order(d.bvh_order, d.bvh_order + d.n_objects);
            ifx::finalize_scene_given(s, nodes, order, bounds[0], bounds[1]);
        } else {
            ifx::finalize_scene(s, bounds);
        }
    } catch (const std::runtime_error& e) {
        err = e.what();
        return IFX_E_INVALID;
    }
    return IFX_OK;
}
}  // namespace

ifx_status ifx_set_scene(ifx_ctx* ctx, const ifx_scene_desc* desc) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!desc) {
            ctx->err = "null scene descriptor";
            return IFX_E_INVALID;
        }
        ifx::HostScene s;
        if (ifx_status rc = scene_from(*desc, s, ctx->err)) return rc;
        int rc = ctx->eng.install_scene(std::move(s), ctx->err);
        ctx->have_scene = rc == 0;
        return rc;
    });
}

ifx_status ifx_scene_digest(ifx_ctx* ctx, uint64_t* digest) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (!digest) return IFX_E_INVALID;
        *digest = ifx::scene_digest(ctx->eng.scene);
        return IFX_OK;
    });
}

ifx_status ifx_get_scene_info(ifx_ctx* ctx, ifx_scene_info* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (!o) return IFX_E_INVALID;
        const ifx::HostScene& s = ctx->eng.scene;
        std::memset(o, 0, sizeof(*o));
        o->n_objects = (uint32_t)s.objects.size();
        o->n_planes = (uint32_t)s.planes.size();
        o->n_quad_lights = (uint32_t)s.quads.size();
        o->n_materials = (uint32_t)s.mats.size() - 1;
        o->n_bvh_nodes = (uint32_t)(s.node_rec.size() / 2);
        o->n_sphere_lights = (uint32_t)s.spheres.size();
        o->n_point_lights = (uint32_t)s.points.size();
        o->max_stack = s.max_stack;
        o->width = s.cfg.width;
        o->height = s.cfg.height;
        o->bounces = s.cfg.bounces;
        o->surfels = s.cfg.surfels;
        o->exposure = s.cfg.exposure;
        o->blend_k = s.blend_k;
        const float* v[5] = {&s.lo.x, &s.hi.x, &s.cam.pos.x, &s.cam.look.x, &s.cam.up.x};
        float* d[5] = {o->lo, o->hi, o->cam_pos, o->cam_look, o->cam_up};
        for (int k = 0; k < 5; ++k) std::memcpy(d[k], v[k], 12);
        o->cam_fov = s.cam.fov;
        return IFX_OK;
    });
}

ifx_status ifx_set_render(ifx_ctx* ctx, uint32_t width, uint32_t height, uint64_t surfels, uint32_t bounces) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        ifx::HCfg c = ctx->eng.scene.cfg;
        c.width = width;
        c.height = height;
        c.surfels = surfels;
        c.bounces = bounces;
        if (!ifx::validate_cfg(c, ctx->err)) return IFX_E_INVALID;
        ctx->eng.scene.cfg = c;
        return IFX_OK;
    });
}

ifx_status ifx_set_use_bvh(ifx_ctx* ctx, int32_t on) {
    return guarded(ctx, [&]() -> ifx_status {
        ctx->eng.use_bvh = on != 0;
        return IFX_OK;
    });
}

ifx_status ifx_generate_surfels(ifx_ctx* ctx, uint64_t target, uint64_t* placed) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (target == 0 || target > 2000000ull) {
            ctx->err = "surfel target must be in 1..2000000";
            return IFX_E_INVALID;
        }
        std::vector<ifx::HSurfel> s;
        if (ctx->eng.has_device() && ctx->eng.place_on_gpu) {
            int rc = ctx->eng.generate_surfels_gpu(target, s, ctx->err);
            if (rc) return rc;
        } else {
            s = ifx::generate_surfels(ctx->eng.scene, target, host_threads());
        }
        if (placed) *placed = s.size();
        return ctx->eng.set_surfels(s, ctx->err);
    });
}

ifx_status ifx_set_surfels(ifx_ctx* ctx, const float* p, const float* n, const float* albedo, const float* area,
                           uint64_t count) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (count && (!p || !n || !albedo || !area)) return IFX_E_INVALID;
        std::vector<ifx::HSurfel> s(count);
        for (uint64_t i = 0; i < count; ++i) {
            s[i].p = ifx::mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
            s[i].n = ifx::mk(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
            s[i].albedo = ifx::mk(albedo[3 * i], albedo[3 * i + 1], albedo[3 * i + 2]);
            s[i].area = area[i];
        }
        ifx::surfel_materials(ctx->eng.scene, s, host_threads());
        return ctx->eng.set_surfels(s, ctx->err);
    });
}

ifx_status ifx_set_surfel_materials(ifx_ctx* ctx, const uint32_t* mat, uint64_t count) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        std::vector<ifx::HSurfel> s = ctx->eng.surfels;
        if (count != s.size() || (count && !mat)) {
            ctx->err = "need one material per surfel (" + std::to_string(s.size()) + ")";
            return IFX_E_INVALID;
        }
        for (uint64_t i = 0; i < count; ++i) {
            if (mat[i] >= ctx->eng.scene.mats.size()) {
                ctx->err = "material index " + std::to_string(mat[i]) + " beyond the scene's materials";
                return IFX_E_INVALID;
            }
            s[i].mat = mat[i];
        }
        return ctx->eng.set_surfels(s, ctx->err);
    });
}

ifx_status ifx_num_surfels(ifx_ctx* ctx, uint64_t* count) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!count) return IFX_E_INVALID;
        *count = ctx->eng.surfels.size();
        return IFX_OK;
    });
}

ifx_status ifx_get_surfels(ifx_ctx* ctx, float* p, float* n, float* albedo, float* area) {
    return guarded(ctx, [&]() -> ifx_status {
        const auto& s = ctx->eng.surfels;
        for (size_t i = 0; i < s.size(); ++i) {
            if (p) std::memcpy(p + 3 * i, &s[i].p, 12);
            if (n) std::memcpy(n + 3 * i, &s[i].n, 12);
            if (albedo) std::memcpy(albedo + 3 * i, &s[i].albedo, 12);
            if (area) area[i] = s[i].area;
        }
        return IFX_OK;
    });
}

ifx_status ifx_transfer(ifx_ctx* ctx, int32_t cluster_res, ifx_transfer_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        auto t0 = std::chrono::steady_clock::now();
        ifx::TransferOut t;
        int rc = ctx->eng.transfer(cluster_res, t, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->clusters = t.clusters;
            o->candidates = t.candidates;
            o->links = t.links;
            o->evals = t.evals;
            o->mean_row_sum = t.mean_row_sum;
            o->ms = (float)ms_since(t0);
        }
        return IFX_OK;
    });
}

ifx_status ifx_get_transfer(ifx_ctx* ctx, uint32_t* starts, uint32_t* idx, float* w, float* sky, float* row_sum) {
    return guarded(ctx, [&]() -> ifx_status { return ctx->eng.download_transfer(starts, idx, w, sky, row_sum, ctx->err); });
}

ifx_status ifx_get_clusters(ifx_ctx* ctx, float* p, float* n, float* area, float* radius, uint32_t* member_starts,
                            uint32_t* members) {
    return guarded(ctx, [&]() -> ifx_status {
        return ctx->eng.download_clusters(p, n, area, radius, member_starts, members, ctx->err);
    });
}

ifx_status ifx_solve(ifx_ctx* ctx, uint32_t max_bounces, ifx_solve_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (max_bounces == 0 || max_bounces > 4096) {
            ctx->err = "bounces must be in 1..4096";
            return IFX_E_INVALID;
        }
        auto t0 = std::chrono::steady_clock::now();
        ifx::SolveOut r;
        int rc = ctx->eng.solve(max_bounces, r, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->iterations = r.iterations;
            o->residual = r.residual;
            o->evals = r.evals;
            o->ms = (float)ms_since(t0);
        }
        return IFX_OK;
    });
}

ifx_status ifx_solve_noalloc(ifx_ctx* ctx, uint32_t max_bounces, float* d_direct, float* d_irr, float* d_rad,
                             ifx_solve_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        auto t0 = std::chrono::steady_clock::now();
        ifx::SolveOut r;
        int rc = ctx->eng.solve_noalloc(max_bounces, (float4*)d_direct, (float4*)d_irr, (float4*)d_rad, r, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->iterations = r.iterations;
            o->residual = r.residual;
            o->evals = r.evals;
            o->ms = (float)ms_since(t0);
        }
        return IFX_OK;
    });
}

ifx_status ifx_get_radiosity(ifx_ctx* ctx, float* direct, float* irr, float* rad) {
    return guarded(ctx, [&]() -> ifx_status { return ctx->eng.download_radiosity(direct, irr, rad, ctx->err); });
}

ifx_status ifx_set_radiosity(ifx_ctx* ctx, const float* direct, const float* irr, const float* rad) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!ctx->eng.have_transfer) {
            ctx->err = "set_radiosity needs surfels uploaded by a transfer";
            return IFX_E_STATE;
        }
        return ctx->eng.upload_radiosity(direct, irr, rad, ctx->err);
    });
}

ifx_status ifx_shade(ifx_ctx* ctx, const ifx_camera* cam, uint32_t width, uint32_t height, int32_t mode,
                     float* film_rgb, uint8_t* rgb8, ifx_shade_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        ifx::HCamera c = ctx->eng.scene.cam;
        if (cam) {
            c.pos = ifx::mk(cam->pos[0], cam->pos[1], cam->pos[2]);
            c.look = ifx::mk(cam->look[0], cam->look[1], cam->look[2]);
            c.up = ifx::mk(cam->up[0], cam->up[1], cam->up[2]);
            c.fov = cam->fov;
        }
        auto t0 = std::chrono::steady_clock::now();
        uint64_t ev = 0;
        int rc = ctx->eng.shade(c, width, height, mode, film_rgb, rgb8, ev, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->evals = ev;
            o->ms = (float)ms_since(t0);
        }
        return IFX_OK;
    });
}

ifx_status ifx_shade_views(ifx_ctx* ctx, const ifx_camera* cams, uint32_t n, uint32_t width, uint32_t height,
                           int32_t mode, float* film_rgb, uint8_t* rgb8, ifx_shade_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (n && !cams) return IFX_E_INVALID;
        std::vector<ifx::HCamera> v(n);
        for (uint32_t i = 0; i < n; ++i) {
            v[i].pos = ifx::mk(cams[i].pos[0], cams[i].pos[1], cams[i].pos[2]);
            v[i].look = ifx::mk(cams[i].look[0], cams[i].look[1], cams[i].look[2]);
            v[i].up = ifx::mk(cams[i].up[0], cams[i].up[1], cams[i].up[2]);
            v[i].fov = cams[i].fov;
        }
        auto t0 = std::chrono::steady_clock::now();
        uint64_t ev = 0;
        int rc = ctx->eng.shade_views(v, width, height, mode, film_rgb, rgb8, ev, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->evals = ev;
            o->ms = (float)ms_since(t0);
        }
        return IFX_OK;
    });
}

ifx_status ifx_turntable_cameras(ifx_ctx* ctx, uint32_t n, ifx_camera* out) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (n && !out) return IFX_E_INVALID;
        std::vector<ifx::HCamera> v = ctx->eng.turntable(n);
        for (uint32_t i = 0; i < n; ++i) {
            out[i].pos[0] = v[i].pos.x, out[i].pos[1] = v[i].pos.y, out[i].pos[2] = v[i].pos.z;
            out[i].look[0] = v[i].look.x, out[i].look[1] = v[i].look.y, out[i].look[2] = v[i].look.z;
            out[i].up[0] = v[i].up.x, out[i].up[1] = v[i].up.y, out[i].up[2] = v[i].up.z;
            out[i].fov = v[i].fov;
        }
        return IFX_OK;
    });
}

ifx_status ifx_dist_mat(ifx_ctx* ctx, const float* pts, uint64_t n, float* d, uint32_t* mat, float* nrm) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (n && (!pts || !d || !mat)) return IFX_E_INVALID;
        return ctx->eng.probe_dist(pts, n, d, mat, nrm, ctx->err);
    });
}

ifx_status ifx_trace_rays(ifx_ctx* ctx, const float* ro, const float* rd, uint64_t n, float tmin, float tmax,
                          float* t, float* nrm, uint32_t* mat, int32_t* hit, uint32_t* steps) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (n && (!ro || !rd || !t || !nrm || !mat || !hit || !steps)) return IFX_E_INVALID;
        return ctx->eng.probe_trace(ro, rd, n, tmin, tmax, t, nrm, mat, hit, steps, ctx->err);
    });
}

ifx_status ifx_soft_shadows(ifx_ctx* ctx, const float* ro, const float* rd, uint64_t n, float tmin,
                            const float* tmax, const float* k, float* vis, uint32_t* steps) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        if (n && (!ro || !rd || !tmax || !k || !vis || !steps)) return IFX_E_INVALID;
        return ctx->eng.probe_shadow(ro, rd, n, tmin, tmax, k, vis, steps, ctx->err);
    });
}

ifx_status ifx_set_profiling(ifx_ctx* ctx, int32_t on) {
    return guarded(ctx, [&]() -> ifx_status {
        ctx->eng.set_profiling(on != 0);
        return IFX_OK;
    });
}

ifx_status ifx_kernel_time(ifx_ctx* ctx, const char* name, uint64_t* launches, double* total_ms) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!name) return IFX_E_INVALID;
        const auto& m = ctx->eng.kernel_stats();
        auto it = m.find(name);
        if (launches) *launches = it == m.end() ? 0 : it->second.launches;
        if (total_ms) *total_ms = it == m.end() ? 0.0 : it->second.ms;
        return IFX_OK;
    });
}

ifx_status ifx_reset_kernel_times(ifx_ctx* ctx) {
    return guarded(ctx, [&]() -> ifx_status {
        ctx->eng.reset_kernel_stats();
        return IFX_OK;
    });
}

ifx_status ifx_set_counting(ifx_ctx* ctx, int32_t on) {
    return guarded(ctx, [&]() -> ifx_status {
        if (on < 0 || on > 2) {
            ctx->err = "counting mode must be 0, 1 or 2";
            return IFX_E_INVALID;
        }
        ctx->eng.set_counting(on);
        return IFX_OK;
    });
}

ifx_status ifx_set_tuning(ifx_ctx* ctx, const char* key, int64_t value) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!key) return IFX_E_INVALID;
        std::string k(key);
        const int64_t hi = k == "link_budget" ? (int64_t)1 << 50 : (int64_t)1 << 30;
        if (value < 0 || value > hi) {
            ctx->err = "tuning value out of range";
            return IFX_E_INVALID;
        }
        if (k == "lds") ctx->eng.use_lds = value != 0;
        else if (k == "link_budget" || k == "adj_tile_rows" || k == "adj_seg") {
            if (k == "link_budget") ctx->eng.link_budget = (uint64_t)value;
            else if (k == "adj_tile_rows") ctx->eng.adj_tile_rows = (uint32_t)value;
            else ctx->eng.adj_seg = (uint32_t)value;
            ctx->eng.invalidate_adjoint();
        }
        else if (k == "vis_blocks") ctx->eng.vis_blocks = (uint32_t)value;
        else if (k == "vis_chunk") ctx->eng.vis_chunk = (uint32_t)value;
        else if (k == "vis_min_idle") {
            if (value < 1 || value > 64) {
                ctx->err = "vis_min_idle must be in 1..64";
                return IFX_E_INVALID;
            }
            ctx->eng.vis_min_idle = (uint32_t)value;
        } else if (k == "vis_tail") {
            if (value < 0 || value > 32) {
                ctx->err = "vis_tail must be in 0..32";
                return IFX_E_INVALID;
            }
            ctx->eng.vis_tail = (uint32_t)value;
        } else if (k == "vis_tail_items") {
            ctx->eng.vis_tail_items = (uint32_t)value;
        } else if (k == "closure_min_idle") {
            if (value < 1 || value > 64) {
                ctx->err = "closure_min_idle must be in 1..64";
                return IFX_E_INVALID;
            }
            ctx->eng.closure_min_idle = (uint32_t)value;
        }
        else if (k == "ff_blocks") ctx->eng.ff_blocks = (uint32_t)value;
        else if (k == "jacobi_blocks") ctx->eng.jacobi_blocks = (uint32_t)value;
        else if (k == "jacobi_rows") ctx->eng.jacobi_rows = (uint32_t)value;
        else if (k == "surfels_gpu") ctx->eng.place_on_gpu = value != 0;
        else if (k == "side_stream") ctx->eng.overlap_direct = value != 0;
        else if (k == "flat_query") ctx->eng.flat_query = value != 0;
        else if (k == "spp") {
            if (value < 1) {
                ctx->err = "spp must be >= 1";
                return IFX_E_INVALID;
            }
            ctx->eng.spp = (uint32_t)value;
        }

        else {
            ctx->err = "unknown tuning key `" + k + "`";
            return IFX_E_INVALID;
        }
        return IFX_OK;
    });
}

ifx_status ifx_set_material_albedo(ifx_ctx* ctx, uint32_t material, const float* rgb) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        return ctx->eng.set_material_albedo(material, rgb, ctx->err);
    });
}

ifx_status ifx_set_light_emit(ifx_ctx* ctx, uint32_t light, const float* rgb) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        return ctx->eng.set_light_emit(light, rgb, ctx->err);
    });
}

ifx_status ifx_set_solve_tolerance(ifx_ctx* ctx, float tol) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!(tol >= 0.0f) || !(tol < 1e30f)) {
            ctx->err = "tolerance must be finite and >= 0";
            return IFX_E_INVALID;
        }
        ctx->eng.solve_tol = tol;
        return IFX_OK;
    });
}

ifx_status ifx_set_adjoint(ifx_ctx* ctx, uint32_t max_iters, float rel_tol) {
    return guarded(ctx, [&]() -> ifx_status {
        if (max_iters == 0 || max_iters > 100000 || !(rel_tol >= 0.0f) || !(rel_tol < 1.0f)) {
            ctx->err = "adjoint settings out of range (1..100000 sweeps, 0 <= rel_tol < 1)";
            return IFX_E_INVALID;
        }
        ctx->eng.adj_max_iters = max_iters;
        ctx->eng.adj_tol = rel_tol;
        return IFX_OK;
    });
}

ifx_status ifx_backward(ifx_ctx* ctx, const float* dL_dfilm, float* dL_dalbedo, float* dL_demit,
                        ifx_backward_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        auto t0 = std::chrono::steady_clock::now();
        ifx::BackwardOut r;
        int rc = ctx->eng.backward(dL_dfilm, dL_dalbedo, dL_demit, r, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->iterations = r.iterations;
            o->residual = r.residual;
            o->ms = (float)ms_since(t0);
            o->link_tiles = r.link_tiles;
        }
        return IFX_OK;
    });
}

ifx_status ifx_backward_surfels(ifx_ctx* ctx, const float* dL_dirr, float* dL_dalbedo, float* dL_demit, float* mu,
                                ifx_backward_stats* o) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        auto t0 = std::chrono::steady_clock::now();
        ifx::BackwardOut r;
        int rc = ctx->eng.backward_surfels(dL_dirr, dL_dalbedo, dL_demit, mu, r, ctx->err);
        if (rc) return rc;
        if (o) {
            std::memset(o, 0, sizeof(*o));
            o->iterations = r.iterations;
            o->residual = r.residual;
            o->ms = (float)ms_since(t0);
            o->link_tiles = r.link_tiles;
        }
        return IFX_OK;
    });
}

ifx_status ifx_get_surfel_materials(ifx_ctx* ctx, uint32_t* mat) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!mat) return IFX_E_INVALID;
        const auto& s = ctx->eng.surfels;
        for (size_t i = 0; i < s.size(); ++i) mat[i] = s[i].mat;
        return IFX_OK;
    });
}

ifx_status ifx_get_direct_geometry(ifx_ctx* ctx, float* g) {
    return guarded(ctx, [&]() -> ifx_status {
        if (ifx_status s = need_scene(ctx)) return s;
        return ctx->eng.download_direct_geometry(g, ctx->err);
    });
}

ifx_status ifx_batch_create(ifx_batch** out, int32_t device) {
    if (!out) return IFX_E_INVALID;
    *out = nullptr;
    ifx_batch* b = new (std::nothrow) ifx_batch();
    if (!b) return IFX_E_NOMEM;
    int rc = b->b.init(device, b->err);
    *out = b;  // kept on failure so the caller can read the message
    return rc;
}

ifx_status ifx_batch_destroy(ifx_batch* b) {
    delete b;
    return IFX_OK;
}

const char* ifx_batch_last_error(const ifx_batch* b) { return b ? b->err.c_str() : "null batch"; }

ifx_status ifx_batch_add_scene(ifx_batch* b, const char* rad_text, int32_t cluster_res, uint32_t* index) {
    return bguarded(b, [&]() -> ifx_status {
        if (!rad_text) return IFX_E_INVALID;
        uint32_t i = 0;
        int rc = b->b.add_scene(rad_text, cluster_res, i, b->err);
        if (rc) return rc;
        if (index) *index = i;
        return IFX_OK;
    });
}

ifx_status ifx_batch_generate_surfels(ifx_batch* b) {
    return bguarded(b, [&]() -> ifx_status { return b->b.generate_surfels(b->err); });
}

ifx_status ifx_batch_set_graph(ifx_batch* b, int32_t on) {
    return bguarded(b, [&]() -> ifx_status {
        b->b.set_graph(on != 0);
        return IFX_OK;
    });
}

ifx_status ifx_batch_set_tuning(ifx_batch* b, uint32_t blocks_per_cu, uint32_t chunk_div, uint32_t wave_rows) {
    return bguarded(b, [&]() -> ifx_status {
        if (blocks_per_cu < 1 || blocks_per_cu > 64 || chunk_div < 1 || chunk_div > 256 || wave_rows > 1) {
            b->err = "tuning out of range (blocks_per_cu 1..64, chunk_div 1..256, wave_rows 0..1)";
            return IFX_E_INVALID;
        }
        b->b.tuning = ifx::BatchTuning{blocks_per_cu, chunk_div, wave_rows};
        b->b.set_graph(b->b.graph_active());  // a captured graph holds the old shapes
        return IFX_OK;
    });
}

ifx_status ifx_batch_autotune(ifx_batch* b, float* best_ms) {
    return bguarded(b, [&]() -> ifx_status {
        double ms = 0;
        int rc = b->b.autotune(&ms, b->err);
        if (rc) return rc;
        if (best_ms) *best_ms = (float)ms;
        return IFX_OK;
    });
}

ifx_status ifx_batch_run(ifx_batch* b, float* ms) {
    return bguarded(b, [&]() -> ifx_status {
        double t = 0;
        int rc = b->b.run(t, b->err);
        if (rc) return rc;
        if (ms) *ms = (float)t;
        return IFX_OK;
    });
}

ifx_status ifx_batch_get_stats(ifx_batch* b, uint32_t scene, ifx_batch_scene_stats* o) {
    return bguarded(b, [&]() -> ifx_status {
        if (!o || scene >= b->b.size()) return IFX_E_INVALID;
        const ifx::BatchSceneStats& s = b->b.stats(scene);
        o->surfels = s.surfels;
        o->candidates = s.candidates;
        o->links = s.links;
        o->evals_transfer = s.evals_transfer;
        o->evals_solve = s.evals_solve;
        o->evals_shade = s.evals_shade;
        o->clusters = s.clusters;
        o->iterations = s.iterations;
        o->width = s.width;
        o->height = s.height;
        o->residual = s.residual;
        o->mean_row_sum = s.mean_row_sum;
        return IFX_OK;
    });
}

ifx_status ifx_batch_get_image(ifx_batch* b, uint32_t scene, float* film_rgb, uint8_t* rgb8) {
    return bguarded(b, [&]() -> ifx_status { return b->b.download_image(scene, film_rgb, rgb8, b->err); });
}

ifx_status ifx_batch_get_radiosity(ifx_batch* b, uint32_t scene, float* direct, float* irr, float* rad) {
    return bguarded(b, [&]() -> ifx_status { return b->b.download_radiosity(scene, direct, irr, rad, b->err); });
}

ifx_status ifx_batch_get_transfer(ifx_batch* b, uint32_t scene, uint32_t* starts, uint32_t* idx, float* w) {
    return bguarded(b, [&]() -> ifx_status { return b->b.download_transfer(scene, starts, idx, w, b->err); });
}

ifx_status ifx_kernel_ops(ifx_ctx* ctx, const char* name, uint64_t* ops) {
    return guarded(ctx, [&]() -> ifx_status {
        if (!name || !ops) return IFX_E_INVALID;
        const auto& m = ctx->eng.kernel_ops();
        auto it = m.find(name);
        *ops = it == m.end() ? 0 : it->second;
        return IFX_OK;
    });
}

}  // extern "C"
