// ORACLE — test infrastructure only (see ovec.h).
// Restates src/trace.rs, src/ltc.rs, src/shade.rs; EXT light terms marked.
#include "olight.h"

#include <cfloat>

namespace orc {

// ---------------------------------------------------------------- trace.rs:27-58
Hit trace(const Scene& sc, V3 ro, V3 rd, float tmin, float tmax, uint32_t& steps_out) {
    float t = tmin;
    uint32_t steps = 0;
    while (steps < MAX_STEPS) {
        V3 p = ro + rd * t;
        DistHit h = sc.dist_mat(p);
        steps += 1;
        float eps = 1e-4f * fmax_(t, 1.0f);
        if (h.d < eps) {
            V3 n = sc.normal(p);
            steps_out = steps + 4;
            return Hit{true, t, p, n, h.mat};
        }
        t += h.d;
        if (t > tmax) break;
    }
    steps_out = steps;
    return Hit{false, 0.0f, splat(0.0f), splat(0.0f), 0};
}

// ---------------------------------------------------------------- trace.rs:84-124
float soft_shadow(const Scene& sc, V3 ro, V3 rd, float tmin, float tmax, float k, uint32_t& steps_out) {
    float res = 1.0f;
    float t = tmin;
    uint32_t steps = 0;
    float ph = 1e20f;
    while (steps < SHADOW_STEPS && t < tmax) {
        float h = sc.dist(ro + rd * t);
        steps += 1;
        if (h < 1e-4f) {
            steps_out = steps;
            return 0.0f;
        }
        float y, d;
        if (h < ph) {
            y = h * h / (2.0f * ph);
            d = std::sqrt(fmax_(h * h - y * y, 0.0f));
        } else {
            y = 0.0f;
            d = h;
        }
        res = fmin_(res, k * d / fmax_(t - y, 1e-4f));
        if (res < 1e-3f) {
            steps_out = steps;
            return 0.0f;
        }
        ph = h;
        t += clampf(h, 0.006f, 0.12f);
    }
    steps_out = steps;
    return clampf(res, 0.0f, 1.0f);
}

// ---------------------------------------------------------------- trace.rs:128-148
bool intersect_quad(const V3 verts[4], V3 ro, V3 rd, float& t_out) {
    V3 e1 = verts[1] - verts[0];
    V3 e2 = verts[3] - verts[0];
    V3 n = cross(e1, e2);
    float denom = dot(n, rd);
    if (std::fabs(denom) < 1e-9f) return false;
    float t = dot(verts[0] - ro, n) / denom;
    if (t < 1e-4f) return false;
    V3 p = (ro + rd * t) - verts[0];
    float u = dot(p, e1) / dot(e1, e1);
    float v = dot(p, e2) / dot(e2, e2);
    if (0.0f <= u && u <= 1.0f && 0.0f <= v && v <= 1.0f) {
        t_out = t;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- trace.rs:188-199
float radical_inverse(uint32_t i, uint32_t base) {
    float f = 1.0f;
    float r = 0.0f;
    float inv = 1.0f / (float)base;
    while (i > 0) {
        f *= inv;
        r += f * (float)(i % base);
        i /= base;
    }
    return r;
}

// ---------------------------------------------------------------- trace.rs:165-186
float hemisphere_closure(const Scene& sc, V3 p, V3 n, float tmax, uint32_t dirs, uint32_t& steps_out) {
    V3 t, b;
    frame(n, t, b);
    uint32_t open = 0, steps = 0;
    for (uint32_t k = 0; k < dirs; ++k) {
        float u1 = radical_inverse(k + 1, 2);
        float u2 = radical_inverse(k + 1, 3);
        float r = std::sqrt(u1);
        float phi = TAU_F * u2;
        V3 d = (t * (r * std::cos(phi)) + b * (r * std::sin(phi))) + n * std::sqrt(fmax_(1.0f - u1, 0.0f));
        uint32_t s = 0;
        Hit h = trace(sc, p + n * 1e-3f, normalize(d), 1e-3f, tmax, s);
        steps += s;
        if (!h.hit) open += 1;
    }
    steps_out = steps;
    return (float)open / (float)dirs;
}

// ---------------------------------------------------------------- ltc.rs:18-127
static const float INV_2PI = 0.15915494f;

static int clip_to_horizon(const V3 l[4], V3 out[5]) {
    int n = 0;
    for (int i = 0; i < 4; ++i) {
        V3 a = l[i];
        V3 b = l[(i + 1) % 4];
        float da = a.z, db = b.z;
        if (da >= 0.0f) out[n++] = a;
        if ((da >= 0.0f) != (db >= 0.0f)) {
            float t = da / (da - db);
            out[n++] = a + (b - a) * t;
        }
        if (n >= 5) break;
    }
    return n;
}

float quad_form_factor(const V3 l[4]) {
    V3 c[5];
    for (int i = 0; i < 5; ++i) c[i] = splat(0.0f);
    int n = clip_to_horizon(l, c);
    if (n < 3) return 0.0f;
    for (int i = 0; i < n; ++i) {
        float len = length(c[i]);
        if (len > 1e-9f) c[i] /= len;
    }
    float sum = 0.0f;
    for (int i = 0; i < n; ++i) {
        V3 a = c[i];
        V3 b = c[(i + 1) % n];
        float cos_t = clampf(dot(a, b), -1.0f, 1.0f);
        float theta = std::acos(cos_t);
        V3 cr = cross(a, b);
        float len = length(cr);
        if (len > 1e-9f) sum += theta * (cr.z / len);
    }
    return fmax_(sum * INV_2PI, 0.0f);
}

void frame(V3 n, V3& t, V3& b) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float bb = n.x * n.y * a;
    t = v3(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
    b = v3(bb, s + n.y * n.y * a, -n.y);
}

void to_local(const V3 verts[4], V3 p, V3 t, V3 b, V3 n, V3 out[4]) {
    for (int i = 0; i < 4; ++i) {
        V3 d = verts[3 - i] - p;
        out[i] = v3(dot(d, t), dot(d, b), dot(d, n));
    }
}

V3 closest_on_quad(const V3 verts[4], V3 q) {
    V3 o = verts[0];
    V3 e1 = verts[1] - o;
    V3 e2 = verts[3] - o;
    V3 d = q - o;
    float u = clampf(dot(d, e1) / dot(e1, e1), 0.0f, 1.0f);
    float v = clampf(dot(d, e2) / dot(e2, e2), 0.0f, 1.0f);
    return (o + e1 * u) + e2 * v;
}

// ---------------------------------------------------------------- shade.rs:10-43
V3 fresnel_schlick(float cos_theta, V3 f0) {
    float m = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    float m2 = m * m;
    return f0 + (splat(1.0f) - f0) * (m2 * m2 * m);
}

float d_ggx(float n_dot_h, float alpha) {
    float a2 = alpha * alpha;
    float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
    return a2 / fmax_(PI_F * d * d, 1e-9f);
}

float v_smith(float n_dot_v, float n_dot_l, float alpha) {
    float a2 = alpha * alpha;
    float gv = n_dot_l * std::sqrt(n_dot_v * n_dot_v * (1.0f - a2) + a2);
    float gl = n_dot_v * std::sqrt(n_dot_l * n_dot_l * (1.0f - a2) + a2);
    return 0.5f / fmax_(gv + gl, 1e-9f);
}

V3 f0_of(const Material& m) { return lerp(splat(0.04f), m.albedo, m.metallic); }
V3 diffuse_albedo(const Material& m) { return m.albedo * (1.0f - m.metallic); }

// EXT --------------------------------------------------------------------------
float sphere_form_factor(float cos_t, float s2) {
    float ill;
    if (cos_t * cos_t > s2) {
        ill = PI_F * s2 * clampf(cos_t, 0.0f, 1.0f);
    } else {
        float sin_t = std::sqrt(fmax_(1.0f - cos_t * cos_t, 0.0f));
        float x = std::sqrt(1.0f / s2 - 1.0f);
        float y = -x * (cos_t / sin_t);
        float sty = sin_t * std::sqrt(fmax_(1.0f - y * y, 0.0f));
        ill = (cos_t * std::acos(clampf(y, -1.0f, 1.0f)) - x * sty) * s2 + std::atan(sty / x);
    }
    return fmax_(ill, 0.0f) / PI_F;
}

// Shared geometry of one EXT emitter as seen from p: direction, distance,
// form factor, shadow radius.  Returns false when it cannot light p.
struct ExtLightGeo {
    V3 ldir;
    float dist, ff, radius, inv_d2;
};

static bool sphere_light_geo(const SphereLight& l, V3 p, V3 n, ExtLightGeo& g) {
    V3 to = l.c - p;
    float dist = length(to);
    if (!(dist > l.r)) return false;
    g.ldir = to / dist;
    g.dist = dist;
    float cos_t = dot(n, g.ldir);
    float s2 = (l.r * l.r) / (dist * dist);
    g.ff = sphere_form_factor(cos_t, s2);
    g.radius = fmax_(l.r, 1e-4f);
    g.inv_d2 = 0.0f;
    return g.ff > 1e-6f;
}

static bool point_light_geo(const PointLight& l, V3 p, V3 n, ExtLightGeo& g) {
    V3 to = l.pos - p;
    float d2 = dot(to, to);
    float dist = std::sqrt(d2);
    if (!(dist > 1e-6f)) return false;
    g.ldir = to / dist;
    g.dist = dist;
    float cos_t = dot(n, g.ldir);
    if (cos_t <= 0.0f) return false;
    g.ff = cos_t / (PI_F * d2);
    g.radius = 1e-4f;
    g.inv_d2 = 1.0f / d2;
    return true;
}

static float ext_shadow(const Scene& sc, V3 p, V3 n, const ExtLightGeo& g, uint32_t& steps) {
    float k = clampf(g.dist / g.radius, 1.5f, 96.0f);
    float shadow_end = fmax_(g.dist - g.radius * 1.25f, 0.05f);
    uint32_t s = 0;
    float vis = soft_shadow(sc, p + n * 3e-3f, g.ldir, 0.015f, shadow_end, k, s);
    steps += s;
    return vis;
}

// ---------------------------------------------------------------- shade.rs:50-125
V3 direct(const Scene& sc, V3 p, V3 n, V3 v, const Material& m, uint32_t& steps_out) {
    V3 out = splat(0.0f);
    uint32_t steps = 0;
    V3 tan, bit;
    frame(n, tan, bit);
    float n_dot_v = fmax_(dot(n, v), 1e-4f);
    float alpha = fmax_(m.roughness * m.roughness, 1e-3f);
    V3 f0 = f0_of(m);
    V3 kd = diffuse_albedo(m);

    for (const QuadLight& light : sc.lights) {
        V3 lc = light.center();
        V3 to_light = lc - p;
        if (dot(light.normal(), -normalize(to_light)) <= 0.0f) continue;
        V3 local[4];
        to_local(light.verts, p, tan, bit, n, local);
        float ff = quad_form_factor(local);
        if (ff <= 1e-6f) continue;
        float dist = length(to_light);
        float radius = fmax_(std::sqrt(light.area() / PI_F), 1e-4f);
        float k = clampf(dist / radius, 1.5f, 96.0f);
        V3 ldir = to_light / dist;
        float shadow_end = fmax_(dist - radius * 1.25f, 0.05f);
        uint32_t s = 0;
        float vis = soft_shadow(sc, p + n * 3e-3f, ldir, 0.015f, shadow_end, k, s);
        steps += s;
        if (vis <= 1e-4f) continue;
        V3 radiance = light.emit * vis;
        out += (kd * radiance) * ff;

        V3 r = (-v) - n * (2.0f * dot(n, -v));
        V3 plane_n = light.normal();
        float denom = dot(r, plane_n);
        V3 rep = lc;
        if (std::fabs(denom) > 1e-6f) {
            float t = dot(light.verts[0] - p, plane_n) / denom;
            if (t > 0.0f) rep = closest_on_quad(light.verts, p + r * t);
        }
        V3 l = normalize(rep - p);
        float n_dot_l = dot(n, l);
        if (n_dot_l <= 0.0f) continue;
        float rep_dist = fmax_(length(rep - p), 1e-4f);
        float wide = clampf(alpha + radius / (2.0f * rep_dist), alpha, 1.0f);
        float energy = (alpha / wide) * (alpha / wide);
        V3 h = normalize(l + v);
        float d = d_ggx(fmax_(dot(n, h), 0.0f), wide);
        float vis_s = v_smith(n_dot_v, n_dot_l, wide);
        V3 f = fresnel_schlick(fmax_(dot(l, h), 0.0f), f0);
        out += (f * (d * vis_s * energy * n_dot_l)) * radiance;
    }

    // EXT: sphere lights (representative point on the sphere, Karis 2013) and
    // point lights (the point itself; intensity falls off as 1/d^2).
    for (const SphereLight& light : sc.sphere_lights) {
        ExtLightGeo g;
        if (!sphere_light_geo(light, p, n, g)) continue;
        float vis = ext_shadow(sc, p, n, g, steps);
        if (vis <= 1e-4f) continue;
        V3 radiance = light.emit * vis;
        out += (kd * radiance) * g.ff;
        V3 r = (-v) - n * (2.0f * dot(n, -v));
        V3 to = light.c - p;
        V3 ctr = r * dot(to, r) - to;
        float lc = length(ctr);
        V3 rep_v = lc > 1e-9f ? to + ctr * clampf(light.r / lc, 0.0f, 1.0f) : to;
        V3 l = normalize(rep_v);
        float n_dot_l = dot(n, l);
        if (n_dot_l <= 0.0f) continue;
        float rep_dist = fmax_(length(rep_v), 1e-4f);
        float wide = clampf(alpha + g.radius / (2.0f * rep_dist), alpha, 1.0f);
        float energy = (alpha / wide) * (alpha / wide);
        V3 h = normalize(l + v);
        float d = d_ggx(fmax_(dot(n, h), 0.0f), wide);
        float vis_s = v_smith(n_dot_v, n_dot_l, wide);
        V3 f = fresnel_schlick(fmax_(dot(l, h), 0.0f), f0);
        out += (f * (d * vis_s * energy * n_dot_l)) * radiance;
    }
    for (const PointLight& light : sc.point_lights) {
        ExtLightGeo g;
        if (!point_light_geo(light, p, n, g)) continue;
        float vis = ext_shadow(sc, p, n, g, steps);
        if (vis <= 1e-4f) continue;
        out += (kd * light.emit) * (g.ff * vis);
        V3 l = g.ldir;
        float n_dot_l = dot(n, l);
        V3 h = normalize(l + v);
        float d = d_ggx(fmax_(dot(n, h), 0.0f), alpha);
        float vis_s = v_smith(n_dot_v, n_dot_l, alpha);
        V3 f = fresnel_schlick(fmax_(dot(l, h), 0.0f), f0);
        out += (f * (d * vis_s * n_dot_l)) * (light.emit * (vis * g.inv_d2));
    }
    steps_out = steps;
    return out;
}

// ---------------------------------------------------------------- shade.rs:133-164
V3 direct_diffuse(const Scene& sc, V3 p, V3 n, V3 albedo, uint32_t& steps_out) {
    V3 out = splat(0.0f);
    uint32_t steps = 0;
    V3 tan, bit;
    frame(n, tan, bit);
    for (const QuadLight& light : sc.lights) {
        V3 lc = light.center();
        V3 to_light = lc - p;
        if (dot(light.normal(), -normalize(to_light)) <= 0.0f) continue;
        V3 local[4];
        to_local(light.verts, p, tan, bit, n, local);
        float ff = quad_form_factor(local);
        if (ff <= 1e-6f) continue;
        float dist = length(to_light);
        float radius = fmax_(std::sqrt(light.area() / PI_F), 1e-4f);
        float k = clampf(dist / radius, 1.5f, 96.0f);
        float shadow_end = fmax_(dist - radius * 1.25f, 0.05f);
        uint32_t s = 0;
        float vis = soft_shadow(sc, p + n * 3e-3f, to_light / dist, 0.015f, shadow_end, k, s);
        steps += s;
        if (vis <= 1e-4f) continue;
        out += (albedo * light.emit) * (ff * vis);
    }
    for (const SphereLight& light : sc.sphere_lights) {  // EXT
        ExtLightGeo g;
        if (!sphere_light_geo(light, p, n, g)) continue;
        float vis = ext_shadow(sc, p, n, g, steps);
        if (vis <= 1e-4f) continue;
        out += (albedo * light.emit) * (g.ff * vis);
    }
    for (const PointLight& light : sc.point_lights) {  // EXT
        ExtLightGeo g;
        if (!point_light_geo(light, p, n, g)) continue;
        float vis = ext_shadow(sc, p, n, g, steps);
        if (vis <= 1e-4f) continue;
        out += (albedo * light.emit) * (g.ff * vis);
    }
    steps_out = steps;
    return out;
}

// ---------------------------------------------------------------- shade.rs:168-181
bool emitter_hit(const Scene& sc, V3 ro, V3 rd, float tmax, V3& emit) {
    bool found = false;
    float bt = 0.0f;
    for (const QuadLight& l : sc.lights) {
        float t;
        if (intersect_quad(l.verts, ro, rd, t)) {
            if (t < tmax && dot(l.normal(), rd) < 0.0f && (!found || t < bt)) {
                found = true;
                bt = t;
                emit = l.emit;
            }
        }
    }
    // EXT: sphere emitters are visible from outside (nearest front hit).
    for (const SphereLight& l : sc.sphere_lights) {
        V3 oc = ro - l.c;
        float b = dot(oc, rd);
        float c = dot(oc, oc) - l.r * l.r;
        float disc = b * b - c;
        if (c <= 0.0f || disc < 0.0f) continue;
        float t = -b - std::sqrt(disc);
        if (t < 1e-4f) continue;
        if (t < tmax && (!found || t < bt)) {
            found = true;
            bt = t;
            emit = l.emit;
        }
    }
    return found;
}

}  // namespace orc
