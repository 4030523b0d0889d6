// ORACLE — test infrastructure only (see ovec.h).
// Restates src/trace.rs, src/ltc.rs and src/shade.rs of the reference, plus
// the EXT sphere / point light terms.
#pragma once
#include "oscene.h"

namespace orc {

static const uint32_t MAX_STEPS = 128;    // src/trace.rs:10
static const uint32_t SHADOW_STEPS = 96;  // src/trace.rs:11

struct Hit {
    bool hit;
    float t;
    V3 p, n;
    uint32_t mat;
};

// src/trace.rs:27-58
Hit trace(const Scene& sc, V3 ro, V3 rd, float tmin, float tmax, uint32_t& steps);
// src/trace.rs:84-124
float soft_shadow(const Scene& sc, V3 ro, V3 rd, float tmin, float tmax, float k, uint32_t& steps);
// src/trace.rs:128-148
bool intersect_quad(const V3 verts[4], V3 ro, V3 rd, float& t);
// src/trace.rs:165-186
float hemisphere_closure(const Scene& sc, V3 p, V3 n, float tmax, uint32_t dirs, uint32_t& steps);
// src/trace.rs:188-199 (also src/surfel.rs:44-54)
float radical_inverse(uint32_t i, uint32_t base);

// src/ltc.rs
float quad_form_factor(const V3 l[4]);
void frame(V3 n, V3& t, V3& b);
void to_local(const V3 verts[4], V3 p, V3 t, V3 b, V3 n, V3 out[4]);
V3 closest_on_quad(const V3 verts[4], V3 q);

// src/shade.rs
V3 fresnel_schlick(float cos_theta, V3 f0);
// src/shade.rs:17-31
float d_ggx(float n_dot_h, float alpha);
float v_smith(float n_dot_v, float n_dot_l, float alpha);
V3 f0_of(const Material& m);
V3 diffuse_albedo(const Material& m);
V3 direct(const Scene& sc, V3 p, V3 n, V3 v, const Material& m, uint32_t& steps);
V3 direct_diffuse(const Scene& sc, V3 p, V3 n, V3 albedo, uint32_t& steps);
bool emitter_hit(const Scene& sc, V3 ro, V3 rd, float tmax, V3& emit);

// EXT: form factor (cosine-weighted solid angle / pi) of a sphere of
// sin^2(angular radius) s2 whose centre is at cos(elevation) cos_t,
// horizon-clipped (Snyder 1996, as given in Lagarde & de Rousiers 2014).
float sphere_form_factor(float cos_t, float s2);

}  // namespace orc
