// ORACLE — test infrastructure only (see ovec.h).  CPU restatement of the
// reference's scene model, SDF primitives, BVH and `.rad` loader:
//   src/sdf.rs, src/bvh.rs, src/scene.rs
// plus the documented config extensions (subtract_cylinder, blend groups,
// sphere and point lights) that the reference lacks; each extension is
// marked EXT and is inert on scenes written in the reference's own grammar.
#pragma once
#include <string>
#include <vector>

#include "ovec.h"

namespace orc {

// src/sdf.rs:9-48
enum PrimKind { SPHERE = 0, ROUNDBOX = 1, PLANE = 2, CAPSULE = 3, TORUS = 4, CYLINDER = 5, CONE = 6 };

struct Prim {
    PrimKind kind;
    V3 c, a, b, n;       // centre / endpoints / box half-size / plane normal
    float r = 0, h = 0;  // radius / plane height
    float major = 0, minor = 0, ra = 0, rb = 0;
    float dist(V3 p) const;                 // src/sdf.rs:50-125
    bool bounds(V3& lo, V3& hi) const;      // src/sdf.rs:127-148
};

// Operation tally of the roofline's counting rule (test instrumentation, no
// counterpart in the reference): FP32 operations per primitive kind, +2 per
// CSG subtraction, 18 per AABB distance (the root once, then both children
// of every interior node expanded) and 15 per scene-distance sample, plus
// the number of primitive evaluations.  Accumulated on the calling thread
// while g_tally points at a tally (TallyScope sets it when tallying is on).
struct OpTally {
    uint64_t ops = 0, prims = 0;
    void add(const OpTally& o) {
        ops += o.ops;
        prims += o.prims;
    }
};
extern thread_local OpTally* g_tally;
extern bool g_tally_on;
static const uint32_t kTallyPrimOps[7] = {10u, 23u, 6u, 24u, 13u, 40u, 57u};
static const uint32_t kTallyAabbOps = 18u, kTallyEvalOps = 15u;
struct TallyScope {
    OpTally* prev;
    explicit TallyScope(OpTally& t) : prev(g_tally) { g_tally = g_tally_on ? &t : nullptr; }
    ~TallyScope() { g_tally = prev; }
};

// src/sdf.rs:151-173
struct Object {
    Prim prim;
    std::vector<Prim> subtract;
    uint32_t mat = 0;
    float blend = 0.0f;
    float dist(V3 p) const {
        float d = prim.dist(p);
        if (g_tally) g_tally->ops += kTallyPrimOps[prim.kind], g_tally->prims += 1;
        for (const Prim& s : subtract) {
            d = fmax_(d, -s.dist(p));
            if (g_tally) g_tally->ops += kTallyPrimOps[s.kind] + 2u, g_tally->prims += 1;
        }
        return d;
    }
};

// src/scene.rs:11-28
struct Material {
    V3 albedo = splat(0.8f);
    float roughness = 0.8f;
    float metallic = 0.0f;
    V3 emissive = splat(0.0f);
};

// src/scene.rs:32-56
struct QuadLight {
    V3 verts[4];
    V3 emit;
    V3 center() const { return (((verts[0] + verts[1]) + verts[2]) + verts[3]) * 0.25f; }
    V3 normal() const { return normalize(cross(verts[1] - verts[0], verts[3] - verts[0])); }
    float area() const {
        float a = length(cross(verts[1] - verts[0], verts[3] - verts[0]));
        float b = length(cross(verts[3] - verts[2], verts[1] - verts[2]));
        return 0.5f * (a + b);
    }
};

// EXT: sphere and point emitters (configs with sphere / point lights).
struct SphereLight {
    V3 c;
    float r;
    V3 emit;
};
struct PointLight {
    V3 pos;
    V3 emit;
};

// src/scene.rs:58-75
struct Camera {
    V3 pos = v3(0.0f, 2.0f, 6.0f);
    V3 look = v3(0.0f, 1.0f, 0.0f);
    V3 up = v3(0.0f, 1.0f, 0.0f);
    float fov = 40.0f;
};

// src/scene.rs:77-161
struct RenderCfg {
    uint32_t width = 800, height = 600;
    uint64_t surfels = 20000;
    uint32_t bounces = 12;
    float exposure = 1.0f;
    bool validate(std::string& err) const;
};

struct DistHit {
    float d;
    uint32_t mat;
};

// src/bvh.rs:26-36
struct BvhNode {
    V3 lo, hi;
    uint32_t start = 0, count = 0, right = 0;
};

// src/bvh.rs:38-163
struct Bvh {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> order;
    bool empty = true;
    void build(const std::vector<Object>& objects);
    void query(V3 p, const std::vector<Object>& objects, DistHit& best) const;
    void query_linear(V3 p, const std::vector<Object>& objects, DistHit& best) const;
    // EXT: blend-group variants (smooth union of the two nearest blended
    // objects); identical to query/query_linear when no object blends.
    void query_blend(V3 p, const std::vector<Object>& objects, float k, DistHit& hard, DistHit& b1,
                     float& b2, bool linear) const;
};

// src/scene.rs:163-178
struct Scene {
    std::vector<Object> objects;
    std::vector<Object> planes;
    std::vector<Material> materials;
    std::vector<QuadLight> lights;
    std::vector<SphereLight> sphere_lights;  // EXT
    std::vector<PointLight> point_lights;    // EXT
    Camera camera;
    RenderCfg cfg;
    Bvh bvh;
    bool use_bvh = true;
    V3 lo, hi;           // world bounds
    float blend_k = 0;   // EXT: largest blend radius; 0 = reference behaviour

    DistHit dist_mat(V3 p) const;   // src/scene.rs:191-208
    float dist(V3 p) const { return dist_mat(p).d; }
    V3 normal(V3 p) const;          // src/scene.rs:218-236
    Material material(uint32_t id) const {
        return id < materials.size() ? materials[id] : Material{};
    }
    float diag() const { return length(hi - lo); }
};

// src/scene.rs:365-571.  Returns false and sets err on malformed input.
bool parse_scene(const std::string& src, Scene& out, std::string& err);

// EXT: polynomial smooth minimum (Quilez), k > 0.
static inline float smin_poly(float a, float b, float k) {
    float h = fmax_(k - std::fabs(a - b), 0.0f) / k;
    return fmin_(a, b) - h * h * k * 0.25f;
}

}  // namespace orc
