// Host-side scene model of the product library: the `.rad` loader, the
// binned-SAH BVH build and the flattening of objects, lights and materials
// into the device records of ifx_sdf.h / ifx_shade.h.
//
// The reference keeps these on its host too (src/scene.rs:247-601,
// src/bvh.rs:52-91,166-309); they run once per scene and are not on the
// timed transfer -> solve -> shade path.
#pragma once
#include <string>
#include <vector>

#include "ifx_shade.h"

namespace ifx {

struct HPrim {
    uint32_t kind = K_SPHERE;
    f3 c{0, 0, 0}, a{0, 0, 0}, b{0, 0, 0}, n{0, 1, 0};
    float r = 0, h = 0, major = 0, minor = 0, ra = 0, rb = 0;
};

struct HObject {
    HPrim prim;
    std::vector<HPrim> subs;
    uint32_t mat = 0;
    float blend = 0;
};

struct HMaterial {
    f3 albedo{0.8f, 0.8f, 0.8f};
    float roughness = 0.8f, metallic = 0.0f;
    f3 emissive{0, 0, 0};
};

struct HQuad {
    f3 v[4];
    f3 emit;
};
struct HSphereLight {
    f3 c;
    float r;
    f3 emit;
};
struct HPointLight {
    f3 pos, emit;
};

struct HCamera {
    f3 pos{0, 2, 6}, look{0, 1, 0}, up{0, 1, 0};
    float fov = 40.0f;
};

struct HCfg {
    uint32_t width = 800, height = 600;
    uint64_t surfels = 20000;
    uint32_t bounces = 12;
    float exposure = 1.0f;
};

struct HostScene {
    std::vector<HObject> objects, planes;
    std::vector<HMaterial> mats;
    std::vector<HQuad> quads;
    std::vector<HSphereLight> spheres;
    std::vector<HPointLight> points;
    HCamera cam;
    HCfg cfg;

    // Derived by finalize().
    f3 lo, hi;
    float blend_k = 0;
    uint32_t max_stack = 0;            // deepest BVH traversal stack
    std::vector<float4> obj_rec;       // 4 per object, BVH leaf order
    std::vector<float4> plane_rec;     // 4 per plane
    std::vector<float4> sub_rec;       // 4 per subtraction primitive
    std::vector<float4> node_rec;      // 2 per BVH node
    std::vector<float4> obox_rec;      // 2 per object (lo, hi of its BVH box), BVH leaf order
    std::vector<uint32_t> order;       // BVH leaf slot -> object index
    std::vector<uint32_t> lin;         // object index -> leaf slot, then slot -> BVH leaf node
    std::vector<float4> quad_rec, sph_rec, pnt_rec, mat_rec;

    float diag() const { return len3(hi - lo); }
    SceneView host_view() const;
    LightsView host_lights() const;
};

// Parses `.rad` text (src/scene.rs:365-571) and finalizes the scene.
bool load_rad(const std::string& src, HostScene& out, std::string& err);
// Render-settings limits (src/scene.rs:94-148).
bool validate_cfg(const HCfg& c, std::string& err);
// Rebuilds the BVH, bounds (unless given: lo, hi) and device records from
// the object lists.  Throws std::runtime_error when the BVH cannot be walked.
void finalize_scene(HostScene& s, const f3* bounds = nullptr);
// A caller's BVH node (src/bvh.rs:26-36).
struct HBvhNode {
    f3 lo, hi;
    uint32_t start, count, right;
};
// Device records from the object lists with a given BVH (nodes over `order`,
// object indices in leaf order) and bounds; throws std::runtime_error when
// the BVH does not cover every object once or cannot be walked.
void finalize_scene_given(HostScene& s, const std::vector<HBvhNode>& nodes, const std::vector<uint32_t>& order,
                          f3 lo, f3 hi);
// 64-bit digest of the flattened records, BVH, bounds, camera and settings.
uint64_t scene_digest(const HostScene& s);
// The four device records of one material (albedo+roughness, emissive+metallic, f0, kd).
void material_records(const HMaterial& m, float4 out[4]);

}  // namespace ifx
