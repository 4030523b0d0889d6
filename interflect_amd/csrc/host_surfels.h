// Host-side surfel placement (the pipeline input; src/surfel.rs:147-255).
#pragma once
#include <vector>

#include "host_scene.h"

namespace ifx {

struct HSurfel {
    f3 p, n, albedo;
    float area;
    uint32_t mat = 0xffffffffu;  // material index when known (surfel placement), else unknown
};

std::vector<HSurfel> generate_surfels(const HostScene& sc, uint64_t target, int threads);
// Each surfel's material from the scene at its position (Scene::dist_mat,
// src/scene.rs:191-208), as surfel placement assigns it.
void surfel_materials(const HostScene& sc, std::vector<HSurfel>& s, int threads);

}  // namespace ifx
