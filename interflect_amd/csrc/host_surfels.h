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

}  // namespace ifx
