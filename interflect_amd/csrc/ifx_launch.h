// Launchers of the camera-ray shading kernels and the probe kernels.  Their
// kernels are the largest code objects of the library (every march step of
// shade_pixel inlines the scene query), so each group sits in a translation
// unit of its own (shade_*.hip, probe.hip) and the build compiles them in
// parallel; the engine calls them through these plain functions.
#pragma once
#include "engine.h"

namespace ifx {

// One launch of k_shade (cams == nullptr) or k_shade_views (grid z = view,
// camera blockIdx.z of `cams`).
struct ShadeLaunch {
    SceneView S;
    LightsView L;
    LdsPlan P;
    ShadeArgs A;
    const CamV* cams;
    float4* film;
    unsigned long long* evals;
    unsigned long long* ops;
    dim3 grid;
    size_t lds_bytes;
    bool staged;  // scene records staged in LDS
    hipStream_t stream;
};

// Counting instantiation (per-lane traversal, CountStack): shade_count.hip.
void launch_shade_count(const ShadeLaunch& a);
// Timed, per-lane traversal (LaneStack): shade_lane.hip.
void launch_shade_lane(const ShadeLaunch& a);
// Timed, wavefront-shared queries (LdsStack), scene in LDS / in global memory:
// shade_wave_lds.hip, shade_wave_glb.hip.
void launch_shade_wave_lds(const ShadeLaunch& a);
void launch_shade_wave_glb(const ShadeLaunch& a);

// Probe kernels (probe.hip): one query per point / ray, per-lane traversal.
struct ProbeLaunch {
    SceneView S;
    LdsPlan P;
    uint32_t blocks;
    size_t lds_bytes;
    bool staged;
    hipStream_t stream;
};
void launch_probe_dist(const ProbeLaunch& a, const float* pts, uint32_t n, float* d, uint32_t* mat, float* nrm);
void launch_probe_trace(const ProbeLaunch& a, const float* ro, const float* rd, uint32_t n, float tmin, float tmax,
                        float* t, float* nrm, uint32_t* mat, int32_t* hit, uint32_t* steps);
void launch_probe_shadow(const ProbeLaunch& a, const float* ro, const float* rd, uint32_t n, float tmin,
                         const float* tmax, const float* k, float* vis, uint32_t* steps);

}  // namespace ifx
