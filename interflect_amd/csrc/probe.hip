// Probe kernels: Scene::dist_mat + normal, trace and soft_shadow
// (src/scene.rs:187-236, src/trace.rs:27-124) at caller-given points and
// rays, one lane each with the per-lane traversal.
#include "ifx_kern.h"
#include "ifx_launch.h"

namespace ifx {

template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_probe_dist(SceneView Sg, LightsView Lg, LdsPlan P, const float* pts,
                                                      uint32_t n, float* d, uint32_t* mat, float* nrm) {
    IFX_SETUP(LDS, false, false);
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    f3 p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    DistHit h = dist_mat(S, p, stk);
    d[i] = h.d;
    mat[i] = h.mat;
    if (nrm) {
        f3 q = scene_normal(S, p, stk);
        nrm[3 * i] = q.x, nrm[3 * i + 1] = q.y, nrm[3 * i + 2] = q.z;
    }
}

template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_probe_trace(SceneView Sg, LightsView Lg, LdsPlan P, const float* ro,
                                                       const float* rd, uint32_t n, float tmin, float tmax, float* t,
                                                       float* nrm, uint32_t* mat, int32_t* hit, uint32_t* steps) {
    IFX_SETUP(LDS, false, false);
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t s = 0;
    Hit h = trace<true>(S, mk(ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]), mk(rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]),
                        tmin, tmax, s, stk);
    t[i] = h.t;
    nrm[3 * i] = h.n.x, nrm[3 * i + 1] = h.n.y, nrm[3 * i + 2] = h.n.z;
    mat[i] = h.mat;
    hit[i] = h.hit ? 1 : 0;
    steps[i] = s;
}

template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_probe_shadow(SceneView Sg, LightsView Lg, LdsPlan P, const float* ro,
                                                        const float* rd, uint32_t n, float tmin, const float* tmax,
                                                        const float* k, float* vis, uint32_t* steps) {
    IFX_SETUP(LDS, false, false);
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t s = 0;
    vis[i] = soft_shadow(S, mk(ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]), mk(rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]),
                         tmin, tmax[i], k[i], s, stk);
    steps[i] = s;
}

void launch_probe_dist(const ProbeLaunch& a, const float* pts, uint32_t n, float* d, uint32_t* mat, float* nrm) {
    const LightsView Lg{};
    hipStream_t stream = a.stream;  // the launch macro's stream
    LAUNCH_L(k_probe_dist, a.staged, a.blocks, a.lds_bytes, a.S, Lg, a.P, pts, n, d, mat, nrm);
}

void launch_probe_trace(const ProbeLaunch& a, const float* ro, const float* rd, uint32_t n, float tmin, float tmax,
                        float* t, float* nrm, uint32_t* mat, int32_t* hit, uint32_t* steps) {
    const LightsView Lg{};
    hipStream_t stream = a.stream;
    LAUNCH_L(k_probe_trace, a.staged, a.blocks, a.lds_bytes, a.S, Lg, a.P, ro, rd, n, tmin, tmax, t, nrm, mat, hit,
             steps);
}

void launch_probe_shadow(const ProbeLaunch& a, const float* ro, const float* rd, uint32_t n, float tmin,
                         const float* tmax, const float* k, float* vis, uint32_t* steps) {
    const LightsView Lg{};
    hipStream_t stream = a.stream;
    LAUNCH_L(k_probe_shadow, a.staged, a.blocks, a.lds_bytes, a.S, Lg, a.P, ro, rd, n, tmin, tmax, k, vis, steps);
}

}  // namespace ifx
