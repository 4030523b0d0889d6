// Host-side surfel placement for the product library (the pipeline input;
// src/surfel.rs:147-255 of the reference).  Halton seeds are projected onto
// the zero level set by Newton steps in parallel batches (each seed is
// independent), then accepted strictly in seed order by Poisson-disc
// rejection, so the surfel set is a pure function of the scene and target.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "host_surfels.h"
#include "ifx_surfel.h"

namespace ifx {

namespace {
struct Seed {
    f3 p, n;
    uint32_t mat;
    bool ok;
};
}  // namespace

std::vector<HSurfel> generate_surfels(const HostScene& sc, uint64_t target, int threads) {
    std::vector<HSurfel> out;
    if (target == 0) return out;
    const SceneView S = sc.host_view();
    const LightsView L = sc.host_lights();
    const SurfelParams P = surfel_params(sc.lo, sc.hi, target);
    const int32_t* dim = P.dim;
    const float d_min_sq = P.d_min_sq;
    const float area = P.area;
    const uint32_t seeds = P.seeds;

    auto project = [&](uint32_t i, Seed& sd) {
        ArrStack stk;
        f3 p;
        sd.ok = project_seed(S, P, i, p, stk);
        sd.p = p;
        if (sd.ok) {
            sd.mat = dist_mat(S, p, stk).mat;
            sd.n = scene_normal(S, p, stk);
        }
    };

    std::vector<std::vector<uint32_t>> occ((size_t)dim[0] * dim[1] * dim[2]);
    const uint32_t batch = 16384;
    std::vector<Seed> buf(batch);
    int nt = std::max(1, threads);
    for (uint32_t b0 = 0; b0 < seeds && out.size() < target; b0 += batch) {
        uint32_t nb = std::min(batch, seeds - b0);
        auto work = [&](int t) {
            for (uint32_t k = (uint32_t)t; k < nb; k += (uint32_t)nt) project(b0 + k, buf[k]);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();

        for (uint32_t k = 0; k < nb; ++k) {
            if (out.size() >= target) break;
            const Seed& sd = buf[k];
            if (!sd.ok) continue;
            int32_t c[3];
            surfel_cell(P, sd.p, c);
            bool close = false;
            for (int32_t z = std::max(c[2] - 1, 0); z <= std::min(c[2] + 1, dim[2] - 1) && !close; ++z)
                for (int32_t y = std::max(c[1] - 1, 0); y <= std::min(c[1] + 1, dim[1] - 1) && !close; ++y)
                    for (int32_t x = std::max(c[0] - 1, 0); x <= std::min(c[0] + 1, dim[0] - 1) && !close; ++x)
                        for (uint32_t j : occ[(size_t)(x + dim[0] * (y + dim[1] * z))]) {
                            f3 dq = out[j].p - sd.p;
                            if (dot3(dq, dq) < d_min_sq) {
                                close = true;
                                break;
                            }
                        }
            if (close) continue;
            occ[(size_t)(c[0] + dim[0] * (c[1] + dim[1] * c[2]))].push_back((uint32_t)out.size());
            MatV m = load_mat(L, sd.mat);
            out.push_back(HSurfel{sd.p, sd.n, m.kd, area, sd.mat});
        }
    }
    return out;
}

void surfel_materials(const HostScene& sc, std::vector<HSurfel>& s, int threads) {
    const SceneView S = sc.host_view();
    const int nt = std::max(1, threads);
    auto work = [&](int t) {
        ArrStack stk;
        for (size_t i = (size_t)t; i < s.size(); i += (size_t)nt) s[i].mat = dist_mat(S, s[i].p, stk).mat;
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
}

}  // namespace ifx
