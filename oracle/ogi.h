// ORACLE — test infrastructure only (see ovec.h).
// Restates src/surfel.rs, src/formfactor.rs, src/solve.rs and the frame
// pipeline of src/main.rs (Cam, render_frame) and src/film.rs.
#pragma once
#include <atomic>
#include <functional>
#include <thread>
#include <vector>

#include "olight.h"

namespace orc {

// src/surfel.rs:22-38
struct Surfel {
    V3 p, n, albedo;
    float area;
    V3 direct, irr, rad;
};

// Order-preserving parallel loop: each index writes only its own slot, so
// the result is identical for any thread count (the reference's
// par_iter().collect()).
void par_for(size_t n, int threads, const std::function<void(size_t)>& f);

// src/surfel.rs:61-141
struct Grid {
    V3 lo;
    float inv_cell;
    int32_t dim[3];
    std::vector<uint32_t> starts, items;
    void build(const std::vector<V3>& points, V3 lo, V3 hi, float cell);
    void cell_of(V3 p, int32_t c[3]) const;
};

// src/surfel.rs:147-255
std::vector<Surfel> generate(const Scene& sc, size_t target);
// src/surfel.rs:262-293
V3 gather(const std::vector<Surfel>& s, const Grid& g, V3 p, V3 n, float radius);

// src/formfactor.rs:31-39
struct Cluster {
    V3 p, n;
    float area, radius;
    std::vector<uint32_t> members;
};

// src/formfactor.rs:48-66
struct Transfer {
    std::vector<Cluster> clusters;
    float mean_row_sum = 0;
    std::vector<uint32_t> starts, idx;
    std::vector<float> w;
    size_t nc = 0, links = 0;
    // Per-row closure sky fraction and final row sum (diagnostics only).
    std::vector<float> sky, row_sum;
    OpTally vis_tally;  // the link shadow marches (when tallying is on)
};

int normal_bucket(V3 n);
std::vector<Cluster> build_clusters(const Scene& sc, const std::vector<Surfel>& s, int32_t res);
// src/formfactor.rs:153-277.  Rows [row_begin, row_end) only when a sample
// is asked for (CPU-baseline timing); the full matrix otherwise.
uint64_t build_transfer(const Scene& sc, const std::vector<Surfel>& s, std::vector<Cluster> clusters,
                        Transfer& out, int threads, size_t row_begin = 0, size_t row_end = SIZE_MAX);

// src/solve.rs:27-108
uint64_t light_surfels(const Scene& sc, std::vector<Surfel>& s, int threads, OpTally* tally = nullptr);
struct SolveStats {
    uint32_t iterations;
    float residual;
};
SolveStats solve(std::vector<Surfel>& s, const Transfer& t, uint32_t max_bounces, int threads);

// src/main.rs:25-51
enum Mode { BEAUTY = 0, DIRECT = 1, INDIRECT = 2, NORMALS = 3, DEPTH = 4, STEPS = 5, REFERENCE = 6 };

// src/reference.rs:24-46: the path tracer's PCG32 (output of the advanced
// state, XSH-RR, 24-bit float).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 6364136223846793005ull + 1ull) {}
    float next_f32() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        uint64_t x = ((s >> 18) ^ s) >> 27;
        uint32_t rot = (uint32_t)(s >> 59);
        uint32_t xv = (uint32_t)x;
        uint32_t v = (xv >> rot) | (xv << ((32u - rot) & 31u));
        return (float)(v >> 8) * (1.0f / 16777216.0f);
    }
};

// src/reference.rs:61-137: one path (quad lights; next-event estimation,
// cosine-weighted bounces, Russian roulette after depth 3).
V3 ref_path(const Scene& sc, V3 ro, V3 rd, float tmax, uint32_t max_depth, Rng& rng);

// src/main.rs:188-218
struct Cam {
    V3 origin, fwd, right, up;
    float tan_half, aspect;
    static Cam make(V3 pos, V3 look, V3 up_hint, float fov, uint32_t w, uint32_t h);
    V3 ray(uint32_t x, uint32_t y, uint32_t w, uint32_t h) const;
};

struct Gi {
    const std::vector<Surfel>* surfels;
    Grid grid;
    float radius;
};

// src/main.rs:229-367, all modes (REFERENCE: `spp` paths per pixel, seeded
// per pixel, src/main.rs:249-258).  Rows [y0, y1) of the film only when a
// sample is asked for.
uint64_t render_frame(const Scene& sc, const Cam& cam, Mode mode, const Gi* gi, float tmax, uint32_t w,
                      uint32_t h, std::vector<V3>& film, int threads, uint32_t y0 = 0, uint32_t y1 = UINT32_MAX,
                      uint32_t spp = 512, OpTally* tally = nullptr);

// src/film.rs:8-61
uint8_t srgb8(float c);
V3 aces(V3 x);
void encode(const std::vector<V3>& film, float exposure, bool tonemap, std::vector<uint8_t>& out);

}  // namespace orc
