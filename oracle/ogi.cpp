// ORACLE — test infrastructure only (see ovec.h).
// Restates src/surfel.rs, src/formfactor.rs, src/solve.rs, src/main.rs
// (Cam, render_frame) and src/film.rs.
#include "ogi.h"

#include <cfloat>
#include <map>
#include <tuple>

namespace orc {

void par_for(size_t n, int threads, const std::function<void(size_t)>& f) {
    if (threads <= 1 || n < 2) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    const size_t chunk = 16;
    auto worker = [&]() {
        for (;;) {
            size_t b = next.fetch_add(chunk);
            if (b >= n) break;
            size_t e = std::min(n, b + chunk);
            for (size_t i = b; i < e; ++i) f(i);
        }
    };
    std::vector<std::thread> pool;
    int nt = (int)std::min<size_t>((size_t)threads, (n + chunk - 1) / chunk);
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------- surfel.rs:71-140
void Grid::build(const std::vector<V3>& points, V3 lo_, V3 hi, float cell) {
    lo = lo_;
    inv_cell = 1.0f / cell;
    V3 ext = hi - lo;
    dim[0] = clampi(as_i32(std::ceil(ext.x * inv_cell)), 1, 512);
    dim[1] = clampi(as_i32(std::ceil(ext.y * inv_cell)), 1, 512);
    dim[2] = clampi(as_i32(std::ceil(ext.z * inv_cell)), 1, 512);
    size_t ncells = (size_t)(dim[0] * dim[1] * dim[2]);
    std::vector<uint32_t> counts(ncells + 1, 0);
    std::vector<uint32_t> cidx(points.size());
    for (size_t i = 0; i < points.size(); ++i) {
        int32_t c[3];
        cell_of(points[i], c);
        cidx[i] = (uint32_t)(c[0] + dim[0] * (c[1] + dim[1] * c[2]));
        counts[cidx[i] + 1] += 1;
    }
    for (size_t i = 0; i < ncells; ++i) counts[i + 1] += counts[i];
    starts = counts;
    std::vector<uint32_t> cursor = counts;
    items.assign(points.size(), 0);
    for (size_t i = 0; i < points.size(); ++i) items[cursor[cidx[i]]++] = (uint32_t)i;
}

void Grid::cell_of(V3 p, int32_t c[3]) const {
    c[0] = clampi(as_i32((p.x - lo.x) * inv_cell), 0, dim[0] - 1);
    c[1] = clampi(as_i32((p.y - lo.y) * inv_cell), 0, dim[1] - 1);
    c[2] = clampi(as_i32((p.z - lo.z) * inv_cell), 0, dim[2] - 1);
}

// ---------------------------------------------------------------- surfel.rs:147-255
std::vector<Surfel> generate(const Scene& sc, size_t target) {
    V3 lo = sc.lo, hi = sc.hi;
    float diag = length(hi - lo);
    float d_min = fmax_(diag / std::sqrt((float)target), 1e-4f);
    float cell = d_min;
    std::vector<V3> pts;
    std::vector<Surfel> out;
    pts.reserve(target);
    out.reserve(target);
    float inv_cell = 1.0f / cell;
    int32_t dim[3] = {clampi(as_i32(std::ceil((hi.x - lo.x) * inv_cell)), 1, 512),
                      clampi(as_i32(std::ceil((hi.y - lo.y) * inv_cell)), 1, 512),
                      clampi(as_i32(std::ceil((hi.z - lo.z) * inv_cell)), 1, 512)};
    size_t ncells = (size_t)(dim[0] * dim[1] * dim[2]);
    std::vector<std::vector<uint32_t>> occ(ncells);
    auto cell_idx = [&](V3 p, int32_t c[3]) {
        c[0] = clampi(as_i32((p.x - lo.x) * inv_cell), 0, dim[0] - 1);
        c[1] = clampi(as_i32((p.y - lo.y) * inv_cell), 0, dim[1] - 1);
        c[2] = clampi(as_i32((p.z - lo.z) * inv_cell), 0, dim[2] - 1);
    };
    float eps = d_min * 0.02f;
    float d_min_sq = d_min * d_min;
    uint64_t s24 = (uint64_t)(uint32_t)target * 24ull;
    uint32_t seeds = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(s24, UINT32_MAX), 6000000ull);

    for (uint32_t i = 0; i < seeds; ++i) {
        if (out.size() >= target) break;
        V3 s = v3(radical_inverse(i + 1, 2), radical_inverse(i + 1, 3), radical_inverse(i + 1, 5));
        V3 p = lo + (hi - lo) * s;
        bool ok = false;
        for (int it = 0; it < 10; ++it) {
            float d = sc.dist(p);
            if (std::fabs(d) < eps) {
                ok = true;
                break;
            }
            if (!std::isfinite(d) || std::fabs(d) > diag) break;
            p -= sc.normal(p) * d;
        }
        if (!ok || std::fabs(sc.dist(p)) >= eps) continue;
        if (p.x < lo.x || p.y < lo.y || p.z < lo.z || p.x > hi.x || p.y > hi.y || p.z > hi.z) continue;
        int32_t c[3];
        cell_idx(p, c);
        bool too_close = false;
        for (int32_t z = std::max(c[2] - 1, 0); z <= std::min(c[2] + 1, dim[2] - 1) && !too_close; ++z)
            for (int32_t y = std::max(c[1] - 1, 0); y <= std::min(c[1] + 1, dim[1] - 1) && !too_close; ++y)
                for (int32_t x = std::max(c[0] - 1, 0); x <= std::min(c[0] + 1, dim[0] - 1) && !too_close; ++x) {
                    size_t ci = (size_t)(x + dim[0] * (y + dim[1] * z));
                    for (uint32_t j : occ[ci]) {
                        if (length_sq(pts[j] - p) < d_min_sq) {
                            too_close = true;
                            break;
                        }
                    }
                }
        if (too_close) continue;
        DistHit hit = sc.dist_mat(p);
        Material m = sc.material(hit.mat);
        V3 n = sc.normal(p);
        size_t ci = (size_t)(c[0] + dim[0] * (c[1] + dim[1] * c[2]));
        occ[ci].push_back((uint32_t)out.size());
        pts.push_back(p);
        Surfel sf;
        sf.p = p;
        sf.n = n;
        sf.albedo = diffuse_albedo(m);
        sf.area = d_min * d_min / 0.65f;
        sf.direct = sf.irr = sf.rad = splat(0.0f);
        out.push_back(sf);
    }
    return out;
}

// ---------------------------------------------------------------- surfel.rs:262-293
V3 gather(const std::vector<Surfel>& s, const Grid& g, V3 p, V3 n, float radius) {
    V3 acc = splat(0.0f);
    float wsum = 0.0f;
    float r2 = radius * radius;
    int32_t c[3];
    g.cell_of(p, c);
    for (int32_t z = std::max(c[2] - 1, 0); z <= std::min(c[2] + 1, g.dim[2] - 1); ++z)
        for (int32_t y = std::max(c[1] - 1, 0); y <= std::min(c[1] + 1, g.dim[1] - 1); ++y)
            for (int32_t x = std::max(c[0] - 1, 0); x <= std::min(c[0] + 1, g.dim[0] - 1); ++x) {
                size_t ci = (size_t)(x + g.dim[0] * (y + g.dim[1] * z));
                for (uint32_t k = g.starts[ci]; k < g.starts[ci + 1]; ++k) {
                    const Surfel& sf = s[g.items[k]];
                    float d2 = length_sq(sf.p - p);
                    if (d2 > r2) continue;
                    float align = dot(sf.n, n);
                    if (align <= 0.15f) continue;
                    float t = 1.0f - d2 / r2;
                    float w = t * t * align;
                    acc += sf.irr * w;
                    wsum += w;
                }
            }
    if (wsum > 1e-6f) return acc / wsum;
    return splat(0.0f);
}

// ---------------------------------------------------------------- formfactor.rs:70-147
int normal_bucket(V3 n) {
    V3 a = vabs(n);
    if (a.x >= a.y && a.x >= a.z) return n.x >= 0.0f ? 0 : 1;
    if (a.y >= a.z) return n.y >= 0.0f ? 2 : 3;
    return n.z >= 0.0f ? 4 : 5;
}

std::vector<Cluster> build_clusters(const Scene& sc, const std::vector<Surfel>& surfels, int32_t res) {
    V3 lo = sc.lo, hi = sc.hi;
    V3 ext = hi - lo;
    V3 inv = v3((float)res / fmax_(ext.x, 1e-6f), (float)res / fmax_(ext.y, 1e-6f), (float)res / fmax_(ext.z, 1e-6f));
    std::map<std::tuple<int32_t, int32_t, int32_t, int>, std::vector<uint32_t>> map;
    for (size_t i = 0; i < surfels.size(); ++i) {
        V3 c = (surfels[i].p - lo) * inv;
        auto key = std::make_tuple(clampi(as_i32(c.x), 0, res - 1), clampi(as_i32(c.y), 0, res - 1),
                                   clampi(as_i32(c.z), 0, res - 1), normal_bucket(surfels[i].n));
        map[key].push_back((uint32_t)i);
    }
    std::vector<Cluster> out;
    out.reserve(map.size());
    for (auto& kv : map) {
        V3 p = splat(0.0f), n = splat(0.0f);
        float area = 0.0f;
        for (uint32_t m : kv.second) {
            const Surfel& s = surfels[m];
            p += s.p * s.area;
            n += s.n * s.area;
            area += s.area;
        }
        if (area <= 1e-12f) continue;
        V3 nn = length_sq(n) > 1e-12f ? normalize(n) : v3(0.0f, 1.0f, 0.0f);
        Cluster cl;
        cl.p = p / area;
        cl.n = nn;
        cl.area = area;
        cl.radius = std::sqrt(area / PI_F);
        cl.members = std::move(kv.second);
        out.push_back(std::move(cl));
    }
    return out;
}

// ---------------------------------------------------------------- formfactor.rs:153-277
static const float MAX_CALIBRATION = 8.0f;

uint64_t build_transfer(const Scene& sc, const std::vector<Surfel>& surfels, std::vector<Cluster> clusters,
                        Transfer& out, int threads, size_t row_begin, size_t row_end) {
    size_t nc = clusters.size();
    float diag = length(sc.hi - sc.lo);
    size_t N = surfels.size();
    row_end = std::min(row_end, N);
    struct Row {
        std::vector<uint32_t> idx;
        std::vector<float> w;
        uint64_t evals = 0;
        float sum = 0, sky = 0;
        OpTally vis;
    };
    std::vector<Row> rows(N);
    par_for(row_end > row_begin ? row_end - row_begin : 0, threads, [&](size_t rr) {
        size_t i = row_begin + rr;
        const Surfel& s = surfels[i];
        Row& row = rows[i];
        float sum = 0.0f;
        TallyScope tally_vis(row.vis);
        for (size_t ci = 0; ci < nc; ++ci) {
            const Cluster& c = clusters[ci];
            V3 d = c.p - s.p;
            float r2 = length_sq(d);
            if (r2 < 1e-9f) continue;
            float r = std::sqrt(r2);
            V3 dir = d / r;
            float cos_i = dot(s.n, dir);
            float cos_c = dot(c.n, -dir);
            if (cos_i <= 1e-3f || cos_c <= 1e-3f) continue;
            float ff = cos_i * cos_c * c.area / (PI_F * r2 + c.area);
            if (ff < 2e-5f) continue;
            float k = clampf(r / fmax_(c.radius, 1e-4f), 1.5f, 32.0f);
            uint32_t e = 0;
            float vis = soft_shadow(sc, s.p + s.n * (diag * 5e-4f), dir, diag * 1e-3f, r - c.radius * 0.75f, k, e);
            row.evals += e;
            if (vis <= 1e-3f) continue;
            float v = ff * vis;
            row.idx.push_back((uint32_t)ci);
            row.w.push_back(v);
            sum += v;
        }
        g_tally = tally_vis.prev;  // the closure is not part of the link tally
        uint32_t e = 0;
        float sky = hemisphere_closure(sc, s.p, s.n, diag * 2.0f, 16, e);
        row.evals += e;
        row.sky = sky;
        float target = clampf(1.0f - sky, 0.0f, 1.0f);
        if (sum > 1e-5f) {
            float scale = fmin_(target / sum, MAX_CALIBRATION);
            for (float& v : row.w) v *= scale;
            sum *= scale;
        }
        row.sum = fmin_(sum, 1.0f);
    });

    size_t links = 0;
    uint64_t evals = 0;
    size_t interior = 0;
    out.vis_tally = OpTally();
    for (const Row& r : rows) {
        out.vis_tally.add(r.vis);
        links += r.idx.size();
        evals += r.evals;
        if (!r.idx.empty()) interior += 1;
    }
    float acc = -0.0f;  // Rust's f32 Sum folds from -0.0
    for (const Row& r : rows)
        if (!r.idx.empty()) acc += r.sum;
    out.mean_row_sum = interior == 0 ? 0.0f : acc / (float)interior;
    out.starts.clear();
    out.idx.clear();
    out.w.clear();
    out.starts.reserve(N + 1);
    out.idx.reserve(links);
    out.w.reserve(links);
    out.sky.resize(N);
    out.row_sum.resize(N);
    out.starts.push_back(0);
    for (size_t i = 0; i < N; ++i) {
        out.idx.insert(out.idx.end(), rows[i].idx.begin(), rows[i].idx.end());
        out.w.insert(out.w.end(), rows[i].w.begin(), rows[i].w.end());
        out.starts.push_back((uint32_t)out.idx.size());
        out.sky[i] = rows[i].sky;
        out.row_sum[i] = rows[i].sum;
    }
    out.clusters = std::move(clusters);
    out.nc = nc;
    out.links = links;
    return evals;
}

// ---------------------------------------------------------------- solve.rs:27-39
uint64_t light_surfels(const Scene& sc, std::vector<Surfel>& surfels, int threads, OpTally* tally) {
    std::vector<uint64_t> ev(surfels.size(), 0);
    std::vector<OpTally> tl(tally ? surfels.size() : 0);
    par_for(surfels.size(), threads, [&](size_t i) {
        Surfel& s = surfels[i];
        OpTally none;
        TallyScope ts(tally ? tl[i] : none);
        uint32_t e = 0;
        V3 c = direct_diffuse(sc, s.p, s.n, s.albedo, e);
        s.direct = c;
        s.rad = c;
        s.irr = splat(0.0f);
        ev[i] = e;
    });
    uint64_t tot = 0;
    for (uint64_t e : ev) tot += e;
    if (tally)
        for (const OpTally& t : tl) tally->add(t);
    return tot;
}

// ---------------------------------------------------------------- solve.rs:47-108
SolveStats solve(std::vector<Surfel>& surfels, const Transfer& t, uint32_t max_bounces, int threads) {
    size_t nc = t.nc;
    if (nc == 0) return SolveStats{0, 0.0f};
    std::vector<V3> cluster_rad(nc, splat(0.0f));
    std::vector<float> deltas(surfels.size(), 0.0f);
    float residual = 0.0f;
    uint32_t iterations = 0;
    for (uint32_t it = 0; it < max_bounces; ++it) {
        for (size_t c = 0; c < nc; ++c) {
            V3 acc = splat(0.0f);
            for (uint32_t m : t.clusters[c].members) {
                const Surfel& s = surfels[m];
                acc += s.rad * s.area;
            }
            cluster_rad[c] = acc / fmax_(t.clusters[c].area, 1e-12f);
        }
        par_for(surfels.size(), threads, [&](size_t i) {
            Surfel& s = surfels[i];
            V3 irr = splat(0.0f);
            for (uint32_t k = t.starts[i]; k < t.starts[i + 1]; ++k) irr += cluster_rad[t.idx[k]] * t.w[k];
            V3 new_rad = s.direct + s.albedo * irr;
            deltas[i] = max_elem(vabs(new_rad - s.rad));
            s.irr = irr;
            s.rad = new_rad;
        });
        residual = 0.0f;
        for (float d : deltas) residual = fmax_(residual, d);
        iterations = it + 1;
        if (residual < 1e-4f) break;
    }
    return SolveStats{iterations, residual};
}

// ---------------------------------------------------------------- main.rs:188-218
Cam Cam::make(V3 pos, V3 look, V3 up_hint, float fov, uint32_t w, uint32_t h) {
    Cam c;
    c.origin = pos;
    c.fwd = normalize(look - pos);
    c.right = normalize(cross(c.fwd, up_hint));
    c.up = cross(c.right, c.fwd);
    const float rads_per_deg = PI_F / 180.0f;
    c.tan_half = std::tan(fov * rads_per_deg * 0.5f);
    c.aspect = (float)w / (float)h;
    return c;
}

V3 Cam::ray(uint32_t x, uint32_t y, uint32_t w, uint32_t h) const {
    float px = (2.0f * ((float)x + 0.5f) / (float)w - 1.0f) * aspect * tan_half;
    float py = (1.0f - 2.0f * ((float)y + 0.5f) / (float)h) * tan_half;
    return normalize((fwd + right * px) + up * py);
}

static const V3 BG = {0.02f, 0.03f, 0.05f};

// ---------------------------------------------------------------- main.rs:229-367
// src/reference.rs:48-58
static V3 cosine_hemisphere(V3 n, Rng& rng) {
    V3 t, b;
    frame(n, t, b);
    float u1 = rng.next_f32();
    float u2 = rng.next_f32();
    float r = std::sqrt(u1);
    float phi = TAU_F * u2;
    return normalize((t * (r * std::cos(phi)) + b * (r * std::sin(phi))) + n * std::sqrt(fmax_(1.0f - u1, 0.0f)));
}

// Quad emitter seen along a ray (src/shade.rs:166-181 over sc.lights only,
// as the path tracer calls it).
static bool quad_emitter_hit(const Scene& sc, V3 ro, V3 rd, float tmax, V3& emit) {
    bool found = false;
    float bt = 0.0f;
    for (const QuadLight& l : sc.lights) {
        float t;
        if (intersect_quad(l.verts, ro, rd, t) && t < tmax && dot(l.normal(), rd) < 0.0f && (!found || t < bt)) {
            found = true;
            bt = t;
            emit = l.emit;
        }
    }
    return found;
}

V3 ref_path(const Scene& sc, V3 ro, V3 rd, float tmax, uint32_t max_depth, Rng& rng) {
    V3 radiance = splat(0.0f), throughput = splat(1.0f);
    V3 origin = ro, dir = rd;
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        uint32_t steps = 0;
        Hit hit = trace(sc, origin, dir, 1e-3f, tmax, steps);
        float surf_t = hit.hit ? hit.t : FLT_MAX;
        V3 e;
        if (quad_emitter_hit(sc, origin, dir, fmin_(surf_t, tmax), e)) {
            if (depth == 0) radiance += throughput * e;
            break;
        }
        if (!hit.hit) break;
        Material m = sc.material(hit.mat);
        V3 albedo = diffuse_albedo(m);
        radiance += throughput * m.emissive;
        for (const QuadLight& light : sc.lights) {
            V3 e1 = light.verts[1] - light.verts[0];
            V3 e2 = light.verts[3] - light.verts[0];
            float r1 = rng.next_f32();
            V3 a = light.verts[0] + e1 * r1;
            float r2 = rng.next_f32();
            V3 lp = a + e2 * r2;
            V3 to = lp - hit.p;
            float dist2 = length_sq(to);
            if (dist2 < 1e-9f) continue;
            float dist = std::sqrt(dist2);
            V3 l = to / dist;
            float cos_s = dot(hit.n, l);
            float cos_l = dot(light.normal(), -l);
            if (cos_s <= 0.0f || cos_l <= 0.0f) continue;
            uint32_t s2 = 0;
            Hit occ = trace(sc, hit.p + hit.n * 1e-3f, l, 1e-3f, dist - 1e-2f, s2);
            if (occ.hit) continue;
            float pdf = dist2 / (cos_l * light.area());
            radiance += ((throughput * albedo) * light.emit) * (cos_s / (PI_F * pdf));
        }
        throughput = throughput * albedo;
        dir = cosine_hemisphere(hit.n, rng);
        origin = hit.p + hit.n * 1e-3f;
        if (depth >= 3) {
            float q = clampf(max_elem(throughput), 0.05f, 0.95f);
            if (rng.next_f32() > q) break;
            throughput = throughput / q;
        }
    }
    return radiance;
}

uint64_t render_frame(const Scene& sc, const Cam& cam, Mode mode, const Gi* gi, float tmax, uint32_t w,
                      uint32_t h, std::vector<V3>& film, int threads, uint32_t y0, uint32_t y1, uint32_t spp,
                      OpTally* tally) {
    film.assign((size_t)w * h, splat(0.0f));
    y1 = std::min(y1, h);
    std::vector<uint64_t> row_evals(h, 0);
    std::vector<OpTally> row_tally(tally ? h : 0);
    bool tonemapped = mode == BEAUTY || mode == DIRECT || mode == INDIRECT || mode == REFERENCE;
    par_for(y1 > y0 ? y1 - y0 : 0, threads, [&](size_t yy) {
        uint32_t y = y0 + (uint32_t)yy;
        uint64_t local = 0;
        OpTally none;
        TallyScope ts(tally ? row_tally[y] : none);
        for (uint32_t x = 0; x < w; ++x) {
            V3& px = film[(size_t)y * w + x];
            V3 rd = cam.ray(x, y, w, h);
            if (mode == REFERENCE) {  // src/main.rs:249-258
                Rng rng(((uint64_t)y << 32) | (uint64_t)x);
                V3 acc = splat(0.0f);
                for (uint32_t k = 0; k < spp; ++k) acc += ref_path(sc, cam.origin, rd, tmax, 12, rng);
                px = acc / (float)spp;
                continue;
            }
            uint32_t steps = 0;
            Hit hh = trace(sc, cam.origin, rd, 1e-3f, tmax, steps);
            local += steps;
            float surf_t = hh.hit ? hh.t : FLT_MAX;
            V3 e;
            if (emitter_hit(sc, cam.origin, rd, fmin_(surf_t, tmax), e)) {
                px = tonemapped ? e : splat(0.9f);
                continue;
            }
            if (!hh.hit) {
                px = BG;
                continue;
            }
            if (mode == NORMALS) {
                px = hh.n * 0.5f + splat(0.5f);
                continue;
            }
            if (mode == DEPTH) {
                float d = 1.0f - clampf(hh.t / tmax, 0.0f, 1.0f);
                px = splat(d * d);
                continue;
            }
            if (mode == STEPS) {
                float f = (float)steps / (float)MAX_STEPS;
                px = v3(f, 1.0f - f, 0.2f);
                continue;
            }
            Material m = sc.material(hh.mat);
            V3 c = splat(0.0f);
            if (mode != INDIRECT) {
                uint32_t s = 0;
                V3 d = direct(sc, hh.p, hh.n, -rd, m, s);
                local += s;
                c += d + m.emissive;
            }
            if (gi) {
                V3 irr = gather(*gi->surfels, gi->grid, hh.p, hh.n, gi->radius);
                c += diffuse_albedo(m) * irr;
                if (m.roughness < 0.45f || m.metallic > 0.05f) {
                    V3 refl = rd - hh.n * (2.0f * dot(hh.n, rd));
                    uint32_t s = 0;
                    Hit rh = trace(sc, hh.p + hh.n * 1e-3f, refl, 1e-3f, tmax, s);
                    local += s;
                    float rt = rh.hit ? rh.t : FLT_MAX;
                    V3 refl_col;
                    V3 e2;
                    if (emitter_hit(sc, hh.p, refl, fmin_(rt, tmax), e2)) {
                        refl_col = e2;
                    } else if (rh.hit) {
                        Material m2 = sc.material(rh.mat);
                        uint32_t s2 = 0;
                        V3 d2 = direct(sc, rh.p, rh.n, -refl, m2, s2);
                        local += s2;
                        V3 i2 = gather(*gi->surfels, gi->grid, rh.p, rh.n, gi->radius);
                        refl_col = (d2 + diffuse_albedo(m2) * i2) + m2.emissive;
                    } else {
                        refl_col = BG;
                    }
                    V3 f = fresnel_schlick(fmax_(dot(hh.n, -rd), 0.0f), f0_of(m));
                    float gloss = clampf(1.0f - m.roughness / 0.45f, 0.0f, 1.0f);
                    gloss = gloss * gloss;
                    c += (f * refl_col) * gloss;
                }
            }
            px = c;
        }
        row_evals[y] = local;
    });
    uint64_t tot = 0;
    for (uint64_t e : row_evals) tot += e;
    if (tally)
        for (const OpTally& t : row_tally) tally->add(t);
    return tot;
}

// ---------------------------------------------------------------- film.rs:8-61
V3 aces(V3 x) {
    const float A = 2.51f, B = 0.03f, C = 2.43f, D = 0.59f, E = 0.14f;
    V3 num = x * (x * A + splat(B));
    V3 den = x * (x * C + splat(D)) + splat(E);
    V3 r = num / den;
    return vmin(vmax(r, splat(0.0f)), splat(1.0f));
}

uint8_t srgb8(float c) {
    c = clampf(c, 0.0f, 1.0f);
    float v = c <= 0.0031308f ? c * 12.92f : 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
    return as_u8(v * 255.0f + 0.5f);
}

void encode(const std::vector<V3>& film, float exposure, bool tonemap, std::vector<uint8_t>& out) {
    out.resize(film.size() * 3);
    for (size_t i = 0; i < film.size(); ++i) {
        V3 scaled = film[i] * exposure;
        V3 m = tonemap ? aces(scaled) : scaled;
        out[3 * i + 0] = srgb8(m.x);
        out[3 * i + 1] = srgb8(m.y);
        out[3 * i + 2] = srgb8(m.z);
    }
}

}  // namespace orc
