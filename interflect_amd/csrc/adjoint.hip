// Backward pass of the frozen-geometry pipeline: gradients of a loss on the
// linear beauty film (before exposure and tone mapping) with respect to the
// material albedos and the light emissions.
//
// With the geometry fixed, the link weights W, the cluster aggregation P
// (area-weighted member mean, src/solve.rs:60-72), the per-light factors
// g_il = form factor x visibility of light_surfels (src/shade.rs:137-162) and
// the gather weights (src/surfel.rs:262-293) do not depend on the parameters.
// The solve's fixed point is
//     irr = M (E + rho (.) irr),   M = W P,   E_i = rho_i (.) D_i,   D_i = sum_l e_l g_il,
// and with ubar = dL/dirr the implicit adjoint is
//     mu = M^T (ubar + rho (.) mu),
//     dL/drho_i = mu_i (.) (D_i + irr_i),   dL/de_l = sum_i mu_i (.) rho_i g_il.
// The film side re-traces each pixel's primary ray: a pixel that sees light l
// gives dL/de_l += u; a diffuse hit adds u (.) (kd a_l + s_l) to dL/de_l,
// u (.) (e_l a_l + gathered irr) to dL/dkd of its material, and scatters
// gamma_ps kd (.) u to dL/dirr_s over the gather weights gamma.
//
// Every sum has a fixed order, so results are bit-identical across runs and
// launch shapes: transposed links are summed per fixed-size segment by a
// fixed lane tree, the gather scatter adds 64-bit fixed-point integers, and
// parameter sums are per-block f64 trees that the host adds in block order.
#include <cmath>

#include "ifx_kern.h"

namespace ifx {
namespace {

constexpr uint32_t kNoMat = 0xffffffffu;
constexpr uint32_t kParamChunk = 1024;  // items per block of the parameter sums

// Power-of-two scale of the fixed-point scatter.  A pixel spreads kd (.) u
// over gather weights that sum to one, so no accumulator exceeds
// npx * max|u| * max kd; the scale keeps that bound below 2^60.
__host__ __device__ inline double fix_scale(float umax, uint64_t npx, float kd_max) {
    double bound = (double)npx * (double)umax * (double)kd_max;
    if (!(bound > 0.0) || !(bound < 1e300)) return 1099511627776.0;  // 2^40
    int e = 0;
    (void)frexp(bound, &e);  // bound < 2^e
    return ldexp(1.0, 60 - e);
}

__device__ __forceinline__ float comp(const float4& v, uint32_t c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

// ---------------------------------------------------------------- film side
// max |u| as non-negative float bits: grid-stride, then one atomic per block.
__global__ __launch_bounds__(256) void k_absmax(const float* __restrict__ u, uint64_t n, uint32_t* out) {
    __shared__ uint32_t wm[4];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t b = 0;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = fbits(fabsf(u[i]));
        b = v > b ? v : b;
    }
    b = wave_max(b);
    if (lane_id() == 0) wm[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = wm[0];
        for (int w = 1; w < 4; ++w) m = wm[w] > m ? wm[w] : m;
        if (m) atomicMax(out, m);
    }
}

struct BwdArgs {
    const float* u;            // [npx*3] dL/dfilm
    const uint32_t* umax;      // bits of max |u|
    float kd_max;
    uint32_t* pmat;            // [npx] material of the diffuse hit, kNoMat otherwise
    float4* pdr;               // [npx] contribution to dL/dkd of that material
    float4* pde;               // [nl][npx] contribution to dL/de
    unsigned long long* uacc;  // [N*3] fixed-point dL/dirr
};

// gather() of src/surfel.rs:262-293 followed by the scatter of its adjoint:
// the same cells, items and weights, each surfel receiving w/wsum * ku.
__device__ __forceinline__ f3 gather_scatter(const GridView& g, const float4* __restrict__ sp,
                                             const float4* __restrict__ sn, const float4* __restrict__ irr, f3 p, f3 n,
                                             float radius, f3 ku, double scale, unsigned long long* uacc) {
    f3 acc = spl(0.0f);
    float wsum = 0.0f;
    const float r2 = radius * radius;
    int32_t c[3];
    grid_cell(g, p, c);
    const int32_t z0 = c[2] - 1 > 0 ? c[2] - 1 : 0, z1 = c[2] + 1 < g.dim[2] - 1 ? c[2] + 1 : g.dim[2] - 1;
    const int32_t y0 = c[1] - 1 > 0 ? c[1] - 1 : 0, y1 = c[1] + 1 < g.dim[1] - 1 ? c[1] + 1 : g.dim[1] - 1;
    const int32_t x0 = c[0] - 1 > 0 ? c[0] - 1 : 0, x1 = c[0] + 1 < g.dim[0] - 1 ? c[0] + 1 : g.dim[0] - 1;
    for (int pass = 0; pass < 2; ++pass) {
        for (int32_t z = z0; z <= z1; ++z)
            for (int32_t y = y0; y <= y1; ++y)
                for (int32_t x = x0; x <= x1; ++x) {
                    uint32_t ci = (uint32_t)(x + g.dim[0] * (y + g.dim[1] * z));
                    uint32_t e = g.starts[ci + 1];
                    for (uint32_t k = g.starts[ci]; k < e; ++k) {
                        uint32_t it = g.items[k];
                        f3 dq = ld3(sp[it]) - p;
                        float d2 = dot3(dq, dq);
                        if (d2 > r2) continue;
                        float align = dot3(ld3(sn[it]), n);
                        if (align <= 0.15f) continue;
                        float t = 1.0f - d2 / r2;
                        float w = t * t * align;
                        if (pass == 0) {
                            acc = acc + ld3(irr[it]) * w;
                            wsum += w;
                        } else {
                            f3 cv = ku * (w / wsum);
                            long long q0 = __double2ll_rn((double)cv.x * scale);
                            long long q1 = __double2ll_rn((double)cv.y * scale);
                            long long q2 = __double2ll_rn((double)cv.z * scale);
                            if (q0) atomicAdd(&uacc[3ull * it], (unsigned long long)q0);
                            if (q1) atomicAdd(&uacc[3ull * it + 1], (unsigned long long)q1);
                            if (q2) atomicAdd(&uacc[3ull * it + 2], (unsigned long long)q2);
                        }
                    }
                }
        if (!(wsum > 1e-6f)) return spl(0.0f);
    }
    return acc / wsum;
}

// Adjoint of render_frame (src/main.rs:247-361) in beauty mode for diffuse
// materials (metallic 0, roughness >= 0.45: no mirror branch, f0 constant).
// 16x16 pixel blocks, one 8x8 tile per wavefront, as the forward shade.
template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_shade_bwd(SceneView Sg, LightsView Lg, LdsPlan P, ShadeArgs A, BwdArgs B) {
    IFX_SETUP(LDS, true, false);
    const uint32_t wv = threadIdx.x >> 6, l = threadIdx.x & 63u;
    const uint32_t x = blockIdx.x * 16u + (wv & 1u) * 8u + (l & 7u);
    const uint32_t y = blockIdx.y * 16u + (wv >> 1) * 8u + (l >> 3);
    if (x >= A.W || y >= A.H) return;
    const size_t npx = (size_t)A.W * A.H;
    const size_t p = (size_t)y * A.W + x;
    const uint32_t nl = n_lights(L);
    const f3 u = mk(B.u[3 * p], B.u[3 * p + 1], B.u[3 * p + 2]);
    for (uint32_t li = 0; li < nl; ++li) B.pde[li * npx + p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t mat = kNoMat;
    f3 dr = spl(0.0f);
    if (u.x != 0.0f || u.y != 0.0f || u.z != 0.0f) {
        const CamV& cam = A.cam;
        float px = (2.0f * ((float)x + 0.5f) / (float)A.W - 1.0f) * cam.aspect * cam.tan_half;
        float py = (1.0f - 2.0f * ((float)y + 0.5f) / (float)A.H) * cam.tan_half;
        f3 rd = nrm3((cam.fwd + cam.right * px) + cam.up * py);
        uint32_t steps = 0;
        Hit hh = trace<true>(S, cam.origin, rd, 1e-3f, A.tmax, steps, stk);
        float surf_t = hh.hit ? hh.t : kFltMax;
        f3 e;
        uint32_t which = 0;
        if (emitter_hit(L, cam.origin, rd, fmn(surf_t, A.tmax), e, &which)) {
            B.pde[which * npx + p] = st4(u);
        } else if (hh.hit) {
            const MatV m = load_mat(L, hh.mat);
            mat = hh.mat;
            f3 tan, bit;
            frame(hh.n, tan, bit);
            const f3 v = -rd;
            const float n_dot_v = fmx(dot3(hh.n, v), 1e-4f);
            const float alpha = fmx(m.roughness * m.roughness, 1e-3f);
            for (uint32_t li = 0; li < nl; ++li) {
                float a;
                f3 sp, em;
                if (light_coeff(S, L, li, hh.p, hh.n, tan, bit, v, n_dot_v, alpha, m.f0, a, sp, em, stk)) {
                    B.pde[li * npx + p] = st4(u * (m.kd * a + sp));
                    dr = dr + (u * em) * a;
                }
            }
            if (A.use_gi) {
                const double scale = fix_scale(bitsf(*B.umax), npx, B.kd_max);
                f3 irr = gather_scatter(A.grid, A.spos, A.snrm, A.irr, hh.p, hh.n, A.radius, m.kd * u, scale, B.uacc);
                dr = dr + u * irr;
            }
        }
    }
    B.pmat[p] = mat;
    B.pdr[p] = st4(dr);
}

__global__ void k_fix_to_float(uint32_t N, const unsigned long long* __restrict__ uacc, const uint32_t* umax,
                               uint64_t npx, float kd_max, float4* ubar) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double s = fix_scale(bitsf(*umax), npx, kd_max);
    float v[3];
    for (int c = 0; c < 3; ++c) v[c] = (float)((double)(long long)uacc[3ull * i + c] / s);
    ubar[i] = make_float4(v[0], v[1], v[2], 0.0f);
}

// ---------------------------------------------------------------- transposed links
// Rows row0 .. row0 + N of a CSR (starts lstart): each link's row and its
// own position.
__global__ void k_link_rows(uint32_t N, const uint32_t* __restrict__ lstart, uint32_t row0, uint32_t* lrow,
                            uint32_t* iota) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (uint32_t k = lstart[i], e = lstart[i + 1]; k < e; ++k) {
        lrow[k] = row0 + i;
        iota[k] = k;
    }
}

// Links per (cluster, row tile): cnt2[c * T + row / R].
__global__ void k_tile_col_count(uint64_t L, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ lrow,
                                 uint32_t R, uint32_t T, uint32_t* cnt2) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L) atomicAdd(&cnt2[(uint64_t)idx[k] * T + lrow[k] / R], 1u);
}

// One tile's column counts into column t of cnt2.
__global__ void k_tile_col_put(uint32_t C, uint32_t T, uint32_t t, const uint32_t* __restrict__ colcnt,
                               uint32_t* cnt2) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C) cnt2[(uint64_t)c * T + t] = colcnt[c];
}

__global__ void k_col_count(uint64_t L, const uint32_t* __restrict__ idx, uint32_t* cnt) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L) atomicAdd(&cnt[idx[k]], 1u);
}

__global__ void k_csc_fill(uint64_t L, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ lrow,
                           const float* __restrict__ w, uint32_t* crow, float* cw) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L) return;
    uint32_t s = perm[k];
    crow[k] = lrow[s];
    cw[k] = w[s];
}

// Segments: every (cluster, row tile) run of the transposed links is cut
// into pieces of S entries, so a segment never spans two tiles and the
// summation order is the same whether the links stay resident or are
// recomputed tile by tile.
__global__ void k_seg_count(uint64_t n, const uint32_t* __restrict__ cnt2, uint32_t S, uint32_t* segn) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) segn[q] = (cnt2[q] + S - 1) / S;
}

// segfirst[c] = first segment of cluster c (c = C: the total).
__global__ void k_seg_first(uint32_t C, uint32_t T, const uint32_t* __restrict__ segoff, uint32_t* segfirst) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= C) segfirst[c] = segoff[(uint64_t)c * T];
}

// Resident links: the ranges of every segment, cluster by cluster, tile by tile.
__global__ void k_seg_make(uint32_t C, uint32_t T, const uint32_t* __restrict__ colstart,
                           const uint32_t* __restrict__ cnt2, const uint32_t* __restrict__ segoff, uint32_t S,
                           uint2* seg) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t b = colstart[c];
    for (uint32_t t = 0; t < T; ++t) {
        const uint64_t q = (uint64_t)c * T + t;
        const uint32_t n = cnt2[q];
        uint32_t s = segoff[q];
        for (uint32_t k = 0; k < n; k += S, ++s) seg[s] = make_uint2(b + k, n - k < S ? b + n : b + k + S);
        b += n;
    }
}

// Surfel -> cluster and the aggregation weight a_m / max(A_c, 1e-12).
__global__ void k_surfel_cluster(uint32_t N, uint32_t C, const uint2* __restrict__ range,
                                 const uint32_t* __restrict__ members, const float4* __restrict__ spos,
                                 const float4* __restrict__ cpos, uint32_t* scl, float* sscale) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const float inv = 1.0f / fmx(cpos[c].w, 1e-12f);
    const uint2 r = range[c];
    for (uint32_t k = r.x; k < r.y; ++k) {
        uint32_t m = members[k];
        if (m < N) {
            scl[m] = c;
            sscale[m] = spos[m].w * inv;
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- adjoint Jacobi
struct AdjArgs {
    uint32_t N, C, nseg, max_iters;
    float tol;
    const uint2* seg;          // [nseg] ranges of the transposed links
    const uint32_t* segfirst;  // [C+1] first segment of each cluster
    const uint32_t* crow;      // transposed links: row (surfel) ...
    const float* cw;           // ... and weight, cluster-major, rows ascending
    const uint32_t* scl;       // [N] cluster of each surfel (C = none)
    const float* sscale;       // [N] a_m / max(A_c, 1e-12)
    const float4* ubar;        // [N] dL/dirr
    const float4* salb;        // [N] rho
    float4* mu;                // [N] adjoint iterate
    float4* v;                 // [N] ubar + rho (.) mu
    float4* segsum;            // [nseg] partial sums
    uint32_t* dmax;            // [2*max_iters] max |delta mu| and max |mu| bits per sweep
};

// Sweep `it` has converged once max |delta mu| <= tol * max |mu| (sweep 0,
// from mu = 0, never counts).  A sweep that did not run left both maxima
// zero, which reads as converged, so later launches fall through.
__device__ __forceinline__ bool adj_converged(const AdjArgs& A, uint32_t it) {
    return it > 0 && bitsf(A.dmax[2 * it]) <= A.tol * bitsf(A.dmax[2 * it + 1]);
}

// mu_{k+1} = P^T W^T (ubar + rho (.) mu_k), mu_0 = 0, two launches per sweep.
// Sweep part 1: each surfel sums its cluster's segment partials in segment
// order (t_c = (W^T v)_c), forms mu and v, and folds max |delta| and max |mu|
// into integer maxima.
__global__ __launch_bounds__(kBlock) void k_adj_surfels(AdjArgs A, uint32_t it) {
    if (it > 0 && adj_converged(A, it - 1)) return;
    const uint32_t gt = blockIdx.x * kBlock + threadIdx.x, nt = gridDim.x * kBlock;
    uint32_t dm = 0, mm = 0;
    for (uint32_t i = gt; i < A.N; i += nt) {
        const uint32_t c = A.scl[i];
        f3 t = spl(0.0f);
        if (it > 0 && c < A.C)
            for (uint32_t s = A.segfirst[c], e = A.segfirst[c + 1]; s < e; ++s) t = t + ld3(A.segsum[s]);
        const f3 m = t * A.sscale[i];
        const f3 old = it > 0 ? ld3(A.mu[i]) : spl(0.0f);
        const uint32_t d = fbits(maxe(abs3(m - old))), a = fbits(maxe(abs3(m)));
        dm = d > dm ? d : dm;
        mm = a > mm ? a : mm;
        A.mu[i] = st4(m);
        A.v[i] = st4(ld3(A.ubar[i]) + ld3(A.salb[i]) * m);
    }
    dm = wave_max(dm);
    mm = wave_max(mm);
    if (lane_id() == 0) {
        if (dm) atomicMax(&A.dmax[2 * it], dm);
        if (mm) atomicMax(&A.dmax[2 * it + 1], mm);
    }
}

// Sum of w * v[row] over the transposed entries b .. e (at most S of them):
// entry b + o goes to lane o % 64 in entry order, then a fixed xor tree.
__device__ __forceinline__ f3 seg_sum(const float4* __restrict__ v, const uint32_t* __restrict__ crow,
                                      const float* __restrict__ cw, uint32_t b, uint32_t e, uint32_t lane) {
    f3 acc = spl(0.0f);
    for (uint32_t k = b + lane; k < e; k += 64u) {
        const f3 x = ld3(v[crow[k]]);
        const float w = cw[k];
        acc.x = fmaf(x.x, w, acc.x);
        acc.y = fmaf(x.y, w, acc.y);
        acc.z = fmaf(x.z, w, acc.z);
    }
    for (int o = 32; o > 0; o >>= 1) {
        acc.x += __shfl_xor(acc.x, o, 64);
        acc.y += __shfl_xor(acc.y, o, 64);
        acc.z += __shfl_xor(acc.z, o, 64);
    }
    return acc;
}

// Sweep part 2, resident links: one wavefront per segment.
__global__ __launch_bounds__(kBlock) void k_adj_segments(AdjArgs A, uint32_t it) {
    if (adj_converged(A, it)) return;
    const uint32_t lane = lane_id();
    const uint32_t gw = (blockIdx.x * kBlock + threadIdx.x) >> 6, nw = (gridDim.x * kBlock) >> 6;
    for (uint32_t s = gw; s < A.nseg; s += nw) {
        const uint2 r = A.seg[s];
        const f3 acc = seg_sum(A.v, A.crow, A.cw, r.x, r.y, lane);
        if (lane == 0) A.segsum[s] = st4(acc);
    }
}

// Sweep part 2, one recomputed tile t of T: one wavefront per cluster walks
// the tile's transposed links of that cluster (tcolstart) in pieces of S,
// the same segments as the resident layout, written at segoff[c * T + t].
__global__ __launch_bounds__(kBlock) void k_adj_tile(AdjArgs A, uint32_t it, uint32_t S, uint32_t T, uint32_t t,
                                                     const uint32_t* __restrict__ tcolstart,
                                                     const uint32_t* __restrict__ tcrow,
                                                     const float* __restrict__ tcw,
                                                     const uint32_t* __restrict__ segoff) {
    if (adj_converged(A, it)) return;
    const uint32_t lane = lane_id();
    const uint32_t gw = (blockIdx.x * kBlock + threadIdx.x) >> 6, nw = (gridDim.x * kBlock) >> 6;
    for (uint32_t c = gw; c < A.C; c += nw) {
        const uint32_t b = tcolstart[c], e = tcolstart[c + 1];
        uint32_t s = segoff[(uint64_t)c * T + t];
        for (uint32_t k = b; k < e; k += S, ++s) {
            const f3 acc = seg_sum(A.v, tcrow, tcw, k, e - k < S ? e : k + S, lane);
            if (lane == 0) A.segsum[s] = st4(acc);
        }
    }
}

// ---------------------------------------------------------------- parameter sums
// Surfel share: dL/drho_i = mu (.) (D + irr), dL/de_l = mu (.) rho g_il.
__global__ void k_surfel_bwd(uint32_t N, LightsView Lv, const float4* __restrict__ mu, const float4* __restrict__ irr,
                             const float4* __restrict__ salb, const float* __restrict__ gterm, float4* sdr,
                             float4* sde) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t nl = n_lights(Lv);
    const f3 m = ld3(mu[i]), rho = ld3(salb[i]);
    f3 D = spl(0.0f);
    for (uint32_t li = 0; li < nl; ++li) {
        float g = gterm[(size_t)li * N + i];
        D = D + light_emit(Lv, li) * g;
        sde[(size_t)li * N + i] = st4((m * rho) * g);
    }
    sdr[i] = st4(m * (D + ld3(irr[i])));
}

// Per-block f64 sums of every parameter's contributions over kParamChunk
// items: albedo slots 3m+c take dr.c of items whose material is m, emission
// slots 3(nm+l)+c take de[l].c.  Fixed tree order within the block.
__global__ __launch_bounds__(256) void k_param_partials(uint32_t n, uint32_t nm, uint32_t nl,
                                                        const uint32_t* __restrict__ mat,
                                                        const float4* __restrict__ dr,
                                                        const float4* __restrict__ de, double* partial) {
    __shared__ double red[256];
    const uint32_t np = 3u * (nm + nl);
    const uint32_t b0 = blockIdx.x * kParamChunk;
    for (uint32_t q = 0; q < np; ++q) {
        double acc = 0.0;
        for (uint32_t j = 0; j < kParamChunk / 256u; ++j) {
            uint32_t i = b0 + j * 256u + threadIdx.x;
            if (i >= n) break;
            float val;
            if (q < 3u * nm)
                val = mat[i] == q / 3u ? comp(dr[i], q % 3u) : 0.0f;
            else
                val = comp(de[(size_t)((q - 3u * nm) / 3u) * n + i], q % 3u);
            acc += (double)val;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (uint32_t s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * np + q] = red[0];
        __syncthreads();
    }
}

}  // namespace

// ====================================================================== host side
int Engine::check_backward(std::string& err) const {
    if (int rc0 = need_device(err)) return rc0;
    if (!have_transfer || !have_solve || !have_gterm_) {
        err = "backward needs transfer and solve first";
        return -3;
    }
    for (const HMaterial& m : scene.mats)
        if (m.metallic != 0.0f || m.roughness < 0.45f) {
            err = "backward supports diffuse materials only (metallic 0, roughness >= 0.45)";
            return -1;
        }
    for (const HSurfel& s : surfels)
        if (s.mat >= scene.mats.size()) {
            err = "backward needs the material of every surfel (place them with ifx_generate_surfels)";
            return -3;
        }
    return 0;
}

int Engine::set_material_albedo(uint32_t m, const float* rgb, std::string& err) {
    if (m >= scene.mats.size() || !rgb) {
        err = "material index out of range";
        return -1;
    }
    HMaterial& mt = scene.mats[m];
    mt.albedo = mk(rgb[0], rgb[1], rgb[2]);
    material_records(mt, &scene.mat_rec[4ull * m]);
    const f3 kd = ld3(scene.mat_rec[4ull * m + 3]);
    for (HSurfel& s : surfels)
        if (s.mat == m) s.albedo = kd;
    have_solve = have_gterm_ = have_beauty_ = false;
    if (!has_device_) return 0;
    CK(hipSetDevice(device_));
    // the side stream's precomputed direct terms used the old value
    if (side_) CK(hipStreamSynchronize(side_));
    join_side();
    direct_pre_ = false;
    CK(hipMemcpy(d_mat.p, scene.mat_rec.data(), 16ull * scene.mat_rec.size(), hipMemcpyHostToDevice));
    if (!surfels_dirty && !surfels.empty()) {
        std::vector<float4> ha(surfels.size());
        for (size_t i = 0; i < surfels.size(); ++i) ha[i] = st4(surfels[i].albedo);
        CK(hipMemcpy(d_salb.p, ha.data(), 16ull * ha.size(), hipMemcpyHostToDevice));
    }
    return 0;
}

int Engine::set_light_emit(uint32_t li, const float* rgb, std::string& err) {
    const size_t nq = scene.quads.size(), ns = scene.spheres.size(), np = scene.points.size();
    if (li >= nq + ns + np || !rgb) {
        err = "light index out of range";
        return -1;
    }
    const f3 e = mk(rgb[0], rgb[1], rgb[2]);
    if (li < nq) {
        scene.quads[li].emit = e;
        scene.quad_rec[7ull * li + 4] = st4(e, scene.quad_rec[7ull * li + 4].w);
    } else if (li < nq + ns) {
        scene.spheres[li - nq].emit = e;
        scene.sph_rec[2ull * (li - nq) + 1] = st4(e);
    } else {
        scene.points[li - nq - ns].emit = e;
        scene.pnt_rec[2ull * (li - nq - ns) + 1] = st4(e);
    }
    have_solve = have_gterm_ = have_beauty_ = false;
    if (!has_device_) return 0;
    CK(hipSetDevice(device_));
    // the side stream's precomputed direct terms used the old value
    if (side_) CK(hipStreamSynchronize(side_));
    join_side();
    direct_pre_ = false;
    if (nq) CK(hipMemcpy(d_quad.p, scene.quad_rec.data(), 16ull * scene.quad_rec.size(), hipMemcpyHostToDevice));
    if (ns) CK(hipMemcpy(d_sph.p, scene.sph_rec.data(), 16ull * scene.sph_rec.size(), hipMemcpyHostToDevice));
    if (np) CK(hipMemcpy(d_pnt.p, scene.pnt_rec.data(), 16ull * scene.pnt_rec.size(), hipMemcpyHostToDevice));
    return 0;
}

int Engine::download_direct_geometry(float* g, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    if (!have_gterm_) {
        err = "no solve has run on this transfer";
        return -3;
    }
    const uint64_t n = (uint64_t)n_lights(dev_lights()) * surfels.size();
    if (n && g) CK(hipMemcpy(g, d_gterm.p, 4ull * n, hipMemcpyDeviceToHost));
    return 0;
}

int Engine::tile_transpose(uint32_t r0, uint32_t nr, uint32_t L, std::string& err) {
    const uint32_t C = n_clusters;
    CK(d_tcolcnt.ensure(4ull * (C + 1)));
    CK(d_tcolstart.ensure(4ull * (C + 1)));
    CK(d_tlrow.ensure(4ull * L));
    CK(d_tperm.ensure(4ull * L));
    CK(d_tperm2.ensure(4ull * L));
    CK(d_tkeys.ensure(4ull * L));
    CK(d_tcrow.ensure(4ull * L));
    CK(d_tcw.ensure(4ull * L));
    CK(hipMemsetAsync(d_tcolcnt.p, 0, 4ull * (C + 1), stream));
    if (L) {
        k_link_rows<<<blocks_for(nr, 256), 256, 0, stream>>>(nr, d_tlstart.as<uint32_t>(), r0, d_tlrow.as<uint32_t>(),
                                                             d_tperm.as<uint32_t>());
        k_col_count<<<blocks_for(L, 256), 256, 0, stream>>>(L, d_tidx.as<uint32_t>(), d_tcolcnt.as<uint32_t>());
        if (int rc = sort_u32(d_tidx.as<uint32_t>(), d_tperm.as<uint32_t>(), d_tkeys.as<uint32_t>(),
                              d_tperm2.as<uint32_t>(), L, bits_for(C ? C - 1 : 1), err))
            return rc;
        k_csc_fill<<<blocks_for(L, 256), 256, 0, stream>>>(L, d_tperm2.as<uint32_t>(), d_tlrow.as<uint32_t>(),
                                                           d_tw.as<float>(), d_tcrow.as<uint32_t>(),
                                                           d_tcw.as<float>());
    }
    if (int rc = scan_u32(d_tcolcnt.as<uint32_t>(), d_tcolstart.as<uint32_t>(), (C + 1), err)) return rc;
    CK(hipGetLastError());
    return 0;
}

void Engine::release_link_tiles() {
    for (DevBuf* b : {&d_tcount, &d_tstart, &d_tci, &d_trow, &d_tv, &d_tkeep, &d_tlstart, &d_tidx, &d_tw, &d_tlrow,
                      &d_tperm, &d_tperm2, &d_tkeys, &d_tcrow, &d_tcw, &d_tcolcnt, &d_tcolstart, &d_tqueue})
        b->release();
    tile_m_.clear();
    tile_l_.clear();
}

// The backward pass's link layout, once per transfer.  Resident: the links
// transposed (cluster-major, rows ascending within a cluster), about 32 bytes
// per link while they are built.  When link_budget is set and that exceeds
// it, nothing is kept: the surfel rows are cut into tiles of adj_tile_rows
// and every sweep recomputes each tile's links from the geometry with the
// transfer's kernels (bit-identical to the transfer's links) and transposes
// them; this pass records each tile's candidate and link counts.  Either way
// the segments are the (cluster, row tile) runs cut into adj_seg entries, so
// both layouts sum the same entries in the same order.
int Engine::adjoint_prepare(std::string& err) {
    if (adj_gen_ == transfer_gen_) return 0;
    const uint32_t N = (uint32_t)surfels.size(), C = n_clusters;
    const uint64_t L = n_links;
    const uint32_t S = adj_seg ? adj_seg : 1024u;
    const uint32_t R = adj_tile_rows ? adj_tile_rows : 4096u;
    const uint32_t T = N ? (N + R - 1) / R : 1u;
    const uint64_t CT = (uint64_t)C * T;
    if (CT + 1 >= (1ull << 31)) {
        err = "clusters x row tiles exceeds 2^31 (raise adj_tile_rows)";
        return -6;
    }
    CK(d_scl.ensure(4ull * N));
    CK(d_sscale.ensure(4ull * N));
    CK(d_segfirst.ensure(4ull * (C + 1)));
    CK(d_cnt2.ensure(4ull * (CT + 1)));
    CK(d_segn.ensure(4ull * (CT + 1)));
    CK(d_segoff.ensure(4ull * (CT + 1)));
    if (N) {
        launch("adj_prep", [&] {
            k_fill_u32<<<blocks_for(N, 256), 256, 0, stream>>>(d_scl.as<uint32_t>(), N, C);
            k_fill_u32<<<blocks_for(N, 256), 256, 0, stream>>>((uint32_t*)d_sscale.p, N, 0u);
            if (C)
                k_surfel_cluster<<<blocks_for(C, 256), 256, 0, stream>>>(
                    N, C, d_crange.as<uint2>(), d_cval2.as<uint32_t>(), d_spos.as<float4>(), d_cpos.as<float4>(),
                    d_scl.as<uint32_t>(), d_sscale.as<float>());
        });
    }
    CK(hipMemsetAsync(d_cnt2.p, 0, 4ull * (CT + 1), stream));
    const bool tiled = link_budget && L && 32ull * L > link_budget;
    if (!tiled) {
        release_link_tiles();
        CK(d_colcnt.ensure(4ull * (C + 1)));
        CK(d_colstart.ensure(4ull * (C + 1)));
        CK(d_lrow.ensure(4ull * L));
        CK(d_perm.ensure(4ull * L));
        CK(d_perm2.ensure(4ull * L));
        CK(d_ckeyt.ensure(4ull * L));
        CK(d_crow.ensure(4ull * L));
        CK(d_cw.ensure(4ull * L));
        CK(hipMemsetAsync(d_colcnt.p, 0, 4ull * (C + 1), stream));
        if (L) {
            launch("adj_prep", [&] {
                k_link_rows<<<blocks_for(N, 256), 256, 0, stream>>>(N, d_lstart.as<uint32_t>(), 0u,
                                                                    d_lrow.as<uint32_t>(), d_perm.as<uint32_t>());
                k_col_count<<<blocks_for(L, 256), 256, 0, stream>>>(L, d_idx.as<uint32_t>(),
                                                                    d_colcnt.as<uint32_t>());
                k_tile_col_count<<<blocks_for(L, 256), 256, 0, stream>>>(L, d_idx.as<uint32_t>(),
                                                                         d_lrow.as<uint32_t>(), R, T,
                                                                         d_cnt2.as<uint32_t>());
            });
            const int kb = (int)bits_for(C ? C - 1 : 1);
            if (int rc_s = sort_u32(d_idx.as<uint32_t>(), d_perm.as<uint32_t>(), d_ckeyt.as<uint32_t>(),
                                    d_perm2.as<uint32_t>(), L, (uint32_t)kb, err))
                return rc_s;
            launch("adj_prep", [&] {
                k_csc_fill<<<blocks_for(L, 256), 256, 0, stream>>>(L, d_perm2.as<uint32_t>(), d_lrow.as<uint32_t>(),
                                                                   d_w.as<float>(), d_crow.as<uint32_t>(),
                                                                   d_cw.as<float>());
            });
        }
        if (int rc_s = scan_u32(d_colcnt.as<uint32_t>(), d_colstart.as<uint32_t>(), (C + 1), err)) return rc_s;
    } else {
        for (DevBuf* b : {&d_lrow, &d_perm, &d_perm2, &d_ckeyt, &d_crow, &d_cw, &d_seg}) b->release();
        tile_m_.assign(T, 0u);
        tile_l_.assign(T, 0u);
        uint64_t ws_max = 0;
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t r0 = t * R, nr = std::min(R, N - r0);
            uint32_t M = 0, Lt = 0;
            if (int rc = tile_links(r0, nr, M, Lt, false, err)) return rc;
            // one tile's working set: row arrays, candidates, links, transpose and sort
            const uint64_t ws = 16ull * (R + 1) + 12ull * M + 40ull * Lt + 8ull * (C + 1);
            ws_max = std::max(ws_max, ws);
            if (ws_max > link_budget) {
                release_link_tiles();
                err = "link_budget " + std::to_string(link_budget) + " bytes is below one link tile's " +
                      std::to_string(ws_max) + " bytes (lower adj_tile_rows)";
                return -5;
            }
            tile_m_[t] = M;
            tile_l_[t] = Lt;
            if (int rc = tile_transpose(r0, nr, Lt, err)) return rc;
            if (C)
                k_tile_col_put<<<blocks_for(C, 256), 256, 0, stream>>>(C, T, t, d_tcolcnt.as<uint32_t>(),
                                                                       d_cnt2.as<uint32_t>());
        }
    }
    launch("adj_prep", [&] {
        k_seg_count<<<blocks_for(CT + 1, 256), 256, 0, stream>>>(CT + 1, d_cnt2.as<uint32_t>(), S,
                                                                 d_segn.as<uint32_t>());
    });
    if (int rc_s = scan_u32(d_segn.as<uint32_t>(), d_segoff.as<uint32_t>(), CT + 1, err)) return rc_s;
    launch("adj_prep", [&] {
        k_seg_first<<<blocks_for(C + 1, 256), 256, 0, stream>>>(C, T, d_segoff.as<uint32_t>(),
                                                                d_segfirst.as<uint32_t>());
    });
    uint32_t nseg = 0;
    CK(hipMemcpyAsync(&nseg, d_segoff.as<uint32_t>() + CT, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    CK(d_segsum.ensure(16ull * (nseg + 1)));
    if (!tiled) {
        CK(d_seg.ensure(8ull * (nseg + 1)));
        if (C && nseg)
            launch("adj_prep", [&] {
                k_seg_make<<<blocks_for(C, 256), 256, 0, stream>>>(C, T, d_colstart.as<uint32_t>(),
                                                                   d_cnt2.as<uint32_t>(), d_segoff.as<uint32_t>(),
                                                                   S, d_seg.as<uint2>());
            });
    }
    CK(hipGetLastError());
    adj_tiled_ = tiled;
    adj_tiles_ = T;
    adj_nseg_ = nseg;
    adj_gen_ = transfer_gen_;
    return 0;
}

int Engine::adjoint_solve(BackwardOut& out, std::string& err) {
    const uint32_t N = (uint32_t)surfels.size(), C = n_clusters;
    const uint32_t iters = adj_max_iters ? adj_max_iters : 1u;
    CK(d_mu.ensure(16ull * N));
    CK(d_v.ensure(16ull * N));
    CK(d_dmax.ensure(8ull * iters));
    CK(hipMemsetAsync(d_dmax.p, 0, 8ull * iters, stream));
    out = BackwardOut();
    if (N == 0) return 0;
    AdjArgs A;
    A.N = N;
    A.C = C;
    A.nseg = adj_nseg_;
    A.max_iters = iters;
    A.tol = adj_tol;
    A.seg = d_seg.as<uint2>();
    A.segfirst = d_segfirst.as<uint32_t>();
    A.crow = d_crow.as<uint32_t>();
    A.cw = d_cw.as<float>();
    A.scl = d_scl.as<uint32_t>();
    A.sscale = d_sscale.as<float>();
    A.ubar = d_ubar.as<float4>();
    A.salb = d_salb.as<float4>();
    A.mu = d_mu.as<float4>();
    A.v = d_v.as<float4>();
    A.segsum = d_segsum.as<float4>();
    A.dmax = d_dmax.as<uint32_t>();
    const uint32_t g_s = jacobi_blocks ? jacobi_blocks : blocks_for(N, kBlock);
    const uint32_t g_seg = jacobi_blocks ? jacobi_blocks : std::max<uint32_t>(1u, blocks_for(64ull * adj_nseg_, kBlock));
    // Sweeps are enqueued in groups; between groups the host reads the
    // maxima and stops enqueuing once the adjoint has converged.
    std::vector<uint32_t> dm(2ull * iters, 0u);
    out.iterations = iters;
    out.link_tiles = adj_tiled_ ? adj_tiles_ : 0u;
    if (adj_tiled_) {
        // Recomputed links: the host checks convergence after every sweep, so
        // no tile is rebuilt for a sweep that will not run.
        const uint32_t S = adj_seg ? adj_seg : 1024u;
        const uint32_t R = adj_tile_rows ? adj_tile_rows : 4096u;
        const uint32_t T = adj_tiles_;
        const uint32_t g_c = jacobi_blocks ? jacobi_blocks : std::max<uint32_t>(1u, blocks_for(64ull * C, kBlock));
        for (uint32_t it = 0; it < iters; ++it) {
            launch("adjoint", [&] { k_adj_surfels<<<g_s, kBlock, 0, stream>>>(A, it); });
            CK(hipMemcpyAsync(dm.data() + 2ull * it, A.dmax + 2ull * it, 8, hipMemcpyDeviceToHost, stream));
            CK(hipStreamSynchronize(stream));
            if (it >= 1 && bitsf(dm[2 * it]) <= adj_tol * bitsf(dm[2 * it + 1])) {
                out.iterations = it;
                break;
            }
            if (it + 1 == iters) break;
            for (uint32_t t = 0; t < T; ++t) {
                const uint32_t r0 = t * R, nr = std::min(R, N - r0);
                uint32_t M = tile_m_[t], Lt = tile_l_[t];
                int rc = 0;
                launch("adj_tiles", [&] {
                    rc = tile_links(r0, nr, M, Lt, true, err);
                    if (!rc) rc = tile_transpose(r0, nr, Lt, err);
                    if (!rc)
                        k_adj_tile<<<g_c, kBlock, 0, stream>>>(A, it, S, T, t, d_tcolstart.as<uint32_t>(),
                                                               d_tcrow.as<uint32_t>(), d_tcw.as<float>(),
                                                               d_segoff.as<uint32_t>());
                });
                if (rc) return rc;
            }
        }
        const uint32_t last = out.iterations < iters ? out.iterations : iters - 1;
        const float mmax = bitsf(dm[2 * last + 1]);
        out.residual = mmax > 0.0f ? bitsf(dm[2 * last]) / mmax : 0.0f;
        return 0;
    }
    const uint32_t group = 8;
    for (uint32_t it0 = 0; it0 < iters; it0 += group) {
        const uint32_t it1 = std::min(iters, it0 + group);
        launch("adjoint", [&] {
            for (uint32_t it = it0; it < it1; ++it) {
                k_adj_surfels<<<g_s, kBlock, 0, stream>>>(A, it);
                if (it + 1 < iters) k_adj_segments<<<g_seg, kBlock, 0, stream>>>(A, it);
            }
        });
        CK(hipMemcpyAsync(dm.data() + 2ull * it0, A.dmax + 2ull * it0, 8ull * (it1 - it0), hipMemcpyDeviceToHost,
                          stream));
        CK(hipStreamSynchronize(stream));
        bool done = false;
        for (uint32_t it = std::max(it0, 1u); it < it1; ++it) {
            if (bitsf(dm[2 * it]) <= adj_tol * bitsf(dm[2 * it + 1])) {
                out.iterations = it;
                done = true;
                break;
            }
        }
        if (done) break;
    }
    const uint32_t last = out.iterations < iters ? out.iterations : iters - 1;
    const float mmax = bitsf(dm[2 * last + 1]);
    out.residual = mmax > 0.0f ? bitsf(dm[2 * last]) / mmax : 0.0f;
    return 0;
}

int Engine::adjoint_params(bool pixels, uint64_t npx, float* d_albedo, float* d_emit, std::string& err) {
    const uint32_t N = (uint32_t)surfels.size();
    const LightsView Lv = dev_lights();
    const uint32_t nl = n_lights(Lv), nm = (uint32_t)scene.mats.size();
    const uint32_t np = 3u * (nm + nl);
    CK(d_sdr.ensure(16ull * N));
    CK(d_sde.ensure(16ull * N * (nl ? nl : 1)));
    const uint32_t nb_s = blocks_for(N, kParamChunk);
    const uint32_t nb_p = pixels ? blocks_for(npx, kParamChunk) : 0u;
    CK(d_partial.ensure(8ull * np * (nb_s + nb_p) + 8));
    double* part = (double*)d_partial.p;
    if (N && np) {
        launch("adj_params", [&] {
            k_surfel_bwd<<<blocks_for(N, 256), 256, 0, stream>>>(N, Lv, d_mu.as<float4>(), d_irr.as<float4>(),
                                                                 d_salb.as<float4>(), d_gterm.as<float>(),
                                                                 d_sdr.as<float4>(), d_sde.as<float4>());
            k_param_partials<<<nb_s, 256, 0, stream>>>(N, nm, nl, d_smat.as<uint32_t>(), d_sdr.as<float4>(),
                                                       d_sde.as<float4>(), part);
        });
    }
    if (nb_p && np) {
        launch("adj_params", [&] {
            k_param_partials<<<nb_p, 256, 0, stream>>>((uint32_t)npx, nm, nl, d_pmat.as<uint32_t>(),
                                                       d_pdr.as<float4>(), d_pde.as<float4>(),
                                                       part + (size_t)np * nb_s);
        });
    }
    std::vector<double> h((size_t)np * (nb_s + nb_p));
    if (!h.empty()) CK(hipMemcpyAsync(h.data(), part, 8ull * h.size(), hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    drain_events();
    std::vector<double> tot(np, 0.0);
    for (uint32_t b = 0; b < nb_s + nb_p; ++b)
        for (uint32_t q = 0; q < np; ++q) tot[q] += h[(size_t)b * np + q];
    if (d_albedo)
        for (uint32_t q = 0; q < 3u * nm; ++q) d_albedo[q] = (float)tot[q];
    if (d_emit)
        for (uint32_t q = 0; q < 3u * nl; ++q) d_emit[q] = (float)tot[3u * nm + q];
    return 0;
}

int Engine::backward(const float* dL_dfilm, float* d_albedo, float* d_emit, BackwardOut& out, std::string& err) {
    if (int rc = check_backward(err)) return rc;
    if (!have_beauty_) {
        err = "backward needs a beauty shade after the solve";
        return -3;
    }
    if (!dL_dfilm) {
        err = "dL_dfilm is null";
        return -1;
    }
    CK(hipSetDevice(device_));
    ShadeArgs A;
    if (int rc = shade_setup(beauty_cam_, beauty_w_, beauty_h_, M_BEAUTY, A, err)) return rc;
    if (int rc = adjoint_prepare(err)) return rc;
    const uint32_t N = (uint32_t)surfels.size();
    const uint64_t npx = (uint64_t)beauty_w_ * beauty_h_;
    const LightsView Lv = dev_lights();
    const uint32_t nl = n_lights(Lv);
    float kd_max = 0.0f;
    for (size_t m = 0; m < scene.mats.size(); ++m) {
        const float4 kd = scene.mat_rec[4 * m + 3];
        kd_max = fmx(kd_max, fmx(fabsf(kd.x), fmx(fabsf(kd.y), fabsf(kd.z))));
    }
    CK(d_ufilm.ensure(12ull * npx));
    CK(d_umax.ensure(16));
    CK(d_uacc.ensure(24ull * N));
    CK(d_ubar.ensure(16ull * N));
    CK(d_pmat.ensure(4ull * npx));
    CK(d_pdr.ensure(16ull * npx));
    CK(d_pde.ensure(16ull * npx * (nl ? nl : 1)));
    CK(hipMemcpyAsync(d_ufilm.p, dL_dfilm, 12ull * npx, hipMemcpyHostToDevice, stream));
    CK(hipMemsetAsync(d_umax.p, 0, 16, stream));
    CK(hipMemsetAsync(d_uacc.p, 0, 24ull * N, stream));
    BwdArgs B;
    B.u = d_ufilm.as<float>();
    B.umax = d_umax.as<uint32_t>();
    B.kd_max = kd_max;
    B.pmat = d_pmat.as<uint32_t>();
    B.pdr = d_pdr.as<float4>();
    B.pde = d_pde.as<float4>();
    B.uacc = d_uacc.as<unsigned long long>();
    launch("shade_bwd", [&] {
        k_absmax<<<std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(3 * npx, 256), 512u)), 256, 0, stream>>>(
            B.u, 3 * npx, d_umax.as<uint32_t>());
        const SceneView S = dev_view();
        const Lds ld = plan_lds(true);
        dim3 grid((beauty_w_ + 15) / 16, (beauty_h_ + 15) / 16);
        LAUNCH_L(k_shade_bwd, ld.staged, grid, ld.bytes, S, Lv, ld.plan, A, B);
        if (N)
            k_fix_to_float<<<blocks_for(N, 256), 256, 0, stream>>>(N, B.uacc, B.umax, npx, kd_max,
                                                                   d_ubar.as<float4>());
    });
    if (int rc = adjoint_solve(out, err)) return rc;
    return adjoint_params(true, npx, d_albedo, d_emit, err);
}

int Engine::backward_surfels(const float* dL_dirr, float* d_albedo, float* d_emit, float* mu, BackwardOut& out,
                             std::string& err) {
    if (int rc = check_backward(err)) return rc;
    if (!dL_dirr) {
        err = "dL_dirr is null";
        return -1;
    }
    CK(hipSetDevice(device_));
    if (int rc = adjoint_prepare(err)) return rc;
    const uint64_t N = surfels.size();
    CK(d_ubar.ensure(16ull * N));
    if (N) {
        CK(hipMemsetAsync(d_ubar.p, 0, 16ull * N, stream));
        if (int rc = copy3_h2d(d_ubar.p, dL_dirr, N, err)) return rc;
    }
    if (int rc = adjoint_solve(out, err)) return rc;
    if (int rc = adjoint_params(false, 0, d_albedo, d_emit, err)) return rc;
    if (mu && N) {
        if (int rc = copy3_d2h(mu, d_mu.p, N, err)) return rc;
        CK(hipStreamSynchronize(stream));
    }
    return 0;
}

}  // namespace ifx
