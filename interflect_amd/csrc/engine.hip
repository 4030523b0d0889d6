// gfx950 kernels of the transfer -> solve -> shade path and the engine that
// drives them.  Every floating-point reduction is either serial per row /
// cluster / pixel in the reference's order or a max (order-free); integer
// atomics are used only for counters, so results are bit-identical across
// runs and launch configurations.
//
// Stage map (reference):
//   transfer: build_clusters  src/formfactor.rs:91-147   -> k_cluster_*
//             build           src/formfactor.rs:153-277  -> k_ff<>, k_visibility, k_closure, k_row_sum, k_compact
//   solve:    light_surfels   src/solve.rs:27-39         -> k_direct_terms, k_direct_sum
//             solve           src/solve.rs:47-108        -> k_jacobi_agg, k_jacobi_rows
//   shade:    Grid::build     src/surfel.rs:71-114       -> k_grid_keys + radix sort
//             render_frame    src/main.rs:229-367        -> k_shade
//             Film::encode    src/film.rs:51-61          -> k_encode
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "ifx_kern.h"
#include "ifx_launch.h"
#include "ifx_sort.h"

namespace ifx {

// ====================================================================== memory
// Scratch of the allocation-free entry point (solve_noalloc): one array per
// device, held until the process ends.  Host code reaches it only through
// the device address hipGetSymbolAddress returns on that device (the
// array's name is a host-side stand-in); scratch_device() looks that address
// up once per device, at context creation.
static __device__ __align__(256) unsigned char scratch[Engine::kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_at;

// The scratch array's address on `device`, which must be current.
hipError_t scratch_device(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_at.find(device);
    if (it != g_scratch_at.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_at[device] = (unsigned char*)p;
    *out = (unsigned char*)p;
    return hipSuccess;
}
}  // namespace

hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (bytes <= cap) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = bytes;
    return hipSuccess;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

hipError_t PinBuf::ensure(size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (bytes <= cap) return hipSuccess;
    release();
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = bytes;
    return hipSuccess;
}

void PinBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

// ====================================================================== transfer: clusters
// Key of (voxel x, y, z, normal bucket), packed so integer order equals the
// reference's sorted tuple order (src/formfactor.rs:101-116).
__global__ void k_cluster_keys(const float4* __restrict__ spos, const float4* __restrict__ snrm, uint32_t n, f3 lo,
                               f3 inv, int32_t res, uint64_t* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 c = (ld3(spos[i]) - lo) * inv;
    uint64_t kx = (uint64_t)clampi(sat_i32(c.x), 0, res - 1);
    uint64_t ky = (uint64_t)clampi(sat_i32(c.y), 0, res - 1);
    uint64_t kz = (uint64_t)clampi(sat_i32(c.z), 0, res - 1);
    uint64_t r = (uint64_t)res;
    keys[i] = ((kx * r + ky) * r + kz) * 6ull + normal_bucket(ld3(snrm[i]));
    vals[i] = i;
}

__global__ void k_seg_flags(const uint64_t* keys, uint32_t n, uint32_t* flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_seg_begin(const uint32_t* flags, const uint32_t* seg, uint32_t n, uint32_t* begin) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) begin[seg[i]] = i;
    if (i == n - 1) begin[seg[i] + flags[i]] = n;
}

// Serial area-weighted sums in ascending member order (src/formfactor.rs:120-145).
// One wavefront per cluster: the members' records are loaded 64 at a time
// and their weighted terms added in member order through readlane (the
// serial sums, bit for bit).
__global__ void k_cluster_build(const uint32_t* begin, uint32_t nseg, const uint32_t* members,
                                const float4* __restrict__ spos, const float4* __restrict__ snrm, float4* cpos,
                                float4* cnrm, uint32_t* valid) {
    const uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
    if (c >= nseg) return;
    f3 p = spl(0.0f), n = spl(0.0f);
    float area = 0.0f;
    for (uint32_t b = begin[c], e = begin[c + 1]; b < e; b += 64u) {
        const uint32_t k = b + lane;
        f3 tp = spl(0.0f), tn = spl(0.0f);
        float ta = 0.0f;
        if (k < e) {
            const uint32_t m = members[k];
            const float4 P = spos[m];
            tp = ld3(P) * P.w;
            tn = ld3(snrm[m]) * P.w;
            ta = P.w;
        }
        const uint32_t cnt = e - b < 64u ? e - b : 64u;
        for (uint32_t j = 0; j < cnt; ++j) {
            p = p + mk(bitsf(__builtin_amdgcn_readlane(fbits(tp.x), j)),
                       bitsf(__builtin_amdgcn_readlane(fbits(tp.y), j)),
                       bitsf(__builtin_amdgcn_readlane(fbits(tp.z), j)));
            n = n + mk(bitsf(__builtin_amdgcn_readlane(fbits(tn.x), j)),
                       bitsf(__builtin_amdgcn_readlane(fbits(tn.y), j)),
                       bitsf(__builtin_amdgcn_readlane(fbits(tn.z), j)));
            area += bitsf(__builtin_amdgcn_readlane(fbits(ta), j));
        }
    }
    if (lane) return;
    valid[c] = area <= 1e-12f ? 0u : 1u;
    f3 nn = dot3(n, n) > 1e-12f ? nrm3(n) : mk(0.0f, 1.0f, 0.0f);
    cpos[c] = st4(p / area, area);
    cnrm[c] = st4(nn, sqrtf(area / kPi));
}

__global__ void k_cluster_compact(uint32_t nseg, const uint32_t* valid, const uint32_t* vidx, const uint32_t* begin,
                                  const float4* cpos_raw, const float4* cnrm_raw, float4* cpos, float4* cnrm,
                                  uint2* range) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nseg || !valid[c]) return;
    uint32_t j = vidx[c];
    cpos[j] = cpos_raw[c];
    cnrm[j] = cnrm_raw[c];
    range[j] = make_uint2(begin[c], begin[c + 1]);
}

// ====================================================================== transfer: links
// One wavefront per surfel row (rows row0 .. row0 + n; count and start are
// indexed from row0); lanes sweep the clusters 64 at a time and a ballot
// keeps survivors in ascending cluster order.  EMIT=false counts, EMIT=true
// writes candidates at the row's scanned offset.  With `masks`, the counting
// pass stores each row's ballots (ceil(C/64) words per row) and the emitting
// pass reads them instead of repeating the culls.  With `key`, the emitting
// pass also writes the visiting-order sort's input with the rows in Morton
// order (row i's candidates from mstart[mrank[i]]): key = cluster, value =
// the candidate's row-major slot, or with jbits set the row's Morton rank
// and the candidate's index in its row packed as (rank << jbits) | index, so
// a stable sort by cluster alone leaves each cluster's candidates in Morton
// order of their rows.  cand_row is written when given.
template <bool EMIT>
__global__ __launch_bounds__(kBlock) void k_ff(const float4* __restrict__ spos, const float4* __restrict__ snrm,
                                               uint32_t row0, uint32_t n, const float4* __restrict__ cpos,
                                               const float4* __restrict__ cnrm, uint32_t C, uint32_t* count,
                                               const uint32_t* __restrict__ start, uint32_t* cand_ci,
                                               uint32_t* cand_row, const uint32_t* __restrict__ mrank,
                                               const uint32_t* __restrict__ mstart, uint32_t* key, uint32_t* val,
                                               uint64_t* masks, uint32_t jbits) {
    uint32_t lane = lane_id();
    uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    uint32_t nwaves = (gridDim.x * kBlock) >> 6;
    const uint32_t groups = (C + 63u) / 64u;
    for (uint32_t j = wave; j < n; j += nwaves) {
        const uint32_t i = row0 + j;
        f3 sp = ld3(spos[i]), sn = ld3(snrm[i]);
        uint32_t base = EMIT ? start[j] : 0u;
        const uint32_t mr = (EMIT && key) ? mrank[i] : 0u;
        const uint32_t mbase = (EMIT && key) ? mstart[mr] : 0u;
        uint32_t run = 0;
        uint32_t mlo = 0, mhi = 0;  // emission: 64 groups' stored ballots, one per lane
        for (uint32_t c0 = 0; c0 < C; c0 += 64) {
            uint32_t c = c0 + lane;
            uint64_t m;
            if (EMIT && masks) {
                // the row's ballots are loaded 64 at a time, so the groups do
                // not each wait on a load of their own
                const uint32_t g = c0 >> 6;
                if ((g & 63u) == 0u) {
                    const uint32_t gl = g + lane;
                    const uint64_t w = gl < groups ? masks[(size_t)j * groups + gl] : 0ull;
                    mlo = (uint32_t)w;
                    mhi = (uint32_t)(w >> 32);
                }
                m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)mhi, (int)(g & 63u)) << 32) |
                    (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)mlo, (int)(g & 63u));
            } else {
                bool pass = false;
                if (c < C) pass = ff_pass(sp, sn, cpos[c], cnrm[c]);
                m = __ballot(pass);
                if (!EMIT && masks && lane == 0) masks[(size_t)j * groups + (c0 >> 6)] = m;
            }
            const bool pass = (m >> lane) & 1ull;
            if (EMIT && pass) {
                const uint32_t r = run + lanes_below(m);
                const uint32_t o = base + r;
                cand_ci[o] = c;
                if (cand_row) cand_row[o] = i;
                if (key) {
                    key[mbase + r] = c;
                    val[mbase + r] = jbits ? (mr << jbits) | r : o;
                }
            }
            run += (uint32_t)__popcll(m);
        }
        if (!EMIT && lane == 0) count[j] = run;
    }
}

// Candidate counts of the rows in Morton order (rorder[r]: the r-th row in
// Morton order); their exclusive sum places each row's candidates in the
// visiting-order sort's input.
__global__ void k_morton_counts(uint32_t N, const uint32_t* __restrict__ rorder, const uint32_t* __restrict__ count,
                                uint32_t* mcount) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) mcount[r] = count[rorder[r]];
}

__global__ void k_morton(const float4* __restrict__ spos, uint32_t N, f3 lo, f3 scale, uint32_t* code,
                         uint32_t* iota) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    f3 q = (ld3(spos[i]) - lo) * scale;
    uint32_t x = (uint32_t)clampi(sat_i32(q.x), 0, 1023), y = (uint32_t)clampi(sat_i32(q.y), 0, 1023),
             z = (uint32_t)clampi(sat_i32(q.z), 0, 1023);
    code[i] = spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
    iota[i] = i;
}

// Surfel records in Morton order (rank r holds surfel rorder[r]) and each
// surfel's rank: the tracing kernels visit surfels in this order, so their
// record loads are coalesced.
__global__ void k_morton_copy(uint32_t N, const uint32_t* __restrict__ rorder, const float4* __restrict__ spos,
                              const float4* __restrict__ snrm, float4* mpos, float4* mnrm, uint32_t* mrank) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const uint32_t i = rorder[r];
    mpos[r] = spos[i];
    mnrm[r] = snrm[i];
    mrank[i] = r;
}

// Cone-traced link visibility (src/formfactor.rs:194-214) as a persistent
// kernel: every lane marches one candidate's shadow ray a sample at a time
// and takes the next candidate from the queue as soon as its ray ends, so
// short and long marches no longer idle each other.  Candidates are taken
// in visiting order (cluster-major, surfels in Morton order: near-parallel
// rays from nearby origins share a wavefront), and their inputs and results
// are stored in that order too, so every index read and result write is
// coalesced: item q reads its row crow[q] and cluster ckey[q] >> shift and
// writes vis[q] = ff * vis for a kept link, 0 for a culled one (k_vis_order
// and k_vis_scatter move between visiting and row-major order).  Each
// sample's distances are answered for the whole wavefront by one shared BVH
// traversal (wave_dist; the scene stays in global memory and is read
// through the scalar cache).  Results sit at fixed positions, so the work
// order never changes one.
// At least kMarchWaves wavefronts per SIMD (72 VGPRs; a few bytes spill to
// scratch outside the sample loop): latency hiding is worth more than the
// registers (stress scene: 96.5 ms at 5 waves, 88.7 at 6, 84.5 at 7, 85.5
// at 8).  The persistent grid is one resident wave of blocks.
//
// Parked tails (VisTail).  A lockstep batch (64 consecutive items, taken
// when the whole wavefront is idle; batch b starts at item 64 b) whose
// marching rays drop to `at` or fewer parks them: each stores its march
// state in st (four float4s: ro | t, rd | tmax, k ph res ff, steps hint
// slot) at b * at + its rank, lane 0 stores their number in cnt[b], and the
// wavefront takes the next batch instead of stepping with most lanes idle.
// A second launch (idx set) marches the parked rays, taken in batch order,
// to their ends: consecutive batches look at one cluster from neighbouring
// surfels, so its wavefronts stay coherent.  A ray's march is the same
// sample sequence wherever it runs, so results and step counts do not
// change.
struct VisTail {
    float4* st = nullptr;
    uint32_t* cnt = nullptr;          // first launch: rays parked per batch (0 at the batch's start)
    const uint32_t* idx = nullptr;    // second launch: the record of each item (null: first launch)
    const uint32_t* total = nullptr;  // second launch: the number of items
    uint32_t at = 0;                  // first launch: park at this many marching rays or fewer (0: never)
};

template <int MODE, bool LDS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kMarchWaves))) void k_visibility(SceneView Sg, LightsView Lg, LdsPlan P,
                                                       const float4* __restrict__ spos,
                                                       const float4* __restrict__ snrm,
                                                       const float4* __restrict__ cpos,
                                                       const float4* __restrict__ cnrm,
                                                       const uint32_t* __restrict__ ckey, uint32_t shift,
                                                       const uint32_t* __restrict__ crow, uint32_t rshift, uint32_t M, float off,
                                                       float tmin, float* vis_out, uint32_t* queue, uint32_t chunk,
                                                       uint32_t min_idle, VisTail tl, unsigned long long* evals,
                                                       unsigned long long* ops) {
    IFX_SETUP(LDS, false, MODE);
    uint32_t acc = 0;
    bool live = false;
    const bool resume = tl.idx != nullptr;
    const uint32_t total = resume ? *tl.total : M;
    WaveQueue wq{0u, 0u, chunk, false, min_idle};
    uint32_t slot = 0, wsteps = 0, batch = 0;
    float ff = 0.0f;
    ShadowMarch m;
    m.begin(spl(0.0f), spl(0.0f), 0.0f, 0.0f, 0.0f);
    for (;;) {
        uint32_t q = 0;
        const bool got = refill(live, wq, queue, total, q);
        if (got && resume) {
            const float4* rec = tl.st + 4ull * tl.idx[q];
            const float4 a = rec[0], b = rec[1], c = rec[2], d = rec[3];
            m.ro = ld3(a), m.t = a.w;
            m.rd = ld3(b), m.tmax = b.w;
            m.k = c.x, m.ph = c.y, m.res = c.z, ff = c.w;
            m.steps = fbits(d.x), m.hint = fbits(d.y), slot = fbits(d.z);
            live = true;
        } else if (got) {
            slot = q;
            const uint32_t i = crow[q] >> rshift, c = ckey[q] >> shift;
            f3 sp = ld3(spos[i]), sn = ld3(snrm[i]);
            float4 CN = cnrm[c];
            float r = 0.0f;
            f3 dir = spl(0.0f);
            ff_geom(sp, sn, cpos[c], CN, r, dir, ff);
            float kk = clmp(r / fmx(CN.w, 1e-4f), 1.5f, 32.0f);
            m.begin(sp + sn * off, dir, tmin, r - CN.w * 0.75f, kk);
            live = true;
        }
        if (tl.at && __ballot(got) != 0ull) {  // a new batch: lane 0 holds its first item
            batch = (uint32_t)__builtin_amdgcn_readlane((int)q, 0) >> 6;
            if (lane_id() == 0) tl.cnt[batch] = 0u;
        }
        if (__ballot(live) == 0) break;
        if constexpr (MODE == 2) ++wsteps;
        float vis = 0.0f;
        if (shadow_step(S, m, live, vis, stk)) {
            vis_out[slot] = vis <= 1e-3f ? 0.0f : ff * vis;
            acc += m.steps;
            live = false;
        }
        if (tl.at) {
            const uint64_t lm = __ballot(live);
            const uint32_t na = (uint32_t)__popcll(lm);
            if (na != 0u && na <= tl.at) {
                if (live) {
                    float4* rec = tl.st + 4ull * (batch * tl.at + lanes_below(lm));
                    rec[0] = make_float4(m.ro.x, m.ro.y, m.ro.z, m.t);
                    rec[1] = make_float4(m.rd.x, m.rd.y, m.rd.z, m.tmax);
                    rec[2] = make_float4(m.k, m.ph, m.res, ff);
                    rec[3] = make_float4(bitsf(m.steps), bitsf(m.hint), bitsf(slot), 0.0f);
                    live = false;
                }
                if (lane_id() == 0) tl.cnt[batch] = na;
            }
        }
    }
    count_steps(evals, acc);
    stk_flush(stk, ops);
    if constexpr (MODE == 2)
        if (lane_id() == 0 && wsteps) atomicAdd(ops + 12, 64ull * wsteps);  // lane slots of the marches
}

// Candidates up to which the link visibility may share the GPU with the
// side stream's closure and direct-light launches (measured, see transfer()).
constexpr uint32_t kVisOverlapMax = 1u << 23;

// Visiting order <-> row-major candidates: the Morton rank of each visited
// candidate's row, and its result back to the candidate's row-major slot.
__global__ void k_vis_order(uint32_t M, const uint32_t* __restrict__ order, const uint32_t* __restrict__ cand_row,
                            const uint32_t* __restrict__ mrank, uint32_t* vrow) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < M) vrow[q] = mrank[cand_row[order[q]]];
}

__global__ void k_vis_scatter(uint32_t M, const uint32_t* __restrict__ order, const float* __restrict__ vis,
                              float* cand_v) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < M) cand_v[order[q]] = vis[q];
}

// The same from packed sort values (Morton rank << jbits | index in the
// row): the slot is the row's candidate start (mbase, by Morton rank) plus
// the index.
__global__ void k_vis_scatter_packed(uint32_t M, const uint32_t* __restrict__ packed, uint32_t jbits,
                                     const uint32_t* __restrict__ mbase, const float* __restrict__ vis,
                                     float* cand_v) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= M) return;
    const uint32_t v = packed[q];
    cand_v[mbase[v >> jbits] + (v & ((1u << jbits) - 1u))] = vis[q];
}

// Each Morton rank's row-major candidate start.
__global__ void k_morton_base(uint32_t N, const uint32_t* __restrict__ rorder, const uint32_t* __restrict__ cstart,
                              uint32_t* mbase) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) mbase[r] = cstart[rorder[r]];
}

// The parked rays' records in batch order: batch b's cnt[b] records from
// position start[b] (the exclusive sum of the counts).
__global__ void k_tail_index(uint32_t nb, uint32_t at, const uint32_t* __restrict__ cnt,
                             const uint32_t* __restrict__ start, uint32_t* idx) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nb * at) return;
    const uint32_t b = (uint32_t)(t / at), r = (uint32_t)(t % at);
    if (r < cnt[b]) idx[start[b] + r] = (uint32_t)t;
}

// Hemisphere closure (src/trace.rs:165-186), persistent like k_visibility.
// Item q = direction-major over surfels in Morton order, so a wavefront
// traces one Halton direction from 64 neighbouring surfels (whose records
// mpos / mnrm hold in that order, read coalesced); a missed ray adds one to
// its surfel's integer open count.  dirs[k] = (r cos phi, r sin phi,
// sqrt(1 - u1)).
template <int MODE, bool LDS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kMarchWaves))) void k_closure(SceneView Sg, LightsView Lg, LdsPlan P,
                                                    const float4* __restrict__ mpos,
                                                    const float4* __restrict__ mnrm, uint32_t N,
                                                    const uint32_t* __restrict__ row_order,
                                                    const float4* __restrict__ dirs, float tmax, uint32_t* open,
                                                    uint32_t* queue, uint32_t chunk, uint32_t min_idle,
                                                    unsigned long long* evals, unsigned long long* ops) {
    IFX_SETUP(LDS, false, MODE);
    uint32_t acc = 0, wsteps = 0;
    bool live = false;
    WaveQueue wq{0u, 0u, chunk, false, min_idle};
    uint32_t row = 0;
    TraceMarch m;
    m.begin(spl(0.0f), spl(0.0f), 0.0f, 0.0f);
    const uint32_t total = 16u * N;
    for (;;) {
        uint32_t q;
        if (refill(live, wq, queue, total, q)) {
            const uint32_t k = q / N, r = q - k * N;
            row = row_order[r];
            f3 p = ld3(mpos[r]), n = ld3(mnrm[r]);
            f3 t, b;
            frame(n, t, b);
            float4 D = dirs[k];
            f3 d = (t * D.x + b * D.y) + n * D.z;
            m.begin(p + n * 1e-3f, nrm3(d), 1e-3f, tmax);
            live = true;
        }
        if (__ballot(live) == 0) break;
        if constexpr (MODE == 2) ++wsteps;
        const int st = miss_step(S, m, live, stk);
        if (st) {
            if (st == 2) atomicAdd(&open[row], 1u);
            acc += m.steps;
            live = false;
        }
    }
    count_steps(evals, acc);
    stk_flush(stk, ops);
    if constexpr (MODE == 2)
        if (lane_id() == 0 && wsteps) atomicAdd(ops + 12, 64ull * wsteps);
}

// Serial per-row link sum in cluster order plus the closure calibration
// (src/formfactor.rs:207-240).  row_sum = -1 marks a row without links.
__global__ void k_row_sum(uint32_t N, const uint32_t* __restrict__ cstart, const float* __restrict__ cand_v,
                          const uint32_t* __restrict__ open, float* sky_out, uint32_t* lcount, float* scale,
                          float* row_sum) {
    // one wavefront per row: its candidates are read coalesced, four
    // 64-candidate chunks in flight at a time, and the kept values are added
    // in candidate order (readlane), the serial sum.  Every lane's value is
    // added, a dropped candidate's as +0.0f: the sum starts at +0.0f and
    // only positive values join it, so adding +0.0f leaves its bits as they
    // are, and the fixed lane sequence unrolls without per-value branches.
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
    if (i >= N) return;
    float sum = 0.0f;
    uint32_t cnt = 0;
    for (uint32_t b = cstart[i], e = cstart[i + 1]; b < e; b += 256u) {
        float v[4];
#pragma unroll
        for (uint32_t c = 0; c < 4u; ++c) {
            const uint32_t k = b + 64u * c + lane;
            v[c] = k < e ? cand_v[k] : 0.0f;
        }
#pragma unroll
        for (uint32_t c = 0; c < 4u; ++c) {
            const uint32_t k = b + 64u * c + lane;
            if (b + 64u * c >= e) break;  // uniform: no candidates left in this row
            const bool keep = k < e && v[c] > 0.0f;
            cnt += (uint32_t)__popcll(__ballot(keep));
            const int32_t x = keep ? (int32_t)fbits(v[c]) : 0;
#pragma unroll
            for (int32_t j = 0; j < 64; ++j) sum += bitsf((uint32_t)__builtin_amdgcn_readlane(x, j));
        }
    }
    if (lane) return;
    float sky = (float)open[i] / 16.0f;
    sky_out[i] = sky;
    float target = clmp(1.0f - sky, 0.0f, 1.0f);
    float sc = 1.0f;
    if (sum > 1e-5f) {
        sc = fmn(target / sum, 8.0f);
        sum *= sc;
    }
    lcount[i] = cnt;
    scale[i] = sc;
    row_sum[i] = cnt ? fmn(sum, 1.0f) : -1.0f;
}

// One wavefront per row: kept candidates in candidate order (ballot ranks),
// four 64-candidate chunks in flight at a time.
__global__ void k_compact(uint32_t N, const uint32_t* __restrict__ cstart, const float* __restrict__ cand_v,
                          const uint32_t* __restrict__ cand_ci, const uint32_t* __restrict__ lstart,
                          const float* __restrict__ scale, uint32_t* idx, float* w) {
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
    if (i >= N) return;
    uint32_t pos = lstart[i];
    const float sc = scale[i];
    for (uint32_t b = cstart[i], e = cstart[i + 1]; b < e; b += 256u) {
        float v[4];
#pragma unroll
        for (uint32_t c = 0; c < 4u; ++c) {
            const uint32_t k = b + 64u * c + lane;
            v[c] = k < e ? cand_v[k] : 0.0f;
        }
#pragma unroll
        for (uint32_t c = 0; c < 4u; ++c) {
            const uint32_t k = b + 64u * c + lane;
            const bool keep = k < e && v[c] > 0.0f;
            const uint64_t m = __ballot(keep);
            if (keep) {
                const uint32_t at = pos + lanes_below(m);
                idx[at] = cand_ci[k];
                w[at] = v[c] * sc;
            }
            pos += (uint32_t)__popcll(m);
        }
    }
}

// Kept candidates (v > 0) of each row, the count k_row_sum reports.
__global__ void k_keep_count(uint32_t n, const uint32_t* __restrict__ cstart, const float* __restrict__ cand_v,
                             uint32_t* keep) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t k = 0;
    for (uint32_t q = cstart[j], e = cstart[j + 1]; q < e; ++q) k += cand_v[q] > 0.0f ? 1u : 0u;
    keep[j] = k;
}

// ====================================================================== solve
// Direct light E, one lane per (light, surfel) pair: the cone shadow of each
// light runs in parallel and k_direct_sum adds the kept terms per surfel in
// the reference's light order (src/shade.rs:137-162, src/solve.rs:27-39).
// Lanes take the surfels in Morton order (positions and normals from the
// Morton-ordered copies mpos / mnrm); terms are stored by surfel index.
template <bool COUNT, bool LDS>
__global__ __launch_bounds__(kBlock) void k_direct_terms(SceneView Sg, LightsView Lg, LdsPlan P,
                                                         const float4* __restrict__ mpos,
                                                         const float4* __restrict__ mnrm,
                                                         const float4* __restrict__ salb, uint32_t N, uint32_t nl,
                                                         const uint32_t* __restrict__ row_order, float4* terms,
                                                         float* gterm, unsigned long long* evals,
                                                         unsigned long long* ops) {
    IFX_SETUP(LDS, true, COUNT);
    // every lane of the grid calls diffuse_term (tail lanes with in = false)
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const bool in = t < N * nl;
    uint32_t steps = 0;
    const uint32_t tc = in ? t : 0u;
    const uint32_t li = tc / N, r = tc - li * N, i = row_order[r];
    f3 p = ld3(mpos[r]), n = ld3(mnrm[r]);
    f3 tan, bit;
    frame(n, tan, bit);
    f3 term = spl(0.0f);
    float g = 0.0f;
    const bool on = diffuse_term(S, L, li, p, n, tan, bit, ld3(salb[i]), term, g, steps, stk, in);
    if (in) {
        terms[(size_t)li * N + i] = on ? st4(term, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        gterm[(size_t)li * N + i] = on ? g : 0.0f;
    }
    count_steps(evals, steps);
    stk_flush(stk, ops);
}

__global__ void k_direct_sum(uint32_t N, uint32_t nl, const float4* __restrict__ terms, float4* direct, float4* rad,
                             float4* irr) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    f3 out = spl(0.0f);
    for (uint32_t li = 0; li < nl; ++li) {
        float4 t = terms[(size_t)li * N + i];
        if (t.w != 0.0f) out = out + ld3(t);
    }
    direct[i] = st4(out);
    rad[i] = st4(out);
    irr[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Jacobi bounces B = E + rho * W (G B) (src/solve.rs:60-102), two launches
// per bounce: the cluster radiosity G B, one wavefront per cluster summing
// its members strictly in member order; then the row gather in link order
// with the max-norm update folded into an integer max of non-negative float
// bits.  Kernel boundaries order the phases (on this multi-XCD part a grid
// barrier would need an L2 write-back and invalidate per block).  Bounce
// `it` runs only while every earlier bounce's update was >= tol, so the
// host may enqueue bounces ahead of the convergence test.
struct JacobiArgs {
    uint32_t N, C, wave_rows;
    float tol;  // stop once the max-norm update is below tol (the reference's 1e-4)
    const uint2* range;
    const uint32_t* members;
    const float4* spos;
    const float4* cpos;
    const uint32_t* lstart;
    const uint32_t* idx;
    const float* w;
    const float4* direct;
    const float4* salb;
    float4* rad;
    float4* irr;
    float4* bc;
    uint32_t* resid;  // [bounces] max |delta| bits
    const uint32_t* sbase;  // sliced layout: per 64-row slice, its first column / 64 (null = CSR)
    const uint32_t* sidx;
    const float* sw;
    const uint32_t* sperm;  // sliced layout: the row at each position (rows by descending length)
};

__device__ __forceinline__ bool jacobi_done(const JacobiArgs& A, uint32_t it) {
    return it > 0 && bitsf(A.resid[it - 1]) < A.tol;
}

__global__ __launch_bounds__(kBlock) void k_jacobi_agg(JacobiArgs A, uint32_t it) {
    if (jacobi_done(A, it)) return;
    const uint32_t lane = lane_id();
    const uint32_t gw = (blockIdx.x * kBlock + threadIdx.x) >> 6, nw = (gridDim.x * kBlock) >> 6;
    for (uint32_t c = gw; c < A.C; c += nw) {
        uint2 r = A.range[c];
        f3 acc = spl(0.0f);
        for (uint32_t b = r.x; b < r.y; b += 64u) {
            uint32_t k = b + lane;
            f3 v = spl(0.0f);
            if (k < r.y) {
                uint32_t m = A.members[k];
                v = ld3(A.rad[m]) * A.spos[m].w;
            }
            // every lane in order; lanes past the cluster's end add +0.0f,
            // which leaves acc's bits as they are (acc is never -0.0f)
#pragma unroll
            for (int32_t j = 0; j < 64; ++j) {
                f3 vj = mk(bitsf(__builtin_amdgcn_readlane(fbits(v.x), j)),
                           bitsf(__builtin_amdgcn_readlane(fbits(v.y), j)),
                           bitsf(__builtin_amdgcn_readlane(fbits(v.z), j)));
                acc = acc + vj;
            }
        }
        if (lane == 0) A.bc[c] = st4(acc / fmx(A.cpos[c].w, 1e-12f));
    }
}

__global__ __launch_bounds__(kBlock) void k_jacobi_rows(JacobiArgs A, uint32_t it) {
    if (jacobi_done(A, it)) return;
    const uint32_t lane = lane_id();
    const uint32_t gt = blockIdx.x * kBlock + threadIdx.x, nt = gridDim.x * kBlock;
    const uint32_t gw = gt >> 6, nw = nt >> 6;
    uint32_t delta = 0;
    if (A.wave_rows) {
        // One wavefront per row: coalesced link loads, the products summed
        // in link order through readlane.
        for (uint32_t i = gw; i < A.N; i += nw) {
            const uint32_t s0 = A.lstart[i], e = A.lstart[i + 1];
            f3 g = spl(0.0f);
            for (uint32_t b = s0; b < e; b += 64u) {
                uint32_t k = b + lane;
                f3 term = spl(0.0f);
                if (k < e) term = ld3(A.bc[A.idx[k]]) * A.w[k];
                uint32_t cnt = e - b < 64u ? e - b : 64u;
                for (uint32_t j = 0; j < cnt; ++j) {
                    f3 tj = mk(bitsf(__builtin_amdgcn_readlane(fbits(term.x), j)),
                               bitsf(__builtin_amdgcn_readlane(fbits(term.y), j)),
                               bitsf(__builtin_amdgcn_readlane(fbits(term.z), j)));
                    g = g + tj;
                }
            }
            if (lane == 0) {
                f3 nr = ld3(A.direct[i]) + ld3(A.salb[i]) * g;
                uint32_t d = fbits(maxe(abs3(nr - ld3(A.rad[i]))));
                delta = d > delta ? d : delta;
                A.irr[i] = st4(g);
                A.rad[i] = st4(nr);
            }
        }
    } else if (A.sbase) {
        // Sliced rows: rows ordered by descending length (sperm), slice q/64
        // stores its rows' k-th links side by side, so lane q%64 reads
        // column k coalesced with its neighbours, and a wavefront's rows have
        // near-equal lengths.  Each row is still summed alone in link order.
        for (uint32_t q = gt; q < A.N; q += nt) {
            const uint32_t i = A.sperm[q];
            f3 g = spl(0.0f);
            const uint32_t len = A.lstart[i + 1] - A.lstart[i];
            const uint64_t b = 64ull * A.sbase[q >> 6] + (q & 63u);
            uint32_t k = 0;
            // the link stream is read once per bounce: non-temporal loads,
            // which leave the cluster radiosities in the caches
            for (; k + kSliceBatch <= len; k += kSliceBatch) {
                uint32_t ix[kSliceBatch];
                float wv[kSliceBatch];
                f3 bv[kSliceBatch];
#pragma unroll
                for (uint32_t j = 0; j < kSliceBatch; ++j) {
                    ix[j] = __builtin_nontemporal_load(A.sidx + b + 64ull * (k + j));
                    wv[j] = __builtin_nontemporal_load(A.sw + b + 64ull * (k + j));
                }
#pragma unroll
                for (uint32_t j = 0; j < kSliceBatch; ++j) bv[j] = ld3(A.bc[ix[j]]);
#pragma unroll
                for (uint32_t j = 0; j < kSliceBatch; ++j) g = g + bv[j] * wv[j];
            }
            for (; k < len; ++k) g = g + ld3(A.bc[A.sidx[b + 64ull * k]]) * A.sw[b + 64ull * k];
            f3 nr = ld3(A.direct[i]) + ld3(A.salb[i]) * g;
            uint32_t d = fbits(maxe(abs3(nr - ld3(A.rad[i]))));
            delta = d > delta ? d : delta;
            A.irr[i] = st4(g);
            A.rad[i] = st4(nr);
        }
    } else {
        for (uint32_t i = gt; i < A.N; i += nt) {
            f3 g = spl(0.0f);
            const uint32_t e = A.lstart[i + 1];
            uint32_t k = A.lstart[i];
            for (; k + kRowBatch <= e; k += kRowBatch) {
                uint32_t ix[kRowBatch];
                float wv[kRowBatch];
                f3 bv[kRowBatch];
#pragma unroll
                for (uint32_t j = 0; j < kRowBatch; ++j) {
                    ix[j] = A.idx[k + j];
                    wv[j] = A.w[k + j];
                }
#pragma unroll
                for (uint32_t j = 0; j < kRowBatch; ++j) bv[j] = ld3(A.bc[ix[j]]);
#pragma unroll
                for (uint32_t j = 0; j < kRowBatch; ++j) g = g + bv[j] * wv[j];
            }
            for (; k < e; ++k) g = g + ld3(A.bc[A.idx[k]]) * A.w[k];
            f3 nr = ld3(A.direct[i]) + ld3(A.salb[i]) * g;
            uint32_t d = fbits(maxe(abs3(nr - ld3(A.rad[i]))));
            delta = d > delta ? d : delta;
            A.irr[i] = st4(g);
            A.rad[i] = st4(nr);
        }
    }
    uint32_t m = wave_max(delta);
    if (lane == 0 && m) atomicMax(&A.resid[it], m);
}

// Sliced copy of the CSR: rows sorted by descending length (a stable radix
// sort of C - length; SELL-C-sigma with sigma = all rows), one slice per 64
// consecutive sorted rows with the columns of its longest row: sort keys,
// widths, a scan, then the fill.  On the stress scene the natural row order
// would pad the slices to 2.4x the links; sorted, 1.0004x.
__global__ void k_sell_keys(uint32_t N, uint32_t C, const uint32_t* __restrict__ lstart, uint32_t* key,
                            uint32_t* row) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    key[i] = C - (lstart[i + 1] - lstart[i]);
    row[i] = i;
}

__global__ void k_sell_widths(uint32_t N, const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ perm,
                              uint32_t* width) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t len = 0;
    if (q < N) {
        const uint32_t i = perm[q];
        len = lstart[i + 1] - lstart[i];
    }
    len = wave_max(len);
    if ((q & 63u) == 0 && q < N) width[q >> 6] = len;
}

// One 256-thread block per 64-row slice: the slice's links pass through an
// LDS tile of 64 rows x 64 links, read along each row (wavefront w reads
// rows 16w..16w+15, a row's 64 links per load) and written along each
// column (wavefront w writes columns 16w..16w+15, one row per lane), so both
// sides are coalesced.  Row r of slice s (sorted position q = 64 s + r)
// keeps link k at 64 sbase[s] + 64 k + r; padding cells are not written.
constexpr uint32_t kSellBlock = 256;
__global__ __launch_bounds__(kSellBlock) void k_sell_fill(uint32_t N, const uint32_t* __restrict__ lstart,
                                                          const uint32_t* __restrict__ perm,
                                                          const uint32_t* __restrict__ idx,
                                                          const float* __restrict__ w,
                                                          const uint32_t* __restrict__ sbase, uint32_t* sidx,
                                                          float* sw) {
    __shared__ uint32_t ti[64][65];
    __shared__ float tw[64][65];
    __shared__ uint32_t rs0[64], rlen[64];
    const uint32_t s = blockIdx.x, t = threadIdx.x, lane = t & 63u, r0 = (t >> 6) * 16u;
    if (t < 64u) {
        const uint32_t q = 64u * s + t;
        uint32_t s0 = 0, len = 0;
        if (q < N) {
            const uint32_t i = perm[q];
            s0 = lstart[i];
            len = lstart[i + 1] - s0;
        }
        rs0[t] = s0;
        rlen[t] = len;
    }
    __syncthreads();
    const uint32_t width = sbase[s + 1] - sbase[s];  // the slice's longest row
    const uint64_t base = 64ull * sbase[s];
    const uint32_t my_len = rlen[lane];
    for (uint32_t k0 = 0; k0 < width; k0 += 64u) {
        for (uint32_t r = r0; r < r0 + 16u; ++r) {
            const uint32_t k = k0 + lane;
            if (k < rlen[r]) {
                ti[r][lane] = idx[rs0[r] + k];
                tw[r][lane] = w[rs0[r] + k];
            }
        }
        __syncthreads();
        for (uint32_t c = r0; c < r0 + 16u; ++c) {
            const uint32_t k = k0 + c;
            if (k < my_len) {
                sidx[base + 64ull * k + lane] = ti[lane][c];
                sw[base + 64ull * k + lane] = tw[lane][c];
            }
        }
        __syncthreads();
    }
}

// ====================================================================== shade
__global__ void k_grid_keys(const float4* __restrict__ spos, uint32_t N, f3 lo, float inv_cell, int32_t dx, int32_t dy,
                            int32_t dz, uint32_t* keys, uint32_t* vals, uint32_t* counts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    f3 p = ld3(spos[i]);
    int32_t cx = clampi(sat_i32((p.x - lo.x) * inv_cell), 0, dx - 1);
    int32_t cy = clampi(sat_i32((p.y - lo.y) * inv_cell), 0, dy - 1);
    int32_t cz = clampi(sat_i32((p.z - lo.z) * inv_cell), 0, dz - 1);
    uint32_t c = (uint32_t)(cx + dx * (cy + dy * cz));
    keys[i] = c;
    vals[i] = i;
    atomicAdd(&counts[c], 1u);
}

__global__ void k_encode(const float4* __restrict__ film, uint32_t npx, float exposure, int32_t tonemap,
                         const float* __restrict__ thr, uint8_t* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    encode_px(film[i], exposure, tonemap, thr, out + 3ull * i);
}

// Clears n words (a kernel, not a memset, so the clear can sit in a captured graph).
__global__ void k_zero_words(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Packed xyz triples <-> float4 records, so host copies stay contiguous.
__global__ void k_pack3(const float4* __restrict__ src, uint64_t n, float* dst) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = src[i];
    dst[3 * i] = v.x;
    dst[3 * i + 1] = v.y;
    dst[3 * i + 2] = v.z;
}

__global__ void k_unpack3(const float* __restrict__ src, uint64_t n, float4* dst) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0f);
}

// ====================================================================== host side
namespace {
float host_srgb(float c) {
    c = clmp(c, 0.0f, 1.0f);
    float v = c <= 0.0031308f ? c * 12.92f : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    float q = v * 255.0f + 0.5f;
    return (float)(!(q > 0.0f) ? 0 : (q >= 255.0f ? 255 : (int)q));
}
}  // namespace

Engine::Lds Engine::plan_lds(bool lights, bool stage) const {
    const HostScene& s = scene;
    const uint32_t n_obj = (uint32_t)s.objects.size();
    uint32_t scene_f4 = (uint32_t)(s.obj_rec.size() + s.plane_rec.size() + s.sub_rec.size() + s.node_rec.size()) +
                        (2u * n_obj + 3u) / 4u;
    uint32_t light_f4 =
        lights ? (uint32_t)(s.quad_rec.size() + s.sph_rec.size() + s.pnt_rec.size() + s.mat_rec.size()) : 0u;
    const uint32_t depth = s.max_stack ? s.max_stack : 1u;
    const size_t stack = stack_bytes(depth);
    Lds r;
    size_t total = (size_t)(scene_f4 + light_f4) * 16u + stack;
    if (use_lds && stage && total <= kLdsBudget) {
        r.plan = LdsPlan{scene_f4, light_f4, depth};
        r.bytes = total;
        r.staged = true;
    } else {
        r.plan = LdsPlan{0u, 0u, depth};
        r.bytes = stack;
        r.staged = false;
    }
    return r;
}

int Engine::check(hipError_t e, const char* what, std::string& err) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? -5 : -4;
}


int Engine::scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, std::string& err) {
    CK(d_tmp.ensure(4ull * gsort::scan_scratch_words(n)));
    CK(gsort::exclusive_scan(in, out, n, d_tmp.as<uint32_t>(), stream));
    return 0;
}

int Engine::sort_u32(const uint32_t* k, const uint32_t* v, uint32_t* ko, uint32_t* vo, uint64_t n, uint32_t bits,
                     std::string& err) {
    CK(d_tmp.ensure(4ull * gsort::sort_scratch_words(n)));
    CK(d_sort_k.ensure(4ull * n));
    CK(d_sort_v.ensure(4ull * n));
    CK(gsort::sort_pairs_to<uint32_t>(k, v, ko, vo, d_sort_k.as<uint32_t>(), d_sort_v.as<uint32_t>(), n, bits,
                                      d_tmp.as<uint32_t>(), stream));
    return 0;
}

int Engine::sort_u64(const uint64_t* k, const uint32_t* v, uint64_t* ko, uint32_t* vo, uint64_t n, uint32_t bits,
                     std::string& err) {
    CK(d_tmp.ensure(4ull * gsort::sort_scratch_words(n)));
    CK(d_sort_k.ensure(8ull * n));
    CK(d_sort_v.ensure(4ull * n));
    CK(gsort::sort_pairs_to<uint64_t>(k, v, ko, vo, d_sort_k.as<uint64_t>(), d_sort_v.as<uint32_t>(), n, bits,
                                      d_tmp.as<uint32_t>(), stream));
    return 0;
}

Engine::~Engine() {
    if (stream) {
        (void)hipSetDevice(device_);
        if (side_) (void)hipStreamSynchronize(side_);
        (void)hipStreamSynchronize(stream);
        drain_events();
        for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
        if (fork_) (void)hipEventDestroy(fork_);
        if (join_) (void)hipEventDestroy(join_);
        if (closed_) (void)hipEventDestroy(closed_);
        if (vis_done_) (void)hipEventDestroy(vis_done_);
        if (side_) (void)hipStreamDestroy(side_);
        (void)hipStreamDestroy(stream);
    }
}

int Engine::init(int device, std::string& err) {
    if (device < 0) {  // host-only context: loader and surfel placement, no kernels
        has_device_ = false;
        return 0;
    }
    int n = 0;
    CK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        err = "no HIP device " + std::to_string(device);
        return -1;
    }
    device_ = device;
    CK(hipSetDevice(device));
    CK(hipStreamCreate(&stream));
    CK(hipStreamCreate(&side_));
    CK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&closed_, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&vis_done_, hipEventDisableTiming));
    {
        int cus = 0;
        CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        n_cu_ = cus > 0 ? cus : 1;
    }
    // Closure directions and sRGB thresholds are scene independent.
    float4 dirs[16];
    for (uint32_t k = 0; k < 16; ++k) {
        auto ri = [](uint32_t i, uint32_t base) {
            float f = 1.0f, r = 0.0f, inv = 1.0f / (float)base;
            while (i > 0) f *= inv, r += f * (float)(i % base), i /= base;
            return r;
        };
        float u1 = ri(k + 1, 2), u2 = ri(k + 1, 3);
        float r = sqrtf(u1);
        float phi = kTau * u2;
        dirs[k] = make_float4(r * cosf(phi), r * sinf(phi), sqrtf(fmx(1.0f - u1, 0.0f)), 0.0f);
    }
    CK(d_dirs.ensure(sizeof(dirs)));
    CK(hipMemcpy(d_dirs.p, dirs, sizeof(dirs), hipMemcpyHostToDevice));
    float thr[255];
    for (int k = 1; k <= 255; ++k) {
        uint32_t lo = 0, hi = fbits(1.0f);  // srgb(1.0) = 255
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (host_srgb(bitsf(mid)) >= (float)k)
                hi = mid;
            else
                lo = mid + 1;
        }
        thr[k - 1] = bitsf(lo);
    }
    CK(d_srgb.ensure(sizeof(thr)));
    CK(hipMemcpy(d_srgb.p, thr, sizeof(thr), hipMemcpyHostToDevice));
    CK(d_counter.ensure(256));
    CK(scratch_device(device, &scratch_));
    has_device_ = true;
    return 0;
}

int Engine::need_device(std::string& err) const {
    if (has_device_) return 0;
    err = "context was created without a device";
    return -3;
}

void Engine::drain_events() {
    std::vector<Pending> keep;
    for (Pending& p : pending_) {
        float ms = 0.0f;
        const hipError_t e = hipEventElapsedTime(&ms, p.a, p.b);
        if (e == hipErrorNotReady) {  // still running on the side stream
            (void)hipGetLastError();
            keep.push_back(p);
            continue;
        }
        if (e == hipSuccess) {
            KernelStat& k = kstats_[p.name];
            k.launches += 1;
            k.ms += ms;
        }
        (void)hipGetLastError();
        event_pool_.push_back(p.a);
        event_pool_.push_back(p.b);
    }
    pending_.swap(keep);
}

void Engine::tally_ops(const char* name, const unsigned long long* cv, int slot) {
    const std::string k(name);
    if (count_ops_) {
        kops_[k] += cv[slot];
        kops_[k + ".prims"] += cv[slot + 4];
    } else if (count_exec_) {
        kops_[k + ".exec_prims"] += cv[slot + 8];
        kops_[k + ".lane_slots"] += cv[slot + 12];
        kops_[k + ".flat_queries"] += cv[slot + 16];
        kops_[k + ".walk_queries"] += cv[slot + 20];
        kops_[k + ".listed"] += cv[slot + 24];
    }
}

void Engine::join_side() {
    if (!side_pending_) return;
    (void)hipStreamWaitEvent(stream, join_, 0);
    side_pending_ = false;
}

SceneView Engine::dev_view() const {
    SceneView v = scene.host_view();
    v.obj = d_obj.as<float4>();
    v.planes = d_planes.as<float4>();
    v.subs = d_subs.as<float4>();
    v.nodes = d_nodes.as<float4>();
    v.lin = d_lin.as<uint32_t>();
    v.use_bvh = use_bvh ? 1u : 0u;
    v.keep_global();
    v.g_obox = flat_query && !scene.obox_rec.empty() ? d_obox.as<float4>() : nullptr;
    return v;
}

LightsView Engine::dev_lights() const {
    LightsView l = scene.host_lights();
    l.quad = d_quad.as<float4>();
    l.sph = d_sph.as<float4>();
    l.pnt = d_pnt.as<float4>();
    l.mat = d_mat.as<float4>();
    return l;
}

// xyz triples of n float4 records to the host through a packed staging buffer
// (asynchronous on the stream; the caller synchronizes).
int Engine::copy3_d2h(float* host, const void* dev, uint64_t n, std::string& err) {
    if (!n || !host) return 0;
    CK(d_stage.ensure(12ull * n));
    k_pack3<<<blocks_for(n, 256), 256, 0, stream>>>((const float4*)dev, n, d_stage.as<float>());
    CK(hipMemcpyAsync(host, d_stage.p, 12ull * n, hipMemcpyDeviceToHost, stream));
    return 0;
}

int Engine::stage_images(uint64_t npx, bool rgb8, bool film, std::string& err) {
    const size_t rgb_bytes = rgb8 ? (3ull * npx + 255) & ~255ull : 0;
    if (!rgb8 && !film) return 0;
    CK(h_img_.ensure(rgb_bytes + (film ? 12ull * npx : 0)));
    if (rgb8) CK(hipMemcpyAsync(h_img_.p, d_rgb8.p, 3ull * npx, hipMemcpyDeviceToHost, stream));
    if (film) {
        CK(d_stage.ensure(12ull * npx));
        k_pack3<<<blocks_for(npx, 256), 256, 0, stream>>>(d_film.as<float4>(), npx, d_stage.as<float>());
        CK(hipMemcpyAsync((char*)h_img_.p + rgb_bytes, d_stage.p, 12ull * npx, hipMemcpyDeviceToHost, stream));
    }
    return 0;
}

void Engine::finish_images(uint64_t npx, uint8_t* rgb8, float* film) const {
    const size_t rgb_bytes = rgb8 ? (3ull * npx + 255) & ~255ull : 0;
    if (rgb8) memcpy(rgb8, h_img_.p, 3ull * npx);
    if (film) memcpy(film, (const char*)h_img_.p + rgb_bytes, 12ull * npx);
}

int Engine::copy3_h2d(void* dev, const float* host, uint64_t n, std::string& err) {
    if (!n || !host) return 0;
    CK(d_stage.ensure(12ull * n));
    CK(hipMemcpyAsync(d_stage.p, host, 12ull * n, hipMemcpyHostToDevice, stream));
    k_unpack3<<<blocks_for(n, 256), 256, 0, stream>>>(d_stage.as<float>(), n, (float4*)dev);
    return 0;
}

int Engine::load_scene(const std::string& src, std::string& err) {
    HostScene s;
    if (!load_rad(src, s, err)) return -2;
    return install_scene(std::move(s), err);
}

int Engine::install_scene(HostScene&& s, std::string& err) {
    if (s.max_stack > (uint32_t)kStackCap) {
        err = "BVH needs a traversal stack of " + std::to_string(s.max_stack) + " entries; the kernels hold " +
              std::to_string(kStackCap);
        return -6;
    }
    if (s.node_rec.size() / 2 > 65535) {
        err = "scene has more than 65535 BVH nodes";
        return -6;
    }
    if (side_) (void)hipStreamSynchronize(side_);  // no side work may still read the old records
    direct_pre_ = side_pending_ = false;
    if (!has_device_) {
        scene = std::move(s);
        surfels.clear();
        surfels_dirty = true;
        have_transfer = have_solve = have_gterm_ = false;
        return 0;
    }
    CK(hipSetDevice(device_));
    auto up = [&](DevBuf& b, const void* h, size_t bytes) -> int {
        CK(b.ensure(bytes));
        if (bytes) CK(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    int rc;
    if ((rc = up(d_obj, s.obj_rec.data(), s.obj_rec.size() * 16))) return rc;
    if ((rc = up(d_planes, s.plane_rec.data(), s.plane_rec.size() * 16))) return rc;
    if ((rc = up(d_subs, s.sub_rec.data(), s.sub_rec.size() * 16))) return rc;
    if ((rc = up(d_nodes, s.node_rec.data(), s.node_rec.size() * 16))) return rc;
    if ((rc = up(d_obox, s.obox_rec.data(), s.obox_rec.size() * 16))) return rc;
    {
        std::vector<uint32_t> lin_pad((s.lin.size() + 3) & ~(size_t)3, 0u);
        std::copy(s.lin.begin(), s.lin.end(), lin_pad.begin());
        if ((rc = up(d_lin, lin_pad.data(), lin_pad.size() * 4))) return rc;
    }
    if ((rc = up(d_quad, s.quad_rec.data(), s.quad_rec.size() * 16))) return rc;
    if ((rc = up(d_sph, s.sph_rec.data(), s.sph_rec.size() * 16))) return rc;
    if ((rc = up(d_pnt, s.pnt_rec.data(), s.pnt_rec.size() * 16))) return rc;
    if ((rc = up(d_mat, s.mat_rec.data(), s.mat_rec.size() * 16))) return rc;
    scene = std::move(s);
    surfels.clear();
    surfels_dirty = true;
    have_transfer = have_solve = have_gterm_ = false;
    return 0;
}

int Engine::set_surfels(const std::vector<HSurfel>& s, std::string& err) {
    if (s.size() >= (1ull << 31)) {
        err = "too many surfels";
        return -1;
    }
    surfels = s;
    surfels_dirty = true;
    have_transfer = have_solve = have_gterm_ = false;
    (void)err;
    return 0;
}

int Engine::transfer(int32_t res, TransferOut& out, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    if (res < 1 || res > 65535) {
        err = "cluster resolution must be in 1..65535";
        return -1;
    }
    CK(hipSetDevice(device_));
    join_side();  // earlier side work may still read the surfel records
    direct_pre_ = false;
    const uint32_t N = (uint32_t)surfels.size();
    have_transfer = have_solve = have_beauty_ = have_gterm_ = false;
    ++transfer_gen_;
    out = TransferOut();
    // H2D of the surfel records (part of the transfer stage).
    if (surfels_dirty) {
        std::vector<float4> hp(N), hn(N), ha(N);
        std::vector<uint32_t> hm(N);
        for (uint32_t i = 0; i < N; ++i) {
            hp[i] = st4(surfels[i].p, surfels[i].area);
            hn[i] = st4(surfels[i].n);
            ha[i] = st4(surfels[i].albedo);
            hm[i] = surfels[i].mat;
        }
        CK(d_spos.ensure(16ull * N));
        CK(d_snrm.ensure(16ull * N));
        CK(d_salb.ensure(16ull * N));
        CK(d_smat.ensure(4ull * N));
        if (N) {
            CK(hipMemcpyAsync(d_spos.p, hp.data(), 16ull * N, hipMemcpyHostToDevice, stream));
            CK(hipMemcpyAsync(d_snrm.p, hn.data(), 16ull * N, hipMemcpyHostToDevice, stream));
            CK(hipMemcpyAsync(d_salb.p, ha.data(), 16ull * N, hipMemcpyHostToDevice, stream));
            CK(hipMemcpyAsync(d_smat.p, hm.data(), 4ull * N, hipMemcpyHostToDevice, stream));
            CK(hipStreamSynchronize(stream));
        }
        surfels_dirty = false;
    }
    unsigned long long* ctr = d_counter.as<unsigned long long>();
    CK(hipMemsetAsync(ctr, 0, 256, stream));
    n_clusters = 0;
    n_links = 0;
    if (N == 0) {
        have_transfer = true;
        return 0;
    }
    const float4* spos = d_spos.as<float4>();
    const float4* snrm = d_snrm.as<float4>();

    // ---- work order: Morton cell of every surfel (coherent wavefronts only)
    const float diag = scene.diag();
    const SceneView S = dev_view();
    CK(d_morton.ensure(4ull * N));
    CK(d_mval.ensure(4ull * N));
    CK(d_mkey_s.ensure(4ull * N));
    CK(d_rorder.ensure(4ull * N));
    {
        const f3 ext = scene.hi - scene.lo;
        const f3 sc = mk(1024.0f / fmx(ext.x, 1e-6f), 1024.0f / fmx(ext.y, 1e-6f), 1024.0f / fmx(ext.z, 1e-6f));
        launch("morton", [&] {
            k_morton<<<blocks_for(N, 256), 256, 0, stream>>>(spos, N, scene.lo, sc, d_morton.as<uint32_t>(),
                                                             d_mval.as<uint32_t>());
        });
        if (int rc_s = sort_u32(d_morton.as<uint32_t>(), d_mval.as<uint32_t>(), d_mkey_s.as<uint32_t>(),
                           d_rorder.as<uint32_t>(), N, kMortonBits, err))
            return rc_s;
        CK(d_mpos.ensure(16ull * N));
        CK(d_mnrm.ensure(16ull * N));
        CK(d_mrank.ensure(4ull * N));
        launch("morton", [&] {
            k_morton_copy<<<blocks_for(N, 256), 256, 0, stream>>>(N, d_rorder.as<uint32_t>(), spos, snrm,
                                                                  d_mpos.as<float4>(), d_mnrm.as<float4>(),
                                                                  d_mrank.as<uint32_t>());
        });
    }
    CK(d_queue.ensure(16));
    CK(hipMemsetAsync(d_queue.p, 0, 16, stream));
    uint32_t* queue = d_queue.as<uint32_t>();
    const uint32_t persist = (uint32_t)n_cu_ * kMarchWaves;  // resident blocks of the march kernels

    // ---- direct-light terms and closure rays: they need only the surfels,
    // the scene and the lights, so they run on the side stream while this
    // stream builds clusters and candidates, whose host reads of counts
    // would otherwise leave the device idle.  A long link-visibility launch
    // waits for all of it: sharing the GPU with that persistent launch
    // slowed it by more than the side work's own time (stress scene, 68M
    // candidates: 343 against 313 ms with the terms beside it, 345 with the
    // closure).  A short one (up to kVisOverlapMax candidates) runs beside
    // the side work, which then fills its tail (product scene, 1.4M
    // candidates: frame 4.0 against 4.7 ms).  When the last transfer had a
    // long one, the direct-light terms wait for this transfer's link
    // visibility and run beside the link scan and compaction instead, so the
    // link visibility waits for the closure alone (stress scene: it starts
    // once the candidate sort is done).
    CK(d_sky.ensure(4ull * N));
    CK(d_open.ensure(4ull * N));
    CK(hipMemsetAsync(d_open.p, 0, 4ull * N, stream));
    const bool side = overlap_direct && !tally();
    const bool pre = side && n_lights(dev_lights()) > 0;
    const bool defer = pre && last_candidates_ > kVisOverlapMax;
    if (pre) {
        CK(d_dctr.ensure(64));
        CK(hipMemsetAsync(d_dctr.p, 0, 8, stream));
    }
    // The closure rays on the side stream (or this one), after the work
    // this stream has queued so far.
    auto closure_rays = [&]() -> int {
        if (side) {
            CK(hipEventRecord(fork_, stream));
            CK(hipStreamWaitEvent(side_, fork_, 0));
            side_pending_ = true;
        }
        hipStream_t cs = side ? side_ : stream;
        const Lds ld = plan_lds(false, false);
        const LightsView Lg{};
        const uint32_t grid = std::max(1u, std::min<uint32_t>(persist, blocks_for(16ull * N, kBlock)));
        launch_on(cs, "closure", [&] {
            hipStream_t stream = cs;  // the launch macro's stream
            LAUNCH_CM(k_closure, tally_mode(), ld.staged, grid, ld.bytes, S, Lg, ld.plan, d_mpos.as<float4>(),
                      d_mnrm.as<float4>(), N,
                      d_rorder.as<uint32_t>(), d_dirs.as<float4>(), diag * 2.0f, d_open.as<uint32_t>(), queue + 1,
                      wave_chunk(16ull * N, grid), std::min(64u, std::max(1u, closure_min_idle)), ctr,
                      tally() ? ctr + 2 : nullptr);
        });
        if (side) {
            CK(hipEventRecord(closed_, side_));
            CK(hipEventRecord(join_, side_));
        }
        return 0;
    };
    if (side && !defer) {
        CK(hipEventRecord(fork_, stream));
        CK(hipStreamWaitEvent(side_, fork_, 0));
        const int rc = pre ? launch_direct(side_, d_dctr.as<unsigned long long>(), nullptr, err) : 0;
        side_pending_ = true;
        direct_pre_ = pre && rc == 0;
        if (rc) return rc;
    }
    // Deferred terms: the closure starts once the candidate count is read
    // below, so this stream's cluster build, culls and that read do not
    // queue behind it (the persistent closure holds every CU until it ends).
    if (!defer)
        if (int rc = closure_rays()) return rc;

    // ---- clusters: key, stable sort, segments, serial sums, compaction
    const f3 lo = scene.lo, ext = scene.hi - scene.lo;
    const f3 inv = mk((float)res / fmx(ext.x, 1e-6f), (float)res / fmx(ext.y, 1e-6f), (float)res / fmx(ext.z, 1e-6f));
    const uint64_t rr = (uint64_t)res;
    const uint32_t key_bits = bits_for(rr * rr * rr * 6ull - 1ull);
    CK(d_ckey.ensure(8ull * N));
    CK(d_ckey2.ensure(8ull * N));
    CK(d_cval.ensure(4ull * N));
    CK(d_cval2.ensure(4ull * N));
    CK(d_cflag.ensure(4ull * (N + 1)));
    CK(d_cseg.ensure(4ull * (N + 1)));
    launch("cluster_keys", [&] {
        k_cluster_keys<<<blocks_for(N, 256), 256, 0, stream>>>(spos, snrm, N, lo, inv, res, d_ckey.as<uint64_t>(),
                                                               d_cval.as<uint32_t>());
    });
    if (int rc_s = sort_u64(d_ckey.as<uint64_t>(), d_cval.as<uint32_t>(), d_ckey2.as<uint64_t>(),
                       d_cval2.as<uint32_t>(), N, key_bits, err))
        return rc_s;
    const uint32_t* members = d_cval2.as<uint32_t>();
    launch("seg_flags", [&] {
        k_seg_flags<<<blocks_for(N, 256), 256, 0, stream>>>(d_ckey2.as<uint64_t>(), N, d_cflag.as<uint32_t>());
    });
    if (int rc_s = scan_u32(d_cflag.as<uint32_t>(), d_cseg.as<uint32_t>(), N, err))
        return rc_s;
    uint32_t tail[2];
    CK(hipMemcpyAsync(&tail[0], d_cseg.as<uint32_t>() + N - 1, 4, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(&tail[1], d_cflag.as<uint32_t>() + N - 1, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    const uint32_t nseg = tail[0] + tail[1];
    DevBuf& seg_begin = d_ckey;  // reuse: the unsorted keys are dead
    CK(d_cpos.ensure(16ull * nseg));
    CK(d_cnrm.ensure(16ull * nseg));
    CK(d_crange.ensure(8ull * nseg));
    CK(d_cvalid.ensure(4ull * (nseg + 1)));
    CK(d_cidx.ensure(4ull * (nseg + 1)));
    DevBuf& raw_pos = d_craw_pos;
    DevBuf& raw_nrm = d_craw_nrm;
    CK(raw_pos.ensure(16ull * nseg));
    CK(raw_nrm.ensure(16ull * nseg));
    launch("seg_begin", [&] {
        k_seg_begin<<<blocks_for(N, 256), 256, 0, stream>>>(d_cflag.as<uint32_t>(), d_cseg.as<uint32_t>(), N,
                                                            seg_begin.as<uint32_t>());
    });
    launch("cluster_build", [&] {
        k_cluster_build<<<blocks_for(64ull * nseg, 256), 256, 0, stream>>>(seg_begin.as<uint32_t>(), nseg, members, spos, snrm,
                                                                 raw_pos.as<float4>(), raw_nrm.as<float4>(),
                                                                 d_cvalid.as<uint32_t>());
    });
    if (int rc_s = scan_u32(d_cvalid.as<uint32_t>(), d_cidx.as<uint32_t>(), nseg, err))
        return rc_s;
    launch("cluster_compact", [&] {
        k_cluster_compact<<<blocks_for(nseg, 256), 256, 0, stream>>>(
            nseg, d_cvalid.as<uint32_t>(), d_cidx.as<uint32_t>(), seg_begin.as<uint32_t>(), raw_pos.as<float4>(),
            raw_nrm.as<float4>(), d_cpos.as<float4>(), d_cnrm.as<float4>(), d_crange.as<uint2>());
    });
    CK(hipMemcpyAsync(&tail[0], d_cidx.as<uint32_t>() + nseg - 1, 4, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(&tail[1], d_cvalid.as<uint32_t>() + nseg - 1, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    const uint32_t C = tail[0] + tail[1];
    n_clusters = C;
    out.clusters = C;

    // ---- candidates: count, scan, emit
    if ((uint64_t)N * C >= (1ull << 32)) {
        err = "surfels x clusters exceeds 2^32 candidate pairs";
        return -6;
    }
    CK(d_ccount.ensure(4ull * (N + 1)));
    CK(d_cstart.ensure(4ull * (N + 1)));
    CK(hipMemsetAsync(d_ccount.as<uint32_t>() + N, 0, 4, stream));
    CK(d_ffmask.ensure(8ull * N * ((C + 63u) / 64u)));
    const uint32_t ffb = ff_blocks ? ff_blocks : std::min<uint32_t>(blocks_for(N, 4), 8192);
    launch("ff_count", [&] {
        k_ff<false><<<ffb, kBlock, 0, stream>>>(spos, snrm, 0u, N, d_cpos.as<float4>(), d_cnrm.as<float4>(), C,
                                                 d_ccount.as<uint32_t>(), nullptr, nullptr, nullptr, nullptr,
                                                 nullptr, nullptr, nullptr, d_ffmask.as<uint64_t>(), 0u);
    });
    if (int rc_s = scan_u32(d_ccount.as<uint32_t>(), d_cstart.as<uint32_t>(), (N + 1), err))
        return rc_s;
    uint32_t M = 0;
    CK(hipMemcpyAsync(&M, d_cstart.as<uint32_t>() + N, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    if (defer)
        if (int rc = closure_rays()) return rc;
    out.candidates = M;
    if (M >= (1u << 31)) {
        err = "candidate count exceeds 2^31";
        return -6;
    }
    // Cluster-major, Morton-minor visiting order of the candidates: the
    // emission lays the sort's input out with the rows in Morton order, so
    // a stable sort by cluster alone (2 passes of 8 bits up to 65536
    // clusters) gives it.
    const uint32_t cbits = bits_for(C ? C - 1 : 0);
    const uint32_t mbits = 0u;  // the sorted keys hold the cluster alone
    const uint32_t vkey_bits = cbits;
    const bool sort_cand = cbits <= 17u && M > 0;
    // Sort values packed as (row's Morton rank << jbits) | index in the row
    // when both fit in 32 bits: the link visibility then reads its row's rank
    // from the sorted values (no gather pass) and the result scatter finds the
    // slot from the row's candidate start.  Otherwise the values are slots.
    const uint32_t jbits0 = std::max(1u, bits_for(C ? C - 1 : 0));
    const uint32_t jbits = sort_cand && bits_for(N ? N - 1 : 0) + jbits0 <= 32u ? jbits0 : 0u;
    CK(d_cand_ci.ensure(4ull * M));
    CK(d_cand_row.ensure(4ull * M));
    CK(d_cand_v.ensure(4ull * M));
    if (sort_cand) {
        CK(d_vkey.ensure(4ull * M));
        CK(d_vval.ensure(4ull * M));
        CK(d_vkey_s.ensure(4ull * M));
        CK(d_vorder.ensure(4ull * M));
        CK(d_mcount.ensure(4ull * (N + 1)));
        CK(d_mstart.ensure(4ull * (N + 1)));
        CK(hipMemsetAsync(d_mcount.as<uint32_t>() + N, 0, 4, stream));
        k_morton_counts<<<blocks_for(N, 256), 256, 0, stream>>>(N, d_rorder.as<uint32_t>(), d_ccount.as<uint32_t>(),
                                                                d_mcount.as<uint32_t>());
        if (int rc_s = scan_u32(d_mcount.as<uint32_t>(), d_mstart.as<uint32_t>(), N + 1, err)) return rc_s;
    }
    launch("ff_emit", [&] {
        k_ff<true><<<ffb, kBlock, 0, stream>>>(spos, snrm, 0u, N, d_cpos.as<float4>(), d_cnrm.as<float4>(), C,
                                                nullptr,
                                                d_cstart.as<uint32_t>(), d_cand_ci.as<uint32_t>(),
                                                jbits ? nullptr : d_cand_row.as<uint32_t>(), d_mrank.as<uint32_t>(),
                                                d_mstart.as<uint32_t>(), sort_cand ? d_vkey.as<uint32_t>() : nullptr,
                                                sort_cand ? d_vval.as<uint32_t>() : nullptr,
                                                d_ffmask.as<uint64_t>(), jbits);
    });
    if (sort_cand) {
        if (int rc_s = sort_u32(d_vkey.as<uint32_t>(), d_vval.as<uint32_t>(), d_vkey_s.as<uint32_t>(),
                           d_vorder.as<uint32_t>(), M, vkey_bits, err))
            return rc_s;
    }

    // ---- link visibility, in visiting order (the unsorted fallback visits row-major)
    last_candidates_ = M;
    // the side stream is done (the closure alone when the terms are deferred)
    if (side && M > kVisOverlapMax) CK(hipStreamWaitEvent(stream, defer ? closed_ : join_, 0));
    if (M) {
        const uint32_t visb =
            vis_blocks ? vis_blocks : std::max(1u, std::min<uint32_t>(persist, blocks_for(M, kBlock)));
        const Lds ld = plan_lds(false, false);
        const LightsView Lg{};
        // the sort's input keys and values are dead: they hold the visiting-order rows and results
        uint32_t* vrow = d_vkey.as<uint32_t>();
        float* vis = sort_cand ? d_vval.as<float>() : d_cand_v.as<float>();
        const uint32_t* crow = !sort_cand ? d_cand_row.as<uint32_t>() : jbits ? d_vorder.as<uint32_t>() : vrow;
        if (sort_cand && !jbits) {
            launch("vis_order", [&] {
                k_vis_order<<<blocks_for(M, 256), 256, 0, stream>>>(M, d_vorder.as<uint32_t>(),
                                                                    d_cand_row.as<uint32_t>(),
                                                                    d_mrank.as<uint32_t>(), vrow);
            });
        }
        // Parked tails need whole-wavefront batches starting at multiples of 64.
        const uint32_t min_idle = std::min(64u, std::max(1u, vis_min_idle));
        const uint32_t at = min_idle == 64u && M >= vis_tail_items ? std::min(vis_tail, 32u) : 0u;
        uint32_t chunk = vis_chunk ? vis_chunk : wave_chunk(M, visb);
        const uint32_t nb = (M + 63u) / 64u;
        VisTail tl{};
        if (at) {
            chunk = std::max(64u, chunk & ~63u);
            CK(d_tail_st.ensure(64ull * nb * at));
            CK(d_tail_cnt.ensure(4ull * (nb + 1)));
            CK(d_tail_start.ensure(4ull * (nb + 1)));
            CK(d_tail_idx.ensure(4ull * nb * at));
            CK(hipMemsetAsync(d_tail_cnt.as<uint32_t>() + nb, 0, 4, stream));
            tl.st = d_tail_st.as<float4>();
            tl.cnt = d_tail_cnt.as<uint32_t>();
            tl.at = at;
        }
        int rc_t = 0;
        // both launches (and the parked rays' index between them) time as one
        launch("visibility", [&] {
            LAUNCH_CM(k_visibility, tally_mode(), ld.staged, visb, ld.bytes, S, Lg, ld.plan,
                      sort_cand ? d_mpos.as<float4>() : spos, sort_cand ? d_mnrm.as<float4>() : snrm,
                      d_cpos.as<float4>(), d_cnrm.as<float4>(),
                      sort_cand ? d_vkey_s.as<uint32_t>() : d_cand_ci.as<uint32_t>(), sort_cand ? mbits : 0u,
                      crow, jbits, M, diag * 5e-4f, diag * 1e-3f, vis, queue,
                      chunk, min_idle, tl, ctr, tally() ? ctr + 1 : nullptr);
            if (!at) return;
            rc_t = scan_u32(d_tail_cnt.as<uint32_t>(), d_tail_start.as<uint32_t>(), nb + 1, err);
            if (rc_t) return;
            k_tail_index<<<blocks_for((uint64_t)nb * at, 256), 256, 0, stream>>>(
                nb, at, d_tail_cnt.as<uint32_t>(), d_tail_start.as<uint32_t>(), d_tail_idx.as<uint32_t>());
            VisTail rt{};
            rt.st = d_tail_st.as<float4>();
            rt.idx = d_tail_idx.as<uint32_t>();
            rt.total = d_tail_start.as<uint32_t>() + nb;
            LAUNCH_CM(k_visibility, tally_mode(), ld.staged, visb, ld.bytes, S, Lg, ld.plan,
                      sort_cand ? d_mpos.as<float4>() : spos, sort_cand ? d_mnrm.as<float4>() : snrm,
                      d_cpos.as<float4>(), d_cnrm.as<float4>(),
                      sort_cand ? d_vkey_s.as<uint32_t>() : d_cand_ci.as<uint32_t>(), sort_cand ? mbits : 0u,
                      crow, jbits, M, diag * 5e-4f, diag * 1e-3f, vis, queue + 2,
                      64u, 64u, rt, ctr, tally() ? ctr + 1 : nullptr);
        });
        if (rc_t) return rc_t;
        if (sort_cand) {
            launch("vis_scatter", [&] {
                if (jbits) {
                    // the Morton counts are dead after the emission: they hold the starts by rank
                    k_morton_base<<<blocks_for(N, 256), 256, 0, stream>>>(N, d_rorder.as<uint32_t>(),
                                                                          d_cstart.as<uint32_t>(),
                                                                          d_mcount.as<uint32_t>());
                    k_vis_scatter_packed<<<blocks_for(M, 256), 256, 0, stream>>>(
                        M, d_vorder.as<uint32_t>(), jbits, d_mcount.as<uint32_t>(), vis, d_cand_v.as<float>());
                } else
                    k_vis_scatter<<<blocks_for(M, 256), 256, 0, stream>>>(M, d_vorder.as<uint32_t>(), vis,
                                                                          d_cand_v.as<float>());
            });
        }
    }
    if (defer) {  // the deferred direct-light terms, beside the link scan and compaction
        CK(hipEventRecord(vis_done_, stream));
        CK(hipStreamWaitEvent(side_, vis_done_, 0));
        const int rc = launch_direct(side_, d_dctr.as<unsigned long long>(), nullptr, err);
        if (rc) return rc;
        direct_pre_ = true;
        CK(hipEventRecord(join_, side_));
    }

    // ---- calibration, link scan, compaction
    CK(d_lcount.ensure(4ull * (N + 1)));
    CK(d_lstart.ensure(4ull * (N + 1)));
    CK(d_scale.ensure(4ull * N));
    CK(d_rowsum.ensure(4ull * N));
    CK(hipMemsetAsync(d_lcount.as<uint32_t>() + N, 0, 4, stream));
    if (side) CK(hipStreamWaitEvent(stream, closed_, 0));  // the closure's open-ray counts
    launch("row_sum", [&] {
        k_row_sum<<<blocks_for(64ull * N, 256), 256, 0, stream>>>(N, d_cstart.as<uint32_t>(), d_cand_v.as<float>(),
                                                          d_open.as<uint32_t>(), d_sky.as<float>(),
                                                          d_lcount.as<uint32_t>(), d_scale.as<float>(),
                                                          d_rowsum.as<float>());
    });
    if (int rc_s = scan_u32(d_lcount.as<uint32_t>(), d_lstart.as<uint32_t>(), (N + 1), err))
        return rc_s;
    uint32_t L = 0;
    CK(hipMemcpyAsync(&L, d_lstart.as<uint32_t>() + N, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    CK(d_idx.ensure(4ull * L));
    CK(d_w.ensure(4ull * L));
    launch("compact", [&] {
        k_compact<<<blocks_for(64ull * N, 256), 256, 0, stream>>>(N, d_cstart.as<uint32_t>(), d_cand_v.as<float>(),
                                                          d_cand_ci.as<uint32_t>(), d_lstart.as<uint32_t>(),
                                                          d_scale.as<float>(), d_idx.as<uint32_t>(), d_w.as<float>());
    });
    // mean_row_sum: serial f32 sum over rows with links, in row order.
    std::vector<float> rs(N);
    unsigned long long cv[32] = {0};
    CK(hipMemcpyAsync(rs.data(), d_rowsum.p, 4ull * N, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(cv, ctr, 256, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    unsigned long long ev = cv[0];
    tally_ops("visibility", cv, 1);
    tally_ops("closure", cv, 2);
    drain_events();
    float acc = -0.0f;
    uint64_t interior = 0;
    for (uint32_t i = 0; i < N; ++i)
        if (rs[i] >= 0.0f) acc += rs[i], ++interior;
    out.mean_row_sum = interior ? acc / (float)interior : 0.0f;
    out.links = L;
    out.evals = ev;
    n_links = L;
    have_transfer = true;
    return 0;
}

// Direct-light terms (k_direct_terms) on stream s; terms, geometry terms and
// the evaluation count go to d_terms, d_gterm and evals.
int Engine::launch_direct(hipStream_t s, unsigned long long* evals, unsigned long long* ops, std::string& err) {
    const uint32_t N = (uint32_t)surfels.size();
    const LightsView Lv = dev_lights();
    const uint32_t nl = n_lights(Lv);
    CK(d_terms.ensure(16ull * N * (nl ? nl : 1)));
    CK(d_gterm.ensure(4ull * N * (nl ? nl : 1)));
    if (!N || !nl) return 0;
    const SceneView S = dev_view();
    const Lds ld = plan_lds(true);
    hipStream_t stream = s;  // the launch macro's stream
    launch_on(stream, "direct_e", [&] {
        LAUNCH_CL(k_direct_terms, count_ops_, ld.staged, blocks_for((uint64_t)N * nl, kBlock), ld.bytes, S, Lv,
                  ld.plan, d_mpos.as<float4>(), d_mnrm.as<float4>(), d_salb.as<float4>(), N, nl,
                  d_rorder.as<uint32_t>(), d_terms.as<float4>(), d_gterm.as<float>(), evals, ops);
    });
    return 0;
}

int Engine::solve(uint32_t bounces, SolveOut& out, std::string& err) {
    if (!have_transfer) {
        err = "solve called before transfer";
        return -3;
    }
    CK(hipSetDevice(device_));
    out = SolveOut();
    // direct-light terms precomputed by transfer() on the side stream
    const bool pre = direct_pre_ && !tally();
    join_side();
    direct_pre_ = false;
    const uint32_t N = (uint32_t)surfels.size();
    const LightsView Lv = dev_lights();
    const uint32_t nl = n_lights(Lv);
    CK(d_direct.ensure(16ull * N));
    CK(d_irr.ensure(16ull * N));
    CK(d_rad.ensure(16ull * N));
    CK(d_bc.ensure(16ull * (n_clusters + 1)));
    CK(d_resid.ensure(4ull * (bounces + 1)));
    CK(d_terms.ensure(16ull * N * (nl ? nl : 1)));
    CK(d_gterm.ensure(4ull * N * (nl ? nl : 1)));
    unsigned long long* ctr = d_counter.as<unsigned long long>();
    CK(hipMemsetAsync(ctr, 0, 256, stream));
    CK(hipMemsetAsync(d_resid.p, 0, 4ull * (bounces + 1), stream));
    const SceneView S = dev_view();
    if (N) {
        if (nl && !pre) {
            if (int rc = launch_direct(stream, ctr, tally() ? ctr + 3 : nullptr, err)) return rc;
        }
        launch("direct_sum", [&] {
            k_direct_sum<<<blocks_for(N, 256), 256, 0, stream>>>(N, nl, d_terms.as<float4>(), d_direct.as<float4>(),
                                                                 d_rad.as<float4>(), d_irr.as<float4>());
        });
    }
    const uint32_t C = n_clusters;
    if (C > 0 && N > 0) {
        JacobiArgs A;
        A.N = N;
        A.C = C;
        A.tol = solve_tol;
        A.wave_rows = jacobi_rows == 2 ? 1u : 0u;
        A.sbase = nullptr;
        A.sidx = nullptr;
        A.sw = nullptr;
        A.sperm = nullptr;
        // automatic: sliced rows when rows are long enough (>= 64 links on
        // average) for coalescing to repay building the copy once per transfer
        if (jacobi_rows == 3 || (jacobi_rows == 0 && n_links >= 64ull * N)) {
            if (int rc = build_sliced(err)) return rc;
            A.sbase = d_sbase.as<uint32_t>();
            A.sidx = d_sidx.as<uint32_t>();
            A.sw = d_sw.as<float>();
            A.sperm = d_sperm.as<uint32_t>();
        }
        A.range = d_crange.as<uint2>();
        A.members = d_cval2.as<uint32_t>();
        A.spos = d_spos.as<float4>();
        A.cpos = d_cpos.as<float4>();
        A.lstart = d_lstart.as<uint32_t>();
        A.idx = d_idx.as<uint32_t>();
        A.w = d_w.as<float>();
        A.direct = d_direct.as<float4>();
        A.salb = d_salb.as<float4>();
        A.rad = d_rad.as<float4>();
        A.irr = d_irr.as<float4>();
        A.bc = d_bc.as<float4>();
        A.resid = d_resid.as<uint32_t>();
        const uint32_t g_agg = jacobi_blocks ? jacobi_blocks : blocks_for(64ull * C, kBlock);
        const uint32_t g_rows =
            jacobi_blocks ? jacobi_blocks : blocks_for(A.wave_rows ? 64ull * N : (uint64_t)N, kBlock);
        // Bounces are enqueued in groups; between groups the host reads the
        // last residual and stops enqueuing once the solve has converged.
        const uint32_t group = 8;
        uint32_t done_at = 0;
        for (uint32_t it0 = 0; it0 < bounces && !done_at; it0 += group) {
            const uint32_t it1 = std::min(bounces, it0 + group);
            launch("jacobi", [&] {
                for (uint32_t it = it0; it < it1; ++it) {
                    k_jacobi_agg<<<g_agg, kBlock, 0, stream>>>(A, it);
                    k_jacobi_rows<<<g_rows, kBlock, 0, stream>>>(A, it);
                }
            });
            if (it1 < bounces) {
                uint32_t last = 0;
                CK(hipMemcpyAsync(&last, A.resid + it1 - 1, 4, hipMemcpyDeviceToHost, stream));
                CK(hipStreamSynchronize(stream));
                if (bitsf(last) < solve_tol) done_at = it1;
            }
        }
    }
    std::vector<uint32_t> res(bounces + 1, 0);
    unsigned long long cv[32] = {0};
    CK(hipMemcpyAsync(res.data(), d_resid.p, 4ull * bounces, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(cv, ctr, 256, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    drain_events();
    unsigned long long ev = cv[0];
    if (pre) {  // the direct-light evaluations counted on the side stream
        unsigned long long dv = 0;
        CK(hipMemcpy(&dv, d_dctr.p, 8, hipMemcpyDeviceToHost));
        ev += dv;
    }
    tally_ops("direct_e", cv, 3);
    out.evals = ev;
    have_beauty_ = false;
    if (C > 0 && N > 0) {
        out.iterations = bounces;
        for (uint32_t it = 0; it < bounces; ++it) {
            if (bitsf(res[it]) < solve_tol) {
                out.iterations = it + 1;
                break;
            }
        }
        out.residual = bitsf(res[out.iterations - 1]);
    }
    have_solve = have_gterm_ = true;
    return 0;
}

// solve() with caller-owned outputs and every scratch buffer in the static
// scratch array (no allocation): the same kernels in the same order, so the
// results equal solve()'s bit for bit.  Layout of the array (256-byte
// aligned sections): evaluation counter, per-bounce residuals, cluster
// radiosity, direct-light terms, light factors.  Counter and residuals are
// cleared by a kernel at every call; every other section is written before
// it is read.
int Engine::solve_noalloc(uint32_t bounces, float4* direct, float4* irr, float4* rad, SolveOut& out,
                          std::string& err) {
    if (!have_transfer) {
        err = "solve called before transfer";
        return -3;
    }
    if (!scratch_) {
        err = "no scratch array on this device";
        return -3;
    }
    if (bounces == 0 || bounces > 4096) {
        err = "bounces must be in 1..4096";
        return -1;
    }
    const uint32_t N = (uint32_t)surfels.size();
    const LightsView Lv = dev_lights();
    const uint32_t nl = n_lights(Lv);
    const uint32_t C = n_clusters;
    auto al = [](uint64_t b) { return (b + 255ull) & ~255ull; };
    const uint64_t o_res = 256, o_bc = o_res + al(4ull * (bounces + 1)), o_terms = o_bc + al(16ull * (C + 1)),
                   o_g = o_terms + al(16ull * N * nl), total = o_g + al(4ull * N * nl);
    if (total > kScratchBytes) {
        err = "solve_noalloc needs " + std::to_string(total) + " bytes of scratch (" + std::to_string(N) +
              " surfels x " + std::to_string(nl) + " lights, " + std::to_string(C) +
              " clusters); the static scratch array holds " + std::to_string(kScratchBytes);
        return -6;
    }
    if (N && (!direct || !irr || !rad)) {
        err = "null output array";
        return -1;
    }
    CK(hipSetDevice(device_));
    out = SolveOut();
    join_side();  // the side stream's closure / terms of the last transfer
    unsigned long long* ctr = (unsigned long long*)scratch_;
    uint32_t* resid = (uint32_t*)(scratch_ + o_res);
    float4* bc = (float4*)(scratch_ + o_bc);
    float4* terms = (float4*)(scratch_ + o_terms);
    float* gterm = (float*)(scratch_ + o_g);
    const uint32_t clear_words = (uint32_t)(o_bc / 4);
    k_zero_words<<<blocks_for(clear_words, 256), 256, 0, stream>>>((uint32_t*)scratch_, clear_words);
    const SceneView S = dev_view();
    if (N) {
        if (nl) {
            const Lds ld = plan_lds(true);
            LAUNCH_CL(k_direct_terms, false, ld.staged, blocks_for((uint64_t)N * nl, kBlock), ld.bytes, S, Lv,
                      ld.plan, d_mpos.as<float4>(), d_mnrm.as<float4>(), d_salb.as<float4>(), N, nl,
                      d_rorder.as<uint32_t>(), terms, gterm, ctr, nullptr);
        }
        k_direct_sum<<<blocks_for(N, 256), 256, 0, stream>>>(N, nl, terms, direct, rad, irr);
    }
    uint32_t res[4097];
    if (C > 0 && N > 0) {
        JacobiArgs A;
        A.N = N;
        A.C = C;
        A.tol = solve_tol;
        A.wave_rows = jacobi_rows == 2 ? 1u : 0u;
        A.sbase = nullptr;
        A.sidx = nullptr;
        A.sw = nullptr;
        A.sperm = nullptr;
        // the sliced copy only when this transfer's is already built (building it allocates)
        const bool sliced = jacobi_rows == 3 || (jacobi_rows == 0 && n_links >= 64ull * N);
        if (sliced && sliced_gen_ == transfer_gen_ && !A.wave_rows) {
            A.sbase = d_sbase.as<uint32_t>();
            A.sidx = d_sidx.as<uint32_t>();
            A.sw = d_sw.as<float>();
            A.sperm = d_sperm.as<uint32_t>();
        }
        A.range = d_crange.as<uint2>();
        A.members = d_cval2.as<uint32_t>();
        A.spos = d_spos.as<float4>();
        A.cpos = d_cpos.as<float4>();
        A.lstart = d_lstart.as<uint32_t>();
        A.idx = d_idx.as<uint32_t>();
        A.w = d_w.as<float>();
        A.direct = direct;
        A.salb = d_salb.as<float4>();
        A.rad = rad;
        A.irr = irr;
        A.bc = bc;
        A.resid = resid;
        const uint32_t g_agg = jacobi_blocks ? jacobi_blocks : blocks_for(64ull * C, kBlock);
        const uint32_t g_rows =
            jacobi_blocks ? jacobi_blocks : blocks_for(A.wave_rows ? 64ull * N : (uint64_t)N, kBlock);
        const uint32_t group = 8;
        bool done = false;
        for (uint32_t it0 = 0; it0 < bounces && !done; it0 += group) {
            const uint32_t it1 = std::min(bounces, it0 + group);
            for (uint32_t it = it0; it < it1; ++it) {
                k_jacobi_agg<<<g_agg, kBlock, 0, stream>>>(A, it);
                k_jacobi_rows<<<g_rows, kBlock, 0, stream>>>(A, it);
            }
            if (it1 < bounces) {
                uint32_t last = 0;
                CK(hipMemcpyAsync(&last, resid + it1 - 1, 4, hipMemcpyDeviceToHost, stream));
                CK(hipStreamSynchronize(stream));
                done = bitsf(last) < solve_tol;
            }
        }
        CK(hipMemcpyAsync(res, resid, 4ull * bounces, hipMemcpyDeviceToHost, stream));
    }
    CK(hipGetLastError());
    unsigned long long ev = 0;
    CK(hipMemcpyAsync(&ev, ctr, 8, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    out.evals = ev;
    if (C > 0 && N > 0) {
        out.iterations = bounces;
        for (uint32_t it = 0; it < bounces; ++it) {
            if (bitsf(res[it]) < solve_tol) {
                out.iterations = it + 1;
                break;
            }
        }
        out.residual = bitsf(res[out.iterations - 1]);
    }
    return 0;
}

int Engine::tile_links(uint32_t r0, uint32_t nr, uint32_t& M, uint32_t& L, bool sized, std::string& err) {
    const uint32_t C = n_clusters;
    const float diag = scene.diag();
    if (!sized) {
        CK(d_tcount.ensure(4ull * (nr + 1)));
        CK(d_tstart.ensure(4ull * (nr + 1)));
        CK(d_tkeep.ensure(4ull * (nr + 1)));
        CK(d_tlstart.ensure(4ull * (nr + 1)));
        CK(d_tqueue.ensure(16));
    }
    const float4* spos = d_spos.as<float4>();
    const float4* snrm = d_snrm.as<float4>();
    const uint32_t ffb = ff_blocks ? ff_blocks : std::min<uint32_t>(blocks_for(nr, 4), 8192);
    CK(hipMemsetAsync(d_tcount.as<uint32_t>() + nr, 0, 4, stream));
    k_ff<false><<<ffb, kBlock, 0, stream>>>(spos, snrm, r0, nr, d_cpos.as<float4>(), d_cnrm.as<float4>(), C,
                                             d_tcount.as<uint32_t>(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                             nullptr, nullptr, nullptr, 0u);
    if (int rc = scan_u32(d_tcount.as<uint32_t>(), d_tstart.as<uint32_t>(), nr + 1, err)) return rc;
    if (!sized) {
        CK(hipMemcpyAsync(&M, d_tstart.as<uint32_t>() + nr, 4, hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        CK(d_tci.ensure(4ull * M));
        CK(d_trow.ensure(4ull * M));
        CK(d_tv.ensure(4ull * M));
    }
    k_ff<true><<<ffb, kBlock, 0, stream>>>(spos, snrm, r0, nr, d_cpos.as<float4>(), d_cnrm.as<float4>(), C, nullptr,
                                            d_tstart.as<uint32_t>(), d_tci.as<uint32_t>(), d_trow.as<uint32_t>(),
                                            nullptr, nullptr, nullptr, nullptr, nullptr, 0u);
    if (M) {
        const uint32_t persist = (uint32_t)n_cu_ * kMarchWaves;  // resident blocks of the march kernels
        const uint32_t visb = vis_blocks ? vis_blocks : std::max(1u, std::min<uint32_t>(persist, blocks_for(M, kBlock)));
        const Lds ld = plan_lds(false, false);
        const LightsView Lg{};
        const SceneView S = dev_view();
        unsigned long long* ctr = d_counter.as<unsigned long long>();
        CK(hipMemsetAsync(d_tqueue.p, 0, 4, stream));
        LAUNCH_CM(k_visibility, 0, ld.staged, visb, ld.bytes, S, Lg, ld.plan, spos, snrm, d_cpos.as<float4>(),
                  d_cnrm.as<float4>(), d_tci.as<uint32_t>(), 0u, d_trow.as<uint32_t>(), 0u, M, diag * 5e-4f,
                  diag * 1e-3f, d_tv.as<float>(), d_tqueue.as<uint32_t>(), wave_chunk(M, visb),
                  std::min(64u, std::max(1u, vis_min_idle)), VisTail{}, ctr + 20, nullptr);
    }
    CK(hipMemsetAsync(d_tkeep.as<uint32_t>() + nr, 0, 4, stream));
    k_keep_count<<<blocks_for(nr, 256), 256, 0, stream>>>(nr, d_tstart.as<uint32_t>(), d_tv.as<float>(),
                                                          d_tkeep.as<uint32_t>());
    if (int rc = scan_u32(d_tkeep.as<uint32_t>(), d_tlstart.as<uint32_t>(), nr + 1, err)) return rc;
    if (!sized) {
        CK(hipMemcpyAsync(&L, d_tlstart.as<uint32_t>() + nr, 4, hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        CK(d_tidx.ensure(4ull * L));
        CK(d_tw.ensure(4ull * L));
    }
    k_compact<<<blocks_for(64ull * nr, 256), 256, 0, stream>>>(nr, d_tstart.as<uint32_t>(), d_tv.as<float>(),
                                                               d_tci.as<uint32_t>(), d_tlstart.as<uint32_t>(),
                                                               d_scale.as<float>() + r0, d_tidx.as<uint32_t>(),
                                                               d_tw.as<float>());
    CK(hipGetLastError());
    return 0;
}

// Sliced copy of the links for the Jacobi row gather, once per transfer.
int Engine::build_sliced(std::string& err) {
    if (sliced_gen_ == transfer_gen_) return 0;
    const uint32_t N = (uint32_t)surfels.size();
    const uint32_t ns = (N + 63u) / 64u;
    CK(d_sbase.ensure(4ull * (ns + 1)));
    CK(d_sperm.ensure(4ull * N));
    DevBuf& width = d_cand_row;  // transfer scratch, dead after the compaction
    CK(width.ensure(4ull * (ns + 1)));
    CK(d_skey.ensure(8ull * N));
    CK(d_skey2.ensure(8ull * N));
    const uint32_t C = n_clusters;
    if (N) {
        uint32_t* key = d_skey.as<uint32_t>();
        uint32_t* row = key + N;
        k_sell_keys<<<blocks_for(N, 256), 256, 0, stream>>>(N, C, d_lstart.as<uint32_t>(), key, row);
        if (int rc_s = sort_u32(key, row, d_skey2.as<uint32_t>(), d_sperm.as<uint32_t>(), N, bits_for(C), err))
            return rc_s;
    }
    CK(hipMemsetAsync(width.p, 0, 4ull * (ns + 1), stream));
    if (N)
        k_sell_widths<<<blocks_for(64ull * ns, 256), 256, 0, stream>>>(N, d_lstart.as<uint32_t>(),
                                                                       d_sperm.as<uint32_t>(), width.as<uint32_t>());
    if (int rc_s = scan_u32(width.as<uint32_t>(), d_sbase.as<uint32_t>(), (ns + 1), err))
        return rc_s;
    uint32_t cols = 0;
    CK(hipMemcpyAsync(&cols, d_sbase.as<uint32_t>() + ns, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    CK(d_sidx.ensure(256ull * cols + 4));
    CK(d_sw.ensure(256ull * cols + 4));
    if (N)
        k_sell_fill<<<ns, kSellBlock, 0, stream>>>(N, d_lstart.as<uint32_t>(), d_sperm.as<uint32_t>(),
                                                            d_idx.as<uint32_t>(), d_w.as<float>(),
                                                            d_sbase.as<uint32_t>(), d_sidx.as<uint32_t>(),
                                                            d_sw.as<float>());
    sliced_cells_ = 64ull * cols;
    sliced_gen_ = transfer_gen_;
    return 0;
}

// Camera, gather radius and (for the GI modes) the surfel grid of one
// render_frame call; shared by shade() and backward().
int Engine::shade_setup(const HCamera& cam, uint32_t W, uint32_t H, int32_t mode, ShadeArgs& A, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    if (mode < 0 || mode > 6) {
        err = "unknown shading mode";
        return -1;
    }
    HCfg c = scene.cfg;
    c.width = W;
    c.height = H;
    if (!validate_cfg(c, err)) return -1;
    CK(hipSetDevice(device_));
    const bool needs_gi = mode == M_BEAUTY || mode == M_INDIRECT;
    if (needs_gi && !have_solve) {
        err = "beauty / indirect shading needs a solve first";
        return -3;
    }
    const uint32_t N = (uint32_t)surfels.size();
    const float diag = scene.diag();
    A.cam = cam_basis(cam, W, H);
    A.W = W;
    A.H = H;
    A.mode = mode;
    A.use_gi = needs_gi ? 1 : 0;
    A.tmax = diag * 4.0f;
    A.radius = (diag / sqrtf((float)scene.cfg.surfels)) * 2.6f;
    A.spp = spp ? spp : 1u;
    A.spos = d_spos.as<float4>();
    A.snrm = d_snrm.as<float4>();
    A.irr = d_irr.as<float4>();
    A.grid.lo = scene.lo;
    A.grid.inv_cell = 1.0f;
    A.grid.dim[0] = A.grid.dim[1] = A.grid.dim[2] = 1;
    A.grid.starts = nullptr;
    A.grid.items = nullptr;

    if (needs_gi) {
        // Grid::build (src/surfel.rs:71-114): cell keys + counts, scan, stable sort.
        const float inv_cell = 1.0f / A.radius;
        const f3 ext = scene.hi - scene.lo;
        int32_t dx = clampi(sat_i32(ceilf(ext.x * inv_cell)), 1, 512);
        int32_t dy = clampi(sat_i32(ceilf(ext.y * inv_cell)), 1, 512);
        int32_t dz = clampi(sat_i32(ceilf(ext.z * inv_cell)), 1, 512);
        uint64_t ncells = (uint64_t)dx * dy * dz;
        CK(d_gkey.ensure(4ull * N));
        CK(d_gkey2.ensure(4ull * N));
        CK(d_gval.ensure(4ull * N));
        CK(d_gval2.ensure(4ull * N));
        CK(d_gstart.ensure(4ull * (ncells + 1)));
        DevBuf& counts = d_cand_row;  // reuse a transfer scratch array
        CK(counts.ensure(4ull * (ncells + 1)));
        CK(hipMemsetAsync(counts.p, 0, 4ull * (ncells + 1), stream));
        if (N) {
            launch("grid_keys", [&] {
                k_grid_keys<<<blocks_for(N, 256), 256, 0, stream>>>(A.spos, N, scene.lo, inv_cell, dx, dy, dz,
                                                                    d_gkey.as<uint32_t>(), d_gval.as<uint32_t>(),
                                                                    counts.as<uint32_t>());
            });
        }
        if (int rc_s = scan_u32(counts.as<uint32_t>(), d_gstart.as<uint32_t>(), (ncells + 1), err))
            return rc_s;
        if (N) {
            const int kb = (int)bits_for(ncells - 1 ? ncells - 1 : 1);
            if (int rc_s = sort_u32(d_gkey.as<uint32_t>(), d_gval.as<uint32_t>(), d_gkey2.as<uint32_t>(),
                               d_gval2.as<uint32_t>(), N, (uint32_t)kb, err))
                return rc_s;
        }
        A.grid.inv_cell = inv_cell;
        A.grid.dim[0] = dx;
        A.grid.dim[1] = dy;
        A.grid.dim[2] = dz;
        A.grid.starts = d_gstart.as<uint32_t>();
        A.grid.items = d_gval2.as<uint32_t>();
    }
    return 0;
}

// One shading launch: counting runs keep the per-lane traversal; timed runs
// share each query across the wavefront, except for camera rays of a scene
// small enough for the uniform scan when one launch holds kWaveShadeMaxPixels
// or more pixels.
static void launch_shade_group(const ShadeLaunch& a, uint64_t npx, bool count) {
    if (count) launch_shade_count(a);
    else if (a.S.n_obj <= kWaveScanObjs && npx >= kWaveShadeMaxPixels) launch_shade_lane(a);
    else if (a.staged) launch_shade_wave_lds(a);
    else launch_shade_wave_glb(a);
}

int Engine::shade(const HCamera& cam, uint32_t W, uint32_t H, int32_t mode, float* film, uint8_t* rgb8,
                  uint64_t& evals, std::string& err) {
    ShadeArgs A;
    if (int rc = shade_setup(cam, W, H, mode, A, err)) return rc;
    have_beauty_ = false;
    unsigned long long* ctr = d_counter.as<unsigned long long>();
    CK(hipMemsetAsync(ctr, 0, 256, stream));
    const uint64_t npx = (uint64_t)W * H;
    CK(d_film.ensure(16ull * npx));
    CK(d_rgb8.ensure(3ull * npx));
    const SceneView S = dev_view();
    const LightsView Lv = dev_lights();
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    const Lds ld = plan_lds(true);
    launch("shade", [&] {
        launch_shade_group(ShadeLaunch{S, Lv, ld.plan, A, nullptr, d_film.as<float4>(), ctr,
                                       tally() ? ctr + 4 : nullptr, grid, ld.bytes, ld.staged, stream},
                           npx, count_ops_);
    });
    const bool tonemapped = mode_tonemapped(mode);
    if (rgb8) {
        launch("encode", [&] {
            k_encode<<<blocks_for(npx, 256), 256, 0, stream>>>(d_film.as<float4>(), (uint32_t)npx,
                                                               scene.cfg.exposure, tonemapped ? 1 : 0,
                                                               d_srgb.as<float>(), d_rgb8.as<uint8_t>());
        });
    }
    if (int rc = stage_images(npx, rgb8 != nullptr, film != nullptr, err)) return rc;
    unsigned long long cv[32] = {0};
    CK(hipMemcpyAsync(cv, ctr, 256, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    finish_images(npx, rgb8, film);
    drain_events();
    tally_ops("shade", cv, 4);
    evals = cv[0];
    if (mode == M_BEAUTY) {
        have_beauty_ = true;
        beauty_cam_ = cam;
        beauty_w_ = W;
        beauty_h_ = H;
    }
    return 0;
}

CamV Engine::cam_basis(const HCamera& cam, uint32_t W, uint32_t H) const {
    // src/main.rs:199-211, evaluated on the host like the reference.
    CamV c;
    c.origin = cam.pos;
    c.fwd = nrm3(cam.look - cam.pos);
    c.right = nrm3(cross3(c.fwd, cam.up));
    c.up = cross3(c.right, c.fwd);
    const float rads_per_deg = kPi / 180.0f;
    c.tan_half = tanf(cam.fov * rads_per_deg * 0.5f);
    c.aspect = (float)W / (float)H;
    return c;
}

int Engine::shade_views(const std::vector<HCamera>& cams, uint32_t W, uint32_t H, int32_t mode, float* film,
                        uint8_t* rgb8, uint64_t& evals, std::string& err) {
    evals = 0;
    if (cams.empty()) return 0;
    if (cams.size() > 65535) {
        err = "at most 65535 views per call";
        return -1;
    }
    ShadeArgs A;
    if (int rc = shade_setup(cams[0], W, H, mode, A, err)) return rc;
    const uint32_t nv = (uint32_t)cams.size();
    std::vector<CamV> hc(nv);
    for (uint32_t v = 0; v < nv; ++v) hc[v] = cam_basis(cams[v], W, H);
    unsigned long long* ctr = d_counter.as<unsigned long long>();
    CK(hipMemsetAsync(ctr, 0, 256, stream));
    DevBuf& d_cams = d_probe_in;
    CK(d_cams.ensure(sizeof(CamV) * nv));
    CK(hipMemcpyAsync(d_cams.p, hc.data(), sizeof(CamV) * nv, hipMemcpyHostToDevice, stream));
    const uint64_t npx = (uint64_t)W * H * nv;
    CK(d_film.ensure(16ull * npx));
    CK(d_rgb8.ensure(3ull * npx));
    const SceneView S = dev_view();
    const LightsView Lv = dev_lights();
    dim3 grid((W + 15) / 16, (H + 15) / 16, nv);
    const Lds ld = plan_lds(true);
    launch("shade_views", [&] {
        launch_shade_group(ShadeLaunch{S, Lv, ld.plan, A, d_cams.as<CamV>(), d_film.as<float4>(), ctr,
                                       tally() ? ctr + 4 : nullptr, grid, ld.bytes, ld.staged, stream},
                           npx, count_ops_);
    });
    if (rgb8) {
        launch("encode", [&] {
            k_encode<<<blocks_for(npx, 256), 256, 0, stream>>>(d_film.as<float4>(), (uint32_t)npx,
                                                               scene.cfg.exposure, mode_tonemapped(mode) ? 1 : 0,
                                                               d_srgb.as<float>(), d_rgb8.as<uint8_t>());
        });
    }
    if (int rc = stage_images(npx, rgb8 != nullptr, film != nullptr, err)) return rc;
    unsigned long long cv[32] = {0};
    CK(hipMemcpyAsync(cv, ctr, 256, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    finish_images(npx, rgb8, film);
    drain_events();
    tally_ops("shade", cv, 4);
    evals = cv[0];
    return 0;
}

std::vector<HCamera> Engine::turntable(uint32_t n) const {
    // src/main.rs:494-503: orbit the camera about `look` in n equal steps.
    std::vector<HCamera> out(n);
    const HCamera& c = scene.cam;
    const f3 base = c.pos - c.look;
    for (uint32_t i = 0; i < n; ++i) {
        const float a = kTau * (float)i / (float)n;
        const float s = sinf(a), co = cosf(a);
        HCamera v = c;
        v.pos = c.look + mk(base.x * co - base.z * s, base.y, base.x * s + base.z * co);
        out[i] = v;
    }
    return out;
}

// ---------------------------------------------------------------------- probes / downloads
int Engine::probe_dist(const float* pts, uint64_t n, float* d, uint32_t* mat, float* nrm, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    CK(hipSetDevice(device_));
    if (n == 0) return 0;
    CK(d_probe_in.ensure(12ull * n));
    CK(d_probe_out.ensure(20ull * n));
    CK(hipMemcpyAsync(d_probe_in.p, pts, 12ull * n, hipMemcpyHostToDevice, stream));
    float* od = d_probe_out.as<float>();
    uint32_t* om = (uint32_t*)(od + n);
    float* on = nrm ? od + 2 * n : nullptr;
    const SceneView S = dev_view();
    const Lds ld = plan_lds(false);
    launch("probe_dist", [&] {
        launch_probe_dist(ProbeLaunch{S, ld.plan, blocks_for(n, kBlock), ld.bytes, ld.staged, stream},
                          d_probe_in.as<float>(), (uint32_t)n, od, om, on);
    });
    CK(hipMemcpyAsync(d, od, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(mat, om, 4ull * n, hipMemcpyDeviceToHost, stream));
    if (nrm) CK(hipMemcpyAsync(nrm, on, 12ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    drain_events();
    return 0;
}

int Engine::probe_trace(const float* ro, const float* rd, uint64_t n, float tmin, float tmax, float* t, float* nrm,
                        uint32_t* mat, int32_t* hit, uint32_t* steps, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    CK(hipSetDevice(device_));
    if (n == 0) return 0;
    CK(d_probe_in.ensure(24ull * n));
    CK(d_probe_out.ensure(28ull * n));
    float* in = d_probe_in.as<float>();
    CK(hipMemcpyAsync(in, ro, 12ull * n, hipMemcpyHostToDevice, stream));
    CK(hipMemcpyAsync(in + 3 * n, rd, 12ull * n, hipMemcpyHostToDevice, stream));
    float* o = d_probe_out.as<float>();
    float* ot = o;
    float* on = o + n;
    uint32_t* om = (uint32_t*)(o + 4 * n);
    int32_t* oh = (int32_t*)(o + 5 * n);
    uint32_t* os = (uint32_t*)(o + 6 * n);
    const SceneView S = dev_view();
    const Lds ld = plan_lds(false);
    launch("probe_trace", [&] {
        launch_probe_trace(ProbeLaunch{S, ld.plan, blocks_for(n, kBlock), ld.bytes, ld.staged, stream}, in,
                           in + 3 * n, (uint32_t)n, tmin, tmax, ot, on, om, oh, os);
    });
    CK(hipMemcpyAsync(t, ot, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(nrm, on, 12ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(mat, om, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(hit, oh, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(steps, os, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    drain_events();
    return 0;
}

int Engine::probe_shadow(const float* ro, const float* rd, uint64_t n, float tmin, const float* tmax, const float* k,
                         float* vis, uint32_t* steps, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    CK(hipSetDevice(device_));
    if (n == 0) return 0;
    CK(d_probe_in.ensure(32ull * n));
    CK(d_probe_out.ensure(8ull * n));
    float* in = d_probe_in.as<float>();
    CK(hipMemcpyAsync(in, ro, 12ull * n, hipMemcpyHostToDevice, stream));
    CK(hipMemcpyAsync(in + 3 * n, rd, 12ull * n, hipMemcpyHostToDevice, stream));
    CK(hipMemcpyAsync(in + 6 * n, tmax, 4ull * n, hipMemcpyHostToDevice, stream));
    CK(hipMemcpyAsync(in + 7 * n, k, 4ull * n, hipMemcpyHostToDevice, stream));
    float* ov = d_probe_out.as<float>();
    uint32_t* os = (uint32_t*)(ov + n);
    const SceneView S = dev_view();
    const Lds ld = plan_lds(false);
    launch("probe_shadow", [&] {
        launch_probe_shadow(ProbeLaunch{S, ld.plan, blocks_for(n, kBlock), ld.bytes, ld.staged, stream}, in,
                            in + 3 * n, (uint32_t)n, tmin, in + 6 * n, in + 7 * n, ov, os);
    });
    CK(hipMemcpyAsync(vis, ov, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipMemcpyAsync(steps, os, 4ull * n, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    drain_events();
    return 0;
}

int Engine::download_transfer(uint32_t* starts, uint32_t* idx, float* w, float* sky, float* row_sum,
                              std::string& err) {
    if (!have_transfer) {
        err = "no transfer matrix";
        return -3;
    }
    const uint64_t N = surfels.size();
    if (N == 0) {
        if (starts) starts[0] = 0;
        return 0;
    }
    if (starts) CK(hipMemcpyAsync(starts, d_lstart.p, 4 * (N + 1), hipMemcpyDeviceToHost, stream));
    if (idx && n_links) CK(hipMemcpyAsync(idx, d_idx.p, 4 * n_links, hipMemcpyDeviceToHost, stream));
    if (w && n_links) CK(hipMemcpyAsync(w, d_w.p, 4 * n_links, hipMemcpyDeviceToHost, stream));
    if (sky) CK(hipMemcpyAsync(sky, d_sky.p, 4 * N, hipMemcpyDeviceToHost, stream));
    if (row_sum) CK(hipMemcpyAsync(row_sum, d_rowsum.p, 4 * N, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    return 0;
}

int Engine::download_clusters(float* p, float* n, float* area, float* radius, uint32_t* mstart, uint32_t* members,
                              std::string& err) {
    if (!have_transfer) {
        err = "no clusters";
        return -3;
    }
    const uint32_t C = n_clusters;
    const uint64_t N = surfels.size();
    std::vector<float4> hp(C), hn(C);
    std::vector<uint2> hr(C);
    std::vector<uint32_t> mem(N);
    if (C) {
        CK(hipMemcpyAsync(hp.data(), d_cpos.p, 16ull * C, hipMemcpyDeviceToHost, stream));
        CK(hipMemcpyAsync(hn.data(), d_cnrm.p, 16ull * C, hipMemcpyDeviceToHost, stream));
        CK(hipMemcpyAsync(hr.data(), d_crange.p, 8ull * C, hipMemcpyDeviceToHost, stream));
        CK(hipMemcpyAsync(mem.data(), d_cval2.p, 4ull * N, hipMemcpyDeviceToHost, stream));
    }
    CK(hipStreamSynchronize(stream));
    uint32_t off = 0;
    for (uint32_t c = 0; c < C; ++c) {
        if (p) p[3 * c] = hp[c].x, p[3 * c + 1] = hp[c].y, p[3 * c + 2] = hp[c].z;
        if (n) n[3 * c] = hn[c].x, n[3 * c + 1] = hn[c].y, n[3 * c + 2] = hn[c].z;
        if (area) area[c] = hp[c].w;
        if (radius) radius[c] = hn[c].w;
        if (mstart) mstart[c] = off;
        for (uint32_t k = hr[c].x; k < hr[c].y; ++k) {
            if (members) members[off] = mem[k];
            ++off;
        }
    }
    if (mstart) mstart[C] = off;
    return 0;
}

int Engine::download_radiosity(float* direct, float* irr, float* rad, std::string& err) {
    if (!have_solve) {
        err = "no solve";
        return -3;
    }
    const uint64_t N = surfels.size();
    if (!N) return 0;
    if (int rc = copy3_d2h(direct, d_direct.p, N, err)) return rc;
    CK(hipStreamSynchronize(stream));
    if (int rc = copy3_d2h(irr, d_irr.p, N, err)) return rc;
    CK(hipStreamSynchronize(stream));
    if (int rc = copy3_d2h(rad, d_rad.p, N, err)) return rc;
    CK(hipStreamSynchronize(stream));
    return 0;
}

int Engine::upload_radiosity(const float* direct, const float* irr, const float* rad, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    const uint64_t N = surfels.size();
    CK(d_direct.ensure(16ull * N));
    CK(d_irr.ensure(16ull * N));
    CK(d_rad.ensure(16ull * N));
    if (N) {
        CK(hipMemsetAsync(d_direct.p, 0, 16ull * N, stream));
        CK(hipMemsetAsync(d_irr.p, 0, 16ull * N, stream));
        CK(hipMemsetAsync(d_rad.p, 0, 16ull * N, stream));
        if (int rc = copy3_h2d(d_direct.p, direct, N, err)) return rc;
        CK(hipStreamSynchronize(stream));
        if (int rc = copy3_h2d(d_irr.p, irr, N, err)) return rc;
        CK(hipStreamSynchronize(stream));
        if (int rc = copy3_h2d(d_rad.p, rad, N, err)) return rc;
    }
    CK(hipStreamSynchronize(stream));
    have_solve = true;
    return 0;
}

}  // namespace ifx
