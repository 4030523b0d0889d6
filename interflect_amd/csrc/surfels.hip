// GPU surfel placement (src/surfel.rs:147-255) with the host path's exact
// result.  The reference accepts Halton seeds strictly in order: a seed is
// kept when no earlier kept seed lies in its 27-cell Poisson neighbourhood
// closer than d_min.  That is the lexicographically first independent set
// of the conflict graph over the projected seeds, which is resolved here in
// parallel rounds:
//   1. every seed is projected onto the surface (one lane per seed);
//   2. the surviving seeds, in seed order, are bucketed by Poisson cell;
//   3. each survivor's earlier conflicting neighbours are listed once;
//   4. each round, an undecided seed becomes rejected if an earlier
//      neighbour is accepted, accepted if every earlier neighbour is
//      rejected, and waits otherwise (rounds read the previous round's
//      states, so the result does not depend on scheduling);
//   5. the accepted seeds are compacted in seed order and cut at the target,
//      which is where the reference's loop stops.
#include "ifx_kern.h"
#include "ifx_surfel.h"

namespace ifx {
namespace {

enum : uint8_t { S_OPEN = 0, S_KEEP = 1, S_DROP = 2 };

// The placement kernels leave lanes at different points of their loops (a
// seed's Newton steps stop when it converges; rejected seeds skip the
// output), so they query the scene with the per-lane traversal (LaneStack):
// the wavefront-shared query is for call sites where all lanes arrive
// together.
template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_seed_project(SceneView Sg, LightsView Lg, LdsPlan P, SurfelParams SP,
                                                         uint32_t b0, uint32_t nb, float4* pos, uint32_t* ok) {
    IFX_SETUP_T(LDS, false, LaneStack);
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nb) return;
    f3 p;
    const bool v = project_seed(S, SP, b0 + k, p, stk);
    pos[k] = st4(p);
    ok[k] = v ? 1u : 0u;
}

// Survivors in seed order: index into the seed batch, position, Poisson cell.
__global__ void k_seed_compact(uint32_t nb, const uint32_t* __restrict__ ok, const uint32_t* __restrict__ at,
                               const float4* __restrict__ pos, SurfelParams SP, uint32_t base, float4* vpos,
                               uint32_t* vcell, uint32_t* counts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb || !ok[k]) return;
    const uint32_t j = base + at[k];
    const float4 p = pos[k];
    vpos[j] = p;
    int32_t c[3];
    surfel_cell(SP, ld3(p), c);
    const uint32_t cell = (uint32_t)(c[0] + SP.dim[0] * (c[1] + SP.dim[1] * c[2]));
    vcell[j] = cell;
    atomicAdd(&counts[cell], 1u);
}

__global__ void k_iota(uint32_t* v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// The conflict graph, once: each survivor's earlier neighbours (lower seed
// index, closer than d_min) from its 27-cell neighbourhood.  `items` lists
// the survivors by cell (seed order within a cell); `starts` indexes it by
// cell.  FILL=false counts them, FILL=true writes them at adj_start[k].
// Threads take the survivors in cell order (t -> items[t]), so a wavefront's
// lanes scan the same neighbour cells; `cpos` holds the positions in that
// order.
__global__ void k_cell_pos(uint32_t nv, const uint32_t* __restrict__ items, const float4* __restrict__ vpos,
                           float4* cpos) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nv) cpos[q] = vpos[items[q]];
}

template <bool FILL>
__global__ void k_earlier_nbrs(uint32_t nv, const float4* __restrict__ cpos, const uint32_t* __restrict__ vcell,
                               const uint32_t* __restrict__ starts, const uint32_t* __restrict__ items,
                               SurfelParams SP, const uint32_t* __restrict__ adj_start, uint32_t* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const uint32_t k = items[t];
    const f3 p = ld3(cpos[t]);
    const uint32_t cell = vcell[k];
    const int32_t dx = SP.dim[0], dy = SP.dim[1], dz = SP.dim[2];
    const int32_t cx = (int32_t)(cell % (uint32_t)dx), cy = (int32_t)((cell / (uint32_t)dx) % (uint32_t)dy),
                  cz = (int32_t)(cell / ((uint32_t)dx * (uint32_t)dy));
    uint32_t n = 0;
    uint32_t* dst = FILL ? out + adj_start[k] : nullptr;
    for (int32_t z = cz - 1 > 0 ? cz - 1 : 0; z <= (cz + 1 < dz - 1 ? cz + 1 : dz - 1); ++z)
        for (int32_t y = cy - 1 > 0 ? cy - 1 : 0; y <= (cy + 1 < dy - 1 ? cy + 1 : dy - 1); ++y)
            for (int32_t x = cx - 1 > 0 ? cx - 1 : 0; x <= (cx + 1 < dx - 1 ? cx + 1 : dx - 1); ++x) {
                const uint32_t c = (uint32_t)(x + dx * (y + dy * z));
                for (uint32_t q = starts[c], e = starts[c + 1]; q < e; ++q) {
                    const uint32_t j = items[q];
                    if (j >= k) break;  // items are in seed order within a cell
                    const f3 dq = ld3(cpos[q]) - p;  // earlier surfel minus candidate
                    if (!(dot3(dq, dq) < SP.d_min_sq)) continue;
                    if (FILL) dst[n] = j;
                    ++n;
                }
            }
    if (!FILL) out[k] = n;
}

// One round over the survivors: an undecided seed is rejected when an
// earlier neighbour is accepted, accepted when all are rejected, and waits
// otherwise.
__global__ void k_mis_round(uint32_t nv, const uint32_t* __restrict__ adj_start, const uint32_t* __restrict__ adj,
                            const uint8_t* __restrict__ st_in, uint8_t* st_out, uint32_t* open) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv) return;
    uint8_t s = st_in[k];
    if (s == S_OPEN) {
        bool wait = false, drop = false;
        for (uint32_t q = adj_start[k], e = adj_start[k + 1]; q < e; ++q) {
            const uint8_t t = st_in[adj[q]];
            if (t == S_KEEP) {
                drop = true;
                break;
            }
            if (t == S_OPEN) wait = true;
        }
        s = drop ? S_DROP : (wait ? S_OPEN : S_KEEP);
        if (s == S_OPEN) atomicAdd(open, 1u);
    }
    st_out[k] = s;
}

__global__ void k_keep_flags(uint32_t nv, const uint8_t* __restrict__ st, uint32_t* flag) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nv) flag[k] = st[k] == S_KEEP ? 1u : 0u;
}

// Accepted surfels in seed order, cut at the target: position, normal,
// material, albedo (kd of the material), area.
template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_surfel_out(SceneView Sg, LightsView Lg, LdsPlan P, uint32_t nv,
                                                       const uint32_t* __restrict__ flag,
                                                       const uint32_t* __restrict__ at,
                                                       const float4* __restrict__ vpos, uint32_t target, float area,
                                                       float4* out_p, float4* out_n, float4* out_a,
                                                       uint32_t* out_m) {
    IFX_SETUP_T(LDS, true, LaneStack);
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nv || !flag[k]) return;
    const uint32_t j = at[k];
    if (j >= target) return;
    const f3 p = ld3(vpos[k]);
    const uint32_t mat = dist_mat(S, p, stk).mat;
    const f3 n = scene_normal(S, p, stk);
    out_p[j] = st4(p, area);
    out_n[j] = st4(n);
    out_a[j] = st4(load_mat(L, mat).kd);
    out_m[j] = mat;
}

__global__ void k_fill_u8(uint8_t* p, uint32_t n, uint8_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

int Engine::generate_surfels_gpu(uint64_t target, std::vector<HSurfel>& out, std::string& err) {
    if (int rc0 = need_device(err)) return rc0;
    out.clear();
    if (target == 0) return 0;
    CK(hipSetDevice(device_));
    const SurfelParams SP = surfel_params(scene.lo, scene.hi, target);
    const uint64_t ncells = (uint64_t)SP.dim[0] * SP.dim[1] * SP.dim[2];
    const uint32_t seeds = SP.seeds;
    const SceneView S = dev_view();
    const LightsView Lv = dev_lights();
    // 1. project every seed; survivors in seed order
    DevBuf d_pos, d_ok, d_at, d_vpos, d_vcell, d_counts, d_starts, d_keys2, d_items, d_items_in, d_st0, d_st1,
        d_open1, d_flag, d_at2, d_op, d_on, d_oa, d_om, d_adjn, d_adjs, d_adj, d_cpos;
    CK(d_pos.ensure(16ull * seeds));
    CK(d_ok.ensure(4ull * (seeds + 1)));
    CK(d_at.ensure(4ull * (seeds + 1)));
    CK(d_vpos.ensure(16ull * seeds));
    CK(d_vcell.ensure(4ull * seeds));
    CK(d_counts.ensure(4ull * (ncells + 1)));
    CK(d_starts.ensure(4ull * (ncells + 1)));
    CK(hipMemsetAsync(d_counts.p, 0, 4ull * (ncells + 1), stream));
    CK(hipMemsetAsync(d_ok.as<uint32_t>() + seeds, 0, 4, stream));
    {
        const Lds ld = plan_lds(false);
        const LightsView Lg{};
        launch("surfel_project", [&] {
            LAUNCH_L(k_seed_project, ld.staged, blocks_for(seeds, kBlock), ld.bytes, S, Lg, ld.plan, SP, 0u, seeds,
                     d_pos.as<float4>(), d_ok.as<uint32_t>());
        });
    }
    if (int rc_s = scan_u32(d_ok.as<uint32_t>(), d_at.as<uint32_t>(), (seeds + 1), err))
        return rc_s;
    uint32_t nv = 0;
    CK(hipMemcpyAsync(&nv, d_at.as<uint32_t>() + seeds, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    if (nv == 0) return 0;
    launch("surfel_cells", [&] {
        k_seed_compact<<<blocks_for(seeds, 256), 256, 0, stream>>>(seeds, d_ok.as<uint32_t>(), d_at.as<uint32_t>(),
                                                                   d_pos.as<float4>(), SP, 0u, d_vpos.as<float4>(),
                                                                   d_vcell.as<uint32_t>(), d_counts.as<uint32_t>());
    });
    // 2. bucket the survivors by cell (stable: seed order within a cell)
    if (int rc_s = scan_u32(d_counts.as<uint32_t>(), d_starts.as<uint32_t>(), (ncells + 1), err))
        return rc_s;
    CK(d_keys2.ensure(4ull * nv));
    CK(d_items.ensure(4ull * nv));
    CK(d_items_in.ensure(4ull * nv));
    k_iota<<<blocks_for(nv, 256), 256, 0, stream>>>(d_items_in.as<uint32_t>(), nv);
    const int kb = (int)bits_for(ncells ? ncells - 1 : 1);
    if (int rc_s = sort_u32(d_vcell.as<uint32_t>(), d_items_in.as<uint32_t>(), d_keys2.as<uint32_t>(),
                       d_items.as<uint32_t>(), nv, (uint32_t)kb, err))
        return rc_s;
    // 3. the conflict graph (earlier neighbours of each survivor)
    CK(d_adjn.ensure(4ull * (nv + 1)));
    CK(d_adjs.ensure(4ull * (nv + 1)));
    CK(d_cpos.ensure(16ull * nv));
    CK(hipMemsetAsync(d_adjn.as<uint32_t>() + nv, 0, 4, stream));
    launch("surfel_graph", [&] {
        k_cell_pos<<<blocks_for(nv, 256), 256, 0, stream>>>(nv, d_items.as<uint32_t>(), d_vpos.as<float4>(),
                                                             d_cpos.as<float4>());
        k_earlier_nbrs<false><<<blocks_for(nv, 256), 256, 0, stream>>>(
            nv, d_cpos.as<float4>(), d_vcell.as<uint32_t>(), d_starts.as<uint32_t>(), d_items.as<uint32_t>(), SP,
            nullptr, d_adjn.as<uint32_t>());
    });
    if (int rc_s = scan_u32(d_adjn.as<uint32_t>(), d_adjs.as<uint32_t>(), (nv + 1), err))
        return rc_s;
    uint32_t edges = 0;
    CK(hipMemcpyAsync(&edges, d_adjs.as<uint32_t>() + nv, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    if (edges >= (1u << 31)) {
        err = "surfel conflict graph exceeds 2^31 edges";
        return -6;
    }
    CK(d_adj.ensure(4ull * edges + 4));
    launch("surfel_graph", [&] {
        k_earlier_nbrs<true><<<blocks_for(nv, 256), 256, 0, stream>>>(
            nv, d_cpos.as<float4>(), d_vcell.as<uint32_t>(), d_starts.as<uint32_t>(), d_items.as<uint32_t>(), SP,
            d_adjs.as<uint32_t>(), d_adj.as<uint32_t>());
    });
    // 4. rounds until no survivor is undecided
    CK(d_st0.ensure(nv));
    CK(d_st1.ensure(nv));
    CK(d_open1.ensure(4ull * 64));
    k_fill_u8<<<blocks_for(nv, 256), 256, 0, stream>>>(d_st0.as<uint8_t>(), nv, S_OPEN);
    uint8_t* st_in = d_st0.as<uint8_t>();
    uint8_t* st_out = d_st1.as<uint8_t>();
    uint32_t rounds = 0;
    for (;;) {
        // 8 rounds per check; open[r] counts the seeds still undecided after round r
        CK(hipMemsetAsync(d_open1.p, 0, 4ull * 8, stream));
        launch("surfel_rounds", [&] {
            for (int r = 0; r < 8; ++r) {
                k_mis_round<<<blocks_for(nv, 256), 256, 0, stream>>>(nv, d_adjs.as<uint32_t>(), d_adj.as<uint32_t>(),
                                                                     st_in, st_out, d_open1.as<uint32_t>() + r);
                std::swap(st_in, st_out);
            }
        });
        uint32_t open[8] = {0};
        CK(hipMemcpyAsync(open, d_open1.p, 4ull * 8, hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        rounds += 8;
        if (open[7] == 0) break;
        // every round settles at least the lowest undecided seed (its earlier
        // neighbours are all decided), so nv rounds always suffice
        if (rounds > nv + 8u) {
            err = "surfel acceptance did not settle within " + std::to_string(nv) + " rounds";
            return -4;
        }
    }
    // 5. accepted in seed order, cut at the target
    CK(d_flag.ensure(4ull * (nv + 1)));
    CK(d_at2.ensure(4ull * (nv + 1)));
    CK(hipMemsetAsync(d_flag.as<uint32_t>() + nv, 0, 4, stream));
    k_keep_flags<<<blocks_for(nv, 256), 256, 0, stream>>>(nv, st_in, d_flag.as<uint32_t>());
    if (int rc_s = scan_u32(d_flag.as<uint32_t>(), d_at2.as<uint32_t>(), (nv + 1), err))
        return rc_s;
    uint32_t kept = 0;
    CK(hipMemcpyAsync(&kept, d_at2.as<uint32_t>() + nv, 4, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    const uint32_t n = (uint32_t)std::min<uint64_t>(kept, target);
    CK(d_op.ensure(16ull * n + 16));
    CK(d_on.ensure(16ull * n + 16));
    CK(d_oa.ensure(16ull * n + 16));
    CK(d_om.ensure(4ull * n + 4));
    {
        const Lds ld = plan_lds(true);
        launch("surfel_out", [&] {
            LAUNCH_L(k_surfel_out, ld.staged, blocks_for(nv, kBlock), ld.bytes, S, Lv, ld.plan, nv,
                     d_flag.as<uint32_t>(), d_at2.as<uint32_t>(), d_vpos.as<float4>(), n, SP.area,
                     d_op.as<float4>(), d_on.as<float4>(), d_oa.as<float4>(), d_om.as<uint32_t>());
        });
    }
    std::vector<float4> hp(n), hn(n), ha(n);
    std::vector<uint32_t> hm(n);
    if (n) {
        CK(hipMemcpyAsync(hp.data(), d_op.p, 16ull * n, hipMemcpyDeviceToHost, stream));
        CK(hipMemcpyAsync(hn.data(), d_on.p, 16ull * n, hipMemcpyDeviceToHost, stream));
        CK(hipMemcpyAsync(ha.data(), d_oa.p, 16ull * n, hipMemcpyDeviceToHost, stream));
        CK(hipMemcpyAsync(hm.data(), d_om.p, 4ull * n, hipMemcpyDeviceToHost, stream));
    }
    CK(hipStreamSynchronize(stream));
    drain_events();
    out.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        out[i].p = ld3(hp[i]);
        out[i].area = hp[i].w;
        out[i].n = ld3(hn[i]);
        out[i].albedo = ld3(ha[i]);
        out[i].mat = hm[i];
    }
    last_place_rounds_ = rounds;
    return 0;
}

}  // namespace ifx
