// Device engine behind the C ABI: owns the HIP stream and device buffers of
// one scene and runs the transfer -> solve -> shade stages on gfx950.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "host_scene.h"
#include "host_surfels.h"

namespace ifx {

// Idle lanes of a persistent march kernel (link visibility, closure rays)
// take new work only when this many lanes of their wavefront are idle: 64
// refills the whole wavefront at once.
constexpr uint32_t kMarchMinIdle = 64;
// Link-visibility batches park their rays once this many or fewer still
// march (tuning key vis_tail), in transfers of at least kVisTailItems
// candidates (vis_tail_items): below that the second launch and its index
// cost more than the idle lanes (product scene, 1.4M candidates: 1.32 ->
// 1.61 ms; stress scene, 68M: 76.2 -> 74.6 ms).
constexpr uint32_t kVisTail = 8;
constexpr uint32_t kVisTailItems = 1u << 23;
// Wavefronts per SIMD the persistent march kernels are compiled for; with 4
// wavefronts per block, also their resident blocks per CU.
constexpr uint32_t kMarchWaves = 7;

struct LdsPlan {
    uint32_t scene_f4, light_f4;  // float4s of scene / light records staged in LDS (0 = global)
    uint32_t stack_depth;         // traversal-stack entries per lane
};

constexpr int kBlock = 256;       // threads per block of every traversal kernel
constexpr int kStackCap = 64;     // deepest traversal stack the kernels accept (the reference's bound)

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const {
        return (T*)p;
    }
    ~DevBuf() { release(); }
};

// Pinned host staging for read-backs of images: a copy into pageable memory
// goes through the runtime's own bounce buffers piece by piece.
struct PinBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    ~PinBuf() { release(); }
};

struct TransferOut {
    uint32_t clusters = 0;
    uint64_t candidates = 0, links = 0, evals = 0;
    float mean_row_sum = 0;
};

struct SolveOut {
    uint32_t iterations = 0;
    float residual = 0;
    uint64_t evals = 0;
};

struct BackwardOut {
    uint32_t iterations = 0;  // adjoint Jacobi sweeps
    float residual = 0;       // last max |delta mu| / max |mu|
    uint32_t link_tiles = 0;  // row tiles recomputed every sweep, 0 = links kept
};

struct CamV {
    f3 origin, fwd, right, up;
    float tan_half, aspect;
};

// Arguments of one render_frame pass (src/main.rs:229-367).
struct ShadeArgs {
    CamV cam;
    uint32_t W, H;
    int32_t mode;
    int32_t use_gi;
    float tmax, radius;
    uint32_t spp;  // paths per pixel of the reference mode
    GridView grid;
    const float4* spos;
    const float4* snrm;
    const float4* irr;
};

enum { M_BEAUTY = 0, M_DIRECT = 1, M_INDIRECT = 2, M_NORMALS = 3, M_DEPTH = 4, M_STEPS = 5, M_REFERENCE = 6 };

// Modes encoded with exposure + ACES (src/main.rs:46-50).
inline bool mode_tonemapped(int32_t m) { return m <= M_INDIRECT || m == M_REFERENCE; }

struct KernelStat {
    uint64_t launches = 0;
    double ms = 0;
};

class Engine {
   public:
    Engine() = default;
    ~Engine();
    int init(int device, std::string& err);

    int load_scene(const std::string& src, std::string& err);
    // A finalized host scene (parsed here or given by the caller) to the device.
    int install_scene(HostScene&& s, std::string& err);
    int set_surfels(const std::vector<HSurfel>& s, std::string& err);
    // surfel::generate on the device, equal to generate_surfels() on the host.
    int generate_surfels_gpu(uint64_t target, std::vector<HSurfel>& out, std::string& err);
    uint32_t place_rounds() const { return last_place_rounds_; }
    int transfer(int32_t res, TransferOut& out, std::string& err);
    int solve(uint32_t bounces, SolveOut& out, std::string& err);
    // solve() into caller-owned device arrays (N float4 records each) with
    // every scratch buffer (direct-light terms, cluster radiosity, residuals,
    // counters) in the library's static device array: nothing is allocated.
    // Calls that overlap on one device race on that array.
    int solve_noalloc(uint32_t bounces, float4* direct, float4* irr, float4* rad, SolveOut& out, std::string& err);
    static constexpr size_t kScratchBytes = 32ull << 20;  // the static array, per device
    // Many cameras over one solve, one launch (film / rgb8: views x h x w x 3).
    int shade_views(const std::vector<HCamera>& cams, uint32_t w, uint32_t h, int32_t mode, float* film,
                    uint8_t* rgb8, uint64_t& evals, std::string& err);
    std::vector<HCamera> turntable(uint32_t n) const;
    int shade(const HCamera& cam, uint32_t w, uint32_t h, int32_t mode, float* film, uint8_t* rgb8, uint64_t& evals,
              std::string& err);

    // Kernel-family probes (one launch each).
    int probe_dist(const float* pts, uint64_t n, float* d, uint32_t* mat, float* nrm, std::string& err);
    int probe_trace(const float* ro, const float* rd, uint64_t n, float tmin, float tmax, float* t, float* nrm,
                    uint32_t* mat, int32_t* hit, uint32_t* steps, std::string& err);
    int probe_shadow(const float* ro, const float* rd, uint64_t n, float tmin, const float* tmax, const float* k,
                     float* vis, uint32_t* steps, std::string& err);

    // Adjoint of the frozen-geometry pipeline (surfel albedo, light emission).
    int set_material_albedo(uint32_t m, const float* rgb, std::string& err);
    int set_light_emit(uint32_t li, const float* rgb, std::string& err);
    int backward(const float* dL_dfilm, float* d_albedo, float* d_emit, BackwardOut& out, std::string& err);
    int backward_surfels(const float* dL_dirr, float* d_albedo, float* d_emit, float* mu, BackwardOut& out,
                         std::string& err);
    int download_direct_geometry(float* g, std::string& err);

    int download_transfer(uint32_t* starts, uint32_t* idx, float* w, float* sky, float* row_sum, std::string& err);
    int download_clusters(float* p, float* n, float* area, float* radius, uint32_t* mstart, uint32_t* members,
                          std::string& err);
    int download_radiosity(float* direct, float* irr, float* rad, std::string& err);
    int upload_radiosity(const float* direct, const float* irr, const float* rad, std::string& err);

    void set_profiling(bool on) { prof_ = on; }
    // Scene-independent device tables and device facts (shared with the batch engine).
    const float4* closure_dirs() const { return d_dirs.as<float4>(); }
    const float* srgb_thresholds() const { return d_srgb.as<float>(); }
    int n_cu() const { return n_cu_; }
    int device() const { return device_; }
    bool has_device() const { return has_device_; }
    // Instrumented launches of the tracing kernels: 1 = the counting
    // instantiation (algorithmic FP32 operations and primitive evaluations of
    // the reference traversal), 2 = the timed kernels tally the primitive
    // evaluations they execute; 0 = off.
    void set_counting(int mode) {
        count_ops_ = mode == 1;
        count_exec_ = mode == 2;
    }
    const std::map<std::string, uint64_t>& kernel_ops() const { return kops_; }
    const std::map<std::string, KernelStat>& kernel_stats() const { return kstats_; }
    void reset_kernel_stats() {
        kstats_.clear();
        kops_.clear();
    }

    HostScene scene;
    std::vector<HSurfel> surfels;
    bool use_bvh = true;
    bool have_transfer = false, have_solve = false;
    uint32_t n_clusters = 0;
    uint64_t n_links = 0;
    hipStream_t stream = nullptr;

    bool use_lds = true;  // stage scene records in LDS when they fit
    uint32_t vis_blocks = 0, ff_blocks = 0, jacobi_blocks = 0;  // 0 = automatic
    uint32_t vis_chunk = 0;     // link-visibility queue chunk per wavefront (0 = automatic)
    uint32_t vis_min_idle = kMarchMinIdle;  // link visibility: idle lanes refill once this many are idle
    uint32_t closure_min_idle = kMarchMinIdle;  // closure rays: the same for k_closure
    // Link visibility: a lockstep batch parks its rays once at most this
    // many still march, and a second launch marches the parked rays (0: off;
    // needs vis_min_idle 64 and at least vis_tail_items candidates).
    uint32_t vis_tail = kVisTail;
    uint32_t vis_tail_items = kVisTailItems;
    // Wavefront-shared queries of marches with a distance bound test each
    // object's box against the wavefront's points instead of walking the BVH.
    bool flat_query = true;
    uint32_t jacobi_rows = 0;  // 0 automatic, 1 lane per row on the CSR, 2 wavefront per row, 3 sliced rows
    float solve_tol = 1e-4f;   // Jacobi stops once max |delta| < solve_tol (src/solve.rs:98)
    bool place_on_gpu = true;  // ifx_generate_surfels on the device when there is one
    // transfer() launches the direct-light terms and the closure rays on a
    // second stream, beside the cluster build and the candidate culls and
    // sort; the link visibility waits for them.
    bool overlap_direct = true;
    uint32_t spp = 512;        // paths per pixel of the reference mode (--spp, src/main.rs:81)
    uint32_t adj_max_iters = 1000;  // adjoint sweeps at most
    float adj_tol = 1e-6f;          // adjoint stops once max |delta mu| <= adj_tol * max |mu|
    uint32_t adj_seg = 1024;        // transposed-link entries per partial sum (fixes the summation order)
    uint32_t adj_tile_rows = 4096;  // surfel rows per link tile (segments never span two tiles)
    // Bytes the backward pass may hold for its own copy of the links (0 = no
    // limit).  Beyond it, every adjoint sweep recomputes the links tile by
    // tile from the geometry instead of keeping a transposed copy.
    uint64_t link_budget = 0;
    // The backward pass's link layout depends on adj_seg, adj_tile_rows and
    // link_budget: call after changing them.
    void invalidate_adjoint() { adj_gen_ = ~0ull; }

   private:
    struct Lds {
        LdsPlan plan;
        size_t bytes;
        bool staged;
    };
    // stage = false: records stay in global memory (kernels whose traversal is
    // wavefront-shared read them through the scalar cache); LDS holds the stacks.
    Lds plan_lds(bool lights, bool stage = true) const;
    int shade_setup(const HCamera& cam, uint32_t W, uint32_t H, int32_t mode, ShadeArgs& A, std::string& err);
    CamV cam_basis(const HCamera& cam, uint32_t W, uint32_t H) const;
    int adjoint_prepare(std::string& err);
    // Links of rows r0 .. r0 + nr recomputed from the geometry with the
    // transfer's kernels: the tile-local CSR d_tlstart / d_tidx / d_tw.
    // sized: the tile buffers already hold this tile (M candidates, L links);
    // else M and L are read back and the buffers grown.
    int tile_links(uint32_t r0, uint32_t nr, uint32_t& M, uint32_t& L, bool sized, std::string& err);
    // The tile's links transposed (cluster-major, rows ascending), with
    // column starts d_tcolstart.
    int tile_transpose(uint32_t r0, uint32_t nr, uint32_t L, std::string& err);
    void release_link_tiles();
    int build_sliced(std::string& err);
    int adjoint_solve(BackwardOut& out, std::string& err);
    int adjoint_params(bool pixels, uint64_t npx, float* d_albedo, float* d_emit, std::string& err);
    int check_backward(std::string& err) const;
    // image read-back: enqueue into h_img_, then copy out once the stream is synchronised
    int stage_images(uint64_t npx, bool rgb8, bool film, std::string& err);
    void finish_images(uint64_t npx, uint8_t* rgb8, float* film) const;
    PinBuf h_img_;
    int copy3_d2h(float* host, const void* dev, uint64_t n, std::string& err);
    int copy3_h2d(void* dev, const float* host, uint64_t n, std::string& err);
    void upload_materials();
    void upload_lights();
    int check(hipError_t e, const char* what, std::string& err);
    // Stable radix sort of (key, u32 value) pairs by the low `bits` key bits
    // into the output arrays (inputs untouched), and an exclusive u32 scan,
    // on `stream`: kernels only (ifx_sort.h), scratch in d_tmp / d_sort_*.
    int sort_u32(const uint32_t* k, const uint32_t* v, uint32_t* ko, uint32_t* vo, uint64_t n, uint32_t bits,
                 std::string& err);
    int sort_u64(const uint64_t* k, const uint32_t* v, uint64_t* ko, uint32_t* vo, uint64_t n, uint32_t bits,
                 std::string& err);
    int scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, std::string& err);
    int need_device(std::string& err) const;
    SceneView dev_view() const;
    LightsView dev_lights() const;
    template <class F>
    void launch(const char* name, F&& f) {
        launch_on(stream, name, static_cast<F&&>(f));
    }
    template <class F>
    void launch_on(hipStream_t s, const char* name, F&& f);
    void drain_events();
    // closure and direct-light terms launched by transfer() on side_ (see overlap_direct)
    int launch_direct(hipStream_t s, unsigned long long* evals, unsigned long long* ops, std::string& err);
    void join_side();  // the main stream waits for side_'s work
    hipStream_t side_ = nullptr;
    hipEvent_t fork_ = nullptr, join_ = nullptr, closed_ = nullptr;  // closed_: the closure kernel is done
    hipEvent_t vis_done_ = nullptr;  // the link visibility launch is done (deferred direct-light terms wait)
    bool direct_pre_ = false;    // terms from the last transfer are pending on side_ and still valid
    bool side_pending_ = false;  // side_ holds work the main stream has not waited for
    uint64_t last_candidates_ = 0;  // candidates of the last transfer (schedules the side stream)
    DevBuf d_dctr;

    unsigned char* scratch_ = nullptr;  // device address of the static scratch array on device_
    int device_ = 0;
    int n_cu_ = 1;
    bool has_device_ = false;
    bool prof_ = false;
    bool count_ops_ = false;
    bool count_exec_ = false;
    bool tally() const { return count_ops_ || count_exec_; }
    int tally_mode() const { return count_ops_ ? 1 : (count_exec_ ? 2 : 0); }
    // kops_[name] (+ ".prims", ".exec_prims") from a kernel's counter slot
    void tally_ops(const char* name, const unsigned long long* cv, int slot);
    std::map<std::string, uint64_t> kops_;
    std::map<std::string, KernelStat> kstats_;
    struct Pending {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> event_pool_;

    // scene
    DevBuf d_obj, d_planes, d_subs, d_nodes, d_obox, d_lin, d_quad, d_sph, d_pnt, d_mat, d_dirs, d_srgb;
    // surfels (float4: pos.xyz+area, nrm, albedo)
    DevBuf d_spos, d_snrm, d_salb;
    bool surfels_dirty = true;
    // clusters
    DevBuf d_ckey, d_ckey2, d_cval, d_cval2, d_cflag, d_cseg, d_cpos, d_cnrm, d_crange, d_cvalid, d_cidx;
    DevBuf d_craw_pos, d_craw_nrm;
    // work order (Morton rows, cluster-major candidates) and queues
    DevBuf d_mpos, d_mnrm, d_mrank;  // surfel positions / normals in Morton order, each surfel's rank
    DevBuf d_morton, d_mval, d_mkey_s, d_rorder, d_vkey, d_vval, d_vkey_s, d_vorder, d_queue, d_open;
    DevBuf d_mcount, d_mstart;  // candidate counts / offsets of the rows in Morton order
    // candidates and links
    DevBuf d_ccount, d_cstart, d_cand_ci, d_cand_row, d_cand_v, d_sky, d_lcount, d_lstart, d_scale, d_rowsum, d_idx, d_w;
    // solve
    DevBuf d_direct, d_irr, d_rad, d_bc, d_resid, d_terms, d_gterm;
    // shade
    DevBuf d_gkey, d_gkey2, d_gval, d_gval2, d_gstart, d_film, d_rgb8;
    // misc
    DevBuf d_counter, d_tmp, d_probe_in, d_probe_out, d_sort_k, d_sort_v;
    // adjoint: surfel materials, transposed links (cluster-major), segments, iterates, parameter partials
    DevBuf d_smat, d_scl, d_sscale, d_lrow, d_perm, d_perm2, d_ckeyt, d_ckeyt2, d_crow, d_cw, d_colcnt, d_colstart,
        d_segcnt, d_segfirst, d_seg, d_segsum, d_ubar, d_uacc, d_umax, d_mu, d_v, d_dmax, d_pmat, d_pdr, d_pde, d_sdr,
        d_sde, d_partial;
    DevBuf d_ufilm, d_stage, d_sbase, d_sidx, d_sw, d_sperm, d_skey, d_skey2, d_ffmask;
    DevBuf d_tail_st, d_tail_cnt, d_tail_start, d_tail_idx;  // parked link-visibility rays
    // link tiles of the budgeted backward pass
    DevBuf d_cnt2, d_segn, d_segoff, d_tcount, d_tstart, d_tci, d_trow, d_tv, d_tkeep, d_tlstart, d_tidx, d_tw,
        d_tlrow, d_tperm, d_tperm2, d_tkeys, d_tcrow, d_tcw, d_tcolcnt, d_tcolstart, d_tqueue;
    bool adj_tiled_ = false;
    uint32_t adj_tiles_ = 0;
    std::vector<uint32_t> tile_m_, tile_l_;
    uint64_t sliced_gen_ = ~0ull, sliced_cells_ = 0;
    uint64_t transfer_gen_ = 0, adj_gen_ = ~0ull;
    uint32_t adj_nseg_ = 0;
    bool have_beauty_ = false, have_gterm_ = false;
    uint32_t last_place_rounds_ = 0;
    HCamera beauty_cam_;
    uint32_t beauty_w_ = 0, beauty_h_ = 0;
    uint32_t gdim_[3] = {1, 1, 1};
};

}  // namespace ifx
