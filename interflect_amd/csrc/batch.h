// Batched transfer -> solve -> shade: many scenes of uneven size per launch.
//
// All scenes' records, surfels, clusters, candidates, links and films live in
// concatenated device arrays addressed through per-scene offsets.  Every
// stage is one launch (or one launch per bounce) over all scenes: per-item
// kernels find their scene from a per-surfel / per-cluster scene index, and
// the persistent tracing kernels walk a list of scenes: a block stages one
// scene in LDS and its lanes refill from that scene's queue until it drains,
// then the block moves on (one-item-per-lane kernels split their grid into
// per-scene block ranges instead).  After one eager frame has fixed the
// sizes, a frame can be captured into a HIP graph and replayed; the launch
// shapes (persistent blocks per CU, queue chunk divisor, row mapping of the
// Jacobi gather) are picked by a small autotuner.  Per-scene results are identical to separate
// single-scene runs.
#pragma once
#include <string>
#include <vector>

#include "engine.h"

namespace ifx {

struct BatchSceneStats {
    uint64_t surfels = 0, candidates = 0, links = 0;
    uint64_t evals_transfer = 0, evals_solve = 0, evals_shade = 0;
    uint32_t clusters = 0, iterations = 0, width = 0, height = 0;
    float residual = 0, mean_row_sum = 0;
};

struct BatchTuning {
    uint32_t blocks_per_cu = kMarchWaves;  // persistent blocks per CU of the closure / visibility kernels
    uint32_t chunk_div = 8;      // queue chunk = even share per wavefront / chunk_div, 64..256 items
    uint32_t wave_rows = 0;      // Jacobi gather: 0 lane per row, 1 wavefront per row
};

class Batch {
   public:
    Batch() = default;
    ~Batch();
    int init(int device, std::string& err);
    int add_scene(const std::string& text, int32_t res, uint32_t& index, std::string& err);
    int generate_surfels(std::string& err);
    // One frame (transfer + solve + shade, beauty, RGB8) of every scene.
    int run(double& ms, std::string& err);
    // Replay frames from a captured graph (captured on the first call after
    // an eager frame).  Off: every frame runs eagerly.
    void set_graph(bool on) {
        use_graph_ = on;
        drop_graph();
    }
    int autotune(double* best_ms, std::string& err);
    int download_image(uint32_t s, float* film, uint8_t* rgb8, std::string& err);
    int download_radiosity(uint32_t s, float* direct, float* irr, float* rad, std::string& err);
    int download_transfer(uint32_t s, uint32_t* starts, uint32_t* idx, float* w, std::string& err);
    uint32_t size() const { return (uint32_t)scenes_.size(); }
    const BatchSceneStats& stats(uint32_t s) const { return stats_[s]; }
    bool graph_active() const { return graph_exec_ != nullptr; }
    BatchTuning tuning;

   private:
    struct SceneIn {
        HostScene sc;
        int32_t res;
        std::vector<HSurfel> surfels;
    };
    int frame(bool eager, std::string& err);
    int upload_static(std::string& err);
    int build_slots(std::string& err);
    int finish_stats(std::string& err);
    void drop_graph();
    int check(hipError_t e, const char* what, std::string& err);
    uint32_t persist_grid() const;

    Engine eng_;  // device, stream and scene-independent tables
    hipStream_t stream_ = nullptr;
    std::vector<SceneIn> scenes_;
    std::vector<BatchSceneStats> stats_;
    bool static_ok_ = false, sized_ = false, use_graph_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;

    // host offsets
    std::vector<uint32_t> s0_, px0_, term0_, cell0_, cl0_, C_, cand0_, M_, nl_;
    uint32_t Ntot_ = 0, Ptot_ = 0, Ttot_ = 0, cells_ = 0, nseg_ = 0, Ctot_ = 0, Mtot_ = 0, Ltot_ = 0, Bmax_ = 0;
    // block partitions per LDS footprint class (see batch.hip) and their LDS bytes
    std::vector<uint32_t> blk_closure_[4], blk_vis_[4], blk_direct_[4], blk_shade_[4], blk_encode_;
    size_t lds_n_[4] = {16, 16, 16, 16}, lds_l_[4] = {16, 16, 16, 16};

    // device
    DevBuf d_rec, d_slots, d_blk, d_ssid, d_spos, d_snrm, d_salb;
    DevBuf d_ckey, d_ckey2, d_cval, d_cval2, d_cflag, d_cseg, d_cbegin, d_craw_pos, d_craw_nrm, d_cvalid, d_cidx;
    DevBuf d_cpos, d_cnrm, d_crange, d_csid, d_ccount;
    DevBuf d_code, d_mkey, d_mkey2, d_mval, d_rorder;
    DevBuf d_open, d_sky, d_counters, d_queues;
    DevBuf d_fcount, d_cstart, d_cand_ci, d_cand_row, d_cand_v, d_vkey, d_vkey2, d_vval, d_vorder;
    DevBuf d_lcount, d_lstart, d_scale, d_rowsum, d_idx, d_w;
    DevBuf d_terms, d_gterm, d_direct, d_rad, d_irr, d_bc, d_resid;
    DevBuf d_gkey, d_gkey2, d_gval, d_gval2, d_gcount, d_gstart, d_film, d_rgb8, d_stage;
    DevBuf d_scratch;
    bool members_alt_ = false, rorder_alt_ = false, vorder_alt_ = false, gitems_alt_ = false;
};

}  // namespace ifx
