/*
 * interflect-mi355x: C ABI of the gfx950 transfer -> solve -> shade path.
 *
 * The reference (Cherie05/interflect v0.1.3) is one Rust binary with no FFI;
 * its stage functions, called in sequence by `main`, are the seams this
 * library replaces.  Each entry point below names the reference function it
 * stands in for.  Conventions:
 *   - every call returns an ifx_status: 0 on success, negative on error, with
 *     a message available from ifx_last_error();
 *   - the caller owns every host buffer; the library owns device memory;
 *   - calls are synchronous on return; a context is used from one thread;
 *   - vectors cross the boundary as packed xyz float triples (N x 3 f32).
 */
#ifndef INTERFLECT_H
#define INTERFLECT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifx_ctx ifx_ctx;
typedef int32_t ifx_status;

#define IFX_OK 0
#define IFX_E_INVALID (-1) /* bad argument */
#define IFX_E_PARSE (-2)   /* scene text rejected (src/scene.rs:365 error paths) */
#define IFX_E_STATE (-3)   /* stage called out of order */
#define IFX_E_DEVICE (-4)  /* HIP runtime error */
#define IFX_E_NOMEM (-5)   /* device allocation failed */
#define IFX_E_LIMIT (-6)   /* scene exceeds a kernel limit */

/* Shading modes of src/main.rs:25-51; REFERENCE is the brute-force path tracer
 * (src/reference.rs, "spp" paths per pixel, quad lights). */
#define IFX_MODE_BEAUTY 0
#define IFX_MODE_DIRECT 1
#define IFX_MODE_INDIRECT 2
#define IFX_MODE_NORMALS 3
#define IFX_MODE_DEPTH 4
#define IFX_MODE_STEPS 5
#define IFX_MODE_REFERENCE 6

typedef struct {
    uint32_t n_objects, n_planes, n_quad_lights, n_materials; /* materials exclude the implicit default */
    uint32_t n_bvh_nodes, n_sphere_lights, n_point_lights, max_stack;
    uint32_t width, height, bounces, reserved;
    uint64_t surfels;
    float exposure, blend_k;
    float lo[3], hi[3];
    float cam_pos[3], cam_look[3], cam_up[3], cam_fov;
} ifx_scene_info;

typedef struct {
    float pos[3], look[3], up[3], fov;
} ifx_camera;

typedef struct {
    uint32_t clusters, reserved;
    uint64_t candidates; /* surfel-cluster pairs that passed the cosine and form-factor culls */
    uint64_t links;
    uint64_t evals;      /* SDF evaluations under the reference's counting rule */
    float mean_row_sum;
    float ms;            /* wall time of the stage, including the upload of surfels set or placed since the
                            last transfer (later transfers on the same surfels upload nothing) */
} ifx_transfer_stats;

typedef struct {
    uint32_t iterations;
    float residual;
    uint64_t evals;
    float ms;
    float reserved;
} ifx_solve_stats;

typedef struct {
    uint64_t evals;
    float ms; /* grid build + per-pixel pass + encode + D2H */
    float reserved;
} ifx_shade_stats;

/* Context lifetime.  `device` is the HIP device ordinal. */
ifx_status ifx_create(ifx_ctx** out, int32_t device);
ifx_status ifx_destroy(ifx_ctx* ctx);
const char* ifx_last_error(const ifx_ctx* ctx);

/* scene::parse + Bvh::build + world_bounds (src/scene.rs:365-571, src/bvh.rs:53). */
ifx_status ifx_load_scene(ifx_ctx* ctx, const char* rad_text);

/* ---- A parsed scene from the caller (the reference's own Scene,
 * src/scene.rs:163-178, after its parser and Bvh::build), for hosts that keep
 * their .rad loader.  Plain arrays; nothing is retained after the call. */
#define IFX_PRIM_SPHERE 0
#define IFX_PRIM_BOX 1
#define IFX_PRIM_PLANE 2
#define IFX_PRIM_CAPSULE 3
#define IFX_PRIM_TORUS 4
#define IFX_PRIM_CYLINDER 5
#define IFX_PRIM_CONE 6

/* Prim (src/sdf.rs:9-48).  a: sphere / box / torus centre, plane normal,
 * capsule / cylinder / cone endpoint a; b: box half extents (Prim::RoundBox b),
 * endpoint b; r: sphere / box / capsule / cylinder radius (box: rounding),
 * plane height, torus major radius, cone radius at a; r2: torus minor radius,
 * cone radius at b. */
typedef struct {
    uint32_t kind;
    float a[3], b[3];
    float r, r2;
} ifx_prim;

/* Object (src/sdf.rs:151-163): base primitive, subtractions
 * subtractions[sub_first .. sub_first + sub_count), material index (0 = the
 * implicit default), smooth-union radius (loader extension; 0 = hard union). */
typedef struct {
    ifx_prim prim;
    uint32_t sub_first, sub_count, mat;
    float blend;
} ifx_object;

/* BVH node (src/bvh.rs:26-36): leaf = [start, start + count) of the order
 * array; interior (count 0) = left child at the next index, right child at
 * `right`. */
typedef struct {
    float lo[3], hi[3];
    uint32_t start, count, right;
} ifx_bvh_node;

/* Material (src/scene.rs:11-28); QuadLight (src/scene.rs:32-56); sphere and
 * point lights (loader extensions). */
typedef struct {
    float albedo[3], roughness, metallic, emissive[3];
} ifx_material;
typedef struct {
    float verts[4][3], emit[3];
} ifx_quad_light;
typedef struct {
    float center[3], radius, emit[3];
} ifx_sphere_light;
typedef struct {
    float pos[3], emit[3];
} ifx_point_light;

typedef struct {
    const ifx_object* objects; /* bounded objects, indexed by the BVH */
    uint32_t n_objects;
    const ifx_object* planes; /* unbounded objects (planes) */
    uint32_t n_planes;
    const ifx_prim* subtractions;
    uint32_t n_subtractions;
    const ifx_bvh_node* bvh_nodes; /* n_bvh_nodes = 0: the library builds the BVH (src/bvh.rs:53) */
    uint32_t n_bvh_nodes;
    const uint32_t* bvh_order; /* n_objects object indices in leaf order, with the nodes */
    const ifx_material* materials; /* index 0 is the implicit default material */
    uint32_t n_materials;
    const ifx_quad_light* quad_lights;
    uint32_t n_quad_lights;
    const ifx_sphere_light* sphere_lights;
    uint32_t n_sphere_lights;
    const ifx_point_light* point_lights;
    uint32_t n_point_lights;
    float lo[3], hi[3]; /* Scene::bounds (src/scene.rs:573-601) */
    ifx_camera camera;
    uint32_t width, height, bounces, reserved;
    uint64_t surfels;
    float exposure;
} ifx_scene_desc;

/* Replaces ifx_load_scene for a caller with a parsed scene.  Validated like
 * the loader's output (render limits, a light, material and subtraction
 * indices, a BVH whose leaves cover every object once).  Equivalent scenes
 * give byte-identical device records whichever way they arrive. */
ifx_status ifx_set_scene(ifx_ctx* ctx, const ifx_scene_desc* desc);
/* 64-bit digest of the scene's flattened device records, BVH, bounds,
 * camera and render settings (no counterpart in the reference; lets a caller
 * check that two routes gave the same scene). */
ifx_status ifx_scene_digest(ifx_ctx* ctx, uint64_t* digest);
ifx_status ifx_get_scene_info(ifx_ctx* ctx, ifx_scene_info* out);
/* CLI overrides of render {} (src/main.rs:394-413); re-validated like the reference. */
ifx_status ifx_set_render(ifx_ctx* ctx, uint32_t width, uint32_t height, uint64_t surfels, uint32_t bounces);
/* Scene::use_bvh, the --no-bvh switch (src/scene.rs:172-174). */
ifx_status ifx_set_use_bvh(ifx_ctx* ctx, int32_t on);

/* surfel::generate (src/surfel.rs:147-255): on the device when the context has one
 * (tuning key "surfels_gpu", default 1), else on the host; both give the same surfels. */
ifx_status ifx_generate_surfels(ifx_ctx* ctx, uint64_t target, uint64_t* placed);
/* Supply surfels directly (the pipeline input "given identical surfels"). */
ifx_status ifx_set_surfels(ifx_ctx* ctx, const float* p, const float* n, const float* albedo, const float* area,
                           uint64_t count);
/* Per-surfel material index (count entries, the surfels of the last
 * ifx_set_surfels).  Without it, ifx_set_surfels takes each surfel's material
 * from the scene at its position (Scene::dist_mat, src/scene.rs:191-208), as
 * surfel::generate does; ifx_backward needs the materials. */
ifx_status ifx_set_surfel_materials(ifx_ctx* ctx, const uint32_t* mat, uint64_t count);
ifx_status ifx_num_surfels(ifx_ctx* ctx, uint64_t* count);
ifx_status ifx_get_surfels(ifx_ctx* ctx, float* p, float* n, float* albedo, float* area);

/* formfactor::build_clusters + formfactor::build (src/formfactor.rs:91, :153). */
ifx_status ifx_transfer(ifx_ctx* ctx, int32_t cluster_res, ifx_transfer_stats* out);
/* Transfer CSR: starts[N+1], idx[links], w[links]; sky[N] closure, row_sum[N] (-1 = no links). */
ifx_status ifx_get_transfer(ifx_ctx* ctx, uint32_t* starts, uint32_t* idx, float* w, float* sky, float* row_sum);
/* Clusters: p[C*3], n[C*3], area[C], radius[C], member_starts[C+1], members[N]. */
ifx_status ifx_get_clusters(ifx_ctx* ctx, float* p, float* n, float* area, float* radius, uint32_t* member_starts,
                            uint32_t* members);

/* solve::light_surfels + solve::solve (src/solve.rs:27, :47). */
ifx_status ifx_solve(ifx_ctx* ctx, uint32_t max_bounces, ifx_solve_stats* out);
ifx_status ifx_get_radiosity(ifx_ctx* ctx, float* direct, float* irr, float* rad);
/* No counterpart in the reference.  ifx_solve for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): it allocates nothing.  Outputs go to caller-owned DEVICE
 * arrays d_direct, d_irr, d_rad of N float4 records each (xyz + one unused
 * float per surfel), equal bit for bit to ifx_solve's; every scratch buffer
 * (direct-light terms, cluster radiosity, residuals, counters) lives in one
 * static 32 MiB device array per GPU, so N x lights x 20 B + clusters x 16 B
 * + 16 KiB must fit in it (IFX_E_LIMIT otherwise).  Needs a transfer in the
 * context; the context's own radiosity (ifx_get_radiosity, ifx_shade) is left
 * as the last ifx_solve made it.  Calls that overlap on one GPU, from any
 * context, race on the shared array. */
ifx_status ifx_solve_noalloc(ifx_ctx* ctx, uint32_t max_bounces, float* d_direct, float* d_irr, float* d_rad,
                             ifx_solve_stats* out);
/* Replace the solved radiosity (e.g. to shade from an externally solved cache). */
ifx_status ifx_set_radiosity(ifx_ctx* ctx, const float* direct, const float* irr, const float* rad);

/* Grid::build + render_frame + Film::encode (src/surfel.rs:71, src/main.rs:229, src/film.rs:51).
 * cam = NULL uses the scene camera.  film_rgb (w*h*3 f32, linear) and rgb8
 * (w*h*3 bytes, exposure + tone map + sRGB) may each be NULL. */
ifx_status ifx_shade(ifx_ctx* ctx, const ifx_camera* cam, uint32_t width, uint32_t height, int32_t mode,
                     float* film_rgb, uint8_t* rgb8, ifx_shade_stats* out);

/* Many cameras over one solve in one launch (the turntable of src/main.rs:492-520 and any
 * other view set); film_rgb / rgb8 hold n x height x width x 3 values, view-major; each view
 * equals ifx_shade with that camera. */
ifx_status ifx_shade_views(ifx_ctx* ctx, const ifx_camera* cams, uint32_t n, uint32_t width, uint32_t height,
                           int32_t mode, float* film_rgb, uint8_t* rgb8, ifx_shade_stats* out);
/* The reference's n turntable cameras: the scene camera orbited about its look point
 * (src/main.rs:494-503). */
ifx_status ifx_turntable_cameras(ifx_ctx* ctx, uint32_t n, ifx_camera* out);

/* Kernel-family probes, one launch each.
 * Scene::dist_mat / Scene::normal (src/scene.rs:191, :218); nrm may be NULL. */
ifx_status ifx_dist_mat(ifx_ctx* ctx, const float* pts, uint64_t n, float* d, uint32_t* mat, float* nrm);
/* trace::trace (src/trace.rs:27). */
ifx_status ifx_trace_rays(ifx_ctx* ctx, const float* ro, const float* rd, uint64_t n, float tmin, float tmax,
                          float* t, float* nrm, uint32_t* mat, int32_t* hit, uint32_t* steps);
/* trace::soft_shadow (src/trace.rs:84), per-ray tmax and k. */
ifx_status ifx_soft_shadows(ifx_ctx* ctx, const float* ro, const float* rd, uint64_t n, float tmin,
                            const float* tmax, const float* k, float* vis, uint32_t* steps);

/* Per-kernel timing with HIP events on the context's stream (no counterpart
 * in the reference; used for the roofline report). */
ifx_status ifx_set_profiling(ifx_ctx* ctx, int32_t on);
/* Kernel `name`'s launch count and summed duration since the last reset. */
ifx_status ifx_kernel_time(ifx_ctx* ctx, const char* name, uint64_t* launches, double* total_ms);
ifx_status ifx_reset_kernel_times(ifx_ctx* ctx);
/* Instrumented modes of the tracing kernels (visibility, closure, direct_e,
 * shade); results are identical, timings are not representative.  No
 * counterpart in the reference; feeds the roofline report.
 *   1: a counting instantiation with the reference's per-lane traversal
 *      tallies algorithmic FP32 operations (ifx_kernel_ops "name") and
 *      primitive evaluations ("name.prims");
 *   2: the timed kernels tally the primitive evaluations they execute, every
 *      lane of a shared traversal counted ("name.exec_prims"), and the
 *      persistent ones (visibility, closure) the lane slots of their march
 *      steps, 64 per wavefront step ("name.lane_slots"; the step count of the
 *      last transfer over it is their lane utilisation);
 *   0: off. */
ifx_status ifx_set_counting(ifx_ctx* ctx, int32_t on);
/* Launch-configuration knobs (no counterpart in the reference; results are
 * identical for every setting, only speed changes):
 *   "lds"           1 = stage scene records in LDS when they fit (default), 0 = global
 *   "vis_blocks"    grid of the link-visibility kernel (0 = one lane per candidate)
 *   "vis_chunk"     candidates a link-visibility wavefront takes per queue access (0 = automatic)
 *   "vis_min_idle"  1..64: idle link-visibility lanes take new candidates once this many
 *                   lanes of their wavefront are idle (default 64)
 *   "closure_min_idle" 1..64: the same for the closure rays (default 64)
 *   "vis_tail"      0..32: a whole-wavefront link-visibility batch parks its rays once this
 *                   many or fewer still march, and a second launch marches the parked rays
 *                   (default 8; 0 = off; needs vis_min_idle 64)
 *   "vis_tail_items" 0..2^30: parking applies from this many candidates up (default 2^23)
 *   "flat_query"    1 = a wavefront's scene queries along marches test every object's box
 *                   against the wavefront's points (default), 0 = they walk the BVH
 *   "ff_blocks"     grid of the candidate-cull kernels (0 = automatic)
 *   "jacobi_blocks" grid of the Jacobi and adjoint sweep kernels (0 = automatic)
 *   "jacobi_rows"   0 = automatic (sliced rows when rows average >= 64 links, else 1),
 *                   1 = one lane per row on the CSR, 2 = one wavefront per row,
 *                   3 = sliced rows (one lane per row, links stored per 64-row slice column by column)
 *   "surfels_gpu"   1 = place surfels on the device (default), 0 = on the host
 *   "side_stream"   1 = closure rays and direct-light terms on a second stream beside the
 *                   link visibility (default), 0 = everything on one stream
 *   "spp"           paths per pixel of IFX_MODE_REFERENCE (--spp, default 512)
 * and of the backward pass's memory (gradients are identical for every
 * link_budget; adj_tile_rows and adj_seg fix its summation order):
 *   "link_budget"   bytes the backward pass may use for its copy of the links
 *                   (0 = no limit, default; up to 2^50).  Keeping them costs about
 *                   32 bytes per link while they are transposed; beyond the budget
 *                   every adjoint sweep recomputes the links tile by tile from the
 *                   geometry instead.  A budget below one tile's working set (about
 *                   16 B per row + 12 B per candidate + 40 B per link of the tile)
 *                   makes ifx_backward return IFX_E_NOMEM.
 *   "adj_tile_rows" surfel rows per link tile (0 = 4096)
 *   "adj_seg"       transposed links per partial sum (0 = 1024) */
ifx_status ifx_set_tuning(ifx_ctx* ctx, const char* key, int64_t value);
ifx_status ifx_kernel_ops(ifx_ctx* ctx, const char* name, uint64_t* ops);

/* ---- Adjoint (no counterpart in the reference, which has no backward pass).
 * Gradients of a loss L on the linear beauty film of the last beauty
 * ifx_shade (before exposure and tone mapping) with respect to each
 * material's albedo and each light's emission, geometry frozen: transfer
 * weights, light visibility and gather weights are held fixed and the solve
 * is differentiated at its fixed point (implicit adjoint).  Diffuse
 * materials only (metallic 0, roughness >= 0.45); surfels carry their
 * material (ifx_generate_surfels, ifx_set_surfels).  Materials are indexed as the scene
 * indexes them: 0 is the implicit default material (src/scene.rs:373), k the
 * k-th material block of the file, so there are n_materials + 1 of them.
 * Lights are indexed quads first, then sphere lights, then point lights. */
typedef struct {
    uint32_t iterations; /* adjoint sweeps */
    float residual;      /* max |delta mu| / max |mu| of the last sweep */
    float ms;            /* wall time of the call */
    uint32_t link_tiles; /* row tiles recomputed every sweep (link_budget), 0 = links kept */
} ifx_backward_stats;

/* Parameter updates that keep the transfer (albedo and emission do not enter it). */
ifx_status ifx_set_material_albedo(ifx_ctx* ctx, uint32_t material, const float* rgb);
ifx_status ifx_set_light_emit(ifx_ctx* ctx, uint32_t light, const float* rgb);
/* Jacobi stopping threshold of ifx_solve (default 1e-4, src/solve.rs:98). */
ifx_status ifx_set_solve_tolerance(ifx_ctx* ctx, float tol);
/* Adjoint sweeps: stop once max |delta mu| <= rel_tol * max |mu| or after max_iters. */
ifx_status ifx_set_adjoint(ifx_ctx* ctx, uint32_t max_iters, float rel_tol);
/* dL_dfilm: width*height*3 floats of the last beauty shade.  Outputs (either may
 * be NULL): dL_dalbedo[(n_materials + 1)*3], dL_demit[lights*3]. */
ifx_status ifx_backward(ifx_ctx* ctx, const float* dL_dfilm, float* dL_dalbedo, float* dL_demit,
                        ifx_backward_stats* out);
/* The solve's share alone, for a given dL/dirr (N*3 floats of surfel
 * irradiance); mu (N*3, may be NULL) receives the adjoint state. */
ifx_status ifx_backward_surfels(ifx_ctx* ctx, const float* dL_dirr, float* dL_dalbedo, float* dL_demit, float* mu,
                                ifx_backward_stats* out);
/* Per-surfel material index (N u32) and the direct-light factors
 * g[light*N + i] = form factor x visibility of the last solve. */
ifx_status ifx_get_surfel_materials(ifx_ctx* ctx, uint32_t* mat);
ifx_status ifx_get_direct_geometry(ifx_ctx* ctx, float* g);

/* ---- Batches: many scenes of uneven size per launch (no counterpart in the
 * reference, which renders one scene per process).  Every stage runs once
 * over all scenes with per-scene offsets; the tracing kernels split their
 * grid into per-scene block ranges.  A frame is transfer + solve (the scene's
 * render.bounces, tolerance 1e-4) + beauty shade + RGB8 encode of every scene,
 * with the scene's own camera and film size; per-scene results equal those of
 * a single-scene context given the same surfels.  With graph mode on, the
 * first frame after an eager frame is captured into a HIP graph and later
 * frames replay it. */
typedef struct ifx_batch ifx_batch;
typedef struct {
    uint64_t surfels, candidates, links;
    uint64_t evals_transfer, evals_solve, evals_shade;
    uint32_t clusters, iterations, width, height;
    float residual, mean_row_sum;
} ifx_batch_scene_stats;

ifx_status ifx_batch_create(ifx_batch** out, int32_t device);
ifx_status ifx_batch_destroy(ifx_batch* b);
const char* ifx_batch_last_error(const ifx_batch* b);
ifx_status ifx_batch_add_scene(ifx_batch* b, const char* rad_text, int32_t cluster_res, uint32_t* index);
/* surfel::generate for every scene (render.surfels requested), on the host. */
ifx_status ifx_batch_generate_surfels(ifx_batch* b);
ifx_status ifx_batch_set_graph(ifx_batch* b, int32_t on);
/* Launch shapes: persistent blocks per CU of the tracing kernels (1..64), wavefront queue chunk
 * divisor, Jacobi row mapping (0 lane, 1 wavefront).  Results do not depend on them. */
ifx_status ifx_batch_set_tuning(ifx_batch* b, uint32_t blocks_per_cu, uint32_t chunk_div, uint32_t wave_rows);
/* Times one eager frame per candidate launch shape and keeps the fastest. */
ifx_status ifx_batch_autotune(ifx_batch* b, float* best_ms);
ifx_status ifx_batch_run(ifx_batch* b, float* ms);
ifx_status ifx_batch_get_stats(ifx_batch* b, uint32_t scene, ifx_batch_scene_stats* out);
ifx_status ifx_batch_get_image(ifx_batch* b, uint32_t scene, float* film_rgb, uint8_t* rgb8);
ifx_status ifx_batch_get_radiosity(ifx_batch* b, uint32_t scene, float* direct, float* irr, float* rad);
/* CSR of one scene with scene-local cluster indices (as ifx_get_transfer). */
ifx_status ifx_batch_get_transfer(ifx_batch* b, uint32_t scene, uint32_t* starts, uint32_t* idx, float* w);

#ifdef __cplusplus
}
#endif

#endif /* INTERFLECT_H */
