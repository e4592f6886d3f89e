/*
 * niq.h - C ABI of the MI355X (gfx950) range-analysis hot path.
 *
 * Plain pointers and sizes only.  Every device pointer is caller-owned HIP
 * device memory; `stream` is a hipStream_t (NULL = default stream).  Calls
 * are asynchronous on `stream` unless documented as synchronising.  A model
 * handle is immutable after creation and may be shared by threads and
 * streams.  Errors: a negative niq_status, with a message from
 * niq_last_error() (thread-local).
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository).
 */
#ifndef NIQ_H
#define NIQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int niq_status;
#define NIQ_OK 0
#define NIQ_ERR_INVALID (-1)      /* bad argument (cf. ValueError in src/mlp.py:131,142) */
#define NIQ_ERR_UNSUPPORTED (-2)  /* shape/mode not supported by any kernel */
#define NIQ_ERR_CAPACITY (-3)     /* an output or workspace buffer is too small */
#define NIQ_ERR_HIP (-4)          /* HIP runtime error */

/* activations (src/mlp.py:283-300; TANH has no reference rule) */
#define NIQ_ACT_NONE 0
#define NIQ_ACT_RELU 1
#define NIQ_ACT_ELU 2
#define NIQ_ACT_SIN 3
#define NIQ_ACT_TANH 4

/* range-analysis modes (src/affine.py:62-76, src/implicit_mlp_utils.py:24-61) */
#define NIQ_MODE_INTERVAL 0
#define NIQ_MODE_AFFINE_FIXED 1
#define NIQ_MODE_AFFINE_TRUNCATE 2
#define NIQ_MODE_AFFINE_APPEND 3
#define NIQ_MODE_AFFINE_ALL 4
#define NIQ_MODE_SLOPE_INTERVAL 5
/* weak-SDF Lipschitz test (src/sdf.py:17-50, the 'sdf' mode of
 * src/implicit_mlp_utils.py:24-30): a box is UNKNOWN where
 * |f(centre)| < sdf_lipschitz * radius, radius = sqrt(sum_v |vec_v|^2);
 * otherwise POSITIVE if f(centre) > offset, NEGATIVE if f(centre) < -offset */
#define NIQ_MODE_SDF 6

/* arithmetic of the error-term contraction */
#define NIQ_PREC_FP32 0
#define NIQ_PREC_BF16 1

/* sign codes (src/implicit_function.py:11-13) */
#define NIQ_SIGN_UNKNOWN 0
#define NIQ_SIGN_POSITIVE 1
#define NIQ_SIGN_NEGATIVE 2

typedef struct niq_model niq_model;

typedef struct niq_eval_cfg {
  int mode;            /* NIQ_MODE_* */
  int precision;       /* NIQ_PREC_* */
  int truncate_count;  /* affine_truncate: rows kept */
  int n_append;        /* affine_append: rows appended per activation */
  float recheck_tol;   /* > 0: FP64 re-test of cells within tol*scale of the threshold */
  int variant;         /* kernel variant, 0 = default; plain FP32: 1 TileCfg,
                          2 TileCfgAlt, 3 TileCfg with 256 registers, 4
                          TileCfgWide (the most waves, the narrowest width
                          tile per wave) (niq_tile_shape); growing-row modes: 2 half the
                          error-term rows per chunk, 3 a 2-waves-per-SIMD
                          register budget; precision mode: 1 streams the
                          packed bf16 weight copies, 2 splits FP32 weights in
                          registers (the default), 3 layer by layer; closest
                          point runs its round as one forward launch only at
                          variant 0.  Every variant gives the same results. */
  float sdf_lipschitz; /* NIQ_MODE_SDF: Lipschitz bound of f (sdf_lipschitz kwarg) */
  int ray_variant;     /* cast_rays tile shape, 0 = default; 1 and 2 are the
                          alternatives the autotuner times (same results) */
} niq_eval_cfg;

/* ---- model ------------------------------------------------------------
 * Replaces generate_implicit_from_file / mlp.load (src/implicit_mlp_utils.py:12-64,
 * src/mlp.py:173-185): the npz op list, lowered by the caller to n_layers dense
 * layers.  Layer i maps dims[i] -> dims[i+1] with W[i] row-major (dims[i],
 * dims[i+1]) (the reference's A, src/mlp.py:218-249, 253-258) and bias b[i]
 * (nullable); acts[i] is applied after layer i (ignored for the last layer, whose
 * output width must be 1: squeeze_last, src/mlp.py:328-332).  Host pointers;
 * weights are copied to the current device.  Synchronising. */
niq_status niq_model_create(int n_layers, const int* dims, const int* acts,
                            const float* const* W, const float* const* b,
                            niq_model** out);
void niq_model_destroy(niq_model* m);
/* bytes of device memory the model holds */
size_t niq_model_device_bytes(const niq_model* m);

/* ---- per-box classification --------------------------------------------
 * Replaces ImplicitFunction.classify_box (src/implicit_function.py:28-37) with
 * AffineImplicitFunction / SlopeIntervalImplicitFunction.classify_general_box
 * (src/affine.py:34-55, src/slope_interval.py:29-51), vmapped over n boxes.
 * lower/upper: (n, d) float32.  codes: (n,) int32 NIQ_SIGN_*.
 * bounds: optional (n, 2) float32 [may_lower, may_upper]. */
niq_status niq_classify_boxes(const niq_model* m, const niq_eval_cfg* cfg,
                              const float* lower, const float* upper, int64_t n, int d,
                              float offset, int32_t* codes, float* bounds, void* stream);

/* General boxes: center (n, d), vecs (n, v, d) (src/affine.py:109-117). */
niq_status niq_classify_general_boxes(const niq_model* m, const niq_eval_cfg* cfg,
                                      const float* center, const float* vecs, int64_t n,
                                      int v, int d, float offset, int32_t* codes,
                                      float* bounds, void* stream);

/* Tile shape of a kernel variant, for autotuners and reports: which 0 is the
 * plain FP32 classify pass (eval cfg variant 0-4), 1 the ray caster
 * (ray_variant 0-2), 2 the growing-row classify (variant 0-3); width a layer
 * width of the model (its template width is used).  *nb: boxes or ray slots
 * per workgroup (which 2: error-term rows per chunk; one box per workgroup),
 * *nw: waves per workgroup, *occ: waves per SIMD the register budget is set
 * for.  Host only. */
niq_status niq_tile_shape(int which, int variant, int width, int* nb, int* nw, int* occ);

/* ---- plain evaluation ----------------------------------------------------
 * Replaces ImplicitFunction.__call__ (src/affine.py:26-28 -> src/mlp.py:96-113,
 * default rules) vmapped over points: x (n, d) -> out (n,). */
niq_status niq_eval_points(const niq_model* m, const float* x, int64_t n, int d,
                           float* out, void* stream);

/* ---- level-set tree --------------------------------------------------------
 * Replaces construct_uniform_unknown_levelset_tree (src/kd_tree.py:103-218) and
 * its level step (src/kd_tree.py:19-100) with bucketing (src/bucketing.py:16-32).
 * All levels run on `stream` with device-side counts; the call synchronises once
 * at the end (and once per level when node_terminate_thresh > 0). */
typedef struct niq_tree_params {
  float lower[3], upper[3];
  int split_depth;                 /* < 0: none (then node_terminate_thresh > 0) */
  int64_t node_terminate_thresh;   /* <= 0: none */
  float offset;
  int batch_process_size;          /* B of the reference; fixes the output order */
  int with_interior_nodes;
  int with_exterior_nodes;
} niq_tree_params;

typedef struct niq_tree_result {
  /* caller-provided device buffers (n, 3) float32 and their capacities (rows) */
  float* unknown_lower; float* unknown_upper; int64_t unknown_capacity;
  float* interior_lower; float* interior_upper; int64_t interior_capacity;
  float* exterior_lower; float* exterior_upper; int64_t exterior_capacity;
  /* filled on return */
  int64_t n_unknown, n_interior, n_exterior;
  int64_t n_classified;            /* boxes bounded */
  int64_t n_recheck;               /* cells re-tested in FP64 */
  int n_levels;
  int32_t level_n[64], level_interior[64], level_exterior[64];
} niq_tree_result;

/* Workspace (device bytes) a tree build needs: level ping-pong lists, codes,
 * per-batch counters. */
niq_status niq_tree_workspace_bytes(const niq_model* m, const niq_eval_cfg* cfg,
                                    const niq_tree_params* p, size_t* bytes);
niq_status niq_tree_build(const niq_model* m, const niq_eval_cfg* cfg,
                          const niq_tree_params* p, void* workspace, size_t workspace_bytes,
                          niq_tree_result* res, void* stream);

/* Asynchronous form of niq_tree_build for serving loops (the same kernels and
 * results as src/extract_cell.py:95-160 with split_depth set): enqueues every
 * level on `stream` and returns without synchronising or allocating, so the
 * call can be captured in a hipGraph.  Requires split_depth >= 0, no
 * node_terminate_thresh and a fused-kernel mode (interval, affine_fixed,
 * slope_interval; FP32 with optional FP64 re-test, or BF16).  The tree's
 * counters stay in the workspace at *counters_offset; after the stream has
 * finished, copy niq_tree_counters_bytes() from there to the host and call
 * niq_tree_decode to fill the counts of `res`. */
niq_status niq_tree_enqueue(const niq_model* m, const niq_eval_cfg* cfg,
                            const niq_tree_params* p, void* workspace, size_t workspace_bytes,
                            niq_tree_result* res, size_t* counters_offset, void* stream);
size_t niq_tree_counters_bytes(void);
niq_status niq_tree_decode(const void* host_counters, const niq_tree_params* p,
                           niq_tree_result* res);

/* ---- ray casting ------------------------------------------------------------
 * Replaces cast_rays / cast_rays_iter (src/queries.py:134-175, 39-132) with the
 * options of get_default_cast_opts (src/queries.py:23-36). */
typedef struct niq_cast_opts {
  float hit_eps, max_dist;
  int n_max_step, n_substeps;
  float safety_factor, interval_grow_fac, interval_shrink_fac;
  float init_step;   /* interval_init_size * max_dist, rounded once to float32 */
} niq_cast_opts;

/* models[n_models] (func ids 1..n_models); roots, dirs (n, 3) float32;
 * out_t (n,) float32, out_hit_id (n,) int32, out_count (n,) int32;
 * workspace >= niq_cast_workspace_bytes.  Synchronising; *n_evals receives
 * the reference's N_evals (evaluations including its bucket padding). */
size_t niq_cast_workspace_bytes(const niq_cast_opts* opts);
niq_status niq_cast_rays(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                         const float* roots, const float* dirs, int64_t n,
                         const niq_cast_opts* opts, float* out_t, int32_t* out_hit_id,
                         int32_t* out_count, void* workspace, size_t workspace_bytes,
                         int64_t* n_evals, void* stream);

/* ---- ray casting without allocation -------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc, processes
 * with graph capture in flight): niq_cast_rays with its workspace held in a
 * static device array of NIQ_CAST_NOALLOC_SCRATCH_BYTES on each GPU instead of
 * a workspace argument.  Calls no hipMalloc or hipFree.  The limit:
 * niq_cast_workspace_bytes(opts) <= NIQ_CAST_NOALLOC_SCRATCH_BYTES, that is
 * n_max_step up to 10,000; larger requests fail with NIQ_ERR_CAPACITY.  Every call
 * on one GPU shares that GPU's array, so calls that overlap on one GPU (from
 * other threads or streams) race.  Synchronising, like niq_cast_rays. */
#define NIQ_CAST_NOALLOC_SCRATCH_BYTES 65536
niq_status niq_cast_rays_noalloc(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                 const float* roots, const float* dirs, int64_t n,
                                 const niq_cast_opts* opts, float* out_t, int32_t* out_hit_id,
                                 int32_t* out_count, int64_t* n_evals, void* stream);

/* ---- frustum ray casting -----------------------------------------------------
 * Replaces cast_rays_frustum (src/queries.py:454-587) with its iteration
 * (:178-364), subdivision (:371-432) and output (:434-452).  The camera is the
 * reference's cam_params (root_pos, look_dir, up_dir, left_dir, fov_x, fov_y,
 * res_x, res_y); outputs out_t, out_hit_id, out_count are (res_x, res_y)
 * row-major, indexed [x][y] as the reference's; *n_evals receives its N_evals
 * (frustum-array sizes summed over iterations).  Synchronising. */
typedef struct niq_frustum_cam {
  float root[3], look[3], up[3], left[3];
  float fov_x, fov_y;       /* degrees */
  int res_x, res_y;
} niq_frustum_cam;

typedef struct niq_frustum_opts {
  niq_cast_opts cast;       /* hit_eps .. init_step as for niq_cast_rays */
  float refine_width_fac;
  int n_side_init;
} niq_frustum_opts;

niq_status niq_cast_rays_frustum(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                 const niq_frustum_cam* cam, const niq_frustum_opts* opts,
                                 float* out_t, int32_t* out_hit_id, int32_t* out_count,
                                 int64_t* n_evals, void* stream);

/* ---- intersection of two functions --------------------------------------------
 * Replaces find_any_intersection (src/kd_tree.py:402-655): a breadth-first
 * kd-tree search over [lower, upper] for a point inside both functions (both
 * values < 0).  *found = 1 and loc = that point, from the first node in the
 * reference's level order that certifies one; else *found = 0 and loc =
 * (-777, -777, -777).  *n_processed counts the nodes searched.  Synchronising
 * (once per level). */
niq_status niq_find_any_intersection(const niq_model* model_a, const niq_eval_cfg* cfg_a,
                                     const niq_model* model_b, const niq_eval_cfg* cfg_b,
                                     const float lower[3], const float upper[3], float eps,
                                     int32_t* found, float loc[3], int64_t* n_processed,
                                     void* stream);

/* ---- closest point ---------------------------------------------------------
 * Replaces closest_point / closest_point_iter (src/kd_tree.py:766-802, 659-763)
 * with the same batch-stack order (B = batch_process_size items per round), so
 * results match the reference's.  query_points (Q, 3); out_dist (Q,) and
 * out_loc (Q, 3) float32 (distance to a spanning node's centre plus its
 * diagonal, and that node's centre; queries that never meet a spanning node
 * keep +inf and (-777, -777, -777)).  Synchronising (every 8 rounds). */
niq_status niq_closest_point(const niq_model* m, const niq_eval_cfg* cfg, const float lower[3],
                             const float upper[3], const float* query_points, int64_t Q, float eps,
                             int batch_process_size, float* out_dist, float* out_loc,
                             int64_t* n_processed, void* stream);

/* ---- marching cubes ----------------------------------------------------------
 * Replaces the extraction half of hierarchical_marching_cubes
 * (src/kd_tree.py:338-399, src/extract_cell.py:314-421): for n_leaves leaf boxes
 * (n, 3), a dense marching-cubes pass on the 2^s-per-side subgrid of each leaf.
 * Writes up to tri_capacity triangles (T, 3, 3) float32 in (leaf, subcell,
 * slot) order; *n_tris receives the full count (NIQ_ERR_CAPACITY if larger
 * than tri_capacity).  Synchronising. */
niq_status niq_mc_extract(const niq_model* m, const float* leaf_lower, const float* leaf_upper,
                          int64_t n_leaves, int n_subcell_depth, float* tri_out, int64_t tri_capacity,
                          int64_t* n_tris, void* stream);
/* The marching-cubes triangle table the kernels use: out[256 * 16] edge ids,
 * -1 padded (host only; for checking). */
int niq_mc_table(signed char* out);

/* ---- level-set sampling and bulk properties ------------------------------------
 * Replace the sampling rounds of sample_surface / sample_surface_uniform
 * (src/kd_tree.py:220-336) and the mass integral of bulk_properties
 * (src/kd_tree.py:804-863).  Candidates come from a counter-based stream,
 * Philox4x32-10 keyed by `seed` (jax.random's threefry cannot be reproduced
 * without JAX): candidate j uses counter (j mod 2^32, j div 2^32, tag, 0) with
 * tag 0x53555246 (surface) or 0x4d415353 (mass); its node is
 * floor(w0 * n_nodes / 2^32) and its position lower + u * (upper - lower) with
 * u_k = (w_{k+1} >> 8) * 2^-24, in float32.  Node lists are (n_nodes, 3)
 * float32 device arrays, 1 <= n_nodes < 2^31. */
size_t niq_sample_workspace_bytes(int64_t n_per_round);

/* Surface samples: candidates first_candidate + r * n_per_round + i of round r
 * are accepted where |f| < width and appended in candidate order to
 * out_points (n_samples, 3) from row *n_found on; *n_found is updated.  Rounds
 * run until n_samples are found (the result is then the first accepted
 * candidates of the stream, whatever n_per_round is) or max_rounds have run;
 * compare *n_found with n_samples to tell which.  *n_rounds receives the
 * rounds run.  Synchronising (once per round). */
niq_status niq_sample_surface_nodes(const niq_model* m, const float* node_lower, const float* node_upper,
                                    int64_t n_nodes, float width, uint64_t seed,
                                    int64_t first_candidate, int64_t n_per_round, int64_t max_rounds,
                                    float* out_points, int64_t n_samples, int64_t* n_found,
                                    int64_t* n_rounds, void* workspace, size_t workspace_bytes,
                                    void* stream);

/* Sums of bulk_properties in float64: the volume and first moment of the
 * interior boxes (src/kd_tree.py:841-853), the volume of the unknown boxes, and
 * for n_samples mass candidates drawn from the unknown boxes the count with
 * f < 0 and the sum of their positions (bulk_properties_sample_mass,
 * src/kd_tree.py:804-833).  Either list may be empty (n = 0); with no unknown
 * boxes no candidates are drawn.  Deterministic for a given device.
 * Synchronising. */
typedef struct niq_bulk_sums {
  double interior_volume;
  double interior_moment[3];
  double unknown_volume;
  double n_inside;
  double inside_moment[3];
} niq_bulk_sums;

niq_status niq_bulk_sums_nodes(const niq_model* m, const float* unknown_lower, const float* unknown_upper,
                               int64_t n_unknown, const float* interior_lower, const float* interior_upper,
                               int64_t n_interior, int64_t n_samples, uint64_t seed, niq_bulk_sums* out,
                               void* stream);

/* ---- normals and shading -------------------------------------------------------
 * The image half of render_image (src/render.py:94-150). */

/* Replaces generate_camera_rays (src/render.py:26-50): the res x res pinhole
 * rays, pixel p = row * res + col, in the reference's float32 operations.
 * look, up, left: the camera frame as generate_camera_rays forms it (up made
 * orthogonal to look and normalised, left = look x up); tan_half =
 * tan(radians(fov) / 2) in float32.  out_roots, out_dirs: (res^2, 3) float32
 * device buffers.  Asynchronous on `stream`. */
niq_status niq_camera_rays(const float eye[3], const float look[3], const float up[3], const float left[3],
                           int res, float tan_half, float* out_roots, float* out_dirs, void* stream);

/* hit_pos = roots + t * dirs (src/render.py:139), float32, unfused; (n, 3). */
niq_status niq_hit_positions(const float* roots, const float* dirs, const float* t, int64_t n,
                             float* out, void* stream);

/* Replaces outward_normals (src/render.py:53-90): per pixel the normalised
 * gradient of function hit_id (1..n_models); (0, 0, 0) where hit_id matches
 * no function.  method NIQ_NORMALS_FINITE_DIFFERENCES: from the 4 tetrahedron
 * samples hit_pos + (+-eps, +-eps, +-eps); NIQ_NORMALS_AUTODIFF: the exact
 * forward-mode gradient (jax.jacfwd), eps unused.  hit_pos, out_normals (n, 3)
 * float32; hit_ids (n,) int32.  Asynchronous on `stream`. */
#define NIQ_NORMALS_FINITE_DIFFERENCES 0
#define NIQ_NORMALS_AUTODIFF 1
niq_status niq_outward_normals(const niq_model* const* models, int n_models, const float* hit_pos,
                               const int32_t* hit_ids, int64_t n, float eps, int method,
                               float* out_normals, void* stream);

#define NIQ_SHADE_NORMAL 0
#define NIQ_SHADE_MATCAP 1

/* Replaces shade_image (src/render.py:160-224) and, with composite / tonemap
 * set, the compositing and tonemap_image of render_image (:139-149, :152-158). */
typedef struct niq_shade_opts {
  int shading;                 /* NIQ_SHADE_NORMAL ('normal') or NIQ_SHADE_MATCAP ('matcap_color') */
  float up_dir[3];
  const float* matcaps[4];     /* device (h, w, 3) float32 images: r, g, b, k */
  int matcap_h, matcap_w;
  float colors[8][3];          /* shading_color_tuple, by hit_id */
  int n_colors;
  const float* color_per_hit;  /* device (n, 3): shading_color_func(hit_pos), or NULL */
  int composite;               /* 1: white where hit_id == 0 */
  int tonemap;                 /* 1: tonemap_image after compositing */
  float gamma, white_level, exposure;
} niq_shade_opts;

/* ray_dirs, normals, out_rgb (n, 3) float32; hit_ids (n,) int32.
 * Asynchronous on `stream`. */
niq_status niq_shade_image(const niq_shade_opts* opts, const float* ray_dirs, const float* normals,
                           const int32_t* hit_ids, int64_t n, float* out_rgb, void* stream);

/* ---- serving loop ------------------------------------------------------------
 * No counterpart in the reference (it runs one query at a time from one Python
 * thread, src/queries.py:134, src/kd_tree.py:103,766).  A server owns
 * n_streams worker threads, each with its own HIP stream at the default
 * priority.  Requests queue in arrival order; a worker takes the oldest one as
 * soon as its device memory fits under mem_cap_bytes beside the work in
 * flight.  Ray requests waiting in the queue with the same eval config and
 * template width are merged into the cast of the request taken (up to 64
 * requests and max_merge_rays rays): each request is a group of rays with its
 * own models, options, inputs and outputs inside one persistent kernel, and
 * every ray is independent, so merged results (N_evals included) equal lone
 * ones bit for bit.  Tree builds in the fused-kernel FP32 modes with a
 * split_depth that wait in the queue with the same eval config, build params
 * (root boxes aside) and template width are merged too (up to 64): one build
 * runs every level of all of them in shared launches, each tree's level as
 * its lone build runs it, so the outputs are a lone build's bit for bit.  A
 * tree build taken alone replays a hipGraph captured once per (model,
 * config, depth, capacities) and shared by the workers (a graph replays on
 * one stream at a time).  Closest-point batches run as they are.
 *
 * Buffers: inputs must be complete on the device when submitted and stay
 * alive, and outputs untouched, until the request has been polled done.
 * Model handles must outlive the requests that use them.  A cached graph is
 * matched to its model by a serial number unique to each niq_model_create,
 * never by handle address, so a model created after another was destroyed
 * never replays the old model's graph (such graphs are evicted unused).
 * Polling one id from several threads is safe: one poll returns the result,
 * the others NIQ_ERR_INVALID. */
typedef struct niq_server niq_server;

typedef struct niq_serve_info {
  int state;             /* 0 queued, 1 running, 2 done */
  niq_status status;     /* the request's own status once done */
  double queue_ms;       /* submit -> start */
  double latency_ms;     /* submit -> done */
  int64_t n_evals;       /* rays: the request's N_evals, as a lone cast gives it */
  int32_t n_batched;     /* rays, trees: requests merged into the launches that
                            served it (1: alone) */
  int32_t graph_replay;  /* trees: 0 built without a graph; 1 served by a graph
                            captured for this request; 2 by replaying a graph
                            cached from an earlier request */
  int64_t n_processed;   /* closest point: nodes processed */
} niq_serve_info;

niq_status niq_serve_create(int n_streams, size_t mem_cap_bytes, int64_t max_merge_rays,
                            niq_server** out);
/* Waits for every accepted request, then stops the workers. */
niq_status niq_serve_destroy(niq_server* srv);

niq_status niq_serve_submit_rays(niq_server* srv, const niq_model* const* models, int n_models,
                                 const niq_eval_cfg* cfg, const float* roots, const float* dirs,
                                 int64_t n, const niq_cast_opts* opts, float* out_t,
                                 int32_t* out_hit_id, int32_t* out_count, uint64_t* request_id);
/* The result's buffers and capacities are those of niq_tree_build; counts are
 * returned by niq_serve_poll. */
niq_status niq_serve_submit_tree(niq_server* srv, const niq_model* m, const niq_eval_cfg* cfg,
                                 const niq_tree_params* p, const niq_tree_result* res,
                                 uint64_t* request_id);
niq_status niq_serve_submit_closest_point(niq_server* srv, const niq_model* m, const niq_eval_cfg* cfg,
                                          const float lower[3], const float upper[3],
                                          const float* query_points, int64_t Q, float eps,
                                          int batch_process_size, float* out_dist, float* out_loc,
                                          uint64_t* request_id);
/* Fills *info (and, for a tree, the counts of *tree_res; may be NULL).  With
 * wait != 0 blocks until the request is done.  A request polled done is
 * forgotten: a later poll of its id returns NIQ_ERR_INVALID.  When the
 * request failed, niq_last_error() on the polling thread gives its message. */
niq_status niq_serve_poll(niq_server* srv, uint64_t request_id, int wait, niq_serve_info* info,
                          niq_tree_result* tree_res);
/* Device bytes the server holds now: admitted requests plus cached graphs. */
size_t niq_serve_bytes_in_use(niq_server* srv);

/* ---- diagnostics -------------------------------------------------------- */
const char* niq_last_error(void);
int niq_version(void);

#ifdef __cplusplus
}
#endif
#endif /* NIQ_H */
