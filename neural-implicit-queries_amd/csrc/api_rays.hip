// Persistent ray caster entry point.
#include "niq_internal.h"
#include "niq_rays_bf16.h"

extern "C" {

// ---------------------------------------------------------- ray casting
// A one-group cast with up to 8 functions (the layout of cast_layout).
size_t niq_cast_workspace_bytes(const niq_cast_opts* opts) {
  CastGroupReq r{};
  r.n_models = 8;
  r.opts.n_max_step = opts ? std::max(opts->n_max_step, 1) : 512;
  return cast_layout(&r, 1, true).total;
}

}  // extern "C"

namespace {

template <int KIND, int V, int W, int VAR, int AS = ACTS_ALL>
niq_status launch_rays_w(const DevModel* models, const RayArgs& a, long long n_rays, hipStream_t s) {
  constexpr int NB = RayTile<W, VAR>::NB, NW = RayTile<W, VAR>::NW;
  constexpr int R = Planes<KIND, V, KIND == 2 ? 3 : 2>::R;
  constexpr int LDSOCC = LdsOcc<R, NB, NW, W>::value;
  constexpr int OCC = RayTile<W, VAR>::OCC ? std::min(RayTile<W, VAR>::OCC, LDSOCC) : LDSOCC;
  if constexpr (!Fits<R, NB, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "ray range form does not fit in LDS at this width");
  } else {
    auto k = k_cast_rays<KIND, V, NB, NW, W, OCC, AS>;
    const long long tiles = (n_rays + NB - 1) / NB;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, models, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int KIND, int V, int VAR, int AS = ACTS_ALL>
niq_status launch_rays_v(int wt, const DevModel* models, const RayArgs& a, long long n_rays, hipStream_t s) {
  switch (wt) {
    case 64: return launch_rays_w<KIND, V, 64, VAR, AS>(models, a, n_rays, s);
    case 128: return launch_rays_w<KIND, V, 128, VAR, AS>(models, a, n_rays, s);
    case 256: return launch_rays_w<KIND, V, 256, VAR, AS>(models, a, n_rays, s);
    case 512: return launch_rays_w<KIND, V, 512, VAR, AS>(models, a, n_rays, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// relu_only: every model of the cast has only ReLU hidden layers; the default
// affine_fixed tile then runs the kernel compiled for ReLU alone
// (bit-identical, fewer registers).
template <int KIND, int V>
niq_status launch_rays(int wt, int variant, bool relu_only, const DevModel* models, const RayArgs& a,
                       long long n_rays, hipStream_t s) {
  if constexpr (KIND == 0 && V == 1) {
    if (variant == 0 && relu_only) return launch_rays_v<KIND, V, 0, ACTS_RELU>(wt, models, a, n_rays, s);
  }
  switch (variant) {
    case 0: return launch_rays_v<KIND, V, 0>(wt, models, a, n_rays, s);
    case 1: return launch_rays_v<KIND, V, 1>(wt, models, a, n_rays, s);
    case 2: return launch_rays_v<KIND, V, 2>(wt, models, a, n_rays, s);
  }
  return fail(NIQ_ERR_INVALID, "ray_variant must be 0, 1 or 2");
}

template <int V, int W, int AS = ACTS_ALL>
niq_status launch_rays_bf16_w(const DevModelBf16* models, const RayArgs& a, long long n_rays, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  if constexpr (!RayPlanesBf16<V, NB, W>::fits) {
    return fail(NIQ_ERR_UNSUPPORTED, "precision-mode ray form does not fit in LDS at this width");
  } else {
    auto k = k_cast_rays_bf16<V, NB, NW, W, AS>;
    const long long tiles = (n_rays + NB - 1) / NB;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, models, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int V>
niq_status launch_rays_bf16(int wt, bool relu_only, const DevModelBf16* models, const RayArgs& a, long long n_rays,
                            hipStream_t s) {
  if (relu_only) {
    switch (wt) {
      case 64: return launch_rays_bf16_w<V, 64, ACTS_RELU>(models, a, n_rays, s);
      case 128: return launch_rays_bf16_w<V, 128, ACTS_RELU>(models, a, n_rays, s);
      case 256: return launch_rays_bf16_w<V, 256, ACTS_RELU>(models, a, n_rays, s);
      case 512: return launch_rays_bf16_w<V, 512, ACTS_RELU>(models, a, n_rays, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "width");
  }
  switch (wt) {
    case 64: return launch_rays_bf16_w<V, 64>(models, a, n_rays, s);
    case 128: return launch_rays_bf16_w<V, 128>(models, a, n_rays, s);
    case 256: return launch_rays_bf16_w<V, 256>(models, a, n_rays, s);
    case 512: return launch_rays_bf16_w<V, 512>(models, a, n_rays, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// Zero the per-group counters, N_evals and histograms: the first `words`
// 32-bit words of the workspace (a kernel, not a memset, so the sequence
// stays correct inside captured graphs).
__global__ void k_cast_init(unsigned char* base, long long words) {
  int* w = reinterpret_cast<int*>(base);
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < words;
       k += (long long)gridDim.x * blockDim.x)
    w[k] = 0;
}

__device__ long long next_bucket_dev(long long s) {
  for (int e = 7; e <= 30; ++e)
    if (s <= (1LL << e)) return 1LL << e;
  return -1;
}

// The reference's N_evals of each group from its histogram of iterations per
// ray: each lock-step iteration evaluates the whole (bucketed) working array
// n_substeps times (src/queries.py:164,170-173).  One thread per group.
__global__ void k_cast_n_evals(const RayGroup* groups, int ng, long long* heads) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const RayGroup G = groups[g];
  long long evals = 0, bucket = G.n, live = G.n;
  for (int it = 1; it <= G.n_max_step + 1 && live > 0; ++it) {
    evals += bucket * G.n_substeps;
    live -= G.iter_hist[it];
    if (live > 0) {
      const long long nbkt = next_bucket_dev(live);
      if (nbkt > 0 && nbkt < bucket) bucket = nbkt;
    }
  }
  heads[2 * g + 1] = evals;
}

}  // namespace

namespace niqi {

// Every check of a cast, done before anything is enqueued.
niq_status cast_check(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg, const float* roots,
                      const float* dirs, int64_t n, const niq_cast_opts* o, const float* out_t,
                      const int32_t* out_hit_id, const int32_t* out_count, int* wt, int* wt16, bool allow_rows) {
  if (!models || n_models < 1 || n_models > 8) return fail(NIQ_ERR_INVALID, "need 1..8 models");
  if (!o) return fail(NIQ_ERR_INVALID, "null opts");
  if (o->n_substeps < 1 || o->n_max_step < 1) return fail(NIQ_ERR_INVALID, "n_substeps and n_max_step must be >= 1");
  if (n < 0 || n > (1LL << 31) - 1) return fail(NIQ_ERR_INVALID, "ray count out of range");
  if (n > 0 && (!roots || !dirs || !out_t || !out_hit_id || !out_count)) return fail(NIQ_ERR_INVALID, "bad buffers");
  *wt = 0; *wt16 = 0;
  for (int i = 0; i < n_models; ++i) {
    if (niq_status st = check_model(models[i], 3)) return st;
    *wt = std::max(*wt, models[i]->wt);
    *wt16 = std::max(*wt16, models[i]->wt16);
  }
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const bool bf16 = cfg && cfg->precision == NIQ_PREC_BF16;
  if (cfg && cfg->precision != NIQ_PREC_FP32 && !bf16) return fail(NIQ_ERR_INVALID, "unknown precision");
  if (cfg && (cfg->ray_variant < 0 || cfg->ray_variant > 2)) return fail(NIQ_ERR_INVALID, "ray_variant must be 0, 1 or 2");
  if (bf16) {
    if (mode != NIQ_MODE_AFFINE_FIXED && mode != NIQ_MODE_INTERVAL)
      return fail(NIQ_ERR_UNSUPPORTED, "precision-mode cast_rays supports interval and affine_fixed");
    for (int i = 0; i < n_models; ++i)
      if (!models[i]->bf16_ok) return fail(NIQ_ERR_UNSUPPORTED, "the precision mode needs weights without FP32 subnormals");
  } else if (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND) {
    if (!allow_rows)
      return fail(NIQ_ERR_UNSUPPORTED, "the growing-row modes allocate per-ray scratch: use niq_cast_rays");
  } else if (mode != NIQ_MODE_AFFINE_FIXED && mode != NIQ_MODE_INTERVAL && mode != NIQ_MODE_SLOPE_INTERVAL &&
             mode != NIQ_MODE_SDF) {
    return fail(NIQ_ERR_INVALID, "unknown range-analysis mode");
  }
  return NIQ_OK;
}

CastLayout cast_layout(const CastGroupReq* reqs, int ng, bool with16) {
  CastLayout L{};
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t off = up(16 * (size_t)ng);   // per group: work counter, N_evals
  int n_models = 0;
  for (int g = 0; g < ng; ++g) {
    L.hist[g] = off;
    off += up(((size_t)std::max(reqs[g].opts.n_max_step, 1) + 2) * sizeof(int));
    n_models += reqs[g].n_models;
  }
  L.zero_end = off;
  L.models = off;
  off = up(off + (size_t)n_models * sizeof(DevModel));
  L.models16 = off;
  if (with16) off = up(off + (size_t)n_models * sizeof(DevModelBf16));
  L.groups = off;
  off = up(off + (size_t)ng * sizeof(RayGroup));
  L.total = off;
  return L;
}

niq_status cast_run_groups(CastGroupReq* reqs, int ng, const niq_eval_cfg* cfg, unsigned char* base, int wt,
                           int wt16, hipStream_t s) {
  if (ng < 1 || ng > kMaxRayGroups) return fail(NIQ_ERR_INVALID, "ray group count out of range");
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const bool bf16 = cfg && cfg->precision == NIQ_PREC_BF16;
  const int rv = cfg ? cfg->ray_variant : 0;
  if (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND) {
    for (int g = 0; g < ng; ++g) {
      CastGroupReq& r = reqs[g];
      if (niq_status st = cast_rays_rows(r.models, r.n_models, cfg, r.roots, r.dirs, r.n, &r.opts, r.out_t, r.out_hit,
                                         r.out_count, &r.n_evals, s))
        return st;
    }
    return NIQ_OK;
  }
  const CastLayout L = cast_layout(reqs, ng, bf16);
  long long* heads = reinterpret_cast<long long*>(base);
  DevModel* table = reinterpret_cast<DevModel*>(base + L.models);
  DevModelBf16* table16 = reinterpret_cast<DevModelBf16*>(base + L.models16);
  RayGroup* gtab = reinterpret_cast<RayGroup*>(base + L.groups);
  // host-side tables: fixed arrays for the common single-group call, a
  // vector only when groups are merged
  DevModel ht_fix[8];
  DevModelBf16 ht16_fix[8];
  RayGroup hg_fix[1];
  std::vector<DevModel> ht_vec;
  std::vector<DevModelBf16> ht16_vec;
  std::vector<RayGroup> hg_vec;
  int n_models = 0;
  for (int g = 0; g < ng; ++g) n_models += reqs[g].n_models;
  DevModel* ht = ht_fix;
  DevModelBf16* ht16 = ht16_fix;
  RayGroup* hg = hg_fix;
  if (n_models > 8 || ng > 1) {
    ht_vec.resize(n_models);
    ht16_vec.resize(bf16 ? n_models : 0);
    hg_vec.resize(ng);
    ht = ht_vec.data(); ht16 = ht16_vec.data(); hg = hg_vec.data();
  }
  long long n_rays = 0;
  int m0 = 0;
  bool relu = true;
  for (int g = 0; g < ng; ++g) {
    const CastGroupReq& r = reqs[g];
    for (int f = 0; f < r.n_models; ++f) {
      relu = relu && r.models[f]->relu_only;
      ht[m0 + f] = r.models[f]->host;
      if (bf16) ht16[m0 + f] = r.models[f]->host16;
    }
    RayGroup G{};
    G.roots = r.roots; G.dirs = r.dirs; G.n = r.n;
    G.model0 = m0; G.n_models = r.n_models;
    G.hit_eps = r.opts.hit_eps; G.max_dist = r.opts.max_dist; G.safety = r.opts.safety_factor;
    G.grow = r.opts.interval_grow_fac; G.shrink = r.opts.interval_shrink_fac; G.init_size = r.opts.init_step;
    G.n_max_step = r.opts.n_max_step; G.n_substeps = r.opts.n_substeps;
    G.out_t = r.out_t; G.out_hit = r.out_hit; G.out_count = r.out_count;
    G.next_ray = reinterpret_cast<unsigned long long*>(heads + 2 * g);
    G.iter_hist = reinterpret_cast<int*>(base + L.hist[g]);
    G.lipschitz = cfg ? cfg->sdf_lipschitz : 1.f;
    hg[g] = G;
    m0 += r.n_models;
    n_rays += r.n;
  }
  niq_status st = NIQ_OK;
  auto hipfail = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  const long long words = (long long)(L.zero_end / 4);
  hipLaunchKernelGGL(k_cast_init, dim3((unsigned)std::min<long long>((words + 255) / 256, 256)), dim3(256), 0, s,
                     base, words);
  hipfail(hipGetLastError(), "k_cast_init");
  if (st == NIQ_OK)
    hipfail(hipMemcpyAsync(table, ht, sizeof(DevModel) * n_models, hipMemcpyHostToDevice, s), "model table");
  if (st == NIQ_OK && bf16)
    hipfail(hipMemcpyAsync(table16, ht16, sizeof(DevModelBf16) * n_models, hipMemcpyHostToDevice, s), "model table");
  if (st == NIQ_OK)
    hipfail(hipMemcpyAsync(gtab, hg, sizeof(RayGroup) * ng, hipMemcpyHostToDevice, s), "group table");
  if (st == NIQ_OK && n_rays > 0) {
    RayArgs a{};
    a.groups = gtab;
    a.n_groups = ng;
    if (bf16) st = mode == NIQ_MODE_AFFINE_FIXED ? launch_rays_bf16<1>(wt16, relu, table16, a, n_rays, s)
                                                 : launch_rays_bf16<0>(wt16, relu, table16, a, n_rays, s);
    else if (mode == NIQ_MODE_AFFINE_FIXED) st = launch_rays<0, 1>(wt, rv, relu, table, a, n_rays, s);
    else if (mode == NIQ_MODE_INTERVAL) st = launch_rays<0, 0>(wt, rv, relu, table, a, n_rays, s);
    else if (mode == NIQ_MODE_SLOPE_INTERVAL) st = launch_rays<1, 1>(wt, rv, relu, table, a, n_rays, s);
    else st = launch_rays<2, 0>(wt, rv, relu, table, a, n_rays, s);
  }
  long long hb_fix[2 * 1];
  std::vector<long long> hb_vec;
  long long* hb = hb_fix;
  if (ng > 1) {
    hb_vec.resize(2 * ng);
    hb = hb_vec.data();
  }
  if (st == NIQ_OK) {
    hipLaunchKernelGGL(k_cast_n_evals, dim3((ng + 63) / 64), dim3(64), 0, s, gtab, ng, heads);
    hipfail(hipGetLastError(), "k_cast_n_evals");
  }
  if (st == NIQ_OK)
    hipfail(hipMemcpyAsync(hb, heads, sizeof(long long) * 2 * ng, hipMemcpyDeviceToHost, s), "N_evals");
  // the host tables on this frame must outlive their copies: wait on every path
  hipfail(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (st == NIQ_OK)
    for (int g = 0; g < ng; ++g) reqs[g].n_evals = hb[2 * g + 1];
  return st;
}

niq_status cast_run(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg, const float* roots,
                    const float* dirs, int64_t n, const niq_cast_opts* o, float* out_t, int32_t* out_hit_id,
                    int32_t* out_count, unsigned char* base, int wt, int wt16, int64_t* n_evals, hipStream_t s) {
  CastGroupReq r{};
  r.models = models; r.n_models = n_models;
  r.roots = roots; r.dirs = dirs; r.n = n; r.opts = *o;
  r.out_t = out_t; r.out_hit = out_hit_id; r.out_count = out_count;
  const niq_status st = cast_run_groups(&r, 1, cfg, base, wt, wt16, s);
  if (st == NIQ_OK && n_evals) *n_evals = r.n_evals;
  return st;
}

}  // namespace niqi

extern "C" niq_status niq_cast_rays(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                    const float* roots, const float* dirs, int64_t n, const niq_cast_opts* o,
                                    float* out_t, int32_t* out_hit_id, int32_t* out_count, void* ws,
                                    size_t ws_bytes, int64_t* n_evals, void* stream) {
  int wt = 0, wt16 = 0;
  if (niq_status st = cast_check(models, n_models, cfg, roots, dirs, n, o, out_t, out_hit_id, out_count, &wt, &wt16))
    return st;
  const size_t need = niq_cast_workspace_bytes(o);
  if (!ws || ws_bytes < need) return fail(NIQ_ERR_CAPACITY, "cast workspace too small: need " + std::to_string(need));
  return cast_run(models, n_models, cfg, roots, dirs, n, o, out_t, out_hit_id, out_count,
                  static_cast<unsigned char*>(ws), wt, wt16, n_evals, (hipStream_t)stream);
}

