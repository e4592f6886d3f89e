// Persistent ray caster entry point.
#include "niq_internal.h"
#include "niq_rays_bf16.h"

extern "C" {

// ---------------------------------------------------------- ray casting
size_t niq_cast_workspace_bytes(const niq_cast_opts* opts) {
  const int n_max = opts ? std::max(opts->n_max_step, 1) : 512;
  size_t b = 256;                                  // work counter
  b += ((size_t)(n_max + 2) * sizeof(int) + 255) / 256 * 256;
  b += 8 * sizeof(DevModel);                       // model table
  b = (b + 255) / 256 * 256;
  b += 8 * sizeof(DevModelBf16);                   // precision-mode model table
  return b;
}

}  // extern "C"

namespace {

template <int KIND, int V, int W>
niq_status launch_rays_w(const DevModel* models, const RayArgs& a, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  constexpr int R = Planes<KIND, V, KIND == 2 ? 3 : 2>::R;
  if constexpr (!Fits<R, NB, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "ray range form does not fit in LDS at this width");
  } else {
    auto k = k_cast_rays<KIND, V, NB, NW, W>;
    const long long tiles = (a.n + NB - 1) / NB;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, models, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int KIND, int V>
niq_status launch_rays(int wt, const DevModel* models, const RayArgs& a, hipStream_t s) {
  switch (wt) {
    case 64: return launch_rays_w<KIND, V, 64>(models, a, s);
    case 128: return launch_rays_w<KIND, V, 128>(models, a, s);
    case 256: return launch_rays_w<KIND, V, 256>(models, a, s);
    case 512: return launch_rays_w<KIND, V, 512>(models, a, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

template <int V, int W>
niq_status launch_rays_bf16_w(const DevModelBf16* models, const RayArgs& a, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  if constexpr (!RayPlanesBf16<V, NB, W>::fits) {
    return fail(NIQ_ERR_UNSUPPORTED, "precision-mode ray form does not fit in LDS at this width");
  } else {
    auto k = k_cast_rays_bf16<V, NB, NW, W>;
    const long long tiles = (a.n + NB - 1) / NB;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, models, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int V>
niq_status launch_rays_bf16(int wt, const DevModelBf16* models, const RayArgs& a, hipStream_t s) {
  switch (wt) {
    case 64: return launch_rays_bf16_w<V, 64>(models, a, s);
    case 128: return launch_rays_bf16_w<V, 128>(models, a, s);
    case 256: return launch_rays_bf16_w<V, 256>(models, a, s);
    case 512: return launch_rays_bf16_w<V, 512>(models, a, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

long long next_bucket(long long s) {
  for (int e = 7; e <= 30; ++e)
    if (s <= (1LL << e)) return 1LL << e;
  return -1;
}

}  // namespace

extern "C" niq_status niq_cast_rays(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                    const float* roots, const float* dirs, int64_t n, const niq_cast_opts* o,
                                    float* out_t, int32_t* out_hit_id, int32_t* out_count, void* ws,
                                    size_t ws_bytes, int64_t* n_evals, void* stream) {
  if (!models || n_models < 1 || n_models > 8) return fail(NIQ_ERR_INVALID, "need 1..8 models");
  if (!o) return fail(NIQ_ERR_INVALID, "null opts");
  if (o->n_substeps < 1 || o->n_max_step < 1) return fail(NIQ_ERR_INVALID, "n_substeps and n_max_step must be >= 1");
  if (n < 0 || n > (1LL << 31) - 1) return fail(NIQ_ERR_INVALID, "ray count out of range");
  if (n > 0 && (!roots || !dirs || !out_t || !out_hit_id || !out_count)) return fail(NIQ_ERR_INVALID, "bad buffers");
  const size_t need = niq_cast_workspace_bytes(o);
  if (!ws || ws_bytes < need) return fail(NIQ_ERR_CAPACITY, "cast workspace too small: need " + std::to_string(need));
  int wt = 0, wt16 = 0;
  for (int i = 0; i < n_models; ++i) {
    if (niq_status st = check_model(models[i], 3)) return st;
    wt = std::max(wt, models[i]->wt);
    wt16 = std::max(wt16, models[i]->wt16);
  }
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const bool bf16 = cfg && cfg->precision == NIQ_PREC_BF16;
  if (cfg && cfg->precision != NIQ_PREC_FP32 && !bf16) return fail(NIQ_ERR_INVALID, "unknown precision");
  if (bf16) {
    if (mode != NIQ_MODE_AFFINE_FIXED && mode != NIQ_MODE_INTERVAL)
      return fail(NIQ_ERR_UNSUPPORTED, "precision-mode cast_rays supports interval and affine_fixed");
    for (int i = 0; i < n_models; ++i)
      if (!models[i]->bf16_ok) return fail(NIQ_ERR_UNSUPPORTED, "the precision mode needs weights without FP32 subnormals");
  }
  hipStream_t s = (hipStream_t)stream;
  unsigned char* base = static_cast<unsigned char*>(ws);
  int* counter = reinterpret_cast<int*>(base);
  int* hist = reinterpret_cast<int*>(base + 256);
  const size_t hist_bytes = ((size_t)(o->n_max_step + 2) * sizeof(int) + 255) / 256 * 256;
  DevModel* table = reinterpret_cast<DevModel*>(base + 256 + hist_bytes);
  const size_t off16 = (256 + hist_bytes + 8 * sizeof(DevModel) + 255) / 256 * 256;
  DevModelBf16* table16 = reinterpret_cast<DevModelBf16*>(base + off16);
  std::vector<DevModel> host_table(n_models);
  std::vector<DevModelBf16> host_table16(bf16 ? n_models : 0);
  for (int i = 0; i < n_models; ++i) host_table[i] = models[i]->host;
  for (int i = 0; i < (int)host_table16.size(); ++i) host_table16[i] = models[i]->host16;
  HIPCHK(hipMemsetAsync(base, 0, 256 + hist_bytes, s));
  HIPCHK(hipMemcpyAsync(table, host_table.data(), sizeof(DevModel) * n_models, hipMemcpyHostToDevice, s));
  if (bf16)
    HIPCHK(hipMemcpyAsync(table16, host_table16.data(), sizeof(DevModelBf16) * n_models, hipMemcpyHostToDevice, s));
  if (n > 0) {
    RayArgs a{};
    a.roots = roots; a.dirs = dirs; a.n = n; a.n_models = n_models;
    a.hit_eps = o->hit_eps; a.max_dist = o->max_dist; a.safety = o->safety_factor;
    a.grow = o->interval_grow_fac; a.shrink = o->interval_shrink_fac; a.init_size = o->init_step;
    a.n_max_step = o->n_max_step; a.n_substeps = o->n_substeps;
    a.out_t = out_t; a.out_hit = out_hit_id; a.out_count = out_count;
    a.next_ray = counter; a.iter_hist = hist;
    niq_status st;
    if (bf16) st = mode == NIQ_MODE_AFFINE_FIXED ? launch_rays_bf16<1>(wt16, table16, a, s)
                                                 : launch_rays_bf16<0>(wt16, table16, a, s);
    else if (mode == NIQ_MODE_AFFINE_FIXED) st = launch_rays<0, 1>(wt, table, a, s);
    else if (mode == NIQ_MODE_INTERVAL) st = launch_rays<0, 0>(wt, table, a, s);
    else if (mode == NIQ_MODE_SLOPE_INTERVAL) st = launch_rays<1, 1>(wt, table, a, s);
    else if (mode == NIQ_MODE_SDF) {
      a.lipschitz = cfg->sdf_lipschitz;
      st = launch_rays<2, 0>(wt, table, a, s);
    } else return fail(NIQ_ERR_UNSUPPORTED, "cast_rays supports interval, affine_fixed, slope_interval and sdf");
    if (st) return st;
  }
  std::vector<int> h(o->n_max_step + 2, 0);
  HIPCHK(hipMemcpyAsync(h.data(), hist, sizeof(int) * h.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // Reproduce the reference's N_evals: each lock-step iteration evaluates the
  // whole (bucketed) working array n_substeps times (src/queries.py:164,170-173).
  if (n_evals) {
    long long evals = 0, bucket = n, live = n;
    for (int it = 1; it <= o->n_max_step + 1 && live > 0; ++it) {
      evals += bucket * o->n_substeps;
      live -= h[it];
      if (live > 0) {
        const long long nbkt = next_bucket(live);
        if (nbkt > 0 && nbkt < bucket) bucket = nbkt;
      }
    }
    *n_evals = evals;
  }
  return NIQ_OK;
}

