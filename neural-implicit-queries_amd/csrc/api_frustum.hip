// Frustum ray casting entry point (src/queries.py:454-587).
#include "niq_internal.h"
#include "niq_frustum.h"

namespace {

long long next_bucket_size(long long s) {
  for (int e = 7; e <= 30; ++e)
    if (s <= (1LL << e)) return 1LL << e;
  return -1;
}

}  // namespace

extern "C" niq_status niq_cast_rays_frustum(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                            const niq_frustum_cam* cam, const niq_frustum_opts* fo, float* out_t,
                                            int32_t* out_hit_id, int32_t* out_count, int64_t* n_evals,
                                            void* stream) {
  if (!models || n_models < 1 || n_models > 8) return fail(NIQ_ERR_INVALID, "need 1..8 models");
  if (!cam || !fo) return fail(NIQ_ERR_INVALID, "null camera or options");
  const niq_cast_opts& o = fo->cast;
  if (o.n_substeps < 1 || o.n_max_step < 1) return fail(NIQ_ERR_INVALID, "n_substeps and n_max_step must be >= 1");
  const int rx = cam->res_x, ry = cam->res_y, ns = fo->n_side_init;
  if (rx < 1 || ry < 1 || (long long)rx * ry > (1LL << 28)) return fail(NIQ_ERR_INVALID, "resolution out of range");
  if (ns < 1 || ns > rx || ns > ry) return fail(NIQ_ERR_INVALID, "n_side_init must be in [1, min(res_x, res_y)]");
  if (!out_t || !out_hit_id || !out_count) return fail(NIQ_ERR_INVALID, "bad buffers");
  for (int i = 0; i < n_models; ++i)
    if (niq_status st = check_model(models[i], 3)) return st;
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  if (cfg && cfg->precision == NIQ_PREC_BF16 &&
      (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND))
    return fail(NIQ_ERR_UNSUPPORTED, "the precision mode runs interval and affine_fixed");
  hipStream_t s = (hipStream_t)stream;
  const long long cap = (long long)rx * ry;   // frusta partition the image

  FrustCam c{};
  for (int k = 0; k < 3; ++k) {
    c.root[k] = cam->root[k]; c.look[k] = cam->look[k]; c.up[k] = cam->up[k]; c.left[k] = cam->left[k];
  }
  // tan(deg2rad(fov) / 2) and deg2rad(fov) / 2 in float32, as the reference forms them
  const float d2r = (float)(3.14159265358979323846 / 180.0);
  const float hx = (cam->fov_x * d2r) / 2.f;
  const float hy = (cam->fov_y * d2r) / 2.f;
  c.half_fov_x = hx; c.half_fov_y = hy;
  c.tan_x = std::tan(hx); c.tan_y = std::tan(hy);
  c.res_x = rx; c.res_y = ry;
  FrustParams P{};
  P.hit_eps = o.hit_eps; P.max_dist = o.max_dist; P.safety = o.safety_factor;
  P.grow = o.interval_grow_fac; P.shrink = o.interval_shrink_fac; P.refine = fo->refine_width_fac;
  P.n_models = n_models;

  // device scratch (stream ordered)
  std::vector<void*> allocs;
  niq_status st = NIQ_OK;
  auto al = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (st == NIQ_OK && hipMallocAsync(&p, (bytes + 255) / 256 * 256, s) != hipSuccess)
      st = fail(NIQ_ERR_HIP, "frustum scratch allocation failed");
    if (p) allocs.push_back(p);
    return p;
  };
  auto mk_list = [&](FrustList& L) {
    L.rng = static_cast<int4*>(al(sizeof(int4) * cap));
    L.t = static_cast<float*>(al(sizeof(float) * cap));
    L.step = static_cast<float*>(al(sizeof(float) * cap));
    L.cnt = static_cast<float*>(al(sizeof(float) * cap));
  };
  FrustList Ls[2], F;
  mk_list(Ls[0]); mk_list(Ls[1]); mk_list(F);
  int* fin_hit = static_cast<int*>(al(sizeof(int) * cap));
  FrustWork W{};
  W.geo = static_cast<float*>(al(sizeof(float) * 14 * cap));
  W.center = static_cast<float*>(al(sizeof(float) * 3 * cap));
  W.vecs = static_cast<float*>(al(sizeof(float) * 9 * cap));
  W.pts = static_cast<float*>(al(sizeof(float) * 6 * cap));
  W.codes = static_cast<int*>(al(sizeof(int) * cap * n_models));
  W.vals = static_cast<float*>(al(sizeof(float) * 2 * cap * n_models));
  W.is_hit = static_cast<int*>(al(sizeof(int) * cap));
  W.hit_id = static_cast<int*>(al(sizeof(int) * cap));
  W.demands = static_cast<int*>(al(sizeof(int) * cap));
  W.nsteps = static_cast<int*>(al(sizeof(int) * cap));
  W.can_step = static_cast<int*>(al(sizeof(int) * cap));
  FrustCounters* C = static_cast<FrustCounters*>(al(sizeof(FrustCounters)));
  int* ticks = static_cast<int*>(al(sizeof(int) * 2 * (ns + 1)));

  // initial tiles: linspace(0, res, n_side + 1) truncated (src/queries.py:497-505)
  std::vector<int> ht(2 * (ns + 1));
  for (int i = 0; i <= ns; ++i) {
    ht[i] = (int)((double)rx * i / ns);
    ht[ns + 1 + i] = (int)((double)ry * i / ns);
  }
  const int n0 = ns * ns;
  long long n = n0, bucket = n0, empty = n0, evals = 0;
  int it = 0, cur = 0;
  if (st == NIQ_OK && (hipMemcpyAsync(ticks, ht.data(), sizeof(int) * ht.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
                       hipMemsetAsync(C, 0, sizeof(FrustCounters), s) != hipSuccess))
    st = fail(NIQ_ERR_HIP, "frustum set-up failed");
  if (st == NIQ_OK) {
    hipLaunchKernelGGL(k_frust_init, dim3((n0 + 255) / 256), dim3(256), 0, s, Ls[0], ns, ticks, ticks + ns + 1,
                       o.init_step);
  }
  FrustCounters h{};
  while (st == NIQ_OK) {
    evals += bucket;
    const FrustList& L = Ls[cur];
    const unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_frust_geom, dim3(g), dim3(256), 0, s, c, L, W, (int)n);
    for (int sub = 0; sub < o.n_substeps && st == NIQ_OK; ++sub) {
      hipLaunchKernelGGL(k_frust_prep, dim3(g), dim3(256), 0, s, c, P, L, W, (int)n);
      for (int f = 0; f < n_models && st == NIQ_OK; ++f) {
        ClassifyArgs a{};
        a.center = W.center; a.vecs = W.vecs; a.vin = 3; a.d = 3; a.n = n; a.offset = 0.f;
        a.codes = W.codes + (size_t)f * cap;
        if ((st = classify_dispatch(models[f], cfg, mode == NIQ_MODE_INTERVAL ? 0 : 3, a, n, s))) break;
        st = points_dev(models[f], W.pts, nullptr, 2, 2 * n, W.vals + (size_t)f * 2 * cap, s);
      }
      if (st) break;
      hipLaunchKernelGGL(k_frust_sub, dim3(g), dim3(256), 0, s, P, L, W, (int)n, (int)cap);
    }
    if (st) break;
    hipLaunchKernelGGL(k_frust_finish, dim3(g), dim3(256), 0, s, c, P, L, W, (int)n, it >= o.n_max_step ? 1 : 0,
                       Ls[cur ^ 1], F, fin_hit, C);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&h, C, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      st = fail(NIQ_ERR_HIP, "frustum iteration failed");
      break;
    }
    it += o.n_substeps;
    if (h.n_valid == 0) break;
    // the reference's padded-array bookkeeping, for N_evals (src/queries.py:557-569)
    const long long nb = next_bucket_size((long long)h.n_valid + h.n_refine);
    if (empty + h.n_refine > bucket || nb < bucket) { bucket = nb; empty = h.n_valid; }
    empty += h.n_refine;
    n = h.n_next;
    cur ^= 1;
    FrustCounters z{0, h.n_fin, 0, 0};
    if (hipMemcpyAsync(C, &z, sizeof(z), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      st = fail(NIQ_ERR_HIP, "frustum counter reset failed");
      break;
    }
  }
  if (st == NIQ_OK && h.n_fin > 0) {
    hipLaunchKernelGGL(k_frust_write, dim3((unsigned)std::min<long long>(h.n_fin, 65536)), dim3(256), 0, s, F,
                       fin_hit, h.n_fin, ry, out_t, out_hit_id, out_count);
    if (hipGetLastError() != hipSuccess) st = fail(NIQ_ERR_HIP, "frustum output write failed");
  }
  for (void* p : allocs) (void)hipFreeAsync(p, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, "frustum stream failed");
  if (st == NIQ_OK && n_evals) *n_evals = evals;
  return st;
}
