// Per-box classification and point evaluation entry points.
#include "niq_internal.h"
#include "niq_bf16.h"
#include "niq_rows.h"

// ------------------------------------------------------------- dispatch
namespace {

// Launch k_classify for (KIND, V) at template width W.
template <int KIND, int V, int W>
niq_status launch_classify_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  constexpr int R = Planes<KIND, V, 0>::R;
  if constexpr (!Fits<R, NB, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "range form does not fit in LDS at this width");
  } else {
    auto k = k_classify<KIND, V, NB, NW, W>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, a);
    HIPCHK(hipGetLastError());
    if (a.recheck_tol > 0.f) {
      auto k64 = k_recheck_fp64<KIND, V, W>;
      const unsigned g64 = grid_for(std::min<long long>(max_n, a.recheck_cap), persistent_blocks(k64, 256));
      hipLaunchKernelGGL(k64, dim3(g64), dim3(256), 0, s, m->dev, a);
      HIPCHK(hipGetLastError());
    }
    return NIQ_OK;
  }
}

template <int KIND, int V>
niq_status launch_classify(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  switch (m->wt) {
    case 64: return launch_classify_w<KIND, V, 64>(m, a, max_n, s);
    case 128: return launch_classify_w<KIND, V, 128>(m, a, max_n, s);
    case 256: return launch_classify_w<KIND, V, 256>(m, a, max_n, s);
    case 512: return launch_classify_w<KIND, V, 512>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

template <int V, int W, int WS>
niq_status launch_bf16_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  if constexpr ((size_t)PlanesBf16<V>::NPL * NB * W * 2 > (size_t)kLdsLimit) {
    return fail(NIQ_ERR_UNSUPPORTED, "precision-mode range form does not fit in LDS at this width");
  } else {
    auto k = k_classify_bf16<V, NB, NW, W, WS>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev16, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int V, int WS>
niq_status launch_bf16_ws(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  switch (m->wt16) {
    case 64: return launch_bf16_w<V, 64, WS>(m, a, max_n, s);
    case 128: return launch_bf16_w<V, 128, WS>(m, a, max_n, s);
    case 256: return launch_bf16_w<V, 256, WS>(m, a, max_n, s);
    case 512: return launch_bf16_w<V, 512, WS>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// Weight source: cfg->variant = 1 streams the four packed bf16 copies, 2 (and
// the default) the FP32 weights split in registers.
template <int V>
niq_status launch_bf16(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                       hipStream_t s) {
  if (!m->bf16_ok) return fail(NIQ_ERR_UNSUPPORTED, "the precision mode needs weights without FP32 subnormals");
  const int ws = cfg ? cfg->variant : 0;
  if (ws == 1) return launch_bf16_ws<V, WSRC_BF16>(m, a, max_n, s);
  return launch_bf16_ws<V, WSRC_FP32>(m, a, max_n, s);
}

// affine_all / affine_truncate / affine_append: one box per workgroup, error
// terms streamed through a private HBM scratch.
template <int W, int MODE>
niq_status launch_rows_w(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                         hipStream_t s) {
  constexpr int NW = RowsCfg<W>::NW;
  auto k = k_classify_rows<W, MODE>;
  const long long d0 = std::max(a.d, a.lower ? 0 : a.vin);
  const int n_keep = cfg ? cfg->truncate_count : 0;
  const int n_append = cfg ? cfg->n_append : 0;
  long long max_rows = d0;
  int n_act = 0;
  for (int i = 0; i + 1 < m->n_layers; ++i)
    if (m->acts[i] != NIQ_ACT_NONE) { max_rows += m->host.L[i].out_pad; ++n_act; }
  if (MODE == ROWS_APPEND) max_rows = d0 + (long long)n_act * n_append;
  if (MODE == ROWS_TRUNCATE) max_rows = d0 + n_keep + 2LL * W;
  max_rows = std::max<long long>(max_rows, 1);
  const size_t per_wg = sizeof(float) * 2 * (size_t)max_rows * W;
  if (max_n == 0) return NIQ_OK;
  long long g = std::min<long long>(max_n, persistent_blocks(k, NW * 64));
  g = std::max<long long>(1, std::min<long long>(g, (long long)((16ull << 30) / per_wg)));
  void* scratch = nullptr;
  HIPCHK(hipMallocAsync(&scratch, per_wg * g, s));
  RowsArgs ra{};
  ra.c = a;
  ra.scratch = static_cast<float*>(scratch);
  ra.max_rows = max_rows;
  ra.n_keep = n_keep;
  ra.n_append = n_append;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(NW * 64), 0, s, m->dev, ra);
  const hipError_t e = hipGetLastError();
  HIPCHK(hipFreeAsync(scratch, s));
  if (e != hipSuccess) return fail(NIQ_ERR_HIP, std::string("k_classify_rows: ") + hipGetErrorString(e));
  return NIQ_OK;
}

template <int MODE>
niq_status launch_rows(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                       hipStream_t s) {
  switch (m->wt) {
    case 64: return launch_rows_w<64, MODE>(m, cfg, a, max_n, s);
    case 128: return launch_rows_w<128, MODE>(m, cfg, a, max_n, s);
    case 256: return launch_rows_w<256, MODE>(m, cfg, a, max_n, s);
    case 512: return launch_rows_w<512, MODE>(m, cfg, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// Route a classification by (mode, precision, number of box vectors).
}  // namespace

niq_status niqi::classify_dispatch(const niq_model* m, const niq_eval_cfg* cfg, int nvec,
                                   const ClassifyArgs& a, long long max_n, hipStream_t s) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const int prec = cfg ? cfg->precision : NIQ_PREC_FP32;
  if (prec == NIQ_PREC_BF16) {
    ClassifyArgs b = a;
    b.recheck_tol = 0.f;   // the precision mode is rigorous by construction
    if (mode == NIQ_MODE_INTERVAL) return launch_bf16<0>(m, cfg, b, max_n, s);
    if (mode == NIQ_MODE_AFFINE_FIXED) {
      switch (nvec) {
        case 1: return launch_bf16<1>(m, cfg, b, max_n, s);
        case 2: return launch_bf16<2>(m, cfg, b, max_n, s);
        case 3: return launch_bf16<3>(m, cfg, b, max_n, s);
      }
    }
    return fail(NIQ_ERR_UNSUPPORTED, "the precision mode supports interval and affine_fixed with 1..3 box vectors");
  }
  if (prec != NIQ_PREC_FP32) return fail(NIQ_ERR_INVALID, "unknown precision");
  if (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND) {
    if (a.recheck_tol > 0.f) return fail(NIQ_ERR_UNSUPPORTED, "the growing-row modes have no FP64 re-test");
    if (!a.lower && a.vin > 64) return fail(NIQ_ERR_UNSUPPORTED, "at most 64 box vectors");
    if (mode == NIQ_MODE_AFFINE_ALL) return launch_rows<ROWS_ALL>(m, cfg, a, max_n, s);
    if (mode == NIQ_MODE_AFFINE_TRUNCATE) {
      if (!cfg || cfg->truncate_count < 1 || cfg->truncate_count > 64)
        return fail(NIQ_ERR_INVALID, "affine_truncate needs 1 <= truncate_count <= 64");
      return launch_rows<ROWS_TRUNCATE>(m, cfg, a, max_n, s);
    }
    if (!cfg || cfg->n_append < 0 || cfg->n_append > 512)
      return fail(NIQ_ERR_INVALID, "affine_append needs 0 <= n_append <= 512");
    return launch_rows<ROWS_APPEND>(m, cfg, a, max_n, s);
  }
  if (mode == NIQ_MODE_INTERVAL) return launch_classify<0, 0>(m, a, max_n, s);
  if (mode == NIQ_MODE_AFFINE_FIXED) {
    switch (nvec) {
      case 1: return launch_classify<0, 1>(m, a, max_n, s);
      case 2: return launch_classify<0, 2>(m, a, max_n, s);
      case 3: return launch_classify<0, 3>(m, a, max_n, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "affine_fixed supports 1..3 box vectors");
  }
  if (mode == NIQ_MODE_SLOPE_INTERVAL) {
    switch (nvec) {
      case 1: return launch_classify<1, 1>(m, a, max_n, s);
      case 2: return launch_classify<1, 2>(m, a, max_n, s);
      case 3: return launch_classify<1, 3>(m, a, max_n, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "slope_interval supports 1..3 box vectors");
  }
  return fail(NIQ_ERR_UNSUPPORTED, "mode not available in the fused kernel");
}

namespace {
template <int W>
niq_status launch_points_w(const niq_model* m, const float* x, int d, long long n, float* out, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_points<NB, NW, W>;
  const long long tiles = (n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, x, d, n, out);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

}  // namespace

namespace niqi {
niq_status classify_with_recheck(const niq_model* m, const niq_eval_cfg* cfg, int nvec, ClassifyArgs a,
                                 long long n, hipStream_t s) {
  const float tol = cfg ? cfg->recheck_tol : 0.f;
  if (!(tol > 0.f) || n == 0) return classify_dispatch(m, cfg, nvec, a, n, s);
  if (n > INT_MAX) return fail(NIQ_ERR_UNSUPPORTED, "FP64 re-test supports at most 2^31-1 boxes per call");
  void* scratch = nullptr;
  HIPCHK(hipMallocAsync(&scratch, sizeof(int) * (size_t)(n + 64), s));
  HIPCHK(hipMemsetAsync(scratch, 0, sizeof(int), s));
  a.recheck_tol = tol;
  a.recheck_cnt = static_cast<int*>(scratch);
  a.recheck_idx = static_cast<int*>(scratch) + 64;
  a.recheck_cap = (int)n;
  niq_status st = classify_dispatch(m, cfg, nvec, a, n, s);
  HIPCHK(hipFreeAsync(scratch, s));
  return st;
}
}  // namespace

extern "C" {

niq_status niq_classify_boxes(const niq_model* m, const niq_eval_cfg* cfg, const float* lower,
                              const float* upper, int64_t n, int d, float offset, int32_t* codes,
                              float* bounds, void* stream) {
  if (niq_status st = check_model(m, d)) return st;
  if (n < 0 || (n > 0 && (!lower || !upper || !codes))) return fail(NIQ_ERR_INVALID, "bad buffers");
  ClassifyArgs a{};
  a.lower = lower; a.upper = upper; a.d = d; a.n = n; a.offset = offset;
  a.codes = codes; a.bounds = bounds;
  return classify_with_recheck(m, cfg, d, a, n, (hipStream_t)stream);
}

niq_status niq_classify_general_boxes(const niq_model* m, const niq_eval_cfg* cfg, const float* center,
                                      const float* vecs, int64_t n, int v, int d, float offset,
                                      int32_t* codes, float* bounds, void* stream) {
  if (niq_status st = check_model(m, d)) return st;
  if (v < 1) return fail(NIQ_ERR_INVALID, "need at least one box vector");
  if (n < 0 || (n > 0 && (!center || !vecs || !codes))) return fail(NIQ_ERR_INVALID, "bad buffers");
  ClassifyArgs a{};
  a.center = center; a.vecs = vecs; a.vin = v; a.d = d; a.n = n; a.offset = offset;
  a.codes = codes; a.bounds = bounds;
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  return classify_with_recheck(m, cfg, mode == NIQ_MODE_INTERVAL ? 0 : v, a, n, (hipStream_t)stream);
}

niq_status niq_eval_points(const niq_model* m, const float* x, int64_t n, int d, float* out, void* stream) {
  if (niq_status st = check_model(m, d)) return st;
  if (n < 0 || (n > 0 && (!x || !out))) return fail(NIQ_ERR_INVALID, "bad buffers");
  hipStream_t s = (hipStream_t)stream;
  switch (m->wt) {
    case 64: return launch_points_w<64>(m, x, d, n, out, s);
    case 128: return launch_points_w<128>(m, x, d, n, out, s);
    case 256: return launch_points_w<256>(m, x, d, n, out, s);
    case 512: return launch_points_w<512>(m, x, d, n, out, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}


}  // extern "C"
