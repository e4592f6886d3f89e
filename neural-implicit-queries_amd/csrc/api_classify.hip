// Per-box classification and point evaluation entry points.
#include "niq_internal.h"

// ------------------------------------------------------------- dispatch
namespace {

// Launch k_classify for (KIND, V) at template width W with tile T, compiled
// for the activation set AS.
template <int KIND, int V, int W, class T = TileCfg<W>, int AS = ACTS_ALL>
niq_status launch_classify_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = T::NB, NW = T::NW;
  constexpr int OCC = T::OCC ? std::min(T::OCC, LdsOcc<Planes<KIND, V, 0>::R, NB, NW, W>::value) : 0;
  constexpr int R = Planes<KIND, V, 0>::R;
  if constexpr (!Fits<R, NB, W>::value) {
    // the planes live in global memory instead (stream-ordered scratch, one
    // slice per workgroup; the grid is capped to bound it)
    auto k = k_classify<KIND, V, NB, NW, W, true>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = grid_for(tiles, std::min<long long>(persistent_blocks(k, NW * 64), 2LL * num_cus()));
    ClassifyArgs b = a;
    HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&b.gx), sizeof(float) * (size_t)R * NB * W * g, s));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, b);
    const hipError_t e = hipGetLastError();
    HIPCHK(hipFreeAsync(b.gx, s));
    if (e != hipSuccess) return fail(NIQ_ERR_HIP, std::string("k_classify (global planes): ") + hipGetErrorString(e));
    if (a.recheck_tol > 0.f) {
      auto k64 = k_recheck_fp64<KIND, V, W>;
      const unsigned g64 = grid_for(std::min<long long>(max_n, a.recheck_cap), persistent_blocks(k64, 256));
      hipLaunchKernelGGL(k64, dim3(g64), dim3(256), 0, s, m->dev, a);
      HIPCHK(hipGetLastError());
    }
    return NIQ_OK;
  } else {
    auto k = k_classify<KIND, V, NB, NW, W, false, OCC, AS>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = tile_grid(tiles, persistent_blocks(k, NW * 64));
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

// Tile variants for the forms trees use most (interval and affine_fixed over
// 3 box vectors): 1 (and 0) TileCfg, 2 TileCfgAlt, 3 TileCfgRegs, 4
// TileCfgWide (16 boxes, the most waves: the narrowest width tile per wave);
// the other forms always take TileCfg.  The waves of a tile split the layer's
// feature tiles among them, so the wave count sets the width tile each wave
// contracts (W / 16 / NW tiles of 16 features).
template <int KIND, int V>
niq_status launch_classify(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s,
                           int variant = 0) {
  // at most two tiles per CU: latency-bound, so the wide tile (same sums in
  // the same order, bit-identical results)
  if (variant == 0 && max_n <= 32LL * num_cus()) {
    switch (m->wt) {
      case 64: return launch_classify_w<KIND, V, 64, TileCfgWide<64>>(m, a, max_n, s);
      case 128: return launch_classify_w<KIND, V, 128, TileCfgWide<128>>(m, a, max_n, s);
      case 256: return launch_classify_w<KIND, V, 256, TileCfgWide<256>>(m, a, max_n, s);
      case 512: return launch_classify_w<KIND, V, 512, TileCfgWide<512>>(m, a, max_n, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "width");
  }
  constexpr bool has_alt = KIND == 0 && (V == 0 || V == 3);
  if constexpr (has_alt) {
    if (variant == 2) {
      switch (m->wt) {
        case 64: return launch_classify_w<KIND, V, 64, TileCfgAlt<64>>(m, a, max_n, s);
        case 128: return launch_classify_w<KIND, V, 128, TileCfgAlt<128>>(m, a, max_n, s);
        case 256: return launch_classify_w<KIND, V, 256, TileCfgAlt<256>>(m, a, max_n, s);
        case 512: return launch_classify_w<KIND, V, 512, TileCfgAlt<512>>(m, a, max_n, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
    if (variant == 3) {
      switch (m->wt) {
        case 64: return launch_classify_w<KIND, V, 64, TileCfgRegs<64>>(m, a, max_n, s);
        case 128: return launch_classify_w<KIND, V, 128, TileCfgRegs<128>>(m, a, max_n, s);
        case 256: return launch_classify_w<KIND, V, 256, TileCfgRegs<256>>(m, a, max_n, s);
        case 512: return launch_classify_w<KIND, V, 512, TileCfgRegs<512>>(m, a, max_n, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
    if (variant == 4) {   // the narrowest width tile: one 16-feature tile per wave (at W <= 256)
      switch (m->wt) {
        case 64: return launch_classify_w<KIND, V, 64, TileCfgWide<64>>(m, a, max_n, s);
        case 128: return launch_classify_w<KIND, V, 128, TileCfgWide<128>>(m, a, max_n, s);
        case 256: return launch_classify_w<KIND, V, 256, TileCfgWide<256>>(m, a, max_n, s);
        case 512: return launch_classify_w<KIND, V, 512, TileCfgWide<512>>(m, a, max_n, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
  }
  if constexpr (KIND == 0 && (V == 0 || V == 3)) {
    // ReLU models, in the forms trees use (the default tile): a kernel
    // compiled for ReLU alone, the same sums and rules (bit-identical), fewer
    // registers
    if (m->relu_only) {
      switch (m->wt) {
        case 64: return launch_classify_w<KIND, V, 64, TileCfg<64>, ACTS_RELU>(m, a, max_n, s);
        case 128: return launch_classify_w<KIND, V, 128, TileCfg<128>, ACTS_RELU>(m, a, max_n, s);
        case 256: return launch_classify_w<KIND, V, 256, TileCfg<256>, ACTS_RELU>(m, a, max_n, s);
        case 512: return launch_classify_w<KIND, V, 512, TileCfg<512>, ACTS_RELU>(m, a, max_n, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
  }
  switch (m->wt) {
    case 64: return launch_classify_w<KIND, V, 64>(m, a, max_n, s);
    case 128: return launch_classify_w<KIND, V, 128>(m, a, max_n, s);
    case 256: return launch_classify_w<KIND, V, 256>(m, a, max_n, s);
    case 512: return launch_classify_w<KIND, V, 512>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// The weak-SDF test: one point plane, any number of box vectors.
template <int W>
niq_status launch_classify_sdf_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_classify_sdf<NB, NW, W>;
  const long long tiles = (max_n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = tile_grid(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, a);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

niq_status launch_classify_sdf(const niq_model* m, const niq_eval_cfg* cfg, ClassifyArgs a, long long max_n,
                               hipStream_t s) {
  if (a.recheck_tol > 0.f) return fail(NIQ_ERR_UNSUPPORTED, "the sdf mode has no FP64 re-test");
  a.lipschitz = cfg->sdf_lipschitz;
  switch (m->wt) {
    case 64: return launch_classify_sdf_w<64>(m, a, max_n, s);
    case 128: return launch_classify_sdf_w<128>(m, a, max_n, s);
    case 256: return launch_classify_sdf_w<256>(m, a, max_n, s);
    case 512: return launch_classify_sdf_w<512>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}
}  // namespace

niq_status niqi::classify_dispatch(const niq_model* m, const niq_eval_cfg* cfg, int nvec,
                                   const ClassifyArgs& a, long long max_n, hipStream_t s) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const int prec = cfg ? cfg->precision : NIQ_PREC_FP32;
  if (mode == NIQ_MODE_SDF) {
    if (prec != NIQ_PREC_FP32) return fail(NIQ_ERR_UNSUPPORTED, "the sdf mode evaluates points in FP32 only");
    return launch_classify_sdf(m, cfg, a, max_n, s);
  }
  if (prec == NIQ_PREC_BF16) return classify_bf16(m, cfg, nvec, a, max_n, s);
  if (prec != NIQ_PREC_FP32) return fail(NIQ_ERR_INVALID, "unknown precision");
  if (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND) {
    if (a.recheck_tol > 0.f) return fail(NIQ_ERR_UNSUPPORTED, "the growing-row modes have no FP64 re-test");
    if (!a.lower && a.vin > 64) return fail(NIQ_ERR_UNSUPPORTED, "at most 64 box vectors");
    return classify_rows(m, cfg, a, max_n, s);
  }
  const int variant = cfg ? cfg->variant : 0;
  if (mode == NIQ_MODE_INTERVAL) return launch_classify<0, 0>(m, a, max_n, s, variant);
  if (mode == NIQ_MODE_AFFINE_FIXED) {
    switch (nvec) {
      case 1: return launch_classify<0, 1>(m, a, max_n, s);
      case 2: return launch_classify<0, 2>(m, a, max_n, s);
      case 3: return launch_classify<0, 3>(m, a, max_n, s, variant);
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
// AS: the activation set of the kernel (ReLU models take the ReLU-only one).
template <int W, int AS = ACTS_ALL>
niq_status launch_points_w(const niq_model* m, const float* x, int d, long long n, float* out, hipStream_t s) {
  if constexpr (AS == ACTS_ALL) {
    if (m->relu_only) return launch_points_w<W, ACTS_RELU>(m, x, d, n, out, s);
  }
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_points<NB, NW, W, AS>;
  const long long tiles = (n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, x, d, n, out);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

}  // namespace

namespace niqi {
bool classify_needs_scratch(const niq_model* m, const niq_eval_cfg* cfg, int nvec) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const int prec = cfg ? cfg->precision : NIQ_PREC_FP32;
  if (prec != NIQ_PREC_FP32) return false;
  int R = 0;
  if (mode == NIQ_MODE_SLOPE_INTERVAL) R = 1 + 2 * nvec;
  else if (mode == NIQ_MODE_AFFINE_FIXED) R = nvec + 2;
  else return false;
  return (size_t)R * 16 * (size_t)m->wt * 4 > (size_t)kLdsLimit;
}

niq_status classify_with_recheck(const niq_model* m, const niq_eval_cfg* cfg, int nvec, ClassifyArgs a,
                                 long long n, hipStream_t s) {
  const float tol = cfg ? cfg->recheck_tol : 0.f;
  if (!(tol > 0.f) || n == 0) return classify_dispatch(m, cfg, nvec, a, n, s);
  if (n > INT_MAX) return fail(NIQ_ERR_UNSUPPORTED, "FP64 re-test supports at most 2^31-1 boxes per call");
  void* scratch = nullptr;
  HIPCHK(hipMallocAsync(&scratch, sizeof(int) * (size_t)(n + 64), s));
  if (const hipError_t e = hipMemsetAsync(scratch, 0, sizeof(int), s); e != hipSuccess) {
    (void)hipFreeAsync(scratch, s);
    return fail(NIQ_ERR_HIP, std::string("re-test counter clear: ") + hipGetErrorString(e));
  }
  a.recheck_tol = tol;
  a.recheck_cnt = static_cast<int*>(scratch);
  a.recheck_idx = static_cast<int*>(scratch) + 64;
  a.recheck_cap = (int)n;
  niq_status st = classify_dispatch(m, cfg, nvec, a, n, s);
  HIPCHK(hipFreeAsync(scratch, s));
  return st;
}
}  // namespace

namespace {
template <class T, int W>
void shape_of(int R, int* nb, int* nw, int* occ) {
  *nb = T::NB; *nw = T::NW;
  const int wgs = (int)((size_t)kLdsLimit / ((size_t)R * T::NB * W * 4));
  const int lds = std::max(1, std::min(4, wgs * T::NW / 4));
  *occ = T::OCC ? std::min(T::OCC, lds) : lds;
}
template <int W>
niq_status tile_shape_w(int which, int variant, int* nb, int* nw, int* occ) {
  if (which == 0) {   // plain FP32 classify, affine_fixed over 3 box vectors (5 planes)
    switch (variant) {
      case 0: case 1: shape_of<TileCfg<W>, W>(5, nb, nw, occ); return NIQ_OK;
      case 2: shape_of<TileCfgAlt<W>, W>(5, nb, nw, occ); return NIQ_OK;
      case 3: shape_of<TileCfgRegs<W>, W>(5, nb, nw, occ); return NIQ_OK;
      case 4: shape_of<TileCfgWide<W>, W>(5, nb, nw, occ); return NIQ_OK;
    }
  } else if (which == 2) {   // growing-row classify: *nb is the error-term rows per chunk
    return rows_tile_shape(variant, W, nb, nw, occ);
  } else if (which == 1) {   // ray caster, affine_fixed (5 planes)
    switch (variant) {
      case 0: shape_of<RayTile<W, 0>, W>(5, nb, nw, occ); return NIQ_OK;
      case 1: shape_of<RayTile<W, 1>, W>(5, nb, nw, occ); return NIQ_OK;
      case 2: shape_of<RayTile<W, 2>, W>(5, nb, nw, occ); return NIQ_OK;
    }
  }
  return fail(NIQ_ERR_INVALID, "unknown kernel or variant");
}
}  // namespace

extern "C" {

niq_status niq_tile_shape(int which, int variant, int width, int* nb, int* nw, int* occ) {
  if (!nb || !nw || !occ) return fail(NIQ_ERR_INVALID, "null output");
  switch (template_width(width)) {
    case 64: return tile_shape_w<64>(which, variant, nb, nw, occ);
    case 128: return tile_shape_w<128>(which, variant, nb, nw, occ);
    case 256: return tile_shape_w<256>(which, variant, nb, nw, occ);
    case 512: return tile_shape_w<512>(which, variant, nb, nw, occ);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width above 512");
}

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
