// Precision-mode (BF16 MFMA, rigorous) classification launches.
#include "niq_internal.h"
#include "niq_bf16.h"

namespace {

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

}  // namespace

niq_status niqi::classify_bf16(const niq_model* m, const niq_eval_cfg* cfg, int nvec, const ClassifyArgs& a,
                               long long max_n, hipStream_t s) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
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
