// Precision-mode (BF16 MFMA, rigorous) classification launches.
#include "niq_internal.h"
#include "niq_bf16.h"
#include "niq_bf16_layer.h"

namespace {

// Boxes per layer-by-layer chunk when the call allocates its own scratch.
constexpr long long kLwChunk = 32768;

bool lw_eligible(const niq_model* m) {
  const DevModelBf16& H = m->host16;
  if (H.n_layers < 2 || H.L[H.n_layers - 1].out_pad != 16 || H.L[0].in_pad != 32) return false;
  for (int l = 0; l + 1 < H.n_layers; ++l)
    if (H.L[l].out_pad % lw::kFeats) return false;
  return true;
}

template <int V, int W, int WS, int AS = ACTS_ALL>
niq_status launch_bf16_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  if constexpr ((size_t)PlanesBf16<V>::NPL * NB * W * 2 > (size_t)kLdsLimit) {
    return fail(NIQ_ERR_UNSUPPORTED, "precision-mode range form does not fit in LDS at this width");
  } else {
    auto k = k_classify_bf16<V, NB, NW, W, WS, AS>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = tile_grid(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev16, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int V, int WS>
niq_status launch_bf16_ws(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  if constexpr (WS == WSRC_FP32) {
    // ReLU models: the kernel compiled for ReLU alone (bit-identical)
    if (m->relu_only) {
      switch (m->wt16) {
        case 64: return launch_bf16_w<V, 64, WS, ACTS_RELU>(m, a, max_n, s);
        case 128: return launch_bf16_w<V, 128, WS, ACTS_RELU>(m, a, max_n, s);
        case 256: return launch_bf16_w<V, 256, WS, ACTS_RELU>(m, a, max_n, s);
        case 512: return launch_bf16_w<V, 512, WS, ACTS_RELU>(m, a, max_n, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
  }
  switch (m->wt16) {
    case 64: return launch_bf16_w<V, 64, WS>(m, a, max_n, s);
    case 128: return launch_bf16_w<V, 128, WS>(m, a, max_n, s);
    case 256: return launch_bf16_w<V, 256, WS>(m, a, max_n, s);
    case 512: return launch_bf16_w<V, 512, WS>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// Layer by layer (variant 3): chunks of up to `nch` boxes, each through a
// prologue, one batched launch per hidden layer and the output layer.
template <int V>
niq_status launch_lw(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  if (max_n <= 0) return NIQ_OK;
  const DevModelBf16& H = m->host16;
  const int wmax = std::max(H.wmax, 32);
  const size_t per_box = lw::scratch_bytes<V>(1, wmax);
  long long nch;
  void* buf = a.lw_buf;
  bool own = false;
  if (buf) {
    nch = (long long)(a.lw_bytes / per_box) / lw::kBoxes * lw::kBoxes;
    if (nch < lw::kBoxes) return fail(NIQ_ERR_CAPACITY, "layer-by-layer scratch too small");
  } else {
    nch = std::min<long long>((max_n + lw::kBoxes - 1) / lw::kBoxes * lw::kBoxes, kLwChunk);
    HIPCHK(hipMallocAsync(&buf, per_box * nch, s));
    own = true;
  }
  nch = std::min<long long>(nch, (max_n + lw::kBoxes - 1) / lw::kBoxes * lw::kBoxes);
  unsigned short* X[2] = {static_cast<unsigned short*>(buf),
                          static_cast<unsigned short*>(buf) + (size_t)nch * lw::Cfg<V>::NPL * wmax};
  niq_status st = NIQ_OK;
  for (long long c0 = 0; c0 < max_n && st == NIQ_OK; c0 += nch) {
    const long long nb = std::min(nch, max_n - c0);
    const long long ntile = (nb + lw::kBoxes - 1) / lw::kBoxes;
    const unsigned gp = (unsigned)std::min<long long>((nb * 32 + 255) / 256, 4096);
    hipLaunchKernelGGL(lw::k_lw_prologue<V>, dim3(gp), dim3(256), 0, s, m->dev16, a, c0, nb, X[0]);
    int cur = 0;
    for (int l = 0; l + 1 < H.n_layers; ++l) {
      const int nfg = H.L[l].out_pad / lw::kFeats;
      const unsigned g = (unsigned)((ntile + 7) / 8 * 8 * nfg);
      hipLaunchKernelGGL(lw::k_lw_layer<V>, dim3(g), dim3(lw::kThreads), 0, s, m->dev16, l, a, c0, nb,
                         X[cur], X[cur ^ 1]);
      cur ^= 1;
    }
    const unsigned gl = (unsigned)(((nb + 15) / 16 + 3) / 4);
    hipLaunchKernelGGL(lw::k_lw_last<V>, dim3(gl), dim3(256), 0, s, m->dev16, H.n_layers - 1, a, c0, nb, X[cur]);
    if (hipGetLastError() != hipSuccess) st = fail(NIQ_ERR_HIP, "layer-by-layer launch failed");
  }
  if (own) HIPCHK(hipFreeAsync(buf, s));
  return st;
}

// Weight source: cfg->variant = 1 streams the four packed bf16 copies, 2 (and
// the default) the FP32 weights split in registers; 3 runs layer by layer
// (hidden widths multiples of 128, interval or three box vectors; otherwise
// the fused kernel of variant 2).
template <int V>
niq_status launch_bf16(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                       hipStream_t s) {
  if (!m->bf16_ok) return fail(NIQ_ERR_UNSUPPORTED, "the precision mode needs weights without FP32 subnormals");
  const int ws = cfg ? cfg->variant : 0;
  if (ws == 1) return launch_bf16_ws<V, WSRC_BF16>(m, a, max_n, s);
  if constexpr (V == 0 || V == 3) {
    if (ws == 3 && lw_eligible(m)) return launch_lw<V>(m, a, max_n, s);
  }
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
