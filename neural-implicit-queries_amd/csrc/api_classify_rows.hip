// Growing-row classification launches (affine_all, affine_truncate,
// affine_append).
#include "niq_internal.h"
#include "niq_rows.h"

namespace {

// affine_all / affine_truncate / affine_append: one box per workgroup, error
// terms streamed through a private HBM scratch.
template <int W, int MODE, int OCC = RowsOcc<W>::value, int CH = RowsCfg<W>::CH>
niq_status launch_rows_w(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                         hipStream_t s) {
  constexpr int NW = RowsCfg<W>::NW;
  auto k = k_classify_rows<W, MODE, OCC, CH>;
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

// Variants (niq_eval_cfg.variant): 0 and 1 the default (RowsCfg chunk, the
// RowsOcc register budget); 2 half the rows per chunk (fewer registers, more
// barriers); 3 at widths up to 128 a two-waves-per-SIMD budget (no spills,
// half the resident workgroups; the wider templates already run at one or
// two).  Every variant sums in the same order: identical results.
template <int MODE>
niq_status launch_rows(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                       hipStream_t s) {
  const int variant = cfg ? cfg->variant : 0;
  if (variant == 2) {
    switch (m->wt) {
      case 64: return launch_rows_w<64, MODE, RowsOcc<64>::value, RowsCfg<64>::CH / 2>(m, cfg, a, max_n, s);
      case 128: return launch_rows_w<128, MODE, RowsOcc<128>::value, RowsCfg<128>::CH / 2>(m, cfg, a, max_n, s);
      case 256: return launch_rows_w<256, MODE, RowsOcc<256>::value, RowsCfg<256>::CH / 2>(m, cfg, a, max_n, s);
      case 512: return launch_rows_w<512, MODE, RowsOcc<512>::value, RowsCfg<512>::CH / 2>(m, cfg, a, max_n, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "width");
  }
  if (variant == 3) {
    switch (m->wt) {
      case 64: return launch_rows_w<64, MODE, 2>(m, cfg, a, max_n, s);
      case 128: return launch_rows_w<128, MODE, 2>(m, cfg, a, max_n, s);
    }
  }
  switch (m->wt) {
    case 64: return launch_rows_w<64, MODE>(m, cfg, a, max_n, s);
    case 128: return launch_rows_w<128, MODE>(m, cfg, a, max_n, s);
    case 256: return launch_rows_w<256, MODE>(m, cfg, a, max_n, s);
    case 512: return launch_rows_w<512, MODE>(m, cfg, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

}  // namespace

niq_status niqi::classify_rows(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a,
                               long long max_n, hipStream_t s) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
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

namespace niqi {
template <int W>
static void rows_shape_w(int variant, int* ch, int* nw, int* occ) {
  *ch = variant == 2 ? RowsCfg<W>::CH / 2 : RowsCfg<W>::CH;
  *nw = RowsCfg<W>::NW;
  *occ = (variant == 3 && W <= 128) ? 2 : RowsOcc<W>::value;
}

niq_status rows_tile_shape(int variant, int wt, int* ch, int* nw, int* occ) {
  if (variant < 0 || variant > 4) return fail(NIQ_ERR_INVALID, "unknown variant");   // 4 runs the default
  switch (wt) {
    case 64: rows_shape_w<64>(variant, ch, nw, occ); return NIQ_OK;
    case 128: rows_shape_w<128>(variant, ch, nw, occ); return NIQ_OK;
    case 256: rows_shape_w<256>(variant, ch, nw, occ); return NIQ_OK;
    case 512: rows_shape_w<512>(variant, ch, nw, occ); return NIQ_OK;
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}
}  // namespace niqi
