#pragma once
// Shared host-side pieces of the C ABI (include/niq.h): error state, launch
// geometry helpers and the model handle.
#include "niq.h"
#include "niq_kernels.h"
#include "niq_bf16.h"



#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace niqi {

extern thread_local std::string g_err;

inline niq_status fail(niq_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(NIQ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

inline int round16(int x) { return (x + 15) / 16 * 16; }
inline int round32(int x) { return (x + 31) / 32 * 32; }

// host bf16 conversions (same bit rules as the device)
inline unsigned short h_bf16_rn(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
inline float h_bf16_f(unsigned short h) {
  unsigned u = (unsigned)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline unsigned short h_bf16_up_nonneg(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  unsigned short h = (unsigned short)(u >> 16);
  if ((u & 0xffffu) && (u & 0x7fffffffu) < 0x7f800000u) h = (unsigned short)(h + 1);
  return h;
}
// gamma_n(u) = n u / (1 - n u), rounded up generously
inline double gamma_n(double n, double u) { return n * u / (1.0 - n * u) * (1.0 + 1e-6); }

// Template width that holds every padded layer width of a model.
inline int template_width(int wmax) {
  if (wmax <= 64) return 64;
  if (wmax <= 128) return 128;
  if (wmax <= 256) return 256;
  if (wmax <= 512) return 512;
  return -1;
}

// closest_point: rounds between host checks of the stack (up to 32 while the
// room they need, 2 K B rows, stays small) and the stack's initial rows.
inline int cp_rounds_per_check(int B) { return (int)std::clamp<long long>((1LL << 16) / B, 2, 32); }
inline long long cp_initial_stack(long long Q, int B) {
  return std::max<long long>(2 * Q, Q + 2LL * cp_rounds_per_check(B) * B);
}

inline int num_cus() {
  static int cus = -1;
  if (cus < 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
  }
  return cus;
}

// Persistent grid: resident workgroups of kernel `f` over the whole device.
template <class F>
inline long long persistent_blocks(F f, int threads) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find((const void*)f);
  int occ = 0;
  if (it == cache.end()) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, threads, 0) != hipSuccess || occ < 1) occ = 1;
    cache[(const void*)f] = occ;
  } else {
    occ = it->second;
  }
  return (long long)occ * num_cus();
}

inline unsigned grid_for(long long work_blocks, long long persist) {
  long long g = std::min(work_blocks, persist);
  return (unsigned)std::max(1LL, g);
}

// Grid of a classification launch (each workgroup strides over box tiles):
// up to kTileGridMult workgroups per resident slot.  Its workgroups then
// retire while the launch runs, so launches on other streams (a concurrent
// tree's count, scan and scatter kernels) get compute units between them
// instead of waiting for the whole launch to drain.
constexpr long long kTileGridMult = 8;
inline unsigned tile_grid(long long tiles, long long persist) { return grid_for(tiles, persist * kTileGridMult); }

}  // namespace niqi

using namespace niqi;

struct niq_model {
  uint64_t serial = 0;         // unique per niq_model_create in this process
  int n_layers = 0;
  int in_dim = 0;
  int wt = 0;                  // template width
  int wt16 = 0;                // template width of the precision-mode copy
  bool bf16_ok = true;         // no FP32 subnormal weights (precision mode)
  bool relu_only = false;      // every hidden activation is ReLU
  bool tanh_only = false;      // every hidden activation is TanH
  DevModelBf16 host16;         // precision-mode copy (bf16 hi/lo/|W| weights)
  DevModelBf16* dev16 = nullptr;
  void* weights16 = nullptr;
  std::vector<int> dims, acts;
  DevModel host;               // with device pointers
  DevModel* dev = nullptr;     // device copy of `host`
  void* weights = nullptr;     // packed weights and biases
  size_t bytes = 0;
};

namespace niqi {

inline niq_status check_model(const niq_model* m, int d) {
  if (!m) return fail(NIQ_ERR_INVALID, "null model");
  if (d != m->in_dim) return fail(NIQ_ERR_INVALID, "point dimension does not match the model input");
  return NIQ_OK;
}

// Classification of up to max_n boxes (count may be device-resident), routed
// by (mode, precision, box vectors).  Defined in api_classify.hip.
// Rows per chunk, waves and register budget (waves per SIMD) of a
// growing-row classify variant at template width wt.
niq_status rows_tile_shape(int variant, int wt, int* ch, int* nw, int* occ);
// Batched level-set trees (the serving loop's merged builds): up to
// kMaxTreeBatch trees with one eval config, one set of params (their root
// boxes aside) and one template width, built level by level in shared
// launches.  Each tree's outputs and counts equal its lone build's.
constexpr int kMaxTreeBatch = 64;
bool tree_batchable(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p);
niq_status tree_batch_bytes(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, int ng,
                            size_t* bytes);
// roots: 6 floats per tree (lower then upper); tree_status (nullable): each
// tree's own status once the build ran (NIQ_ERR_CAPACITY where its lists
// overflowed); synchronises.
niq_status tree_batch_build(const niq_model* const* models, int ng, const niq_eval_cfg* cfg,
                            const niq_tree_params* p, const float* roots, void* ws, size_t ws_bytes,
                            niq_tree_result* const* res, niq_status* tree_status, hipStream_t s);
niq_status classify_dispatch(const niq_model* m, const niq_eval_cfg* cfg, int nvec, const ClassifyArgs& a,
                             long long max_n, hipStream_t s);
// Whether a classification of `nvec` box vectors keeps its planes in global
// scratch (allocated per call) because they exceed LDS.
bool classify_needs_scratch(const niq_model* m, const niq_eval_cfg* cfg, int nvec);
// The precision-mode launches (api_classify_bf16.hip) and the growing-row
// modes (api_classify_rows.hip).
niq_status classify_bf16(const niq_model* m, const niq_eval_cfg* cfg, int nvec, const ClassifyArgs& a,
                         long long max_n, hipStream_t s);
niq_status classify_rows(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                         hipStream_t s);
// ... with the optional FP64 re-test list in stream-ordered scratch.
niq_status classify_with_recheck(const niq_model* m, const niq_eval_cfg* cfg, int nvec, ClassifyArgs a,
                                 long long n, hipStream_t s);
// Plain evaluation of mult * (*n_dev) points (n_dev null: max_n points).
// Defined in api_cp_mc.hip.
niq_status points_dev(const niq_model* m, const float* x, const int* n_dev, int mult, long long max_n, float* out,
                      hipStream_t s);
// The checks of niq_cast_rays (everything validated before any work is
// enqueued; *wt, *wt16 receive the template widths), and the cast over a
// checked request with its workspace at `base`.  Defined in api_rays.hip.
niq_status cast_check(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg, const float* roots,
                      const float* dirs, int64_t n, const niq_cast_opts* o, const float* out_t,
                      const int32_t* out_hit_id, const int32_t* out_count, int* wt, int* wt16,
                      bool allow_rows = true);
niq_status cast_run(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg, const float* roots,
                    const float* dirs, int64_t n, const niq_cast_opts* o, float* out_t, int32_t* out_hit_id,
                    int32_t* out_count, unsigned char* base, int wt, int wt16, int64_t* n_evals, hipStream_t s);
// One request of a cast over a table of ray groups: its functions, rays,
// options and outputs; n_evals receives its N_evals.
struct CastGroupReq {
  const niq_model* const* models;
  int n_models;
  const float* roots;
  const float* dirs;
  int64_t n;
  niq_cast_opts opts;
  float* out_t;
  int32_t* out_hit;
  int32_t* out_count;
  int64_t n_evals;
};
constexpr int kMaxRayGroups = 64;

// Workspace layout of a group cast: per group a work counter and N_evals,
// then the per-group histograms (everything below zero_end starts at zero),
// the FP32 and precision-mode model tables and the group table.
struct CastLayout {
  size_t hist[kMaxRayGroups];
  size_t zero_end, models, models16, groups, total;
};
CastLayout cast_layout(const CastGroupReq* reqs, int ng, bool with16);
// One persistent cast over ng checked requests that share `cfg` (mode,
// precision, variants) and the template widths wt, wt16; synchronises.
// Defined in api_rays.hip.
niq_status cast_run_groups(CastGroupReq* reqs, int ng, const niq_eval_cfg* cfg, unsigned char* base, int wt,
                           int wt16, hipStream_t s);
// cast_rays in the growing-row modes: the reference's lock-step loop over
// the live rays with the rows classifier (stream-ordered scratch, one host
// read per iteration).  Defined in api_rays_rows.hip.
niq_status cast_rays_rows(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg, const float* roots,
                          const float* dirs, int64_t n, const niq_cast_opts* o, float* out_t, int32_t* out_hit_id,
                          int32_t* out_count, int64_t* n_evals, hipStream_t s);
// niq_tree_enqueue with the root box optionally read from the device (6
// floats, lower then upper), so that one captured build serves any root box.
// Defined in api_tree.hip.
niq_status tree_enqueue_async(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, void* ws,
                              size_t ws_bytes, niq_tree_result* res, size_t* counters_offset, const float* root_dev,
                              void* stream);

}  // namespace niqi

