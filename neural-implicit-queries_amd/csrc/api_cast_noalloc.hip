// niq_cast_rays_noalloc: the persistent ray cast of niq_cast_rays with its
// workspace (work counter, N_evals, iteration histogram, model tables) in a
// file-scope device array instead of caller- or library-allocated memory.
#include "niq_internal.h"

// One copy per GPU, held from library load to process end.  Every call on a
// GPU shares it, so calls that overlap on one GPU race; each call clears what
// must start at zero (the counter and histogram) before use.
static __device__ __align__(256) unsigned char scratch[NIQ_CAST_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch`, looked up on the first
// call on each device with that device current.  Host code uses only this
// address: the array's name in host code is a host-side stand-in.
niq_status cast_scratch(unsigned char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(NIQ_ERR_UNSUPPORTED, "device index above the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(NIQ_ERR_HIP, "hipGetSymbolAddress gave no address for the cast scratch");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return NIQ_OK;
}

}  // namespace

extern "C" niq_status niq_cast_rays_noalloc(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                            const float* roots, const float* dirs, int64_t n,
                                            const niq_cast_opts* o, float* out_t, int32_t* out_hit_id,
                                            int32_t* out_count, int64_t* n_evals, void* stream) {
  int wt = 0, wt16 = 0;
  if (niq_status st =
          cast_check(models, n_models, cfg, roots, dirs, n, o, out_t, out_hit_id, out_count, &wt, &wt16, false))
    return st;
  if (niq_cast_workspace_bytes(o) > (size_t)NIQ_CAST_NOALLOC_SCRATCH_BYTES)
    return fail(NIQ_ERR_CAPACITY, "n_max_step too large for niq_cast_rays_noalloc: its scratch holds "
                                  "NIQ_CAST_NOALLOC_SCRATCH_BYTES bytes (niq_cast_workspace_bytes)");
  unsigned char* base = nullptr;
  if (niq_status st = cast_scratch(&base)) return st;
  return cast_run(models, n_models, cfg, roots, dirs, n, o, out_t, out_hit_id, out_count, base, wt, wt16, n_evals,
                  (hipStream_t)stream);
}
