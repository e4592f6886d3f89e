// cast_rays in the growing-row modes (affine_truncate, affine_append,
// affine_all).  Their forms do not fit the persistent ray caster's tiles (the
// error-term matrix grows by one row per feature per activation), so each
// iteration runs the reference's lock-step loop over the live rays
// (src/queries.py:39-132) as device kernels around the rows classifier:
// segment boxes and hit-test points are formed for every live ray, every
// function classifies its segment boxes and evaluates its points, the step
// update and termination are applied per ray, and the survivors are
// compacted.  The host reads the live count once per iteration, as the
// reference does.  Per ray the arithmetic is that of k_cast_rays, so a ray
// gets the same (t, hit_id, count) whichever loop runs it.
#include "niq_internal.h"

namespace {

struct RowsRayState {
  const float* roots; const float* dirs;
  int* live;      // (n,) live ray ids
  int* next;      // (n,) survivors of this iteration
  float* t; float* step; int* cnt;       // by ray id
  int* is_hit; int* hit_id; int* nsteps; int* can_step;   // by live slot
  float* center; float* vec; float* pts;   // (n, 3), (n, 3), (2n, 3) by live slot
  int* codes;     // (n_models, n) by live slot
  float* vals;    // (n_models, 2n)
  int* n_next;    // survivors counter
};

__global__ void k_rr_init(RowsRayState S, int n, float init_step) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    S.live[i] = i;
    S.t[i] = 0.f;
    S.step[i] = init_step;
    S.cnt[i] = 0;
  }
}

// Start of a substep: the step and hit-test geometry of every live ray
// (src/queries.py:46-70), in the operation order of k_cast_rays.
__global__ void k_rr_prep(RowsRayState S, int n, int first_sub, float hit_eps) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = S.live[i];
    if (first_sub) { S.is_hit[i] = 0; S.hit_id[i] = 0; S.nsteps[i] = 0; }
    const int hit = S.is_hit[i];
    S.can_step[i] = !hit;
    S.nsteps[i] += hit ? 0 : 1;
    const float t = S.t[r], hs = 0.5f * S.step[r];
    for (int k = 0; k < 3; ++k) {
      const float ro = S.roots[3 * (size_t)r + k], di = S.dirs[3 * (size_t)r + k];
      const float ps = ro + t * di;
      const float hv = hs * di;
      const float pm = ps + hv;
      const float pe = ro + (t + hit_eps) * di;
      S.center[3 * (size_t)i + k] = pm;
      S.vec[3 * (size_t)i + k] = hv;
      S.pts[6 * (size_t)i + k] = ps;
      S.pts[6 * (size_t)i + 3 + k] = pe;
    }
  }
}

__device__ __forceinline__ float sign_of(float x) {
  return x > 0.f ? 1.f : (x < 0.f ? -1.f : (x == 0.f ? 0.f : x));
}

// One function's verdict on every live ray (src/queries.py:63-76).
__global__ void k_rr_hit(RowsRayState S, int n, int f) {
  const int* codes = S.codes + (size_t)f * n;
  const float* vals = S.vals + (size_t)f * 2 * n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = codes[i];
    S.can_step[i] = S.can_step[i] && (c == CODE_POSITIVE || c == CODE_NEGATIVE);
    const bool this_hit = sign_of(vals[2 * (size_t)i]) != sign_of(vals[2 * (size_t)i + 1]);
    if (this_hit) { S.hit_id[i] = f + 1; S.is_hit[i] = 1; }
  }
}

// End of a substep: the step update (src/queries.py:79-92).
__global__ void k_rr_step(RowsRayState S, int n, float hit_eps, float safety, float grow, float shrink) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = S.live[i];
    const bool can = S.can_step[i];
    float step = S.step[r];
    const float this_step = can ? step : hit_eps;
    if (!S.is_hit[i]) S.t[r] = S.t[r] + this_step * safety;
    step = can ? step * grow : step * shrink;
    S.step[r] = fmaxf(step, hit_eps);
  }
}

// End of an iteration: count, termination, outputs and survivors
// (src/queries.py:113-129, 160-173).
__global__ void k_rr_finish(RowsRayState S, int n, float max_dist, int n_max_step, float* out_t, int* out_hit,
                            int* out_count) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = S.live[i];
    const int cnt = S.cnt[r] + S.nsteps[i];
    S.cnt[r] = cnt;
    const float t = S.t[r];
    const bool term = S.is_hit[i] || (t > max_dist) || (cnt >= n_max_step);
    if (term) {
      out_t[r] = t;
      out_hit[r] = S.hit_id[i];
      out_count[r] = cnt;
    } else {
      S.next[atomicAdd(S.n_next, 1)] = r;
    }
  }
}

long long next_bucket_host(long long s) {
  for (int e = 7; e <= 30; ++e)
    if (s <= (1LL << e)) return 1LL << e;
  return -1;
}

}  // namespace

niq_status niqi::cast_rays_rows(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                const float* roots, const float* dirs, int64_t n, const niq_cast_opts* o,
                                float* out_t, int32_t* out_hit_id, int32_t* out_count, int64_t* n_evals,
                                hipStream_t s) {
  if (n > INT_MAX / 2) return fail(NIQ_ERR_UNSUPPORTED, "the growing-row modes cast at most 2^30 rays per call");
  if (n == 0) {
    if (n_evals) *n_evals = 0;
    return NIQ_OK;
  }
  std::vector<void*> allocs;
  niq_status st = NIQ_OK;
  auto al = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (st == NIQ_OK && hipMallocAsync(&p, (bytes + 255) / 256 * 256, s) != hipSuccess)
      st = fail(NIQ_ERR_HIP, "cast_rays scratch allocation failed");
    if (p) allocs.push_back(p);
    return p;
  };
  const size_t N = (size_t)n;
  RowsRayState S{};
  S.roots = roots; S.dirs = dirs;
  int* lists[2] = {static_cast<int*>(al(4 * N)), static_cast<int*>(al(4 * N))};
  S.t = static_cast<float*>(al(4 * N));
  S.step = static_cast<float*>(al(4 * N));
  S.cnt = static_cast<int*>(al(4 * N));
  S.is_hit = static_cast<int*>(al(4 * N));
  S.hit_id = static_cast<int*>(al(4 * N));
  S.nsteps = static_cast<int*>(al(4 * N));
  S.can_step = static_cast<int*>(al(4 * N));
  S.center = static_cast<float*>(al(12 * N));
  S.vec = static_cast<float*>(al(12 * N));
  S.pts = static_cast<float*>(al(24 * N));
  S.codes = static_cast<int*>(al(4 * N * n_models));
  S.vals = static_cast<float*>(al(8 * N * n_models));
  S.n_next = static_cast<int*>(al(256));
  if (st == NIQ_OK && hipMemsetAsync(S.n_next, 0, sizeof(int), s) != hipSuccess)
    st = fail(NIQ_ERR_HIP, "cast_rays set-up failed");
  const unsigned g0 = (unsigned)std::min<size_t>((N + 255) / 256, 8192);
  if (st == NIQ_OK) {
    S.live = lists[0];
    hipLaunchKernelGGL(k_rr_init, dim3(g0), dim3(256), 0, s, S, (int)n, o->init_step);
  }
  long long live = n, bucket = n, evals = 0;
  int cur = 0;
  while (st == NIQ_OK && live > 0) {
    const int nl = (int)live;
    S.live = lists[cur];
    S.next = lists[cur ^ 1];
    const unsigned g = (unsigned)std::min<long long>((live + 255) / 256, 8192);
    for (int sub = 0; sub < o->n_substeps && st == NIQ_OK; ++sub) {
      hipLaunchKernelGGL(k_rr_prep, dim3(g), dim3(256), 0, s, S, nl, sub == 0 ? 1 : 0, o->hit_eps);
      for (int f = 0; f < n_models && st == NIQ_OK; ++f) {
        RowsRayState Sf = S;
        ClassifyArgs a{};
        a.center = S.center; a.vecs = S.vec; a.vin = 1; a.d = 3; a.n = nl; a.offset = 0.f;
        a.codes = S.codes + (size_t)f * nl;
        if ((st = classify_dispatch(models[f], cfg, 1, a, nl, s))) break;
        if ((st = points_dev(models[f], S.pts, nullptr, 2, 2LL * nl, S.vals + (size_t)f * 2 * nl, s))) break;
        hipLaunchKernelGGL(k_rr_hit, dim3(g), dim3(256), 0, s, Sf, nl, f);
      }
      if (st) break;
      hipLaunchKernelGGL(k_rr_step, dim3(g), dim3(256), 0, s, S, nl, o->hit_eps, o->safety_factor,
                         o->interval_grow_fac, o->interval_shrink_fac);
    }
    if (st) break;
    hipLaunchKernelGGL(k_rr_finish, dim3(g), dim3(256), 0, s, S, nl, o->max_dist, o->n_max_step, out_t,
                       out_hit_id, out_count);
    int nn = 0;
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&nn, S.n_next, sizeof(int), hipMemcpyDeviceToHost, s) ||
        hipMemsetAsync(S.n_next, 0, sizeof(int), s) || hipStreamSynchronize(s)) {
      st = fail(NIQ_ERR_HIP, "cast_rays iteration failed");
      break;
    }
    // the reference's N_evals: the bucketed working array, n_substeps times
    // per iteration (src/queries.py:164, 170-173)
    evals += bucket * o->n_substeps;
    live = nn;
    if (live > 0) {
      const long long nb = next_bucket_host(live);
      if (nb > 0 && nb < bucket) bucket = nb;
    }
    cur ^= 1;
  }
  for (void* p : allocs) (void)hipFreeAsync(p, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, "cast_rays stream failed");
  if (st == NIQ_OK && n_evals) *n_evals = evals;
  return st;
}
