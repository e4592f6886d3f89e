#pragma once
// Frustum ray casting (src/queries.py:178-587): image tiles march forward as
// general boxes (the frustum's step segment: its centre ray and two side
// vectors), split in two along the wider index axis when they grow too wide
// for their step or demand it, and stop when a single-pixel frustum hits, the
// frustum passes max_dist or the iteration budget runs out.  Each frustum
// evolves from its own state only, so the live frusta are kept in a compact
// list (any order); the classify and point passes are the fused kernels.
#include "niq_device.h"

struct FrustCam {
  float root[3], look[3], up[3], left[3];
  float tan_x, tan_y;         // tan(deg2rad(fov) / 2), float32
  float half_fov_x, half_fov_y;  // deg2rad(fov) / 2, float32
  int res_x, res_y;
};

struct FrustParams {
  float hit_eps, max_dist, safety, grow, shrink, refine;
  int n_models;
};

// A list of frusta: pixel ranges [x0, y0, x1, y1) and their march state.
struct FrustList {
  int4* rng;
  float* t;
  float* step;
  float* cnt;
};

struct FrustCounters {
  int n_next;     // entries written to the next list (survivors and B halves)
  int n_fin;      // finished frusta
  int n_valid;    // survivors of this iteration
  int n_refine;   // survivors that split
};

// Per-frustum geometry of one iteration (src/queries.py:275-308) and the
// substep state; all arrays have one entry per live frustum.
struct FrustWork {
  float* geo;     // 14 floats: r_uu, r_lu, r_ul (9), mid (3), expand, single
  float* center;  // (n, 3) general-box centres
  float* vecs;    // (n, 3, 3) general-box vectors
  float* pts;     // (2n, 3) hit-test points
  int* codes;     // (n_models, cap) sign codes
  float* vals;    // (n_models, 2 cap) point values
  int* is_hit;
  int* hit_id;
  int* demands;
  int* nsteps;
  int* can_step;
};

__device__ __forceinline__ void frust_cam_ray(const FrustCam& c, float tx, float ty, float* r) {
  float p[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = c.look[k] + c.left[k] * (tx * c.tan_x) + c.up[k] * (ty * c.tan_y);
  const float nrm = sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) r[k] = p[k] / nrm;
}

static __global__ void k_frust_init(FrustList L, int n_side, const int* xt, const int* yt, float init_step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_side * n_side) return;
  const int a = i % n_side, b = i / n_side;   // tile(x ticks), repeat(y ticks)
  L.rng[i] = make_int4(xt[a], yt[b], xt[a + 1], yt[b + 1]);
  L.t[i] = 0.f;
  L.step[i] = init_step;
  L.cnt[i] = 0.f;
}

static __global__ void k_frust_geom(FrustCam c, FrustList L, FrustWork W, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 r = L.rng[i];
    const float rx1 = (float)c.res_x + 1.f, ry1 = (float)c.res_y + 1.f;
    const float xlo = 2.f * (float)r.x / rx1 - 1.f, xhi = 2.f * (float)(r.z - 1) / rx1 - 1.f;
    const float ylo = 2.f * (float)r.y / ry1 - 1.f, yhi = 2.f * (float)(r.w - 1) / ry1 - 1.f;
    float uu[3], lu[3], ul[3], ll[3];
    frust_cam_ray(c, xhi, yhi, uu);
    frust_cam_ray(c, xlo, yhi, lu);
    frust_cam_ray(c, xhi, ylo, ul);
    frust_cam_ray(c, xlo, ylo, ll);
    float mid[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) mid[k] = 0.5f * (uu[k] + ll[k]);
    const float len = sqrtf(mid[0] * mid[0] + mid[1] * mid[1] + mid[2] * mid[2]);
    float* g = W.geo + 14 * (size_t)i;
#pragma unroll
    for (int k = 0; k < 3; ++k) { g[k] = uu[k]; g[3 + k] = lu[k]; g[6 + k] = ul[k]; g[9 + k] = mid[k] / len; }
    g[12] = 1.f / len;
    g[13] = (r.x + 1 == r.z && r.y + 1 == r.w) ? 1.f : 0.f;
    W.is_hit[i] = 0; W.hit_id[i] = 0; W.demands[i] = 0; W.nsteps[i] = 0;
  }
}

// Start of a substep: the step segment as a general box and the two points.
static __global__ void k_frust_prep(FrustCam c, FrustParams P, FrustList L, FrustWork W, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* g = W.geo + 14 * (size_t)i;
    const float t = L.t[i], step = L.step[i];
    const float t_adj = (t + step) * g[12];
    const int hit = W.is_hit[i];
    W.can_step[i] = !hit;
    W.nsteps[i] += hit ? 0 : 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float mid = g[9 + k];
      W.center[3 * (size_t)i + k] = c.root[k] + 0.5f * (t + t_adj) * mid;
      W.vecs[9 * (size_t)i + k] = 0.5f * (t_adj - t) * mid;
      W.vecs[9 * (size_t)i + 3 + k] = (g[k] - g[3 + k]) * t_adj / 2.f;
      W.vecs[9 * (size_t)i + 6 + k] = (g[k] - g[6 + k]) * t_adj / 2.f;
      W.pts[6 * (size_t)i + k] = c.root[k] + t * mid;
      W.pts[6 * (size_t)i + 3 + k] = c.root[k] + (t + P.hit_eps) * mid;
    }
  }
}

// End of a substep (src/queries.py:214-269) after every function's passes.
static __global__ void k_frust_sub(FrustParams P, FrustList L, FrustWork W, int n, int cap) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int can = W.can_step[i], hit = W.is_hit[i], hid = W.hit_id[i];
    for (int f = 0; f < P.n_models; ++f) {
      const int code = W.codes[(size_t)f * cap + i];
      can = can && (code == CODE_POSITIVE || code == CODE_NEGATIVE);
      const float v0 = W.vals[(size_t)f * 2 * cap + 2 * (size_t)i];
      const float v1 = W.vals[(size_t)f * 2 * cap + 2 * (size_t)i + 1];
      const bool this_hit = sign_f(v0) != sign_f(v1);
      if (this_hit) hid = f + 1;
      hit = hit || this_hit;
    }
    float t = L.t[i], step = L.step[i];
    const float single = W.geo[14 * (size_t)i + 13];
    const float this_step = can ? step : P.hit_eps * single;
    if (!hit) t = t + this_step * P.safety;
    step = can ? step * P.grow : step * P.shrink;
    W.demands[i] = W.demands[i] || (step < P.hit_eps) || hit;
    step = fmaxf(step, P.hit_eps);
    L.t[i] = t; L.step[i] = step;
    W.is_hit[i] = hit; W.hit_id[i] = hid;
  }
}

// End of an iteration (src/queries.py:321-362, 384-425): counts, termination,
// refinement and the split; survivors and B halves go to the next list.
static __global__ void k_frust_finish(FrustCam c, FrustParams P, FrustList L, FrustWork W, int n, int count_term,
                                      FrustList N, FrustList F, int* fin_hit, FrustCounters* C) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 r = L.rng[i];
    const int area = (r.z - r.x) * (r.w - r.y);
    const float cnt = L.cnt[i] + (float)W.nsteps[i] * (1.f / (float)area);
    const float t = L.t[i], step = L.step[i];
    const bool hit = W.is_hit[i] && area == 1;
    const bool term = hit || (t > P.max_dist) || count_term;
    if (term) {
      const int k = atomicAdd(&C->n_fin, 1);
      F.rng[k] = r; F.t[k] = t; F.cnt[k] = cnt;
      fin_hit[k] = W.hit_id[i];
      continue;
    }
    atomicAdd(&C->n_valid, 1);
    const float wx = 2.f * sinf(c.half_fov_x * (float)(r.z - r.x) / (float)c.res_x) * t;
    const float wy = 2.f * sinf(c.half_fov_y * (float)(r.w - r.y) / (float)c.res_y) * t;
    const bool can_subd = (r.z > r.x + 1) || (r.w > r.y + 1);
    const bool needs = ((wx > P.refine * step) || (wy > P.refine * step) || W.demands[i]) && can_subd;
    if (!needs) {
      const int k = atomicAdd(&C->n_next, 1);
      N.rng[k] = r; N.t[k] = t; N.step[k] = step; N.cnt[k] = cnt;
      continue;
    }
    atomicAdd(&C->n_refine, 1);
    const bool sx = (r.z - r.x) >= (r.w - r.y);
    const int xm = (r.x + r.z) / 2, ym = (r.y + r.w) / 2;
    const int4 A = make_int4(r.x, r.y, sx ? xm : r.z, sx ? r.w : ym);
    const int4 B = make_int4(sx ? xm : r.x, sx ? r.y : ym, r.z, r.w);
    const int k = atomicAdd(&C->n_next, 2);
    N.rng[k] = A; N.t[k] = t; N.step[k] = step; N.cnt[k] = cnt;
    N.rng[k + 1] = B; N.t[k + 1] = t; N.step[k + 1] = step; N.cnt[k + 1] = cnt;
  }
}

// Every finished frustum covers its pixels (src/queries.py:571-585); outputs
// are indexed [x, y].
static __global__ void k_frust_write(FrustList F, const int* fin_hit, int n_fin, int res_y, float* out_t,
                                     int* out_hit, int* out_count) {
  for (int i = blockIdx.x; i < n_fin; i += gridDim.x) {
    const int4 r = F.rng[i];
    const int w = r.w - r.y;
    const int np = (r.z - r.x) * w;
    const float t = F.t[i];
    const int h = fin_hit[i];
    const int cn = (int)F.cnt[i];
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const size_t o = (size_t)(r.x + p / w) * res_y + (r.y + p % w);
      out_t[o] = t; out_hit[o] = h; out_count[o] = cn;
    }
  }
}
