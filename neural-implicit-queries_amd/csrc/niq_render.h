#pragma once
// Normals and shading of ray-cast hits (src/render.py:53-224).
//
// Hit pixels are listed per function (its hit_id), so the tetrahedron samples
// of a function are evaluated for its own hits only; the 4 sample points of a
// listed pixel are formed straight in the LDS input plane of the fused point
// pass.  Shading, compositing and the tonemap are one elementwise pass.
#include "niq_kernels.h"

struct ShadeParams {
  int shading;             // 0: normal, 1: matcap_color
  float up[3];
  const float* mat[4];     // r, g, b, k matcaps (h, w, 3)
  int mh, mw;
  float colors[8][3];
  int n_colors;
  const float* color_per_hit;   // (n, 3) or null
  int composite;
  int tonemap;
  float gamma_inv, white2, exposure;
};

// The tetrahedron of sample offsets (src/render.py:67-72): row q, axis k of
// (+,-,-), (-,-,+), (-,+,-), (+,+,+); bit 3q + k of kTetPos is set where +eps.
constexpr unsigned kTetPos = 0x001u | 0x020u | 0x080u | 0xE00u;
__device__ __forceinline__ float tet_off(int q, int k, float eps) {
  return ((kTetPos >> (3 * q + k)) & 1u) ? eps : -eps;
}

// Lists hit pixels by function (hit_id 1..n_models) with one atomic per wave
// and function; other pixels get the zero normal.
static __global__ void __launch_bounds__(256)
k_hit_lists(const int* __restrict__ hit_ids, long long n, int n_models, long long cap, int* __restrict__ lists,
            int* __restrict__ counts, float* __restrict__ normals) {
  const int lane = threadIdx.x & 63;
  for (long long base = (long long)blockIdx.x * 256; base < n; base += (long long)gridDim.x * 256) {
    const long long p = base + threadIdx.x;
    const int h = p < n ? hit_ids[p] : 0;
    const bool listed = h >= 1 && h <= n_models;
    if (p < n && !listed) {
      normals[3 * p] = 0.f; normals[3 * p + 1] = 0.f; normals[3 * p + 2] = 0.f;
    }
    for (int f = 0; f < n_models; ++f) {
      const unsigned long long m = __ballot(listed && h == f + 1);
      if (!m) continue;
      const int leader = __ffsll((long long)m) - 1;
      int slot0 = 0;
      if (lane == leader) slot0 = atomicAdd(&counts[f], __popcll(m));
      slot0 = __shfl(slot0, leader);
      if (listed && h == f + 1) {
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        lists[(size_t)f * cap + slot0 + rank] = (int)p;
      }
    }
  }
}

// Values of the 4 tetrahedron samples around each listed pixel: vals[4 s + q].
template <int NB, int NW, int WMAX, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64)
k_normal_eval(const DevModel* __restrict__ Mp, const int* __restrict__ list, const int* __restrict__ count,
              const float* __restrict__ hit_pos, float eps, float* __restrict__ vals) {
  using PL = Planes<2, 0, 1>;
  constexpr int NT = NW * 64;
  __shared__ float X[NB * WMAX];
  const long long n = 4LL * (*count);
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    for (int e = threadIdx.x; e < NB * 16; e += NT) {
      const int b = e >> 4, k = e & 15;
      const long long j = i0 + b;
      float val = 0.f;
      if (j < n && k < 3) {
        const long long pix = list[j >> 2];
        val = hit_pos[3 * pix + k] + tet_off((int)(j & 3), k, eps);
      }
      X[xoff<NB, WMAX>(0, b, k)] = val;
    }
    mlp_forward<PL, 2, 0, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long j = i0 + b;
      if (j < n) vals[j] = v[0];
    });
  }
}

// Forward-mode gradient (method 'autodiff', jax.jacfwd of the MLP,
// src/render.py:62-64), normalised: one box per listed pixel with the primal
// plane at the hit and three unit tangent planes.
template <int NB, int NW, int WMAX>
__global__ void __launch_bounds__(NW * 64)
k_grad_normals(const DevModel* __restrict__ Mp, const int* __restrict__ list, const int* __restrict__ count,
               const float* __restrict__ hit_pos, float* __restrict__ normals) {
  using PL = Planes<3, 3, 0>;
  constexpr int NT = NW * 64;
  __shared__ float X[PL::R * NB * WMAX];
  const long long n = *count;
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    for (int e = threadIdx.x; e < PL::R * NB * 16; e += NT) {
      const int p = e / (NB * 16), b = (e / 16) % NB, k = e & 15;
      const long long s = i0 + b;
      float val = 0.f;
      if (s < n && k < 3) val = p == 0 ? hit_pos[3 * (long long)list[s] + k] : (p - 1 == k ? 1.f : 0.f);
      X[xoff<NB, WMAX>(p, b, k)] = val;
    }
    mlp_forward<PL, 3, 3, NB, NW, WMAX>(*Mp, X, [&](const float* v, int b) {
      const long long s = i0 + b;
      if (s >= n) return;
      const float nrm = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(v[1], v[1]), __fmul_rn(v[2], v[2])), __fmul_rn(v[3], v[3])));
      const long long pix = list[s];
#pragma unroll
      for (int k = 0; k < 3; ++k) normals[3 * pix + k] = v[1 + k] / nrm;
    });
  }
}

// Tetrahedron gradient, normalised (src/render.py:74-79, src/geometry.py:12).
static __global__ void __launch_bounds__(256)
k_normal_finish(const int* __restrict__ list, const int* __restrict__ count, const float* __restrict__ vals,
                float eps, float* __restrict__ normals) {
  const long long n = *count;
  for (long long s = (long long)blockIdx.x * 256 + threadIdx.x; s < n; s += (long long)gridDim.x * 256) {
    float g[3] = {0.f, 0.f, 0.f};
    for (int q = 0; q < 4; ++q) {
      const float v = vals[4 * s + q];
#pragma unroll
      for (int k = 0; k < 3; ++k) g[k] = __fadd_rn(g[k], __fmul_rn(tet_off(q, k, eps), v));
    }
    const float nrm = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(g[0], g[0]), __fmul_rn(g[1], g[1])), __fmul_rn(g[2], g[2])));
    const long long pix = list[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) normals[3 * pix + k] = g[k] / nrm;
  }
}

// map_coordinates(order=1, mode='nearest') of channel c of an (h, w, 3) image.
__device__ __forceinline__ float matcap_lookup(const float* img, int h, int w, int c, float x, float y) {
  const float x0 = floorf(x), y0 = floorf(y);
  const float wx1 = x - x0, wy1 = y - y0;
  const float wx[2] = {1.f - wx1, wx1}, wy[2] = {1.f - wy1, wy1};
  // clamp before the conversion (NaN and huge coordinates land on an edge)
  const int xl = (int)fminf(fmaxf(x0, -1.f), (float)h), yl = (int)fminf(fmaxf(y0, -1.f), (float)w);
  const int xi[2] = {min(max(xl, 0), h - 1), min(max(xl + 1, 0), h - 1)};
  const int yi[2] = {min(max(yl, 0), w - 1), min(max(yl + 1, 0), w - 1)};
  float acc = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const float term = __fmul_rn(__fmul_rn(wx[a], wy[b]), img[((size_t)xi[a] * w + yi[b]) * 3 + c]);
      acc = (a == 0 && b == 0) ? term : __fadd_rn(acc, term);
    }
  return acc;
}

__device__ __forceinline__ void unit3(float* v) {
  const float n = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(v[0], v[0]), __fmul_rn(v[1], v[1])), __fmul_rn(v[2], v[2])));
  v[0] = v[0] / n; v[1] = v[1] / n; v[2] = v[2] / n;
}

__device__ __forceinline__ float dot3(const float* a, const float* b) {
  return __fadd_rn(__fadd_rn(__fmul_rn(a[0], b[0]), __fmul_rn(a[1], b[1])), __fmul_rn(a[2], b[2]));
}

__device__ __forceinline__ float tonemap1(float x, const ShadeParams& P) {
  x = __fmul_rn(x, P.exposure);
  const float num = __fmul_rn(x, __fadd_rn(1.f, x / P.white2));
  const float den = __fadd_rn(1.f, x);
  return powf(num / den, P.gamma_inv);
}

// shade_image (src/render.py:160-224), then render_image's compositing and
// tonemap (src/render.py:139-149) when asked for.
static __global__ void __launch_bounds__(256)
k_shade(ShadeParams P, const float* __restrict__ ray_dirs, const float* __restrict__ normals,
        const int* __restrict__ hit_ids, long long n, float* __restrict__ out) {
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
    const int h = hit_ids[p];
    float nr[3] = {normals[3 * p], normals[3 * p + 1], normals[3 * p + 2]};
    float col[3];
    if (P.shading == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) col[k] = __fadd_rn(nr[k], 1.f) / 2.f;
    } else {
      float rd[3] = {ray_dirs[3 * p], ray_dirs[3 * p + 1], ray_dirs[3 * p + 2]};
      // ray_up = orthogonal_dir(up, ray_dir) (src/geometry.py:15-19)
      unit3(rd);
      const float d = dot3(P.up, rd);
      float ru[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) ru[k] = P.up[k] - __fmul_rn(d, rd[k]);
      unit3(ru);
      float rdo[3] = {ray_dirs[3 * p], ray_dirs[3 * p + 1], ray_dirs[3 * p + 2]};
      const float rl[3] = {__fmul_rn(rdo[1], ru[2]) - __fmul_rn(rdo[2], ru[1]),
                           __fmul_rn(rdo[2], ru[0]) - __fmul_rn(rdo[0], ru[2]),
                           __fmul_rn(rdo[0], ru[1]) - __fmul_rn(rdo[1], ru[0])};
      const float nl[3] = {-rl[0], -rl[1], -rl[2]};
      const float u = __fmul_rn(dot3(nl, nr), 0.98f);
      const float v = __fmul_rn(dot3(ru, nr), 0.98f);
      const float mx = __fmul_rn(__fadd_rn(u, 1.f) / 2.f, (float)P.mh);
      const float my = __fmul_rn(__fadd_rn(-v, 1.f) / 2.f, (float)P.mw);
      float c[3] = {1.f, 1.f, 1.f};
      if (P.color_per_hit) {
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = P.color_per_hit[3 * p + k];
      } else if (h >= 1 && h <= P.n_colors) {
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = P.colors[h - 1][k];
      }
      const float ck = 1.f - __fadd_rn(__fadd_rn(c[0], c[2]), c[1]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float mr = matcap_lookup(P.mat[0], P.mh, P.mw, k, mx, my);
        const float mg = matcap_lookup(P.mat[1], P.mh, P.mw, k, mx, my);
        const float mb = matcap_lookup(P.mat[2], P.mh, P.mw, k, mx, my);
        const float mk = matcap_lookup(P.mat[3], P.mh, P.mw, k, mx, my);
        float o = __fadd_rn(__fmul_rn(c[0], mr), __fmul_rn(c[2], mb));
        o = __fadd_rn(o, __fmul_rn(c[1], mg));
        col[k] = __fadd_rn(o, __fmul_rn(ck, mk));
      }
    }
    if (P.composite && h == 0) col[0] = col[1] = col[2] = 1.f;
    if (P.tonemap) {
#pragma unroll
      for (int k = 0; k < 3; ++k) col[k] = tonemap1(col[k], P);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * p + k] = col[k];
  }
}
