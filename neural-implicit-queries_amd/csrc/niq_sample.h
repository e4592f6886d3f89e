#pragma once
// Rejection sampling over kd-tree nodes (sample_surface_iter,
// src/kd_tree.py:220-246; bulk_properties_sample_mass, src/kd_tree.py:804-833).
//
// A candidate is a pure function of its index: Philox4x32-10 (Salmon et al.,
// SC 2011) of the counter (j lo, j hi, tag, 0) picks the node and the three
// uniforms, so the evaluation kernel draws candidates straight into the LDS
// input plane of the fused MLP pass and no candidate position goes to HBM.
// Surface rounds write one flag byte per candidate; the stable compaction
// draws the accepted candidates again.  The mass pass reduces in registers.
#include "niq_kernels.h"

struct SampleSrc {
  const float* lower;          // (n_nodes, 3)
  const float* upper;
  long long n_nodes;
  unsigned key0, key1, tag;
  unsigned long long j0;       // index of the first candidate of the launch
};

// Running sums of one workgroup (count or volume, then the first moment).
struct Sum4 {
  double c, x, y, z;
};

struct SampCounters {
  long long found;       // rows of the output filled
  long long base;        // `found` before this round
  long long n_samples;   // output rows
};

constexpr unsigned kTagSurface = 0x53555246u;
constexpr unsigned kTagMass = 0x4d415353u;
constexpr int kSampBlock = 1024;

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const unsigned hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const unsigned hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

// Candidate j: node floor(w0 * n_nodes / 2^32), position lower + u * (upper -
// lower) with u_k = (w_{k+1} >> 8) 2^-24, unfused float32.
__device__ __forceinline__ void sample_candidate(const SampleSrc& S, unsigned long long j, float p[3]) {
  const uint4 w = philox4x32_10(make_uint4((unsigned)j, (unsigned)(j >> 32), S.tag, 0u), S.key0, S.key1);
  const long long node = (long long)(((unsigned long long)w.x * (unsigned long long)S.n_nodes) >> 32);
  const unsigned wk[3] = {w.y, w.z, w.w};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float lo = S.lower[3 * node + k], up = S.upper[3 * node + k];
    const float u = (float)(wk[k] >> 8) * 0x1p-24f;
    p[k] = __fadd_rn(lo, __fmul_rn(u, __fsub_rn(up, lo)));
  }
}

// n candidates S.j0 .. S.j0 + n - 1 through the fused point pass.  MASS:
// per-workgroup count and position sums of the candidates with f < 0 into
// partial[blockIdx.x]; otherwise flags[i] = |f| < width.
template <int NB, int NW, int WMAX, bool MASS, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64)
k_sample_eval(const DevModel* __restrict__ Mp, SampleSrc S, long long n, float width,
              unsigned char* __restrict__ flags, Sum4* __restrict__ partial) {
  using PL = Planes<2, 0, 1>;
  constexpr int NT = NW * 64;
  __shared__ float X[NB * WMAX];
  __shared__ float P[MASS ? NB * 3 : 1];
  double cnt = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    __syncthreads();  // the previous tile's epilogue has read P
    for (int e = threadIdx.x; e < NB * 16; e += NT) {
      const int b = e >> 4, k = e & 15;
      if (k >= 3) X[xoff<NB, WMAX>(0, b, k)] = 0.f;
    }
    for (int b = threadIdx.x; b < NB; b += NT) {
      float p[3] = {0.f, 0.f, 0.f};
      if (i0 + b < n) sample_candidate(S, S.j0 + (unsigned long long)(i0 + b), p);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        X[xoff<NB, WMAX>(0, b, k)] = p[k];
        if constexpr (MASS) P[3 * b + k] = p[k];
      }
    }
    mlp_forward<PL, 2, 0, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i >= n) return;
      if constexpr (MASS) {
        if (v[0] < 0.f) {
          cnt += 1.0;
          sx += (double)P[3 * b];
          sy += (double)P[3 * b + 1];
          sz += (double)P[3 * b + 2];
        }
      } else {
        flags[i] = fabsf(v[0]) < width ? 1 : 0;
      }
    });
  }
  if constexpr (MASS) {
    // the epilogue runs on lanes 0..15 of wave 0
    if (threadIdx.x < 64) {
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) {
        cnt += __shfl_xor(cnt, o);
        sx += __shfl_xor(sx, o);
        sy += __shfl_xor(sy, o);
        sz += __shfl_xor(sz, o);
      }
      if (threadIdx.x == 0) partial[blockIdx.x] = Sum4{cnt, sx, sy, sz};
    }
  }
}

// Accepted candidates per block of kSampBlock flags.
static __global__ void __launch_bounds__(kSampBlock)
k_samp_count(const unsigned char* __restrict__ flags, long long n, int* __restrict__ bcnt) {
  __shared__ int wsum[kSampBlock / 64];
  const long long i = (long long)blockIdx.x * kSampBlock + threadIdx.x;
  const int f = (i < n) ? flags[i] : 0;
  const unsigned long long m = __ballot(f != 0);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kSampBlock / 64; ++w) s += wsum[w];
    bcnt[blockIdx.x] = s;
  }
}

// Exclusive offsets of the blocks and the output rows this round fills.
static __global__ void __launch_bounds__(kSampBlock)
k_samp_scan(const int* __restrict__ bcnt, int nb, long long* __restrict__ bofs, SampCounters* __restrict__ C) {
  __shared__ long long part[kSampBlock];
  const int per = (nb + kSampBlock - 1) / kSampBlock;
  const int b0 = min(nb, (int)threadIdx.x * per), b1 = min(nb, b0 + per);
  long long s = 0;
  for (int b = b0; b < b1; ++b) s += bcnt[b];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long run = 0;
    for (int t = 0; t < kSampBlock; ++t) {
      const long long v = part[t];
      part[t] = run;
      run += v;
    }
    C->base = C->found;
    C->found = min(C->found + run, C->n_samples);
  }
  __syncthreads();
  long long off = part[threadIdx.x];
  for (int b = b0; b < b1; ++b) {
    bofs[b] = off;
    off += bcnt[b];
  }
}

// Accepted candidates to their rows, in candidate order; rows past the
// output are dropped (src/kd_tree.py:242-246).
static __global__ void __launch_bounds__(kSampBlock)
k_samp_scatter(const unsigned char* __restrict__ flags, long long n, const long long* __restrict__ bofs,
               const SampCounters* __restrict__ C, SampleSrc S, float* __restrict__ out) {
  __shared__ int wsum[kSampBlock / 64];
  const long long i = (long long)blockIdx.x * kSampBlock + threadIdx.x;
  const int f = (i < n) ? flags[i] : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(f != 0);
  if (lane == 0) wsum[wave] = __popcll(m);
  __syncthreads();
  if (!f) return;
  int wp = 0;
  for (int w = 0; w < wave; ++w) wp += wsum[w];
  const int lp = __popcll(m & ((1ull << lane) - 1ull));
  const long long row = C->base + bofs[blockIdx.x] + wp + lp;
  if (row >= C->n_samples) return;
  float p[3];
  sample_candidate(S, S.j0 + (unsigned long long)i, p);
#pragma unroll
  for (int k = 0; k < 3; ++k) out[3 * row + k] = p[k];
}

// Volume and first moment of n boxes, one partial per workgroup.
static __global__ void __launch_bounds__(256)
k_box_sums(const float* __restrict__ lower, const float* __restrict__ upper, long long n, Sum4* __restrict__ partial) {
  __shared__ Sum4 red[256];
  Sum4 a{0.0, 0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double l[3], u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      l[k] = (double)lower[3 * i + k];
      u[k] = (double)upper[3 * i + k];
    }
    const double v = (u[0] - l[0]) * (u[1] - l[1]) * (u[2] - l[2]);
    a.c += v;
    a.x += v * 0.5 * (l[0] + u[0]);
    a.y += v * 0.5 * (l[1] + u[1]);
    a.z += v * 0.5 * (l[2] + u[2]);
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[threadIdx.x].c += red[threadIdx.x + o].c;
      red[threadIdx.x].x += red[threadIdx.x + o].x;
      red[threadIdx.x].y += red[threadIdx.x + o].y;
      red[threadIdx.x].z += red[threadIdx.x + o].z;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Fixed-order sum of np partials into *out.
static __global__ void __launch_bounds__(256)
k_sum_partials(const Sum4* __restrict__ partial, int np, Sum4* __restrict__ out) {
  __shared__ Sum4 red[256];
  Sum4 a{0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < np; i += 256) {
    a.c += partial[i].c;
    a.x += partial[i].x;
    a.y += partial[i].y;
    a.z += partial[i].z;
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[threadIdx.x].c += red[threadIdx.x + o].c;
      red[threadIdx.x].x += red[threadIdx.x + o].x;
      red[threadIdx.x].y += red[threadIdx.x + o].y;
      red[threadIdx.x].z += red[threadIdx.x + o].z;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}
