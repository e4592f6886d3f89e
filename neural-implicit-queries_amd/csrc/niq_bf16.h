// Precision mode: the error-term contraction in BF16 on MFMA with FP32
// accumulation, made rigorous again by widening.
//
// Each activation X (FP32) is held as X = hi + lo with hi = bf16(X),
// lo = bf16(X - hi) (|X - hi - lo| <= 2^-16 |X|); weights likewise.  A
// dense layer computes, per plane,  hi.Whi + lo.Whi + hi.Wlo  with
// v_mfma_f32_16x16x32_bf16 (every bf16 x bf16 product is exact in FP32).  The
// dropped terms and the FP32 accumulation are bounded a priori by
//     |X.W - computed| <= c |X|.|W|,   c = 3.1 * 2^-16 + gamma_{3K}(2^-23),
// and summed over the planes of a box this is c (sum_p |X_p|).|W|.  That
// bound rides in one extra non-negative plane e = err + c sum_p |X_p|,
// rounded up to bf16 and multiplied by |W| rounded up to bf16, so the
// computed interval error already contains the widening.  Elementwise FP32
// steps carry a few-ulp margin, and the last radius-versus-threshold test is
// done in FP64.  Only the affine forms (interval, affine_fixed) run here.
#pragma once
#include "niq_device.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Packed bf16 weight fragments of one layer (K padded to 32): for fragment
// (ft, kb) lane l holds 8 values W[32 kb + 8 (l / 16) + s][16 ft + l % 16].
struct DevLayerBf16 {
  int in_pad;    // multiple of 32
  int out_pad;   // multiple of 16
  int act;
  int pad_;
  const bf16x8* Wh;   // bf16(W)
  const bf16x8* Wl;   // bf16(W - bf16(W))
  const bf16x8* Wa;   // |W| rounded up to bf16
  const float* bias;
  float c_split;      // a-priori contraction bound c of this layer
  float e_fix;        // (1 + 3 gamma_K): FP32 accumulation of the e plane
};

struct DevModelBf16 {
  int n_layers;
  int wmax;
  int pad_[2];
  DevLayerBf16 L[NIQ_MAX_LAYERS];
};

// -------------------------------------------------------------- bf16 bits
__device__ __forceinline__ unsigned short bf16_bits_rn(float x) {
  const unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
// smallest bf16 >= x, for x >= 0 (or +inf)
__device__ __forceinline__ unsigned short bf16_bits_up_nonneg(float x) {
  const unsigned u = __float_as_uint(x);
  unsigned short h = (unsigned short)(u >> 16);
  if ((u & 0xffffu) && (u & 0x7fffffffu) < 0x7f800000u) h = (unsigned short)(h + 1);
  return h;
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ __bf16 bits_to_bf16(unsigned short h) { return __builtin_bit_cast(__bf16, h); }
// next float up (x finite and >= 0)
__device__ __forceinline__ float up1(float x) {
  return (x >= 0.f && x < INFINITY) ? __uint_as_float(__float_as_uint(x) + 1u) : x;
}

// LDS layout: X[plane][box][k] as bf16 with the 16-byte chunk index
// XOR-swizzled by box (conflict-free 16 B reads and 8 B writes).
template <int NB, int WMAX>
__device__ __forceinline__ int boff16(int plane, int b, int k) {
  return (plane * NB + b) * WMAX + ((((k >> 3) ^ (b & 7)) << 3) | (k & 7));
}

// Planes: for p in 0..V: hi at plane 2p, lo at plane 2p+1 (base p = 0,
// aff_i p = 1 + i); the e plane is plane 2(V+1).
template <int V> struct PlanesBf16 {
  static constexpr int NS = V + 1;          // split planes
  static constexpr int NPL = 2 * NS + 1;    // bf16 planes in LDS
};

constexpr float kU24 = 5.9604644775390625e-08f;   // 2^-24

// Store one (box, feature) element's post-activation affine form.
template <int V, int NB, int WMAX>
__device__ __forceinline__ void store_split(unsigned short* Xs, int b, int k, const float* s, float err, float c) {
  float mag = 0.f;
#pragma unroll
  for (int p = 0; p <= V; ++p) {
    const unsigned short h = bf16_bits_rn(s[p]);
    const unsigned short l = bf16_bits_rn(s[p] - bf16_to_f(h));
    Xs[boff16<NB, WMAX>(2 * p, b, k)] = h;
    Xs[boff16<NB, WMAX>(2 * p + 1, b, k)] = l;
    mag = mag + fabsf(s[p]);
  }
  const float e = up1(up1(err + c * mag));
  Xs[boff16<NB, WMAX>(2 * (V + 1), b, k)] = bf16_bits_up_nonneg(e);
}

// Rigorous affine rule for one element: v[0] base, v[1..V] aff, v[V+1] err
// (pre-activation, err already contains the contraction widening).
template <int V>
__device__ __forceinline__ void affine_rule_rigorous(int act, float* v) {
  if (act == ACT_NONE) {   // identity: only the rounding of base + bias to absorb
    float amag = 0.f;
#pragma unroll
    for (int i = 0; i < V; ++i) amag = amag + fabsf(v[1 + i]);
    v[V + 1] = up1(up1((v[V + 1] + 4.f * kU24 * (fabsf(v[0]) + amag)) * (1.f + 4.f * kU24)));
    return;
  }
  float rad = 0.f;
#pragma unroll
  for (int i = 0; i < V; ++i) rad = rad + fabsf(v[1 + i]);
  rad = rad + v[V + 1];
  rad = up1(up1((rad + 4.f * kU24 * fabsf(v[0])) * (1.f + 8.f * kU24)));
  const float lo = v[0] - rad, hi = v[0] + rad;
  float a, b, d;
  act_rule(act, lo, hi, a, b, d);
  // the error of the chosen line on [lo, hi], re-evaluated and padded for
  // the rounding of alpha, beta and the transcendental evaluations
  float m;
  if (act == ACT_RELU) {
    const float e0 = fabsf(relu_t(lo) - (a * lo + b));
    const float e1 = fabsf(relu_t(hi) - (a * hi + b));
    d = fmaxf(fmaxf(fabsf(d), fabsf(b)), fmaxf(e0, e1));
    m = 16.f * kU24 * (fabsf(lo) + fabsf(hi));
  } else {
    d = fabsf(d);
    m = 64.f * kU24 * (fabsf(b) + a * (fabsf(lo) + fabsf(hi)) + 1.f);
  }
  const float base = a * v[0] + b;
  float amag = 0.f;
  v[0] = base;
#pragma unroll
  for (int i = 0; i < V; ++i) { v[1 + i] = a * v[1 + i]; amag = amag + fabsf(v[1 + i]); }
  // err' = a err + delta, plus the rounding of base' and aff'
  const float err = a * v[V + 1] + d + m + 4.f * kU24 * (fabsf(base) + amag);
  v[V + 1] = up1(up1(err * (1.f + 4.f * kU24)));
}

// Fused forward in precision mode (affine forms only: KIND 0).  Xs holds
// PlanesBf16<V>::NPL bf16 planes.  fin(v, b) gets base, aff_i, err (FP32,
// err rigorous) of output feature 0.
template <int V, int NB, int NW, int WMAX, class Fin>
__device__ __forceinline__ void mlp_forward_bf16(const DevModelBf16& M, unsigned short* __restrict__ Xs, Fin&& fin) {
  constexpr int NS = V + 1;
  constexpr int R = NS + 1;   // accumulators: split planes + e plane
  constexpr int BT = NB / 16;
  constexpr int FTMAX = (WMAX / 16 + NW - 1) / NW;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ln = lane & 15;
  const int lg = lane >> 4;

  for (int l = 0; l < M.n_layers; ++l) {
    const DevLayerBf16 Ly = M.L[l];
    const int nkb = Ly.in_pad >> 5;
    const int nft = Ly.out_pad >> 4;
    const bool last = (l == M.n_layers - 1);
    floatx4 acc[FTMAX][R][BT];
#pragma unroll
    for (int f = 0; f < FTMAX; ++f)
#pragma unroll
      for (int p = 0; p < R; ++p)
#pragma unroll
        for (int t = 0; t < BT; ++t) acc[f][p][t] = floatx4{0.f, 0.f, 0.f, 0.f};

    __syncthreads();
    for (int kb = 0; kb < nkb; ++kb) {
      bf16x8 wh[FTMAX], wl[FTMAX], wa[FTMAX];
#pragma unroll
      for (int f = 0; f < FTMAX; ++f) {
        const int ft = wave + f * NW;
        if (ft < nft) {
          const size_t o = (size_t)(ft * nkb + kb) * 64 + lane;
          wh[f] = Ly.Wh[o]; wl[f] = Ly.Wl[o]; wa[f] = Ly.Wa[o];
        }
      }
#pragma unroll
      for (int t = 0; t < BT; ++t) {
        const int b = t * 16 + ln;
        const int k0 = kb * 32 + lg * 8;
#pragma unroll
        for (int p = 0; p < NS; ++p) {
          const bf16x8 xh = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(2 * p, b, k0)]);
          const bf16x8 xl = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(2 * p + 1, b, k0)]);
#pragma unroll
          for (int f = 0; f < FTMAX; ++f) {
            const int ft = wave + f * NW;
            if (ft < nft) {
              acc[f][p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[f], xh, acc[f][p][t], 0, 0, 0);
              acc[f][p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[f], xl, acc[f][p][t], 0, 0, 0);
              acc[f][p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[f], xh, acc[f][p][t], 0, 0, 0);
            }
          }
        }
        const bf16x8 xe = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(2 * NS, b, k0)]);
#pragma unroll
        for (int f = 0; f < FTMAX; ++f) {
          const int ft = wave + f * NW;
          if (ft < nft) acc[f][NS][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[f], xe, acc[f][NS][t], 0, 0, 0);
        }
      }
    }

    __syncthreads();

#pragma unroll
    for (int f = 0; f < FTMAX; ++f) {
      const int ft = wave + f * NW;
      if (ft >= nft) continue;
      const floatx4 bias4 = *reinterpret_cast<const floatx4*>(&Ly.bias[ft * 16 + lg * 4]);
#pragma unroll
      for (int t = 0; t < BT; ++t) {
        const int b = t * 16 + ln;
        unsigned short hs[NS][4], ls[NS][4], es[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v[V + 2];
          v[0] = acc[f][0][t][j] + bias4[j];
#pragma unroll
          for (int i = 0; i < V; ++i) v[1 + i] = acc[f][1 + i][t][j];
          v[V + 1] = up1(acc[f][NS][t][j] * Ly.e_fix);
          if (last) {
            if (ft == 0 && lg == 0 && j == 0) fin(v, b);
            continue;
          }
          affine_rule_rigorous<V>(Ly.act, v);
          float mag = 0.f;
#pragma unroll
          for (int p = 0; p < NS; ++p) {
            hs[p][j] = bf16_bits_rn(v[p]);
            ls[p][j] = bf16_bits_rn(v[p] - bf16_to_f(hs[p][j]));
            mag = mag + fabsf(v[p]);
          }
          // next layer's widening rides in the e plane (c of the next layer)
          const float cn = M.L[l + 1].c_split;
          es[j] = bf16_bits_up_nonneg(up1(up1(v[V + 1] + cn * up1(mag * (1.f + 8.f * kU24)))));
        }
        if (!last) {
          const int k = ft * 16 + lg * 4;
#pragma unroll
          for (int p = 0; p < NS; ++p) {
            ushort4 h4 = {hs[p][0], hs[p][1], hs[p][2], hs[p][3]};
            ushort4 l4 = {ls[p][0], ls[p][1], ls[p][2], ls[p][3]};
            *reinterpret_cast<ushort4*>(&Xs[boff16<NB, WMAX>(2 * p, b, k)]) = h4;
            *reinterpret_cast<ushort4*>(&Xs[boff16<NB, WMAX>(2 * p + 1, b, k)]) = l4;
          }
          ushort4 e4 = {es[0], es[1], es[2], es[3]};
          *reinterpret_cast<ushort4*>(&Xs[boff16<NB, WMAX>(2 * NS, b, k)]) = e4;
        }
      }
    }
  }
}

// Final bounds in FP64 from the rigorous FP32 form (base, aff, err).
template <int V>
__device__ __forceinline__ void final_bounds_fp64(const float* v, double& lo, double& hi, double& scale) {
  double rad = 0.0;
#pragma unroll
  for (int i = 0; i < V; ++i) rad += ::fabs((double)v[1 + i]);
  rad += (double)v[V + 1];
  rad += 4.0 * (double)kU24 * ::fabs((double)v[0]);   // rounding of base + bias
  lo = (double)v[0] - rad;
  hi = (double)v[0] + rad;
  scale = ::fabs((double)v[0]) + rad;
}

template <int V, int NB, int NW, int WMAX>
__global__ void __launch_bounds__(NW * 64)
k_classify_bf16(const DevModelBf16* __restrict__ Mp, ClassifyArgs a) {
  using PB = PlanesBf16<V>;
  constexpr int NT = NW * 64;
  __shared__ __attribute__((aligned(16))) unsigned short Xs[PB::NPL * NB * WMAX];
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  const long long ntiles = (n + NB - 1) / NB;
  const float c0 = Mp->L[0].c_split;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    // layer-0 inputs: features 0..31 (K padded to 32)
    for (int e = threadIdx.x; e < NB * 32; e += NT) {
      const int b = e >> 5, k = e & 31;
      const long long i = i0 + b;
      float s[V + 1];
#pragma unroll
      for (int p = 0; p <= V; ++p) s[p] = 0.f;
      float err = 0.f;
      if (i < n && k < a.d) {
        if (a.lower) {
          const float lo = a.lower[i * a.d + k], up = a.upper[i * a.d + k];
          const float c = 0.5f * (lo + up), h = up - c;
          s[0] = c;
          if (V == 0) err = fabsf(h);
#pragma unroll
          for (int p = 1; p <= V; ++p) s[p] = (p - 1 == k) ? h : 0.f;
        } else {
          s[0] = a.center[i * a.d + k];
          if (V == 0) {
            for (int r = 0; r < a.vin; ++r) err = err + fabsf(a.vecs[(i * a.vin + r) * a.d + k]);
            err = up1(err * (1.f + 4.f * kU24));
          }
#pragma unroll
          for (int p = 1; p <= V; ++p) s[p] = a.vecs[(i * a.vin + (p - 1)) * a.d + k];
        }
      }
      store_split<V, NB, WMAX>(Xs, b, k, s, err, c0);
    }
    mlp_forward_bf16<V, NB, NW, WMAX>(*Mp, Xs, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i >= n) return;
      double lo, hi, scale;
      final_bounds_fp64<V>(v, lo, hi, scale);
      a.codes[i] = sign_code<double>(lo, hi, (double)a.offset);
      if (a.bounds) {
        // outward rounding to float
        float flo = (float)lo, fhi = (float)hi;
        if ((double)flo > lo) flo = nextafterf(flo, -INFINITY);
        if ((double)fhi < hi) fhi = nextafterf(fhi, INFINITY);
        a.bounds[2 * i] = flo; a.bounds[2 * i + 1] = fhi;
      }
    });
  }
}
