// Precision mode: the error-term contraction in BF16 on MFMA with FP32
// accumulation, made rigorous again by widening.
//
// Activations and weights are split into bf16 terms: the base row (magnitude
// O(1)) three ways, X = x1 + x2 + x3 (|residual| <= 2^-24 |X|, W likewise),
// the error-term rows (magnitude ~ box size) two ways.  A dense layer then
// sums exact bf16 x bf16 products with v_mfma_f32_16x16x32_bf16, smallest
// terms first:
//     base:  x3 w1 + x2 w2 + x1 w3 + x2 w1 + x1 w2 + x1 w1
//     aff:   a2 w1 + a1 w2 + a1 w1
// into a fresh accumulator per 32-wide K chunk, and the chunk results are
// added in FP32.  The dropped terms plus both rounding stages are bounded a
// priori by |X.W - computed| <= c |X|.|W|, with
//     c = split residual + 1.05 gamma_33(2^-23) + gamma_{K/32}(2^-24)
// (unit 2^-23 covers any FP32 rounding mode inside the MFMA), so c_base is a
// few 1e-6 and c_aff a few 1e-5.  The bound rides in one extra non-negative
// plane e = err + c_base |base| + c_aff sum_i |aff_i|, rounded up to bf16 and
// multiplied by |W| rounded up to bf16: the computed interval error already
// contains the widening.  Elementwise FP32 steps carry a few-ulp margin, and
// the last radius-versus-threshold test is done in FP64.  Only the affine
// forms (interval, affine_fixed) run here.
#pragma once
#include "niq_device.h"
#include "niq_kernels.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Packed bf16 weight fragments of one layer (K padded to 32): for fragment
// (ft, kb) lane l holds 8 values W[32 kb + 8 (l / 16) + s][16 ft + l % 16].
// The same fragments also exist once in FP32 (W32: [fragment][half][lane]
// float4, half h holding s = 4h..4h+3), split into bf16 terms in registers.
struct DevLayerBf16 {
  int in_pad;    // multiple of 32
  int out_pad;   // multiple of 16
  int act;
  int pad_;
  const bf16x8* W1;   // w1 = bf16(W)
  const bf16x8* W2;   // w2 = bf16(W - w1)
  const bf16x8* W3;   // w3 = bf16(W - w1 - w2)
  const bf16x8* Wa;   // |W| rounded up to bf16
  const floatx4* W32; // W itself (FP32), for the in-register split
  const float* bias;
  float c_base;       // a-priori contraction bound of the base row
  float c_aff;        // ... of the error-term rows
  float e_fix;        // 1 + bound of the (non-negative) e-plane accumulation
  float e_fix32;      // the same for e |w1| + e |w2| (in-register split)
};

// Weight sources of the precision-mode kernel.
enum { WSRC_BF16 = 0, WSRC_FP32 = 1 };

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

typedef float floatx2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef unsigned uintx4_t __attribute__((ext_vector_type(4)));

// Two floats to packed bf16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned pk_bf16_rn(float a, float b) {
  const floatx2_t x = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(x, bf16x2_t));
}

// In-register split of 8 FP32 weights into W = w1 + w2 + w3 exactly (each
// term the bf16 rounding of the remainder, as the packed copies hold), plus
// |w1| and |w2|: |W| <= (|w1| + |w2|) / (1 - 2^-16).
__device__ __forceinline__ void split_w8(const floatx4 lo, const floatx4 hi, bf16x8& w1, bf16x8& w2, bf16x8& w3,
                                         bf16x8& a1, bf16x8& a2) {
  uintx4_t u1, u2, u3;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float x0 = q < 2 ? lo[2 * q] : hi[2 * q - 4];
    const float x1 = q < 2 ? lo[2 * q + 1] : hi[2 * q - 3];
    const unsigned h1 = pk_bf16_rn(x0, x1);
    const float r0 = x0 - __uint_as_float(h1 << 16);
    const float r1 = x1 - __uint_as_float(h1 & 0xffff0000u);
    const unsigned h2 = pk_bf16_rn(r0, r1);
    const float s0 = r0 - __uint_as_float(h2 << 16);
    const float s1 = r1 - __uint_as_float(h2 & 0xffff0000u);
    u1[q] = h1;
    u2[q] = h2;
    u3[q] = pk_bf16_rn(s0, s1);
  }
  w1 = __builtin_bit_cast(bf16x8, u1);
  w2 = __builtin_bit_cast(bf16x8, u2);
  w3 = __builtin_bit_cast(bf16x8, u3);
  a1 = __builtin_bit_cast(bf16x8, u1 & 0x7fff7fffu);
  a2 = __builtin_bit_cast(bf16x8, u2 & 0x7fff7fffu);
}

// Base value split three ways for two features at once (packed pairs, low
// half the first feature): x = h1 + h2 + h3 + r, |r| <= 2^-24 |x|.
__device__ __forceinline__ void split3_pair(float a, float b, unsigned& h1, unsigned& h2, unsigned& h3) {
  h1 = pk_bf16_rn(a, b);
  const float ra = a - __uint_as_float(h1 << 16), rb = b - __uint_as_float(h1 & 0xffff0000u);
  h2 = pk_bf16_rn(ra, rb);
  h3 = pk_bf16_rn(ra - __uint_as_float(h2 << 16), rb - __uint_as_float(h2 & 0xffff0000u));
}
// Error-term value split two ways for two features.
__device__ __forceinline__ void split2_pair(float a, float b, unsigned& h1, unsigned& h2) {
  h1 = pk_bf16_rn(a, b);
  h2 = pk_bf16_rn(a - __uint_as_float(h1 << 16), b - __uint_as_float(h1 & 0xffff0000u));
}
// Two non-negative floats rounded up to bf16, packed.
__device__ __forceinline__ unsigned pk_bf16_up_nonneg(float a, float b) {
  return ((__float_as_uint(a) + 0xffffu) >> 16) | ((__float_as_uint(b) + 0xffffu) & 0xffff0000u);
}

// Absolute floor added to every e-plane value: covers the split of values
// below the FP32 normal range (at most 2^-126 each) whatever the conversion's
// subnormal handling.
constexpr float kTinyE = 1e-30f;

// LDS layout: X[plane][box][k] as bf16 with the 16-byte chunk index
// XOR-swizzled by box (conflict-free 16 B reads and 8 B writes).
template <int NB, int WMAX>
__device__ __forceinline__ int boff16(int plane, int b, int k) {
  return (plane * NB + b) * WMAX + ((((k >> 3) ^ (b & 7)) << 3) | (k & 7));
}

// LDS planes: base x1, x2, x3 at 0, 1, 2; aff_i terms a1, a2 at 3 + 2i,
// 4 + 2i; the e plane last.
template <int V> struct PlanesBf16 {
  static constexpr int NS = V + 1;          // affine rows (base + aff)
  static constexpr int NPL = 3 + 2 * V + 1; // bf16 planes in LDS
  static constexpr int E = 3 + 2 * V;       // index of the e plane
};

// Waves per SIMD that a tile's bf16 planes admit (workgroups per CU by LDS x
// waves per workgroup / 4 SIMDs), capped at 4: the fused kernel's register
// budget, so the compiler does not trade registers for occupancy LDS
// already rules out (at width 512 one workgroup per CU: 256 VGPRs a wave).
template <int V, int NB, int NW, int WMAX>
struct BfOcc {
  static constexpr int wgs = (int)((size_t)kLdsLimit / ((size_t)PlanesBf16<V>::NPL * NB * WMAX * 2));
  static constexpr int raw = wgs * NW / 4;
  static constexpr int value = raw < 1 ? 1 : (raw > 4 ? 4 : raw);
};

constexpr float kU24 = 5.9604644775390625e-08f;   // 2^-24

// Split and store one (box, feature) element's affine form: base three ways,
// aff two ways, and the e plane err + c_base |base| + c_aff sum |aff|.
template <int V, int NB, int WMAX>
__device__ __forceinline__ void store_split(unsigned short* Xs, int b, int k, const float* s, float err,
                                            float c_base, float c_aff) {
  const unsigned short x1 = bf16_bits_rn(s[0]);
  const float r1 = s[0] - bf16_to_f(x1);
  const unsigned short x2 = bf16_bits_rn(r1);
  const unsigned short x3 = bf16_bits_rn(r1 - bf16_to_f(x2));
  Xs[boff16<NB, WMAX>(0, b, k)] = x1;
  Xs[boff16<NB, WMAX>(1, b, k)] = x2;
  Xs[boff16<NB, WMAX>(2, b, k)] = x3;
  float amag = 0.f;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const unsigned short h = bf16_bits_rn(s[1 + i]);
    Xs[boff16<NB, WMAX>(3 + 2 * i, b, k)] = h;
    Xs[boff16<NB, WMAX>(4 + 2 * i, b, k)] = bf16_bits_rn(s[1 + i] - bf16_to_f(h));
    amag = amag + fabsf(s[1 + i]);
  }
  const float e = up1(up1(err + c_base * fabsf(s[0]) + up1(c_aff * up1(amag * (1.f + 4.f * kU24))) + kTinyE));
  Xs[boff16<NB, WMAX>(3 + 2 * V, b, k)] = bf16_bits_up_nonneg(e);
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
  if (act == ACT_RELU) {
    // any slope in [0, 1] is sound here: the line's error is re-evaluated
    // below, so the chord slope comes from the reciprocal estimate
    if (lo >= 0.f) { a = 1.f; b = 0.f; d = 0.f; }
    else if (hi <= 0.f) { a = 0.f; b = 0.f; d = 0.f; }
    else {
      a = fminf(fmaxf(hi * __builtin_amdgcn_rcpf(hi - lo), 0.f), 1.f);
      b = -0.5f * (a * lo);
      d = b;
    }
  } else {
    act_rule(act, lo, hi, a, b, d);
  }
  if (b == 0.f && d == 0.f && (a == 1.f || a == 0.f)) {
    // determinate feature (identity or zero): exact in FP32, no margin needed
#pragma unroll
    for (int i = 0; i <= V + 1; ++i) v[i] = a * v[i];
    return;
  }
  // The chosen line's error on [lo, hi], re-evaluated, plus margins for the
  // rounding of alpha x + beta and of the activation values (|act(x)| <= |x|
  // for relu, elu, tanh and sin, so both scale with |lo| + |hi|).
  const float span = fabsf(lo) + fabsf(hi);
  if (act == ACT_RELU) {
    const float e0 = fabsf(relu_t(lo) - (a * lo + b));
    const float e1 = fabsf(relu_t(hi) - (a * hi + b));
    d = fmaxf(fmaxf(fabsf(d), fabsf(b)), fmaxf(e0, e1));
  } else {
    d = fabsf(d);
  }
  const float m = 8.f * kU24 * (span + fabsf(b));
  const float pa = a * v[0];
  const float base = pa + b;
  float amag = 0.f;
  v[0] = base;
#pragma unroll
  for (int i = 0; i < V; ++i) { v[1 + i] = a * v[1 + i]; amag = amag + fabsf(v[1 + i]); }
  // err' = a err + delta + margins for the rounding of base' and aff'
  const float err = a * v[V + 1] + d + m + 2.f * kU24 * (fabsf(pa) + fabsf(b) + amag);
  v[V + 1] = up1(up1(err * (1.f + 4.f * kU24)));
}

// The next layer's bf16 planes of 4 consecutive features of one box, as
// packed pairs: base three ways, aff two ways, and the e plane
// err + c_base |base| + c_aff sum |aff| (+ kTinyE) rounded up.
template <int V> struct OutPlanes {
  unsigned x[3][2];
  unsigned a[V > 0 ? V : 1][2][2];
  unsigned e[2];
};

template <int V, int NV>
__device__ __forceinline__ void split_out4(float (&v)[4][NV], float cb_next, float ca_next, OutPlanes<V>& o) {
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    split3_pair(v[2 * q][0], v[2 * q + 1][0], o.x[0][q], o.x[1][q], o.x[2][q]);
#pragma unroll
    for (int i = 0; i < V; ++i) split2_pair(v[2 * q][1 + i], v[2 * q + 1][1 + i], o.a[i][0][q], o.a[i][1][q]);
    float ev[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float* w = v[2 * q + h];
      float amag = 0.f;
#pragma unroll
      for (int i = 0; i < V; ++i) amag = amag + fabsf(w[1 + i]);
      ev[h] = up1(up1(w[V + 1] + cb_next * fabsf(w[0]) + up1(ca_next * up1(amag * (1.f + 4.f * kU24))) + kTinyE));
    }
    o.e[q] = pk_bf16_up_nonneg(ev[0], ev[1]);
  }
}

// Fused forward in precision mode (affine forms only: KIND 0).  Xs holds
// PlanesBf16<V>::NPL bf16 planes.  fin(v, b) gets base, aff_i, err (FP32,
// err rigorous) of output feature 0.
// NP plain point rows (for the ray caster's hit test) ride after the range
// form, each split NPS ways (planes PB::NPL + r * NPS + s); they get the
// plain activation and reach `fin` in v[V + 2 + r].
template <int V, int NB, int NW, int WMAX, int WS, int NP = 0, int NPS = 3, int AS = ACTS_ALL, class Fin>
__device__ __forceinline__ void mlp_forward_bf16(const DevModelBf16& M, unsigned short* __restrict__ Xs, Fin&& fin) {
  using PB = PlanesBf16<V>;
  constexpr int NS = V + 1;
  constexpr int R = NS + 1;   // accumulators: base, aff_i, e
  constexpr int RA = R + NP;  // ... and the point rows
  constexpr int BT = NB / 16;
  constexpr int FTMAX = (WMAX / 16 + NW - 1) / NW;
  const int lane = threadIdx.x & 63;
  // wave-uniform, so the feature-tile tests below are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ln = lane & 15;
  const int lg = lane >> 4;
  const floatx4 z4 = {0.f, 0.f, 0.f, 0.f};

  for (int l = 0; l < M.n_layers; ++l) {
    const DevLayerBf16 Ly = M.L[l];
    const int nkb = Ly.in_pad >> 5;
    const int nft = Ly.out_pad >> 4;
    const bool last = (l == M.n_layers - 1);
    floatx4 acc[FTMAX][RA][BT];
#pragma unroll
    for (int f = 0; f < FTMAX; ++f)
#pragma unroll
      for (int p = 0; p < RA; ++p)
#pragma unroll
        for (int t = 0; t < BT; ++t) acc[f][p][t] = z4;

    __syncthreads();
    // Weight fragments are fetched one K chunk ahead: FP32 fragments (split
    // here) or the packed bf16 copies.  `FULL`: every feature tile of this
    // wave exists, so the loop carries no per-tile test.
    auto contract = [&](auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
      floatx4 wpf[FTMAX][2];
      bf16x8 wnx[FTMAX][4];
      // global loads (vmcnt only), so the LDS wait before a chunk's first
      // MFMA does not wait for the next chunk's weights as well
      auto fetch = [&](int f, int kb) {
        const int ft = wave + f * NW;
        if constexpr (WS == WSRC_FP32) {
          const auto P = as_global(Ly.W32) + 2 * (size_t)(ft * nkb + kb) * 64 + lane;
          wpf[f][0] = P[0];
          wpf[f][1] = P[64];
        } else {
          const size_t o = (size_t)(ft * nkb + kb) * 64 + lane;
          wnx[f][0] = as_global(Ly.W1)[o]; wnx[f][1] = as_global(Ly.W2)[o];
          wnx[f][2] = as_global(Ly.W3)[o]; wnx[f][3] = as_global(Ly.Wa)[o];
        }
      };
#pragma unroll
      for (int f = 0; f < FTMAX; ++f)
        if (FULL || wave + f * NW < nft) fetch(f, 0);
      for (int kb = 0; kb < nkb; ++kb) {
#pragma unroll
        for (int t = 0; t < BT; ++t) {
          const int b = t * 16 + ln;
          const int k0 = kb * 32 + lg * 8;
          const bf16x8 x1 = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(0, b, k0)]);
          const bf16x8 x2 = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(1, b, k0)]);
          const bf16x8 x3 = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(2, b, k0)]);
          const bf16x8 xe = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(PB::E, b, k0)]);
          bf16x8 xa[V > 0 ? V : 1][2];   // error-term rows, read once for all feature tiles
#pragma unroll
          for (int i = 0; i < V; ++i) {
            xa[i][0] = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(3 + 2 * i, b, k0)]);
            xa[i][1] = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(4 + 2 * i, b, k0)]);
          }
#pragma unroll
          for (int f = 0; f < FTMAX; ++f) {
            const int ft = wave + f * NW;
            if (!FULL && ft >= nft) continue;
            bf16x8 w1, w2, w3, wa1, wa2;
            if constexpr (WS == WSRC_FP32) {
              split_w8(wpf[f][0], wpf[f][1], w1, w2, w3, wa1, wa2);
            } else {
              w1 = wnx[f][0]; w2 = wnx[f][1]; w3 = wnx[f][2]; wa1 = wnx[f][3];
            }
            // the next chunk's fragment goes into the same registers; the
            // last chunk fetches itself again, so the loop body has no branch
            // and the load waits count only what is needed
            if (t + 1 == BT) fetch(f, kb + 1 < nkb ? kb + 1 : kb);
            // base row, smallest products first, fresh accumulator per chunk
            {
              floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, x3, z4, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, x2, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3, x1, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, x2, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, x1, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, x1, c, 0, 0, 0);
              acc[f][0][t] += c;
            }
#pragma unroll
            for (int i = 0; i < V; ++i) {
              const bf16x8 a1 = xa[i][0];
              const bf16x8 a2 = xa[i][1];
              floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, a2, z4, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, a1, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, a1, c, 0, 0, 0);
              acc[f][1 + i][t] += c;
            }
            if constexpr (WS == WSRC_FP32) {
              // non-negative terms: one accumulator over the whole K
              acc[f][NS][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa2, xe, acc[f][NS][t], 0, 0, 0);
              acc[f][NS][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, xe, acc[f][NS][t], 0, 0, 0);
            } else {
              acc[f][NS][t] += __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, xe, z4, 0, 0, 0);
            }
            // plain point rows (hit test): the same split products, no bound
#pragma unroll
            for (int r = 0; r < NP; ++r) {
              const int pl = PB::NPL + r * NPS;
              const bf16x8 q1 = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(pl, b, k0)]);
              const bf16x8 q2 = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(pl + 1, b, k0)]);
              floatx4 c;
              if constexpr (NPS == 3) {
                const bf16x8 q3 = *reinterpret_cast<const bf16x8*>(&Xs[boff16<NB, WMAX>(pl + 2, b, k0)]);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, q3, z4, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, q2, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3, q1, c, 0, 0, 0);
              } else {
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3, q1, z4, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, q2, c, 0, 0, 0);
              }
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, q2, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, q1, c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, q1, c, 0, 0, 0);
              acc[f][R + r][t] += c;
            }
          }
        }
        // keeps the next chunk's weight loads in this iteration (issued one
        // chunk ahead); otherwise they may be moved to the top of the next
        // iteration, just before their use
        __asm__ volatile("" ::: "memory");
      }
    };
    if (wave + (FTMAX - 1) * NW < nft) contract(std::true_type{});
    else contract(std::false_type{});

    __syncthreads();

    const float cb_next = last ? 0.f : M.L[l + 1].c_base;
    const float ca_next = last ? 0.f : M.L[l + 1].c_aff;
#pragma unroll
    for (int f = 0; f < FTMAX; ++f) {
      const int ft = wave + f * NW;
      if (ft >= nft) continue;
      const floatx4 bias4 = *reinterpret_cast<const floatx4*>(&Ly.bias[ft * 16 + lg * 4]);
#pragma unroll
      for (int t = 0; t < BT; ++t) {
        const int b = t * 16 + ln;
        float v[4][V + 2 + NP];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v[j][0] = acc[f][0][t][j] + bias4[j];
#pragma unroll
          for (int i = 0; i < V; ++i) v[j][1 + i] = acc[f][1 + i][t][j];
          v[j][V + 1] = up1(acc[f][NS][t][j] * (WS == WSRC_FP32 ? Ly.e_fix32 : Ly.e_fix));
#pragma unroll
          for (int r = 0; r < NP; ++r) v[j][V + 2 + r] = acc[f][R + r][t][j] + bias4[j];
        }
        if (last) {
          if (ft == 0 && lg == 0) fin(v[0], b);
          continue;
        }
        auto hidden = [&](int act) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            affine_rule_rigorous<V>(act, v[j]);
#pragma unroll
            for (int r = 0; r < NP; ++r) v[j][V + 2 + r] = act_point(act, v[j][V + 2 + r]);
          }
        };
        if constexpr (AS == ACTS_RELU) {
          // the rules with the activation a constant (ReLU, or the identity)
          if (Ly.act == ACT_RELU) hidden(ACT_RELU);
          else hidden(ACT_NONE);
        } else {
          hidden(Ly.act);
        }
        // next layer's planes; its contraction bound rides in the e plane
        OutPlanes<V> o;
        split_out4<V>(v, cb_next, ca_next, o);
        const int k = ft * 16 + lg * 4;
#pragma unroll
        for (int r = 0; r < NP; ++r) {
          const int pl = PB::NPL + r * NPS;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            unsigned h1, h2, h3 = 0;
            if constexpr (NPS == 3) split3_pair(v[2 * q][V + 2 + r], v[2 * q + 1][V + 2 + r], h1, h2, h3);
            else split2_pair(v[2 * q][V + 2 + r], v[2 * q + 1][V + 2 + r], h1, h2);
            *reinterpret_cast<unsigned*>(&Xs[boff16<NB, WMAX>(pl, b, k + 2 * q)]) = h1;
            *reinterpret_cast<unsigned*>(&Xs[boff16<NB, WMAX>(pl + 1, b, k + 2 * q)]) = h2;
            if constexpr (NPS == 3) *reinterpret_cast<unsigned*>(&Xs[boff16<NB, WMAX>(pl + 2, b, k + 2 * q)]) = h3;
          }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q)
          *reinterpret_cast<uint2*>(&Xs[boff16<NB, WMAX>(q, b, k)]) = uint2{o.x[q][0], o.x[q][1]};
#pragma unroll
        for (int i = 0; i < V; ++i) {
          *reinterpret_cast<uint2*>(&Xs[boff16<NB, WMAX>(3 + 2 * i, b, k)]) = uint2{o.a[i][0][0], o.a[i][0][1]};
          *reinterpret_cast<uint2*>(&Xs[boff16<NB, WMAX>(4 + 2 * i, b, k)]) = uint2{o.a[i][1][0], o.a[i][1][1]};
        }
        *reinterpret_cast<uint2*>(&Xs[boff16<NB, WMAX>(PB::E, b, k)]) = uint2{o.e[0], o.e[1]};
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

template <int V, int NB, int NW, int WMAX, int WS, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64, (BfOcc<V, NB, NW, WMAX>::value))
k_classify_bf16(const DevModelBf16* __restrict__ Mp, ClassifyArgs a) {
  using PB = PlanesBf16<V>;
  constexpr int NT = NW * 64;
  __shared__ __attribute__((aligned(16))) unsigned short Xs[PB::NPL * NB * WMAX];
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  const long long ntiles = (n + NB - 1) / NB;
  const float cb0 = Mp->L[0].c_base, ca0 = Mp->L[0].c_aff;
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
      store_split<V, NB, WMAX>(Xs, b, k, s, err, cb0, ca0);
    }
    mlp_forward_bf16<V, NB, NW, WMAX, WS, 0, 3, AS>(*Mp, Xs, [&](const float* v, int b) {
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
