// Device-side building blocks of the range-analysis forward pass (gfx950).
//
// The per-box MLP range analysis of the reference (src/affine_layers.py:11-180,
// src/affine.py:85-193, src/slope_interval_layers.py:11-143) is evaluated for a
// tile of NB boxes by one workgroup.  The range form of a box is a set of R
// "planes" of width w (affine: base, error-term rows, interval error; slope:
// primal, slope centres, slope widths; or plain points).  Each dense layer is
// the batched contraction  OUT[p] = X[p] . W  (or . |W| for the non-negative
// planes) on MFMA (v_mfma_f32_16x16x4_f32: M = 16 output features, N = 16
// boxes, K = 4 inputs), with every plane of a (feature, box) pair landing in
// the same lane, so the linearisation rules are plain per-lane code.
//
// Activations live in LDS as X[plane][box][feature]; the feature index is
// XOR-swizzled by the box so that both the 16-byte operand reads and the
// 16-byte result writes are bank-conflict free.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define NIQ_MAX_LAYERS 24

enum { ACT_NONE = 0, ACT_RELU = 1, ACT_ELU = 2, ACT_SIN = 3, ACT_TANH = 4 };
enum { CODE_UNKNOWN = 0, CODE_POSITIVE = 1, CODE_NEGATIVE = 2 };

typedef float floatx4 __attribute__((ext_vector_type(4)));

// One dense layer, lowered and padded to multiples of 16.
struct DevLayer {
  int in_pad;            // K, multiple of 16
  int out_pad;           // output features, multiple of 16
  int act;               // activation applied after this layer (hidden layers)
  int pad_;
  const floatx4* Wp;     // fragment-packed weights, [out_pad/16][in_pad/16][64 lanes]
  const float* bias;     // [out_pad]
  const float* W;        // row-major (in_pad, out_pad) copy for the FP64 re-test
};

struct DevModel {
  int n_layers;
  int wmax;              // max padded width over all layer inputs/outputs
  int in_dim;            // input dimension (3)
  int pad_;
  DevLayer L[NIQ_MAX_LAYERS];
};

// ------------------------------------------------------------------ helpers
// Math by element type, so each rule is written once for FP32 (the kernels)
// and FP64 (the near-threshold re-test).
// A global-address-space view of a device pointer: loads through it are
// global loads, counted by vmcnt only (flat loads also count against
// lgkmcnt, so an LDS wait would wait for them as well).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

template <class T> struct Fm;
template <> struct Fm<float> {
  static __device__ __forceinline__ float exp(float x) { return expf(x); }
  static __device__ __forceinline__ float expm1(float x) { return expm1f(x); }
  static __device__ __forceinline__ float log(float x) { return logf(x); }
  static __device__ __forceinline__ float sin(float x) { return sinf(x); }
  static __device__ __forceinline__ float cos(float x) { return cosf(x); }
  static __device__ __forceinline__ float acos(float x) { return acosf(x); }
  static __device__ __forceinline__ float tanh(float x) { return tanhf(x); }
  static __device__ __forceinline__ float atanh(float x) { return atanhf(x); }
  static __device__ __forceinline__ float sqrt(float x) { return sqrtf(x); }
  static __device__ __forceinline__ float ceil(float x) { return ceilf(x); }
  static __device__ __forceinline__ float floor(float x) { return floorf(x); }
  static __device__ __forceinline__ float abs(float x) { return fabsf(x); }
  static __device__ __forceinline__ float min(float a, float b) { return fminf(a, b); }
  static __device__ __forceinline__ float max(float a, float b) { return fmaxf(a, b); }
  static constexpr float kMax = 3.4028234663852886e38f;
  static constexpr float kTwoPi = 6.2831854820251465f;    // float32(2 pi)
  static constexpr float kHalfPi = 1.5707963705062866f;   // float32(pi / 2)
};
template <> struct Fm<double> {
  static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
  static __device__ __forceinline__ double expm1(double x) { return ::expm1(x); }
  static __device__ __forceinline__ double log(double x) { return ::log(x); }
  static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
  static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
  static __device__ __forceinline__ double acos(double x) { return ::acos(x); }
  static __device__ __forceinline__ double tanh(double x) { return ::tanh(x); }
  static __device__ __forceinline__ double atanh(double x) { return ::atanh(x); }
  static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double ceil(double x) { return ::ceil(x); }
  static __device__ __forceinline__ double floor(double x) { return ::floor(x); }
  static __device__ __forceinline__ double abs(double x) { return ::fabs(x); }
  static __device__ __forceinline__ double min(double a, double b) { return ::fmin(a, b); }
  static __device__ __forceinline__ double max(double a, double b) { return ::fmax(a, b); }
  static constexpr double kMax = 1.7976931348623157e308;
  static constexpr double kTwoPi = 6.283185307179586;
  static constexpr double kHalfPi = 1.5707963267948966;
};

template <class T> __device__ __forceinline__ T relu_t(T x) { return x > T(0) ? x : T(0); }

// jnp.nan_to_num(x, nan=0): NaN -> 0, +-inf -> +-max finite
template <class T> __device__ __forceinline__ T nan_to_num_t(T x) {
  if (x != x) return T(0);
  if (isinf(x)) return x > T(0) ? Fm<T>::kMax : -Fm<T>::kMax;
  return x;
}
template <class T> __device__ __forceinline__ T clip_t(T x, T lo, T hi) {
  return Fm<T>::min(Fm<T>::max(x, lo), hi);
}
template <class T> __device__ __forceinline__ T elu_t(T x) {
  return x > T(0) ? x : Fm<T>::expm1(Fm<T>::min(x, T(0)));
}

template <class T> __device__ __forceinline__ T act_point(int act, T x) {
  switch (act) {
    case ACT_RELU: return relu_t(x);
    case ACT_ELU: return elu_t(x);
    case ACT_SIN: return Fm<T>::sin(x);
    case ACT_TANH: return Fm<T>::tanh(x);
    default: return x;
  }
}

// Derivative of each activation as jax differentiates it (jax.nn.relu: 0 at
// 0; jax.nn.elu: exp(x) for x <= 0; sin: cos; tanh: 1 - tanh^2).
template <class T> __device__ __forceinline__ T act_deriv(int act, T x) {
  switch (act) {
    case ACT_RELU: return x > T(0) ? T(1) : T(0);
    case ACT_ELU: return x > T(0) ? T(1) : Fm<T>::exp(x);
    case ACT_SIN: return Fm<T>::cos(x);
    case ACT_TANH: { const T t = Fm<T>::tanh(x); return T(1) - t * t; }
    default: return T(1);
  }
}

// ------------------------------------------------ linearisation rules
// Each returns alpha, beta, delta for the output range [lo, hi] of one
// feature, with |act(x) - (alpha x + beta)| <= delta on [lo, hi].

// Chebyshev ReLU, src/affine_layers.py:34-56
template <class T>
__device__ __forceinline__ void relu_rule(T lo, T hi, T& a, T& b, T& d) {
  T al = (relu_t(hi) - relu_t(lo)) / (hi - lo);
  if (lo >= T(0)) al = T(1);
  if (hi < T(0)) al = T(0);
  al = clip_t(nan_to_num_t(al), T(0), T(1));
  T be = (relu_t(lo) - al * lo) / T(2);
  a = al; b = be; d = be;
}

// src/affine_layers.py:59-98
template <class T>
__device__ __forceinline__ void elu_rule(T lo, T hi, T& a, T& b, T& d) {
  T lF = elu_t(lo), uF = elu_t(hi);
  T lS = Fm<T>::min(Fm<T>::exp(lo), T(1)), uS = Fm<T>::min(Fm<T>::exp(hi), T(1));
  T al = (uF - lF) / (hi - lo);
  if (lo >= T(0)) al = T(1);
  al = nan_to_num_t(al);
  al = Fm<T>::min(Fm<T>::max(al, lS), uS);
  T r_up = lF - al * lo;
  T xl = Fm<T>::min(Fm<T>::max(Fm<T>::log(al), lo), hi);
  T r_lo = (al - T(1)) - al * xl;
  T be = T(0.5) * (r_up + r_lo);
  T de = T(0.5) * Fm<T>::abs(r_up - r_lo);
  if (lo >= T(0)) { al = T(1); be = T(0); de = T(0); }
  a = al; b = be; d = de;
}

// src/utils.py:209-231
template <class T>
__device__ __forceinline__ void sin_bound_t(T lo, T hi, T& ol, T& ou) {
  T fl = Fm<T>::sin(lo), fu = Fm<T>::sin(hi);
  T l2 = lo / Fm<T>::kTwoPi, u2 = hi / Fm<T>::kTwoPi;
  bool cmin = Fm<T>::ceil(l2 - T(0.75)) < (u2 - T(0.75));
  bool cmax = Fm<T>::ceil(l2 - T(0.25)) < (u2 - T(0.25));
  ol = cmin ? T(-1) : Fm<T>::min(fl, fu);
  ou = cmax ? T(1) : Fm<T>::max(fl, fu);
}
template <class T>
__device__ __forceinline__ void cos_bound_t(T lo, T hi, T& ol, T& ou) {
  sin_bound_t(lo + Fm<T>::kHalfPi, hi + Fm<T>::kHalfPi, ol, ou);
}

// src/affine_layers.py:100-138
template <class T>
__device__ __forceinline__ void sin_rule(T lo, T hi, T& a, T& b, T& d) {
  const T two_pi = Fm<T>::kTwoPi;
  T sl, su;
  cos_bound_t(lo, hi, sl, su);
  T al = clip_t(T(0.5) * (sl + su), T(-1), T(1));
  T iA = Fm<T>::acos(al), iB = -iA;
  T locs[6];
  locs[0] = lo; locs[1] = hi;
  locs[2] = two_pi * Fm<T>::ceil((lo + iA) / two_pi) - iA;
  locs[3] = two_pi * Fm<T>::floor((hi - iA) / two_pi) + iA;
  locs[4] = two_pi * Fm<T>::ceil((lo + iB) / two_pi) - iB;
  locs[5] = two_pi * Fm<T>::floor((hi - iB) / two_pi) + iB;
  T rl = T(INFINITY), ru = T(-INFINITY);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    T x = Fm<T>::min(Fm<T>::max(locs[i], lo), hi);
    T v = Fm<T>::sin(x) - al * x;
    rl = Fm<T>::min(rl, v); ru = Fm<T>::max(ru, v);
  }
  T be = T(0.5) * (ru + rl);
  a = al; b = be; d = ru - be;
}

// TanH (this project's rule; the reference has none): secant slope, extrema
// of tanh(x) - alpha x at the endpoints and +-atanh(sqrt(1 - alpha)).
template <class T>
__device__ __forceinline__ void tanh_rule(T lo, T hi, T& a, T& b, T& d) {
  T tl = Fm<T>::tanh(lo), tu = Fm<T>::tanh(hi);
  T al = (tu - tl) / (hi - lo);
  T dl = T(1) - tl * tl;
  al = (hi > lo) ? al : dl;
  al = clip_t(nan_to_num_t(al), T(0), T(1));
  T xs = Fm<T>::atanh(Fm<T>::sqrt(T(1) - al));
  T locs[4] = {lo, hi, xs, -xs};
  T rl = T(INFINITY), ru = T(-INFINITY);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T x = Fm<T>::min(Fm<T>::max(locs[i], lo), hi);
    T v = Fm<T>::tanh(x) - al * x;
    rl = Fm<T>::min(rl, v); ru = Fm<T>::max(ru, v);
  }
  a = al; b = T(0.5) * (ru + rl); d = T(0.5) * (ru - rl);
}

template <class T>
__device__ __forceinline__ void act_rule(int act, T lo, T hi, T& a, T& b, T& d) {
  switch (act) {
    case ACT_RELU: relu_rule(lo, hi, a, b, d); break;
    case ACT_ELU: elu_rule(lo, hi, a, b, d); break;
    case ACT_SIN: sin_rule(lo, hi, a, b, d); break;
    case ACT_TANH: tanh_rule(lo, hi, a, b, d); break;
    default: a = T(1); b = T(0); d = T(0); break;
  }
}

template <class T>
__device__ __forceinline__ int sign_code(T lo, T hi, T offset) {
  int c = CODE_UNKNOWN;
  if (lo > offset) c = CODE_POSITIVE;
  if (hi < -offset) c = CODE_NEGATIVE;
  return c;
}

// jnp.sign
__device__ __forceinline__ float sign_f(float x) {
  return x > 0.f ? 1.f : (x < 0.f ? -1.f : (x == 0.f ? 0.f : x));
}

// -------------------------------------------------------- plane layouts
// A plane layout describes the R planes of one range form plus any number of
// plain point planes that ride in the same tile.
//  kind 0: affine   [base, aff_0..aff_{V-1}, err]      R = V + 2 (interval: V = 0)
//  kind 1: slope    [primal, c_0..c_{V-1}, w_0..w_{V-1}] R = 1 + 2V
//  kind 2: none     (points only)
//  kind 3: tangent  [primal, d/dx_0..d/dx_{V-1}]          R = 1 + V (forward-mode gradient)
// followed by P point planes.
template <int KIND, int V, int P>
struct Planes {
  static constexpr int RB = KIND == 0 ? V + 2 : (KIND == 1 ? 1 + 2 * V : (KIND == 3 ? 1 + V : 0));
  static constexpr int R = RB + P;
  // plane multiplies |W| instead of W
  __device__ static constexpr bool absw(int p) {
    return (KIND == 0 && p == V + 1) || (KIND == 1 && p > V && p < RB);
  }
  // plane receives the bias
  __device__ static constexpr bool bias(int p) { return (KIND != 2 && p == 0) || p >= RB; }
};

// affine (interval / affine_fixed) hidden rule, src/affine.py:164-176
template <int V, class T>
__device__ __forceinline__ void affine_fixed_rule(int act, T* v) {
  if (act == ACT_NONE) return;
  T rad = T(0);
#pragma unroll
  for (int i = 0; i < V; ++i) rad = rad + Fm<T>::abs(v[1 + i]);
  rad = rad + v[V + 1];
  T lo = v[0] - rad, hi = v[0] + rad;
  T a, b, d;
  act_rule(act, lo, hi, a, b, d);
  v[0] = a * v[0] + b;
#pragma unroll
  for (int i = 0; i < V; ++i) v[1 + i] = a * v[1 + i];
  v[V + 1] = a * v[V + 1] + Fm<T>::abs(d);
}

// slope-interval hidden rule, src/slope_interval_layers.py:36-109
template <int V, class T>
__device__ __forceinline__ void slope_rule(int act, T* v) {
  if (act == ACT_NONE) return;
  T sl[V], su[V];
  T rad = T(0);
#pragma unroll
  for (int i = 0; i < V; ++i) {
    sl[i] = v[1 + i] - v[1 + V + i];
    su[i] = v[1 + i] + v[1 + V + i];
    rad = rad + Fm<T>::max(su[i], -sl[i]);
  }
  T pl = v[0] - rad, pu = v[0] + rad;
  T dfl, dfu;
  bool general = false;
  if (act == ACT_RELU) {
    dfl = pl > T(0) ? T(1) : T(0);
    dfu = pu < T(0) ? T(0) : T(1);
  } else if (act == ACT_ELU) {
    dfl = Fm<T>::min(Fm<T>::exp(pl), T(1));
    dfu = Fm<T>::min(Fm<T>::exp(pu), T(1));
  } else if (act == ACT_TANH) {
    T tl = Fm<T>::tanh(pl), tu = Fm<T>::tanh(pu);
    T dl = T(1) - tl * tl, du = T(1) - tu * tu;
    dfl = Fm<T>::min(dl, du);
    dfu = (pl <= T(0) && pu >= T(0)) ? T(1) : Fm<T>::max(dl, du);
  } else {  // sin: derivative cos may change sign -> full interval product
    cos_bound_t(pl, pu, dfl, dfu);
    general = true;
  }
#pragma unroll
  for (int i = 0; i < V; ++i) {
    T nl, nu;
    if (!general) {
      nl = Fm<T>::min(sl[i] * dfl, sl[i] * dfu);
      nu = Fm<T>::max(su[i] * dfl, su[i] * dfu);
    } else {
      T p0 = sl[i] * dfl, p1 = sl[i] * dfu, p2 = su[i] * dfl, p3 = su[i] * dfu;
      nl = Fm<T>::min(Fm<T>::min(Fm<T>::min(p0, p1), p2), p3);
      nu = Fm<T>::max(Fm<T>::max(Fm<T>::max(p0, p1), p2), p3);
    }
    T c = T(0.5) * (nl + nu);
    v[1 + i] = c;
    v[1 + V + i] = nu - c;
  }
  v[0] = act_point(act, v[0]);
}

// Final bounds of a range form (output feature 0): lo, hi and the magnitude
// scale |base| + radius.
template <int KIND, int V, class T>
__device__ __forceinline__ void final_bounds(const T* v, T& lo, T& hi, T& scale) {
  if constexpr (KIND == 0) {
    T rad = T(0);
#pragma unroll
    for (int i = 0; i < V; ++i) rad = rad + Fm<T>::abs(v[1 + i]);
    rad = rad + v[V + 1];
    lo = v[0] - rad; hi = v[0] + rad;
    scale = Fm<T>::abs(v[0]) + rad;
  } else if constexpr (KIND == 1) {
    T rad = T(0);
#pragma unroll
    for (int i = 0; i < V; ++i) {
      T sl = v[1 + i] - v[1 + V + i];
      T su = v[1 + i] + v[1 + V + i];
      rad = rad + Fm<T>::max(su, -sl);
    }
    lo = v[0] - rad; hi = v[0] + rad;
    scale = Fm<T>::abs(v[0]) + rad;
  } else {
    lo = hi = v[0]; scale = Fm<T>::abs(v[0]);
  }
}

// Activation sets a kernel is compiled for: every activation, or ReLU or TanH
// alone (models whose hidden layers all use that one; the rules of the others
// and their constants then take no registers).
enum { ACTS_ALL = 0, ACTS_RELU = 1, ACTS_TANH = 2 };

template <class PL, int KIND, int V, int AS = ACTS_ALL, class T>
__device__ __forceinline__ void apply_hidden(int act, T* v) {
  if constexpr (AS == ACTS_RELU || AS == ACTS_TANH) {
    // the same rule with the activation a constant (identity otherwise)
    constexpr int A = AS == ACTS_RELU ? ACT_RELU : ACT_TANH;
    if (act == A) apply_hidden<PL, KIND, V, ACTS_ALL>(A, v);
    return;
  }
  if constexpr (KIND == 0) affine_fixed_rule<V>(act, v);
  else if constexpr (KIND == 1) slope_rule<V>(act, v);
  else if constexpr (KIND == 3) {
    if (act != ACT_NONE) {
      const T d = act_deriv(act, v[0]);
#pragma unroll
      for (int i = 1; i <= V; ++i) v[i] = v[i] * d;
      v[0] = act_point(act, v[0]);
    }
  }
#pragma unroll
  for (int p = PL::RB; p < PL::R; ++p) v[p] = act_point(act, v[p]);
}

// ------------------------------------------------------------ forward pass
// X: LDS activation buffer of R * NB * WMAX floats.  On entry the input
// planes hold the layer-0 inputs (features 0..in_pad-1, zero padded).  The
// tile callback `fin(v, box)` is invoked for output feature 0 of the last
// layer by the lanes that own it (wave 0, lanes 0..15 of each box tile).
template <int NB, int WMAX>
__device__ __forceinline__ int xoff(int p, int b, int k) {
  return (p * NB + b) * WMAX + (k ^ ((b & 7) << 2));
}

template <class PL, int KIND, int V, int NB, int NW, int WMAX, int AS = ACTS_ALL, class Fin>
__device__ __forceinline__ void mlp_forward(const DevModel& M, float* __restrict__ X, Fin&& fin) {
  constexpr int R = PL::R;
  constexpr int BT = NB / 16;
  constexpr int FTMAX = (WMAX / 16 + NW - 1) / NW;
  const int lane = threadIdx.x & 63;
  // wave-uniform, so the feature-tile tests below are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ln = lane & 15;
  const int lg = lane >> 4;

  for (int l = 0; l < M.n_layers; ++l) {
    const DevLayer Ly = M.L[l];
    const int nkb = Ly.in_pad >> 4;
    const int nft = Ly.out_pad >> 4;
    const bool last = (l == M.n_layers - 1);
    floatx4 acc[FTMAX][R][BT];
#pragma unroll
    for (int f = 0; f < FTMAX; ++f)
#pragma unroll
      for (int p = 0; p < R; ++p)
#pragma unroll
        for (int t = 0; t < BT; ++t) acc[f][p][t] = floatx4{0.f, 0.f, 0.f, 0.f};

    __syncthreads();  // layer input complete in X

    // Contraction over k-blocks of 16 inputs.  Each accumulator sums its
    // k-blocks in ascending order; the next k-block's weights (global) and
    // input fragments (LDS, unchanged until the barrier below) are loaded
    // while the current one's MFMAs run.  `FULL`: every feature tile of this
    // wave exists, so the loop carries no per-tile test.
    auto contract = [&](auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
      floatx4 a_nx[FTMAX];
#pragma unroll
      for (int f = 0; f < FTMAX; ++f) {
        const int ft = wave + f * NW;
        if (FULL || ft < nft) a_nx[f] = Ly.Wp[(size_t)(ft * nkb) * 64 + lane];
      }
      for (int kb = 0; kb < nkb; ++kb) {
        floatx4 a[FTMAX];
        floatx4 bx[R][BT];
#pragma unroll
        for (int f = 0; f < FTMAX; ++f) a[f] = a_nx[f];
#pragma unroll
        for (int p = 0; p < R; ++p)
#pragma unroll
          for (int t = 0; t < BT; ++t)
            bx[p][t] = *reinterpret_cast<const floatx4*>(&X[xoff<NB, WMAX>(p, t * 16 + ln, kb * 16 + lg * 4)]);
        if (kb + 1 < nkb) {
#pragma unroll
          for (int f = 0; f < FTMAX; ++f) {
            const int ft = wave + f * NW;
            if (FULL || ft < nft) a_nx[f] = Ly.Wp[(size_t)(ft * nkb + kb + 1) * 64 + lane];
          }
        }
#pragma unroll
        for (int p = 0; p < R; ++p) {
#pragma unroll
          for (int t = 0; t < BT; ++t) {
#pragma unroll
            for (int f = 0; f < FTMAX; ++f) {
              const int ft = wave + f * NW;
              if (FULL || ft < nft) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                  const float av = PL::absw(p) ? fabsf(a[f][j]) : a[f][j];
                  acc[f][p][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bx[p][t][j], acc[f][p][t], 0, 0, 0);
                }
              }
            }
          }
        }
      }
    };
    // A layer with a single output tile (the last) would leave all waves
    // but one idle: when it has enough k-blocks they are split into kKSplit
    // chunks, one per wave, and the partial sums added in chunk order.  The
    // chunking depends on the model only, not on the tile shape, so every
    // tile shape gives the same bits.
    constexpr int kKSplit = 4;
    const bool ksplit = (nft == 1) && nkb >= 2 * kKSplit;
    if (ksplit) {
      if (wave < kKSplit) {
        const int kb0 = wave * nkb / kKSplit, kb1 = (wave + 1) * nkb / kKSplit;
        floatx4 a_nx = {0.f, 0.f, 0.f, 0.f};
        if (kb0 < kb1) a_nx = Ly.Wp[(size_t)kb0 * 64 + lane];
        for (int kb = kb0; kb < kb1; ++kb) {
          const floatx4 a = a_nx;
          floatx4 bx[R][BT];
#pragma unroll
          for (int p = 0; p < R; ++p)
#pragma unroll
            for (int t = 0; t < BT; ++t)
              bx[p][t] = *reinterpret_cast<const floatx4*>(&X[xoff<NB, WMAX>(p, t * 16 + ln, kb * 16 + lg * 4)]);
          if (kb + 1 < kb1) a_nx = Ly.Wp[(size_t)(kb + 1) * 64 + lane];
#pragma unroll
          for (int p = 0; p < R; ++p)
#pragma unroll
            for (int t = 0; t < BT; ++t)
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float av = PL::absw(p) ? fabsf(a[j]) : a[j];
                acc[0][p][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bx[p][t][j], acc[0][p][t], 0, 0, 0);
              }
        }
      }
    } else if (wave + (FTMAX - 1) * NW < nft) {
      contract(std::true_type{});
    } else {
      contract(std::false_type{});
    }

    __syncthreads();  // every wave has consumed X

    if (ksplit) {
      // partial sums through LDS (X is free now); wave 0 adds them in order
      floatx4* Pp = reinterpret_cast<floatx4*>(X);
      if (wave > 0 && wave < kKSplit) {
#pragma unroll
        for (int p = 0; p < R; ++p)
#pragma unroll
          for (int t = 0; t < BT; ++t) Pp[(((wave - 1) * R + p) * BT + t) * 64 + lane] = acc[0][p][t];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int w = 1; w < kKSplit; ++w)
#pragma unroll
          for (int p = 0; p < R; ++p)
#pragma unroll
            for (int t = 0; t < BT; ++t) acc[0][p][t] += Pp[(((w - 1) * R + p) * BT + t) * 64 + lane];
      }
      __syncthreads();  // partials read before the caller reuses X
    }

#pragma unroll
    for (int f = 0; f < FTMAX; ++f) {
      const int ft = wave + f * NW;
      if (ft >= nft) continue;
      const floatx4 bias4 = *reinterpret_cast<const floatx4*>(&Ly.bias[ft * 16 + lg * 4]);
#pragma unroll
      for (int t = 0; t < BT; ++t) {
        const int b = t * 16 + ln;
        floatx4 outv[R];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v[R];
#pragma unroll
          for (int p = 0; p < R; ++p) v[p] = acc[f][p][t][j] + (PL::bias(p) ? bias4[j] : 0.f);
          if (!last) {
            apply_hidden<PL, KIND, V, AS>(Ly.act, v);
          } else if (ft == 0 && lg == 0 && j == 0) {
            fin(v, b);
          }
#pragma unroll
          for (int p = 0; p < R; ++p) outv[p][j] = v[p];
        }
        if (!last) {
#pragma unroll
          for (int p = 0; p < R; ++p)
            *reinterpret_cast<floatx4*>(&X[xoff<NB, WMAX>(p, b, ft * 16 + lg * 4)]) = outv[p];
        }
      }
    }
  }
}
