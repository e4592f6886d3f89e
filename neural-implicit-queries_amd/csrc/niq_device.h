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
};

struct DevModel {
  int n_layers;
  int wmax;              // max padded width over all layer inputs/outputs
  int in_dim;            // input dimension (3)
  int pad_;
  DevLayer L[NIQ_MAX_LAYERS];
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ float relu_f(float x) { return x > 0.f ? x : 0.f; }

// jnp.nan_to_num(x, nan=0): NaN -> 0, +-inf -> +-FLT_MAX
__device__ __forceinline__ float nan_to_num_f(float x) {
  if (x != x) return 0.f;
  if (isinf(x)) return x > 0.f ? 3.4028234663852886e38f : -3.4028234663852886e38f;
  return x;
}
__device__ __forceinline__ float clipf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}
__device__ __forceinline__ float elu_f(float x) { return x > 0.f ? x : expm1f(fminf(x, 0.f)); }

__device__ __forceinline__ float act_point(int act, float x) {
  switch (act) {
    case ACT_RELU: return relu_f(x);
    case ACT_ELU: return elu_f(x);
    case ACT_SIN: return sinf(x);
    case ACT_TANH: return tanhf(x);
    default: return x;
  }
}

// ------------------------------------------------ linearisation rules (FP32)
// Each returns alpha, beta, delta for the output range [lo, hi] of one
// feature, with |act(x) - (alpha x + beta)| <= delta on [lo, hi].

// Chebyshev ReLU, src/affine_layers.py:34-56
__device__ __forceinline__ void relu_rule(float lo, float hi, float& a, float& b, float& d) {
  float al = (relu_f(hi) - relu_f(lo)) / (hi - lo);
  if (lo >= 0.f) al = 1.f;
  if (hi < 0.f) al = 0.f;
  al = clipf(nan_to_num_f(al), 0.f, 1.f);
  float be = (relu_f(lo) - al * lo) / 2.f;
  a = al; b = be; d = be;
}

// src/affine_layers.py:59-98
__device__ __forceinline__ void elu_rule(float lo, float hi, float& a, float& b, float& d) {
  float lF = elu_f(lo), uF = elu_f(hi);
  float lS = fminf(expf(lo), 1.f), uS = fminf(expf(hi), 1.f);
  float al = (uF - lF) / (hi - lo);
  if (lo >= 0.f) al = 1.f;
  al = nan_to_num_f(al);
  al = fminf(fmaxf(al, lS), uS);
  float r_up = lF - al * lo;
  float xl = fminf(fmaxf(logf(al), lo), hi);
  float r_lo = (al - 1.f) - al * xl;
  float be = 0.5f * (r_up + r_lo);
  float de = 0.5f * fabsf(r_up - r_lo);
  if (lo >= 0.f) { al = 1.f; be = 0.f; de = 0.f; }
  a = al; b = be; d = de;
}

// src/utils.py:209-231
__device__ __forceinline__ void sin_bound_f(float lo, float hi, float& ol, float& ou) {
  const float two_pi = 6.2831854820251465f;   // float32(2 pi)
  float fl = sinf(lo), fu = sinf(hi);
  float l2 = lo / two_pi, u2 = hi / two_pi;
  bool cmin = ceilf(l2 - 0.75f) < (u2 - 0.75f);
  bool cmax = ceilf(l2 - 0.25f) < (u2 - 0.25f);
  ol = cmin ? -1.f : fminf(fl, fu);
  ou = cmax ? 1.f : fmaxf(fl, fu);
}
__device__ __forceinline__ void cos_bound_f(float lo, float hi, float& ol, float& ou) {
  const float half_pi = 1.5707963705062866f;  // float32(pi / 2)
  sin_bound_f(lo + half_pi, hi + half_pi, ol, ou);
}

// src/affine_layers.py:100-138
__device__ __forceinline__ void sin_rule(float lo, float hi, float& a, float& b, float& d) {
  const float two_pi = 6.2831854820251465f;
  float sl, su;
  cos_bound_f(lo, hi, sl, su);
  float al = clipf(0.5f * (sl + su), -1.f, 1.f);
  float iA = acosf(al), iB = -iA;
  float locs[6];
  locs[0] = lo; locs[1] = hi;
  locs[2] = two_pi * ceilf((lo + iA) / two_pi) - iA;
  locs[3] = two_pi * floorf((hi - iA) / two_pi) + iA;
  locs[4] = two_pi * ceilf((lo + iB) / two_pi) - iB;
  locs[5] = two_pi * floorf((hi - iB) / two_pi) + iB;
  float rl = INFINITY, ru = -INFINITY;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    float x = fminf(fmaxf(locs[i], lo), hi);
    float v = sinf(x) - al * x;
    rl = fminf(rl, v); ru = fmaxf(ru, v);
  }
  float be = 0.5f * (ru + rl);
  a = al; b = be; d = ru - be;
}

// TanH (this project's rule; the reference has none): secant slope, extrema
// of tanh(x) - alpha x at the endpoints and +-atanh(sqrt(1 - alpha)).
__device__ __forceinline__ void tanh_rule(float lo, float hi, float& a, float& b, float& d) {
  float tl = tanhf(lo), tu = tanhf(hi);
  float al = (tu - tl) / (hi - lo);
  float dl = 1.f - tl * tl;
  al = (hi > lo) ? al : dl;
  al = clipf(nan_to_num_f(al), 0.f, 1.f);
  float xs = atanhf(sqrtf(1.f - al));
  float locs[4] = {lo, hi, xs, -xs};
  float rl = INFINITY, ru = -INFINITY;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float x = fminf(fmaxf(locs[i], lo), hi);
    float v = tanhf(x) - al * x;
    rl = fminf(rl, v); ru = fmaxf(ru, v);
  }
  a = al; b = 0.5f * (ru + rl); d = 0.5f * (ru - rl);
}

__device__ __forceinline__ void act_rule(int act, float lo, float hi, float& a, float& b, float& d) {
  switch (act) {
    case ACT_RELU: relu_rule(lo, hi, a, b, d); break;
    case ACT_ELU: elu_rule(lo, hi, a, b, d); break;
    case ACT_SIN: sin_rule(lo, hi, a, b, d); break;
    case ACT_TANH: tanh_rule(lo, hi, a, b, d); break;
    default: a = 1.f; b = 0.f; d = 0.f; break;
  }
}

__device__ __forceinline__ int sign_code(float lo, float hi, float offset) {
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
// followed by P point planes.
template <int KIND, int V, int P>
struct Planes {
  static constexpr int RB = KIND == 0 ? V + 2 : (KIND == 1 ? 1 + 2 * V : 0);
  static constexpr int R = RB + P;
  // plane multiplies |W| instead of W
  __device__ static constexpr bool absw(int p) {
    return (KIND == 0 && p == V + 1) || (KIND == 1 && p > V && p < RB);
  }
  // plane receives the bias
  __device__ static constexpr bool bias(int p) { return (KIND != 2 && p == 0) || p >= RB; }
};

// affine (interval / affine_fixed) hidden rule, src/affine.py:164-176
template <int V>
__device__ __forceinline__ void affine_fixed_rule(int act, float* v) {
  if (act == ACT_NONE) return;
  float rad = 0.f;
#pragma unroll
  for (int i = 0; i < V; ++i) rad = rad + fabsf(v[1 + i]);
  rad = rad + v[V + 1];
  float lo = v[0] - rad, hi = v[0] + rad;
  float a, b, d;
  act_rule(act, lo, hi, a, b, d);
  v[0] = a * v[0] + b;
#pragma unroll
  for (int i = 0; i < V; ++i) v[1 + i] = a * v[1 + i];
  v[V + 1] = a * v[V + 1] + fabsf(d);
}

// slope-interval hidden rule, src/slope_interval_layers.py:36-109
template <int V>
__device__ __forceinline__ void slope_rule(int act, float* v) {
  if (act == ACT_NONE) return;
  float sl[V], su[V];
  float rad = 0.f;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    sl[i] = v[1 + i] - v[1 + V + i];
    su[i] = v[1 + i] + v[1 + V + i];
    rad = rad + fmaxf(su[i], -sl[i]);
  }
  float pl = v[0] - rad, pu = v[0] + rad;
  float dfl, dfu;
  bool general = false;
  if (act == ACT_RELU) {
    dfl = pl > 0.f ? 1.f : 0.f;
    dfu = pu < 0.f ? 0.f : 1.f;
  } else if (act == ACT_ELU) {
    dfl = fminf(expf(pl), 1.f);
    dfu = fminf(expf(pu), 1.f);
  } else if (act == ACT_TANH) {
    float tl = tanhf(pl), tu = tanhf(pu);
    float dl = 1.f - tl * tl, du = 1.f - tu * tu;
    dfl = fminf(dl, du);
    dfu = (pl <= 0.f && pu >= 0.f) ? 1.f : fmaxf(dl, du);
  } else {  // sin: derivative cos may change sign -> full interval product
    cos_bound_f(pl, pu, dfl, dfu);
    general = true;
  }
#pragma unroll
  for (int i = 0; i < V; ++i) {
    float nl, nu;
    if (!general) {
      nl = fminf(sl[i] * dfl, sl[i] * dfu);
      nu = fmaxf(su[i] * dfl, su[i] * dfu);
    } else {
      float p0 = sl[i] * dfl, p1 = sl[i] * dfu, p2 = su[i] * dfl, p3 = su[i] * dfu;
      nl = fminf(fminf(fminf(p0, p1), p2), p3);
      nu = fmaxf(fmaxf(fmaxf(p0, p1), p2), p3);
    }
    float c = 0.5f * (nl + nu);
    v[1 + i] = c;
    v[1 + V + i] = nu - c;
  }
  v[0] = act_point(act, v[0]);
}

// Final bounds of a range form (output feature 0): lo, hi.
template <int KIND, int V>
__device__ __forceinline__ void final_bounds(const float* v, float& lo, float& hi, float& scale) {
  if constexpr (KIND == 0) {
    float rad = 0.f;
#pragma unroll
    for (int i = 0; i < V; ++i) rad = rad + fabsf(v[1 + i]);
    rad = rad + v[V + 1];
    lo = v[0] - rad; hi = v[0] + rad;
    scale = fabsf(v[0]) + rad;
  } else if constexpr (KIND == 1) {
    float rad = 0.f;
#pragma unroll
    for (int i = 0; i < V; ++i) {
      float sl = v[1 + i] - v[1 + V + i];
      float su = v[1 + i] + v[1 + V + i];
      rad = rad + fmaxf(su, -sl);
    }
    lo = v[0] - rad; hi = v[0] + rad;
    scale = fabsf(v[0]) + rad;
  } else {
    lo = hi = v[0]; scale = fabsf(v[0]);
  }
}

template <class PL, int KIND, int V>
__device__ __forceinline__ void apply_hidden(int act, float* v) {
  if constexpr (KIND == 0) affine_fixed_rule<V>(act, v);
  else if constexpr (KIND == 1) slope_rule<V>(act, v);
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

template <class PL, int KIND, int V, int NB, int NW, int WMAX, class Fin>
__device__ __forceinline__ void mlp_forward(const DevModel& M, float* __restrict__ X, Fin&& fin) {
  constexpr int R = PL::R;
  constexpr int BT = NB / 16;
  constexpr int FTMAX = (WMAX / 16 + NW - 1) / NW;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
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

    floatx4 a_nx[FTMAX];
#pragma unroll
    for (int f = 0; f < FTMAX; ++f) {
      const int ft = wave + f * NW;
      if (ft < nft) a_nx[f] = Ly.Wp[(size_t)(ft * nkb) * 64 + lane];
    }
    for (int kb = 0; kb < nkb; ++kb) {
      floatx4 a[FTMAX];
#pragma unroll
      for (int f = 0; f < FTMAX; ++f) a[f] = a_nx[f];
      if (kb + 1 < nkb) {
#pragma unroll
        for (int f = 0; f < FTMAX; ++f) {
          const int ft = wave + f * NW;
          if (ft < nft) a_nx[f] = Ly.Wp[(size_t)(ft * nkb + kb + 1) * 64 + lane];
        }
      }
#pragma unroll
      for (int p = 0; p < R; ++p) {
#pragma unroll
        for (int t = 0; t < BT; ++t) {
          const int b = t * 16 + ln;
          const floatx4 bx = *reinterpret_cast<const floatx4*>(&X[xoff<NB, WMAX>(p, b, kb * 16 + lg * 4)]);
#pragma unroll
          for (int f = 0; f < FTMAX; ++f) {
            const int ft = wave + f * NW;
            if (ft < nft) {
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float av = PL::absw(p) ? fabsf(a[f][j]) : a[f][j];
                acc[f][p][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bx[j], acc[f][p][t], 0, 0, 0);
              }
            }
          }
        }
      }
    }

    __syncthreads();  // every wave has consumed X

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
            apply_hidden<PL, KIND, V>(Ly.act, v);
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
