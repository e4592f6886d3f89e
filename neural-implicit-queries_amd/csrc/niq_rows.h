// affine_all: range analysis that keeps every error term (src/affine.py:177-181).
//
// After activation l a box carries base (w), the error-term matrix E (n_l x w)
// and no interval part (err stays 0 in this mode).  Each activation appends
// one row per feature (diag(delta)), so n_l = d + sum of the hidden widths:
// up to 1027 rows of 128 for a 3->128x8->1 MLP, 526 KB per box.  That does not
// fit on chip, so a workgroup takes one box at a time and streams E through a
// private HBM scratch, 64 rows per chunk:
//   rows of the previous E, scaled by alpha on the way into LDS, go through the
//   MFMA contraction with the layer's packed weights;
//   the diag(delta) rows added by the previous activation need no contraction:
//   their image is delta_k * A[k, :] (an outer product, on the vector ALUs);
//   |E| column sums (the radius) accumulate in registers while the chunk's
//   output rows are written back, and are reduced across the lanes of a wave
//   in a fixed order, so results do not depend on scheduling.
// The base row and the linearisation run per feature on the vector ALUs.
#pragma once
#include "niq_kernels.h"

template <int WMAX> struct RowsCfg;
template <> struct RowsCfg<64>  { static constexpr int CH = 64, NW = 4; };
template <> struct RowsCfg<128> { static constexpr int CH = 64, NW = 4; };
template <> struct RowsCfg<256> { static constexpr int CH = 32, NW = 4; };
template <> struct RowsCfg<512> { static constexpr int CH = 32, NW = 4; };

struct RowsArgs {
  ClassifyArgs c;          // boxes, count, offset, codes, bounds
  float* scratch;          // per workgroup: 2 * max_rows * WMAX floats
  long long max_rows;      // d + sum of hidden widths
};

template <int WMAX>
__global__ void __launch_bounds__(RowsCfg<WMAX>::NW * 64)
k_classify_all(const DevModel* __restrict__ Mp, RowsArgs ra) {
  constexpr int CH = RowsCfg<WMAX>::CH, NW = RowsCfg<WMAX>::NW;
  constexpr int NT = NW * 64;
  constexpr int BT = CH / 16;
  constexpr int FTMAX = (WMAX / 16 + NW - 1) / NW;
  __shared__ float X[CH * WMAX];        // one chunk of input rows (swizzled)
  __shared__ float s_base[WMAX];        // base of the current layer input
  __shared__ float s_alpha[WMAX];       // column scale of the stored rows
  __shared__ float s_scale[WMAX];       // diag scales of the new rows (h, then delta)
  __shared__ float s_rad[WMAX];
  __shared__ float s_nb[WMAX];          // new base (pre-activation)
  const ClassifyArgs& a = ra.c;
  const DevModel& M = *Mp;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ln = lane & 15, lg = lane >> 4;
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  float* buf[2] = {ra.scratch + (size_t)blockIdx.x * 2 * ra.max_rows * WMAX,
                   ra.scratch + (size_t)blockIdx.x * 2 * ra.max_rows * WMAX + (size_t)ra.max_rows * WMAX};

  for (long long i = blockIdx.x; i < n; i += gridDim.x) {
    // layer-0 input: base = centre, new rows diag(half), no stored rows
    // axis boxes: the d rows diag(half) enter as new rows (their image is
    // half_k * A[k, :]); general boxes: the v box vectors enter as stored rows
    for (int k = threadIdx.x; k < WMAX; k += NT) {
      float c = 0.f, h = 0.f;
      if (k < a.d) {
        if (a.lower) {
          const float lo = a.lower[i * a.d + k], up = a.upper[i * a.d + k];
          c = 0.5f * (lo + up);
          h = up - c;
        } else {
          c = a.center[i * a.d + k];
        }
      }
      s_base[k] = c; s_scale[k] = h; s_alpha[k] = 1.f;
    }
    if (!a.lower) {
      for (int e = threadIdx.x; e < a.vin * WMAX; e += NT) {
        const int r = e / WMAX, k = e % WMAX;
        buf[0][(size_t)r * WMAX + k] = (k < a.d) ? a.vecs[(i * a.vin + r) * a.d + k] : 0.f;
      }
    }
    long long n_stored = a.lower ? 0 : a.vin;   // rows of E held in buf[cur]
    int n_new = a.lower ? a.d : 0;              // diag rows to add at this layer
    int cur = 0;
    __syncthreads();
    for (int l = 0; l < M.n_layers; ++l) {
      const DevLayer Ly = M.L[l];
      const int nkb = Ly.in_pad >> 4, nft = Ly.out_pad >> 4;
      const bool last = (l == M.n_layers - 1);
      float* Ein = buf[cur];
      float* Eout = buf[cur ^ 1];
      // new base row: base . A + b (sequential over k)
      for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
        float acc = 0.f;
        for (int k = 0; k < Ly.in_pad; ++k) acc = acc + s_base[k] * Ly.W[(size_t)k * Ly.out_pad + f];
        s_nb[f] = acc + Ly.bias[f];
      }
      floatx4 csum[FTMAX];
#pragma unroll
      for (int f = 0; f < FTMAX; ++f) csum[f] = floatx4{0.f, 0.f, 0.f, 0.f};
      // stored rows through MFMA, CH rows per chunk
      for (long long r0 = 0; r0 < n_stored; r0 += CH) {
        __syncthreads();
        for (int e = threadIdx.x; e < CH * Ly.in_pad; e += NT) {
          const int rr = e / Ly.in_pad, k = e % Ly.in_pad;
          const long long r = r0 + rr;
          X[xoff<CH, WMAX>(0, rr, k)] = (r < n_stored) ? s_alpha[k] * Ein[r * WMAX + k] : 0.f;
        }
        __syncthreads();
        floatx4 acc[FTMAX][BT];
#pragma unroll
        for (int f = 0; f < FTMAX; ++f)
#pragma unroll
          for (int t = 0; t < BT; ++t) acc[f][t] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < nkb; ++kb) {
          floatx4 w[FTMAX];
#pragma unroll
          for (int f = 0; f < FTMAX; ++f) {
            const int ft = wave + f * NW;
            if (ft < nft) w[f] = Ly.Wp[(size_t)(ft * nkb + kb) * 64 + lane];
          }
#pragma unroll
          for (int t = 0; t < BT; ++t) {
            const floatx4 bx = *reinterpret_cast<const floatx4*>(&X[xoff<CH, WMAX>(0, t * 16 + ln, kb * 16 + lg * 4)]);
#pragma unroll
            for (int f = 0; f < FTMAX; ++f) {
              const int ft = wave + f * NW;
              if (ft < nft) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                  acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[f][j], bx[j], acc[f][t], 0, 0, 0);
              }
            }
          }
        }
#pragma unroll
        for (int f = 0; f < FTMAX; ++f) {
          const int ft = wave + f * NW;
          if (ft >= nft) continue;
#pragma unroll
          for (int t = 0; t < BT; ++t) {
            const long long r = r0 + t * 16 + ln;
            if (r < n_stored) {
              if (!last) *reinterpret_cast<floatx4*>(&Eout[r * WMAX + ft * 16 + lg * 4]) = acc[f][t];
#pragma unroll
              for (int j = 0; j < 4; ++j) csum[f][j] = csum[f][j] + fabsf(acc[f][t][j]);
            }
          }
        }
      }
      // reduce the column sums over the 16 row lanes of each wave (fixed order)
#pragma unroll
      for (int f = 0; f < FTMAX; ++f) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v = csum[f][j];
          v = v + __shfl_xor(v, 1);
          v = v + __shfl_xor(v, 2);
          v = v + __shfl_xor(v, 4);
          v = v + __shfl_xor(v, 8);
          csum[f][j] = v;
        }
        const int ft = wave + f * NW;
        if (ft < nft && ln == 0) {
#pragma unroll
          for (int j = 0; j < 4; ++j) s_rad[ft * 16 + lg * 4 + j] = csum[f][j];
        }
      }
      __syncthreads();
      // the new rows: scale_k * A[k, :], appended after the stored rows
      for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
        float rs = s_rad[f];
        float dsum = 0.f;
        for (int k = 0; k < n_new; ++k) {
          const float v = s_scale[k] * Ly.W[(size_t)k * Ly.out_pad + f];
          if (!last) Eout[(n_stored + k) * WMAX + f] = v;
          dsum = dsum + fabsf(v);
        }
        s_rad[f] = rs + dsum;
      }
      __syncthreads();
      const long long n_out = n_stored + n_new;
      if (last) {
        if (threadIdx.x == 0) {
          const float base = s_nb[0], rad = s_rad[0];
          const float lo = base - rad, hi = base + rad;
          a.codes[i] = sign_code(lo, hi, a.offset);
          if (a.bounds) { a.bounds[2 * i] = lo; a.bounds[2 * i + 1] = hi; }
        }
        __syncthreads();
        break;
      }
      // linearise each feature: alpha scales all rows, delta rows are new
      for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
        const float base = s_nb[f], rad = s_rad[f];
        const float lo = base - rad, hi = base + rad;
        float al = 1.f, be = 0.f, de = 0.f;
        if (Ly.act != ACT_NONE) act_rule(Ly.act, lo, hi, al, be, de);
        s_base[f] = (Ly.act != ACT_NONE) ? al * base + be : base;
        s_alpha[f] = al;
        s_scale[f] = fabsf(de);
      }
      n_stored = n_out;
      n_new = (Ly.act != ACT_NONE) ? Ly.out_pad : 0;
      cur ^= 1;
      __syncthreads();
    }
  }
}
