// Range analysis with a growing set of error terms: affine_all,
// affine_truncate and affine_append (src/affine.py:127-193).
//
// After each activation a box carries base (w), an error-term matrix E
// (n rows x w) and an interval part err (w).  Each activation adds rows: one
// per feature (diag(delta): affine_all and affine_truncate) or the n_append
// largest deltas (affine_append); affine_truncate then keeps the n_keep rows of
// largest sum |row| and folds the rest into err.  E reaches 1027 x 128 floats
// (526 KB) per box for a 3->128x8->1 MLP under affine_all, which does not fit
// on chip, so a workgroup takes one box at a time and streams E through a
// private HBM scratch, CH rows per chunk:
//   stored rows (scaled by the column factor alpha on the way into LDS) go
//   through the MFMA contraction with the layer's packed weights;
//   rows added by the previous activation have a single non-zero each, so
//   their image is scale_r * A[idx_r, :] (on the vector ALUs);
//   |E| column sums (the radius) accumulate in registers while the chunk's
//   output rows are written back, and are reduced across the lanes of a wave
//   in a fixed order, so results do not depend on scheduling.
// The base row, err.|A| and the linearisation run per feature on the vector
// ALUs.
#pragma once
#include "niq_kernels.h"

enum { ROWS_ALL = 0, ROWS_TRUNCATE = 1, ROWS_APPEND = 2 };

template <int WMAX> struct RowsCfg;
template <> struct RowsCfg<64>  { static constexpr int CH = 64, NW = 4; };
template <> struct RowsCfg<128> { static constexpr int CH = 32, NW = 4; };
template <> struct RowsCfg<256> { static constexpr int CH = 32, NW = 4; };
template <> struct RowsCfg<512> { static constexpr int CH = 32, NW = 4; };

struct RowsArgs {
  ClassifyArgs c;          // boxes, count, offset, codes, bounds
  float* scratch;          // per workgroup: 2 * max_rows * WMAX floats
  long long max_rows;      // largest row count of E
  int n_keep;              // affine_truncate
  int n_append;            // affine_append
};

// Block-wide argmax over n values (largest first, lowest index on ties);
// every thread gets the result.  `sv`/`si` are 2 * NT scratch entries.
template <int NT>
__device__ __forceinline__ int block_argmax(const float* vals, const unsigned char* taken, int n, float* sv, int* si,
                                            float& best) {
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  for (int k = threadIdx.x; k < n; k += NT) {
    if (taken[k]) continue;
    const float v = vals[k];
    if (v > bv || (v == bv && k < bi)) { bv = v; bi = k; }
  }
  sv[threadIdx.x] = bv; si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const float v = sv[threadIdx.x + w];
      const int i = si[threadIdx.x + w];
      if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && i < si[threadIdx.x])) { sv[threadIdx.x] = v; si[threadIdx.x] = i; }
    }
    __syncthreads();
  }
  best = sv[0];
  const int r = si[0];
  __syncthreads();
  return r;
}

// Register budget: four waves per SIMD at width <= 128 (LDS admits four
// workgroups per CU or more there).
template <int WMAX> struct RowsOcc { static constexpr int value = WMAX <= 128 ? 4 : 1; };

// OCC: the register budget in waves per SIMD (RowsOcc by default; 2 trades
// resident workgroups for registers).  CH: rows per chunk of the error-term
// stream (a multiple of 16; each row's products and each lane's column sums
// run in row order whatever the chunk, so every CH gives the same bits).
template <int WMAX, int MODE, int OCC = RowsOcc<WMAX>::value, int CH = RowsCfg<WMAX>::CH>
__global__ void __launch_bounds__(RowsCfg<WMAX>::NW * 64, OCC)
k_classify_rows(const DevModel* __restrict__ Mp, RowsArgs ra) {
  static_assert(CH % 16 == 0, "chunks of whole 16-row tiles");
  constexpr int NW = RowsCfg<WMAX>::NW;
  constexpr int NT = NW * 64;
  constexpr int BT = CH / 16;
  constexpr int FTMAX = (WMAX / 16 + NW - 1) / NW;
  constexpr int MAXNEW = WMAX;
  __shared__ float X[CH * WMAX];        // one chunk of input rows (swizzled)
  __shared__ float s_base[WMAX];        // base of the current layer input
  __shared__ float s_alpha[WMAX];       // column scale of the stored rows
  __shared__ float s_err[WMAX];         // interval part of the layer input
  __shared__ float s_scale[MAXNEW];     // rows added by the last activation:
  __shared__ int s_idx[MAXNEW];         //   value s_scale[r] at column s_idx[r]
  __shared__ float s_rad[WMAX];
  __shared__ float s_nb[WMAX];          // new base (pre-activation)
  __shared__ float s_ne[WMAX];          // new err (pre-activation)
  __shared__ float s_mag[2 * WMAX + 64];    // row magnitudes of the candidates
  __shared__ unsigned char s_taken[2 * WMAX + 64];
  __shared__ float s_rv[NT];
  __shared__ float s_pe[NT];
  __shared__ int s_ri[NT];
  __shared__ int s_keep[64];
  const ClassifyArgs& a = ra.c;
  const DevModel& M = *Mp;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // scalar tile tests
  const int ln = lane & 15, lg = lane >> 4;
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  float* buf[2] = {ra.scratch + (size_t)blockIdx.x * 2 * ra.max_rows * WMAX,
                   ra.scratch + (size_t)blockIdx.x * 2 * ra.max_rows * WMAX + (size_t)ra.max_rows * WMAX};

  for (long long i = blockIdx.x; i < n; i += gridDim.x) {
    // Layer-0 input (src/affine.py:109-117).  Axis boxes: the d rows
    // diag(half) enter as added rows (image half_k * A[k, :]); general boxes:
    // the v box vectors enter as stored rows.
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
      s_base[k] = c; s_alpha[k] = 1.f; s_err[k] = 0.f;
      s_scale[k] = h; s_idx[k] = k;
    }
    if (!a.lower) {
      for (int e = threadIdx.x; e < a.vin * WMAX; e += NT) {
        const int r = e / WMAX, k = e % WMAX;
        buf[0][(size_t)r * WMAX + k] = (k < a.d) ? a.vecs[(i * a.vin + r) * a.d + k] : 0.f;
      }
    }
    long long n_stored = a.lower ? 0 : a.vin;
    int n_new = a.lower ? a.d : 0;
    int cur = 0;
    __syncthreads();
    for (int l = 0; l < M.n_layers; ++l) {
      const DevLayer Ly = M.L[l];
      const int nkb = Ly.in_pad >> 4, nft = Ly.out_pad >> 4;
      const bool last = (l == M.n_layers - 1);
      float* Ein = buf[cur];
      float* Eout = buf[cur ^ 1];
      // base . A + b and err . |A|: sequential over k per feature, or, for a
      // layer narrower than the workgroup, over G ranges of k joined in order
      if (Ly.out_pad >= NT) {
        for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
          float acc = 0.f, ea = 0.f;
#pragma unroll 8
          for (int k = 0; k < Ly.in_pad; ++k) {
            const float w = Ly.W[(size_t)k * Ly.out_pad + f];
            acc = acc + s_base[k] * w;
            if (MODE != ROWS_ALL) ea = ea + s_err[k] * fabsf(w);
          }
          s_nb[f] = acc + Ly.bias[f];
          s_ne[f] = ea;
        }
      } else {
        const int G = NT / Ly.out_pad;
        const int grp = (int)threadIdx.x / Ly.out_pad, f = (int)threadIdx.x % Ly.out_pad;
        const int per = (Ly.in_pad + G - 1) / G;
        if (grp < G) {
          float acc = 0.f, ea = 0.f;
          const int k0 = grp * per, k1 = min(Ly.in_pad, k0 + per);
#pragma unroll 8
          for (int k = k0; k < k1; ++k) {
            const float w = Ly.W[(size_t)k * Ly.out_pad + f];
            acc = acc + s_base[k] * w;
            if (MODE != ROWS_ALL) ea = ea + s_err[k] * fabsf(w);
          }
          s_rv[threadIdx.x] = acc;
          s_pe[threadIdx.x] = ea;
        }
        __syncthreads();
        if ((int)threadIdx.x < Ly.out_pad) {
          float acc = 0.f, ea = 0.f;
          for (int g2 = 0; g2 < G; ++g2) {
            acc = acc + s_rv[g2 * Ly.out_pad + threadIdx.x];
            ea = ea + s_pe[g2 * Ly.out_pad + threadIdx.x];
          }
          s_nb[threadIdx.x] = acc + Ly.bias[threadIdx.x];
          s_ne[threadIdx.x] = ea;
        }
      }
      floatx4 csum[FTMAX];
#pragma unroll
      for (int f = 0; f < FTMAX; ++f) csum[f] = floatx4{0.f, 0.f, 0.f, 0.f};
      for (long long r0 = 0; r0 < n_stored; r0 += CH) {
        // the chunk's rows: every 16-byte load issued before any is used
        constexpr int NL = CH * WMAX / 4 / NT;
        floatx4 ld[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
          const int e = 4 * ((int)threadIdx.x + j * NT);
          const int rr = e / Ly.in_pad, k = e % Ly.in_pad;
          const long long r = r0 + rr;
          ld[j] = (e < CH * Ly.in_pad && r < n_stored)
                      ? as_global(reinterpret_cast<const floatx4*>(Ein))[(r * WMAX + k) >> 2]
                      : floatx4{0.f, 0.f, 0.f, 0.f};
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NL; ++j) {
          const int e = 4 * ((int)threadIdx.x + j * NT);
          if (e >= CH * Ly.in_pad) continue;
          const int rr = e / Ly.in_pad, k = e % Ly.in_pad;
          floatx4 v;
#pragma unroll
          for (int q = 0; q < 4; ++q) v[q] = s_alpha[k + q] * ld[j][q];
          *reinterpret_cast<floatx4*>(&X[xoff<CH, WMAX>(0, rr, k)]) = v;
        }
        __syncthreads();
        floatx4 acc[FTMAX][BT];
#pragma unroll
        for (int f = 0; f < FTMAX; ++f)
#pragma unroll
          for (int t = 0; t < BT; ++t) acc[f][t] = floatx4{0.f, 0.f, 0.f, 0.f};
        // weights one k-block ahead of the MFMAs.  `FULL`: every feature
        // tile of this wave exists, so the loop body carries no per-tile
        // test and is one block the scheduler can order freely
        const auto Wp = as_global(Ly.Wp);
        auto kloop = [&](auto full_tag) {
          constexpr bool FULL = decltype(full_tag)::value;
          floatx4 w_nx[FTMAX];
#pragma unroll
          for (int f = 0; f < FTMAX; ++f) {
            const int ft = wave + f * NW;
            if (FULL || ft < nft) w_nx[f] = Wp[(size_t)(ft * nkb) * 64 + lane];
          }
          for (int kb = 0; kb < nkb; ++kb) {
            floatx4 w[FTMAX];
#pragma unroll
            for (int f = 0; f < FTMAX; ++f) w[f] = w_nx[f];
            // the last k-block fetches itself again: no branch in the loop
            const int kn = kb + 1 < nkb ? kb + 1 : kb;
#pragma unroll
            for (int f = 0; f < FTMAX; ++f) {
              const int ft = wave + f * NW;
              if (FULL || ft < nft) w_nx[f] = Wp[(size_t)(ft * nkb + kn) * 64 + lane];
            }
            // issue point of those loads: ahead of this k-block's MFMAs (their
            // wait comes with the copy at the next k-block)
            __asm__ volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < BT; ++t) {
              const floatx4 bx = *reinterpret_cast<const floatx4*>(&X[xoff<CH, WMAX>(0, t * 16 + ln, kb * 16 + lg * 4)]);
#pragma unroll
              for (int f = 0; f < FTMAX; ++f) {
                const int ft = wave + f * NW;
                if (FULL || ft < nft) {
#pragma unroll
                  for (int j = 0; j < 4; ++j)
                    acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[f][j], bx[j], acc[f][t], 0, 0, 0);
                }
              }
            }
          }
        };
        if (wave + (FTMAX - 1) * NW < nft) kloop(std::true_type{});
        else kloop(std::false_type{});
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
      // column sums over the 16 row lanes of each wave, fixed order
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
      // added rows: scale_r * A[idx_r, :], appended after the stored rows;
      // radius = sum |E| + err (src/affine.py:85-91)
      if (Ly.out_pad >= NT) {
        for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
          float dsum = 0.f;
#pragma unroll 8
          for (int r = 0; r < n_new; ++r) {
            const float v = s_scale[r] * Ly.W[(size_t)s_idx[r] * Ly.out_pad + f];
            if (!last) Eout[(n_stored + r) * WMAX + f] = v;
            dsum = dsum + fabsf(v);
          }
          s_rad[f] = (s_rad[f] + dsum) + s_ne[f];
        }
      } else {
        // narrower than the workgroup: G groups of threads split the added
        // rows, and the partial sums join in group order
        const int G = NT / Ly.out_pad;
        const int grp = (int)threadIdx.x / Ly.out_pad, f = (int)threadIdx.x % Ly.out_pad;
        const int per = (n_new + G - 1) / G;
        if (grp < G) {
          float dsum = 0.f;
          const int ra0 = grp * per, ra1 = min(n_new, ra0 + per);
#pragma unroll 8
          for (int r = ra0; r < ra1; ++r) {
            const float v = s_scale[r] * Ly.W[(size_t)s_idx[r] * Ly.out_pad + f];
            if (!last) Eout[(n_stored + r) * WMAX + f] = v;
            dsum = dsum + fabsf(v);
          }
          s_rv[threadIdx.x] = dsum;
        }
        __syncthreads();
        if ((int)threadIdx.x < Ly.out_pad) {
          float dsum = 0.f;
          for (int g2 = 0; g2 < G; ++g2) dsum = dsum + s_rv[g2 * Ly.out_pad + threadIdx.x];
          s_rad[threadIdx.x] = (s_rad[threadIdx.x] + dsum) + s_ne[threadIdx.x];
        }
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
      // linearise each feature (src/affine.py:164-193)
      const bool act = Ly.act != ACT_NONE;
      for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
        const float base = s_nb[f], rad = s_rad[f];
        const float lo = base - rad, hi = base + rad;
        float al = 1.f, be = 0.f, de = 0.f;
        if (act) act_rule(Ly.act, lo, hi, al, be, de);
        s_base[f] = act ? al * base + be : base;
        s_alpha[f] = al;
        s_err[f] = act ? al * s_ne[f] : s_ne[f];
        s_scale[f] = fabsf(de);   // |delta| per feature
        s_idx[f] = f;
      }
      __syncthreads();
      n_stored = n_out;
      n_new = 0;
      if (!act) { cur ^= 1; continue; }
      if (MODE == ROWS_ALL) {
        n_new = Ly.out_pad;          // diag(delta)
        cur ^= 1;
      } else if (MODE == ROWS_APPEND) {
        // top_k(delta, n_append): largest first, lowest index on ties; the
        // rest of sum(delta) goes into every feature's err (src/affine.py:183-191)
        for (int f = threadIdx.x; f < Ly.out_pad; f += NT) { s_mag[f] = s_scale[f]; s_taken[f] = 0; }
        __syncthreads();
        float dsum = 0.f;
        if (threadIdx.x == 0)
          for (int f = 0; f < Ly.out_pad; ++f) dsum = dsum + s_mag[f];
        const int na = min(ra.n_append, Ly.out_pad);
        float ksum = 0.f;
        for (int r = 0; r < na; ++r) {
          float bv;
          const int k = block_argmax<NT>(s_mag, s_taken, Ly.out_pad, s_rv, s_ri, bv);
          if (threadIdx.x == 0) { s_taken[k] = 1; s_keep[r] = k; ksum = ksum + bv; }
          __syncthreads();
        }
        if (threadIdx.x == 0) s_rv[0] = dsum - ksum;
        __syncthreads();
        const float extra = s_rv[0];
        for (int f = threadIdx.x; f < Ly.out_pad; f += NT) s_err[f] = s_err[f] + extra;
        __syncthreads();
        if (threadIdx.x < na) {
          const int k = s_keep[threadIdx.x];
          s_idx[threadIdx.x] = k;
          s_scale[threadIdx.x] = s_mag[k];
        }
        n_new = na;
        cur ^= 1;
      } else {  // ROWS_TRUNCATE
        // candidates: alpha * E rows (in Eout) then diag(|delta|); keep the
        // n_keep of largest sum |row| (stable, decreasing), fold the rest
        // into err (src/affine.py:127-162)
        const long long nc = n_out + Ly.out_pad;
        if (nc > ra.n_keep) {
          for (long long r = threadIdx.x; r < n_out; r += NT) {
            float m = 0.f;
            for (int f = 0; f < Ly.out_pad; ++f) m = m + fabsf(s_alpha[f] * Eout[r * WMAX + f]);
            s_mag[r] = m;
            s_taken[r] = 0;
          }
          for (int f = threadIdx.x; f < Ly.out_pad; f += NT) { s_mag[n_out + f] = s_scale[f]; s_taken[n_out + f] = 0; }
          __syncthreads();
          for (int r = 0; r < ra.n_keep; ++r) {
            float bv;
            const int k = block_argmax<NT>(s_mag, s_taken, (int)nc, s_rv, s_ri, bv);
            if (threadIdx.x == 0) { s_taken[k] = 1; s_keep[r] = k; }
            __syncthreads();
          }
          // dropped rows into err, column-wise, in row order
          for (int f = threadIdx.x; f < Ly.out_pad; f += NT) {
            float add = 0.f;
            for (long long r = 0; r < nc; ++r) {
              if (s_taken[r]) continue;
              const float v = r < n_out ? s_alpha[f] * Eout[r * WMAX + f] : (r - n_out == f ? s_scale[f] : 0.f);
              add = add + fabsf(v);
            }
            s_err[f] = s_err[f] + add;
          }
          // kept rows, in sorted order, into the other buffer (alpha applied)
          for (int e = threadIdx.x; e < ra.n_keep * Ly.out_pad; e += NT) {
            const int r = e / Ly.out_pad, f = e % Ly.out_pad;
            const int k = s_keep[r];
            Ein[(size_t)r * WMAX + f] = k < n_out ? s_alpha[f] * Eout[(size_t)k * WMAX + f]
                                                  : (k - n_out == f ? s_scale[f] : 0.f);
          }
          __syncthreads();
          for (int f = threadIdx.x; f < Ly.out_pad; f += NT) s_alpha[f] = 1.f;
          n_stored = ra.n_keep;
          // cur unchanged: kept rows now live in buf[cur]
        } else {
          // no truncation: every candidate row is kept, alpha-scaled rows then diag
          n_new = Ly.out_pad;
          cur ^= 1;
        }
      }
      __syncthreads();
    }
  }
}
