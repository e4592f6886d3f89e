// Kernels of the range-analysis hot path: box classification, point
// evaluation, the device-resident level-set tree step and the persistent
// ray caster.  Everything here is gfx950-only HIP; host dispatch lives in
// niq_api.hip.
#pragma once
#include "niq_device.h"

// Workgroup shapes per padded max width: boxes per tile (NB) and waves (NW).
// OCC: waves per SIMD the register budget is set for (0: what LDS admits).
template <int WMAX> struct TileCfg;
template <> struct TileCfg<64>  { static constexpr int NB = 32, NW = 4, NBP = 64, OCC = 0; };
template <> struct TileCfg<128> { static constexpr int NB = 32, NW = 4, NBP = 64, OCC = 0; };
template <> struct TileCfg<256> { static constexpr int NB = 16, NW = 8, NBP = 64, OCC = 0; };
template <> struct TileCfg<512> { static constexpr int NB = 16, NW = 8, NBP = 64, OCC = 0; };

// Alternative tiles offered to the autotuner (eval cfg variant 2): twice the
// boxes per workgroup where LDS allows it (weights read half as often per box),
// fewer waves per workgroup at width 512.
template <int WMAX> struct TileCfgAlt;
template <> struct TileCfgAlt<64>  { static constexpr int NB = 64, NW = 4, OCC = 0; };
template <> struct TileCfgAlt<128> { static constexpr int NB = 64, NW = 8, OCC = 0; };
template <> struct TileCfgAlt<256> { static constexpr int NB = 32, NW = 8, OCC = 0; };
template <> struct TileCfgAlt<512> { static constexpr int NB = 16, NW = 4, OCC = 0; };

// The default tile with a 256-register budget (eval cfg variant 3): no
// spills, fewer resident workgroups where LDS would admit more.
template <int WMAX> struct TileCfgRegs {
  static constexpr int NB = TileCfg<WMAX>::NB, NW = TileCfg<WMAX>::NW, OCC = 2;
};

// Tiles for levels too small to fill the chip: a workgroup's latency bounds
// them, so its feature tiles are spread over more waves (one or two each).
template <int WMAX> struct TileCfgWide;
template <> struct TileCfgWide<64>  { static constexpr int NB = 16, NW = 4, OCC = 0; };
template <> struct TileCfgWide<128> { static constexpr int NB = 16, NW = 8, OCC = 0; };
template <> struct TileCfgWide<256> { static constexpr int NB = 16, NW = 16, OCC = 0; };
template <> struct TileCfgWide<512> { static constexpr int NB = 16, NW = 16, OCC = 0; };

constexpr int kLdsLimit = 160 * 1024;

// Ray-caster tiles (eval cfg ray_variant): ray slots per workgroup (NB),
// waves (NW) and the waves per SIMD the register budget is set for (OCC,
// 0 = what LDS admits).  Variant 0 is TileCfg; 1 keeps its shape with a
// 256-register budget (two waves per SIMD: no spills, fewer resident
// workgroups); 2 takes 16-slot tiles at widths up to 128 and 32-slot tiles
// at width 256, also with 256 registers.  Every variant sums each (ray,
// feature) in the same order, so the results are bit-identical.
template <int WMAX, int VAR> struct RayTile {
  static constexpr int NB = TileCfg<WMAX>::NB, NW = TileCfg<WMAX>::NW, OCC = VAR == 0 ? 0 : 2;
};
template <> struct RayTile<64, 2> { static constexpr int NB = 16, NW = 4, OCC = 2; };
template <> struct RayTile<128, 2> { static constexpr int NB = 16, NW = 4, OCC = 2; };
template <> struct RayTile<256, 2> { static constexpr int NB = 32, NW = 8, OCC = 2; };

template <int R, int NB, int WMAX>
struct Fits { static constexpr bool value = (size_t)R * NB * WMAX * 4 <= (size_t)kLdsLimit; };

// Waves per SIMD that the LDS of a tile admits (workgroups per CU x waves per
// workgroup / 4 SIMDs), capped at 4 (128 VGPRs a wave): the register budget
// given to the compiler, so registers do not cut occupancy below what LDS
// allows.
template <int R, int NB, int NW, int WMAX>
struct LdsOcc {
  static constexpr int wgs = (int)((size_t)kLdsLimit / ((size_t)R * NB * WMAX * 4));
  static constexpr int raw = wgs * NW / 4;
  static constexpr int value = raw < 1 ? 1 : (raw > 4 ? 4 : raw);
};

// ----------------------------------------------------------- classification
struct ClassifyArgs {
  const float* lower;      // axis-aligned boxes (n, d) or null
  const float* upper;
  const float* center;     // general boxes (n, d)
  const float* vecs;       // (n, vin, d)
  int vin;                 // vectors per general box
  int d;                   // spatial dimension (<= 16)
  const int* n_dev;        // device-resident count (null: use n)
  long long n;
  float offset;
  int* codes;              // (n,) sign codes
  float* bounds;           // (n, 2) lower/upper, nullable
  float recheck_tol;       // > 0: list boxes within tol*scale of the threshold
  int* recheck_idx;        // (recheck_cap,)
  int* recheck_cnt;        // device counter
  int recheck_cap;
  void* lw_buf;            // precision mode, layer by layer: activation scratch
  size_t lw_bytes;         //   (null: allocated per call, stream ordered)
  float lipschitz;         // weak-SDF test: Lipschitz bound of f
  float* gx;               // forms too large for LDS: activation planes in global
                           //   memory, R * NB * WMAX floats per workgroup
};

// Weak-SDF sign code (src/sdf.py:42-48): UNKNOWN where the value at the centre
// could change sign within the box, else its sign against the offset.
__device__ __forceinline__ int sdf_code(float val, float rad, float lipschitz, float offset) {
  const bool can_change = fabsf(val) - __fmul_rn(rad, lipschitz) < 0.f;
  int code = CODE_UNKNOWN;
  if (!can_change && val > offset) code = CODE_POSITIVE;
  if (!can_change && val < -offset) code = CODE_NEGATIVE;
  return code;
}

// Fill the layer-0 input planes of one tile (features 0..15, zero padded).
template <int KIND, int V, int NB, int NT, int WMAX>
__device__ __forceinline__ void fill_box_inputs(float* X, const ClassifyArgs& a, long long i0, long long n) {
  using PL = Planes<KIND, V, 0>;
  constexpr int RB = PL::RB;
  for (int e = threadIdx.x; e < RB * NB * 16; e += NT) {
    const int p = e / (NB * 16);
    const int b = (e / 16) % NB;
    const int k = e & 15;
    const long long i = i0 + b;
    float val = 0.f;
    if (i < n && k < a.d) {
      float c, h = 0.f;
      if (a.lower) {
        const float lo = a.lower[i * a.d + k], up = a.upper[i * a.d + k];
        c = 0.5f * (lo + up);
        h = up - c;
        if (p == 0) val = c;
        else if (KIND == 0 && V == 0) val = fabsf(h);              // interval err
        else if (p >= 1 && p <= V) val = (p - 1 == k) ? h : 0.f;   // diag(half)
        else val = 0.f;
      } else {
        if (p == 0) val = a.center[i * a.d + k];
        else if (KIND == 0 && V == 0) {
          float s = 0.f;
          for (int r = 0; r < a.vin; ++r) s = s + fabsf(a.vecs[(i * a.vin + r) * a.d + k]);
          val = s;
        } else if (p >= 1 && p <= V) val = a.vecs[(i * a.vin + (p - 1)) * a.d + k];
        else val = 0.f;
      }
    }
    X[xoff<NB, WMAX>(p, b, k)] = val;
  }
}

template <int KIND, int V, int NB, int NW, int WMAX, bool GX = false, int OCC = 0, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64, (OCC ? OCC : (GX ? 2 : LdsOcc<Planes<KIND, V, 0>::R, NB, NW, WMAX>::value)))
k_classify(const DevModel* __restrict__ Mp, ClassifyArgs a) {
  using PL = Planes<KIND, V, 0>;
  constexpr int NT = NW * 64;
  float* X;
  if constexpr (GX) {
    // this workgroup's slice of the global activation planes (L2-resident)
    X = a.gx + (size_t)blockIdx.x * (PL::R * NB * WMAX);
  } else {
    __shared__ float Xs[PL::R * NB * WMAX];
    X = Xs;
  }
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    fill_box_inputs<KIND, V, NB, NT, WMAX>(X, a, i0, n);
    mlp_forward<PL, KIND, V, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i >= n) return;
      float lo, hi, scale;
      final_bounds<KIND, V>(v, lo, hi, scale);
      a.codes[i] = sign_code(lo, hi, a.offset);
      if (a.bounds) { a.bounds[2 * i] = lo; a.bounds[2 * i + 1] = hi; }
      if (a.recheck_tol > 0.f) {
        const float tol = a.recheck_tol * scale;
        if (fabsf(lo - a.offset) <= tol || fabsf(hi + a.offset) <= tol) {
          const int slot = atomicAdd(a.recheck_cnt, 1);
          if (slot < a.recheck_cap) a.recheck_idx[slot] = (int)i;
        }
      }
    });
  }
}

// Weak-SDF classification (src/sdf.py:31-50): f at the box centre against
// lipschitz * radius, radius = sqrt(sum_v |vec_v|^2), |vec| = sqrt(sum_k vec_k^2)
// (axis boxes: vecs = diag(upper - centre), src/implicit_function.py:28-37).
// bounds, when asked for, are val -/+ lipschitz * radius.
template <int NB, int NW, int WMAX>
__global__ void __launch_bounds__(NW * 64)
k_classify_sdf(const DevModel* __restrict__ Mp, ClassifyArgs a) {
  using PL = Planes<2, 0, 1>;
  constexpr int NT = NW * 64;
  __shared__ float X[NB * WMAX];
  __shared__ float Rd[NB];
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    __syncthreads();  // the previous tile's epilogue has read Rd
    for (int e = threadIdx.x; e < NB * 16; e += NT) {
      const int b = e >> 4, k = e & 15;
      const long long i = i0 + b;
      float val = 0.f;
      if (i < n && k < a.d)
        val = a.lower ? 0.5f * (a.lower[i * a.d + k] + a.upper[i * a.d + k]) : a.center[i * a.d + k];
      X[xoff<NB, WMAX>(0, b, k)] = val;
    }
    for (int b = threadIdx.x; b < NB; b += NT) {
      const long long i = i0 + b;
      float r2 = 0.f;
      if (i < n) {
        if (a.lower) {
          for (int k = 0; k < a.d; ++k) {
            const float c = 0.5f * (a.lower[i * a.d + k] + a.upper[i * a.d + k]);
            const float h = a.upper[i * a.d + k] - c;
            const float nv = sqrtf(__fmul_rn(h, h));
            r2 = __fadd_rn(r2, __fmul_rn(nv, nv));
          }
        } else {
          for (int v = 0; v < a.vin; ++v) {
            float ss = 0.f;
            for (int k = 0; k < a.d; ++k) {
              const float x = a.vecs[(i * a.vin + v) * a.d + k];
              ss = __fadd_rn(ss, __fmul_rn(x, x));
            }
            const float nv = sqrtf(ss);
            r2 = __fadd_rn(r2, __fmul_rn(nv, nv));
          }
        }
      }
      Rd[b] = sqrtf(r2);
    }
    mlp_forward<PL, 2, 0, NB, NW, WMAX>(*Mp, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i >= n) return;
      const float val = v[0], rad = Rd[b];
      a.codes[i] = sdf_code(val, rad, a.lipschitz, a.offset);
      if (a.bounds) {
        const float r = __fmul_rn(rad, a.lipschitz);
        a.bounds[2 * i] = val - r;
        a.bounds[2 * i + 1] = val + r;
      }
    });
  }
}

// ------------------------------------------------ FP64 near-threshold re-test
// One workgroup per listed box: the same range-analysis forward in FP64 on the
// vector ALUs, overwriting the FP32 sign code.  Used for the few cells whose
// FP32 bound lies within recheck_tol * scale of the threshold, so that a cell
// is only reported empty when the FP64 restatement agrees.
template <int KIND, int V, int WMAX>
static __global__ void __launch_bounds__(256)
k_recheck_fp64(const DevModel* __restrict__ Mp, ClassifyArgs a) {
  using PL = Planes<KIND, V, 0>;
  constexpr int R = PL::R;
  __shared__ double X[2][R][WMAX];
  const DevModel& M = *Mp;
  const int n = min(*a.recheck_cnt, a.recheck_cap);
  for (int j = blockIdx.x; j < n; j += gridDim.x) {
    const long long i = a.recheck_idx[j];
    for (int e = threadIdx.x; e < R * 16; e += blockDim.x) {
      const int p = e / 16, k = e % 16;
      double val = 0.0;
      if (k < a.d) {
        if (a.lower) {
          const double lo = a.lower[i * a.d + k], up = a.upper[i * a.d + k];
          const double c = 0.5 * (lo + up), h = up - c;
          if (p == 0) val = c;
          else if (KIND == 0 && V == 0) val = ::fabs(h);
          else if (p >= 1 && p <= V) val = (p - 1 == k) ? h : 0.0;
        } else {
          if (p == 0) val = a.center[i * a.d + k];
          else if (KIND == 0 && V == 0) {
            double s = 0.0;
            for (int r = 0; r < a.vin; ++r) s = s + ::fabs((double)a.vecs[(i * a.vin + r) * a.d + k]);
            val = s;
          } else if (p >= 1 && p <= V) val = a.vecs[(i * a.vin + (p - 1)) * a.d + k];
        }
      }
      X[0][p][k] = val;
    }
    __syncthreads();
    int cur = 0;
    for (int l = 0; l < M.n_layers; ++l) {
      const DevLayer& Ly = M.L[l];
      const bool last = (l == M.n_layers - 1);
      for (int f = threadIdx.x; f < Ly.out_pad; f += blockDim.x) {
        double v[R];
#pragma unroll
        for (int p = 0; p < R; ++p) v[p] = 0.0;
        for (int k = 0; k < Ly.in_pad; ++k) {
          const double w = (double)Ly.W[(size_t)k * Ly.out_pad + f];
#pragma unroll
          for (int p = 0; p < R; ++p) v[p] = v[p] + X[cur][p][k] * (PL::absw(p) ? ::fabs(w) : w);
        }
#pragma unroll
        for (int p = 0; p < R; ++p)
          if (PL::bias(p)) v[p] = v[p] + (double)Ly.bias[f];
        if (!last) {
          apply_hidden<PL, KIND, V>(Ly.act, v);
#pragma unroll
          for (int p = 0; p < R; ++p) X[cur ^ 1][p][f] = v[p];
        } else if (f == 0) {
          double lo, hi, scale;
          final_bounds<KIND, V>(v, lo, hi, scale);
          a.codes[i] = sign_code<double>(lo, hi, (double)a.offset);
        }
      }
      __syncthreads();
      cur ^= 1;
    }
  }
}

// ------------------------------------------------------------ point values
template <int NB, int NW, int WMAX, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64)
k_points(const DevModel* __restrict__ Mp, const float* __restrict__ x, int d, long long n,
         float* __restrict__ out) {
  using PL = Planes<2, 0, 1>;
  constexpr int NT = NW * 64;
  __shared__ float X[NB * WMAX];
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    for (int e = threadIdx.x; e < NB * 16; e += NT) {
      const int b = e >> 4, k = e & 15;
      const long long i = i0 + b;
      X[xoff<NB, WMAX>(0, b, k)] = (i < n && k < d) ? x[i * d + k] : 0.f;
    }
    mlp_forward<PL, 2, 0, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i < n) out[i] = v[0];
    });
  }
}

// Point values of n = (*n_dev) * mult points (3D), or of n_host points when
// n_dev is null.
template <int NB, int NW, int WMAX, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64)
k_points_dev(const DevModel* __restrict__ Mp, const float* __restrict__ x, const int* __restrict__ n_dev, int mult,
             long long n_host, float* __restrict__ out) {
  using PL = Planes<2, 0, 1>;
  constexpr int NT = NW * 64;
  __shared__ float X[NB * WMAX];
  const long long n = n_dev ? (long long)(*n_dev) * mult : n_host;
  const long long ntiles = (n + NB - 1) / NB;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long i0 = tile * NB;
    for (int e = threadIdx.x; e < NB * 16; e += NT) {
      const int b = e >> 4, k = e & 15;
      const long long i = i0 + b;
      X[xoff<NB, WMAX>(0, b, k)] = (i < n && k < 3) ? x[i * 3 + k] : 0.f;
    }
    mlp_forward<PL, 2, 0, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i < n) out[i] = v[0];
    });
  }
}

// -------------------------------------------------- level-set tree step
// Device counters of one tree build.  n[l & 1] is the node count of level l.
struct TreeCounters {
  int n[2];
  int n_int, n_ext;
  long long n_classified;
  int overflow;
  int recheck_cnt;       // cells listed for the FP64 re-test in this level
  long long n_recheck;   // over all levels
  int level_n[64];
  int level_int[64];
  int level_ext[64];
};

// Per batch of B nodes: counts of UNKNOWN / POSITIVE / NEGATIVE codes.  The
// bodies of the tree-step kernels take their batch index (bx) as an argument,
// so a batched build of several trees runs them with one tree per grid row.
__device__ __forceinline__ void tree_count_body(const int* __restrict__ codes, const TreeCounters* __restrict__ C,
                                                int level, int B, int* __restrict__ bcnt, int bx) {
  const int n = C->n[level & 1];
  const long long start = (long long)bx * B;
  if (start >= n) return;
  const long long end = min(start + (long long)B, (long long)n);
  int cu = 0, cp = 0, cn = 0;
  for (long long i = start + threadIdx.x; i < end; i += blockDim.x) {
    const int c = codes[i];
    cu += (c == CODE_UNKNOWN); cp += (c == CODE_POSITIVE); cn += (c == CODE_NEGATIVE);
  }
  __shared__ int s[3][256];
  s[0][threadIdx.x] = cu; s[1][threadIdx.x] = cp; s[2][threadIdx.x] = cn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
      s[2][threadIdx.x] += s[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bcnt[bx * 4 + 0] = s[0][0];
    bcnt[bx * 4 + 1] = s[1][0];
    bcnt[bx * 4 + 2] = s[2][0];
  }
}

static __global__ void __launch_bounds__(256)
k_tree_count(const int* __restrict__ codes, const TreeCounters* __restrict__ C, int level, int B,
             int* __restrict__ bcnt) {
  tree_count_body(codes, C, level, B, bcnt, blockIdx.x);
}

// Single workgroup of 1024: exclusive scan of the per-batch counts, level
// totals.  boff[b] = {unknown prefix, exterior offset, interior offset}.
__device__ __forceinline__ void tree_scan_body(const int* __restrict__ bcnt, int* __restrict__ boff,
                                               TreeCounters* C, int level, int B, int split, long long cap_next,
                                               long long cap_int, long long cap_ext) {
  const int n = C->n[level & 1];
  const int nb = (int)((n + (long long)B - 1) / B);
  __shared__ int s[3][1024];
  __shared__ int carry[3];
  if (threadIdx.x == 0) { carry[0] = 0; carry[1] = C->n_ext; carry[2] = C->n_int; }
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int b = base + threadIdx.x;
    int v0 = 0, v1 = 0, v2 = 0;
    if (b < nb) { v0 = bcnt[b * 4 + 0]; v1 = bcnt[b * 4 + 1]; v2 = bcnt[b * 4 + 2]; }
    s[0][threadIdx.x] = v0; s[1][threadIdx.x] = v1; s[2][threadIdx.x] = v2;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1) {   // Hillis-Steele inclusive scan
      int a0 = 0, a1 = 0, a2 = 0;
      if ((int)threadIdx.x >= w) {
        a0 = s[0][threadIdx.x - w]; a1 = s[1][threadIdx.x - w]; a2 = s[2][threadIdx.x - w];
      }
      __syncthreads();
      s[0][threadIdx.x] += a0; s[1][threadIdx.x] += a1; s[2][threadIdx.x] += a2;
      __syncthreads();
    }
    if (b < nb) {
      boff[b * 4 + 0] = carry[0] + s[0][threadIdx.x] - v0;
      boff[b * 4 + 1] = carry[1] + s[1][threadIdx.x] - v1;
      boff[b * 4 + 2] = carry[2] + s[2][threadIdx.x] - v2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry[0] += s[0][1023]; carry[1] += s[1][1023]; carry[2] += s[2][1023];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int U = carry[0];
    C->level_n[level] = n;
    C->n_recheck += C->recheck_cnt;
    C->recheck_cnt = 0;
    C->level_int[level] = C->n_int;
    C->level_ext[level] = C->n_ext;
    C->n_classified += n;
    C->n_ext = carry[1];
    C->n_int = carry[2];
    const long long nn = split ? 2LL * U : (long long)U;
    if (nn > cap_next || carry[1] > cap_ext || carry[2] > cap_int) C->overflow = 1;
    C->n[(level + 1) & 1] = (int)min(nn, cap_next);
  }
}

static __global__ void __launch_bounds__(1024)
k_tree_scan(const int* __restrict__ bcnt, int* __restrict__ boff, TreeCounters* C, int level, int B,
            int split, long long cap_next, long long cap_int, long long cap_ext) {
  tree_scan_body(bcnt, boff, C, level, B, split, cap_next, cap_int, cap_ext);
}

struct TreeScatterArgs {
  const int* codes;
  const float* lower; const float* upper;   // current level (n, 3)
  float* next_lower; float* next_upper;     // next level / final unknown list
  long long cap_next;
  float* int_lower; float* int_upper; long long cap_int;   // nullable
  float* ext_lower; float* ext_upper; long long cap_ext;   // nullable
  // heap indices of the nodes (root 1, children 2h and 2h + 1) while the
  // level lies in the precomputed top of the tree; null below it
  const int* heap; int* next_heap; long long cap_heap;
};

__device__ __forceinline__ void store3(float* p, long long i, float x, float y, float z) {
  p[3 * i] = x; p[3 * i + 1] = y; p[3 * i + 2] = z;
}

// Stable per-batch placement of children and finished leaves, in the
// reference's order (src/kd_tree.py:46-96, src/bucketing.py:16-32).  The batch
// is walked in chunks of 256 consecutive nodes, one per thread (coalesced
// reads and writes); ranks within a chunk come from wave ballots.
__device__ __forceinline__ void tree_scatter_body(const TreeScatterArgs& a, const int* __restrict__ bcnt,
                                                  const int* __restrict__ boff, const TreeCounters* __restrict__ C,
                                                  int level, int B, int split, int bx) {
  const int n = C->n[level & 1];
  const long long start = (long long)bx * B;
  if (start >= n) return;
  const int cnt = (int)min((long long)B, (long long)n - start);
  const long long Ub = boff[bx * 4 + 0];
  const long long nUb = bcnt[bx * 4 + 0];
  const long long Eb = boff[bx * 4 + 1];
  const long long Ib = boff[bx * 4 + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  __shared__ int wsum[3][4];
  int ru0 = 0, rp0 = 0, rn0 = 0;   // nodes of each kind before this chunk
  for (int c0 = 0; c0 < cnt; c0 += 256) {
    const int j = c0 + (int)threadIdx.x;
    const bool in = j < cnt;
    const long long i = start + j;
    const int c = in ? a.codes[i] : -1;
    const unsigned long long mu = __ballot(c == CODE_UNKNOWN);
    const unsigned long long mp = __ballot(c == CODE_POSITIVE);
    const unsigned long long mn = __ballot(c == CODE_NEGATIVE);
    if (lane == 0) { wsum[0][wave] = __popcll(mu); wsum[1][wave] = __popcll(mp); wsum[2][wave] = __popcll(mn); }
    __syncthreads();
    int pu = 0, pp = 0, pn = 0, tu = 0, tp = 0, tn = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (w < wave) { pu += wsum[0][w]; pp += wsum[1][w]; pn += wsum[2][w]; }
      tu += wsum[0][w]; tp += wsum[1][w]; tn += wsum[2][w];
    }
    __syncthreads();   // wsum is rewritten by the next chunk
    if (in) {
      const float l0 = a.lower[3 * i], l1 = a.lower[3 * i + 1], l2 = a.lower[3 * i + 2];
      const float u0 = a.upper[3 * i], u1 = a.upper[3 * i + 1], u2 = a.upper[3 * i + 2];
      if (c == CODE_UNKNOWN) {
        const long long ru = ru0 + pu + __popcll(mu & below);
        if (split) {
          const float e0 = u0 - l0, e1 = u1 - l1, e2 = u2 - l2;
          int dim = 0; float em = e0;
          if (e1 > em) { dim = 1; em = e1; }
          if (e2 > em) { dim = 2; }
          const float m0 = 0.5f * (l0 + u0), m1 = 0.5f * (l1 + u1), m2 = 0.5f * (l2 + u2);
          const long long pa = 2 * Ub + ru, pb = 2 * Ub + nUb + ru;
          if (pa < a.cap_next) {
            store3(a.next_lower, pa, l0, l1, l2);
            store3(a.next_upper, pa, dim == 0 ? m0 : u0, dim == 1 ? m1 : u1, dim == 2 ? m2 : u2);
          }
          if (pb < a.cap_next) {
            store3(a.next_lower, pb, dim == 0 ? m0 : l0, dim == 1 ? m1 : l1, dim == 2 ? m2 : l2);
            store3(a.next_upper, pb, u0, u1, u2);
          }
          if (a.next_heap) {
            const int h = a.heap[i];
            if (pa < a.cap_heap) a.next_heap[pa] = 2 * h;
            if (pb < a.cap_heap) a.next_heap[pb] = 2 * h + 1;
          }
        } else {
          const long long p = Ub + ru;
          if (p < a.cap_next) { store3(a.next_lower, p, l0, l1, l2); store3(a.next_upper, p, u0, u1, u2); }
        }
      } else if (c == CODE_POSITIVE) {
        const long long p = Eb + rp0 + pp + __popcll(mp & below);
        if (a.ext_lower && p < a.cap_ext) { store3(a.ext_lower, p, l0, l1, l2); store3(a.ext_upper, p, u0, u1, u2); }
      } else if (c == CODE_NEGATIVE) {
        const long long p = Ib + rn0 + pn + __popcll(mn & below);
        if (a.int_lower && p < a.cap_int) { store3(a.int_lower, p, l0, l1, l2); store3(a.int_upper, p, u0, u1, u2); }
      }
    }
    ru0 += tu; rp0 += tp; rn0 += tn;
  }
}

static __global__ void __launch_bounds__(256)
k_tree_scatter(TreeScatterArgs a, const int* __restrict__ bcnt, const int* __restrict__ boff,
               const TreeCounters* __restrict__ C, int level, int B, int split) {
  tree_scatter_body(a, bcnt, boff, C, level, B, split, blockIdx.x);
}

// --------------------------------------------------------- ray casting
// One group of rays of a cast: its functions (models[model0 .. model0 +
// n_models)), options, inputs and outputs.  A cast takes a table of groups;
// rays of different groups never share a tile, so groups may differ in
// models and options (not in range-analysis mode or template width).
struct RayGroup {
  const float* roots;      // (n, 3)
  const float* dirs;       // (n, 3)
  long long n;
  int model0, n_models;
  float hit_eps, max_dist, safety, grow, shrink, init_size;
  int n_max_step, n_substeps;
  float* out_t; int* out_hit; int* out_count;
  unsigned long long* next_ray;  // work counter, zero on entry
  int* iter_hist;          // (n_max_step + 2,) rays by iterations lived, nullable
  float lipschitz;         // weak-SDF form (KIND 2): Lipschitz bound of f
};

struct RayArgs {
  const RayGroup* groups;
  int n_groups;
};

// Persistent certified-step ray caster (src/queries.py:39-175).  A workgroup
// owns NB ray slots; every step it bounds the next segment of each live ray
// and samples the two points of the hit test, for each function in turn.
// Terminated slots are refilled from the current group's counter, so the tile
// stays full and no host round trip happens until all rays are done; once a
// group has no rays left and the tile has drained, the workgroup moves to the
// next group (starting from group blockIdx mod n_groups) until it has seen
// every group run dry.  Slot state is replicated in lanes 0..NB-1 of every
// wave (each wave computes the same update), and per-slot results are
// exchanged through LDS.  KIND 2 is the weak-SDF form: three point planes
// (segment centre, start, start + hit_eps) and the segment radius |half_vec|
// kept in LDS per slot.
template <int KIND, int V, int NB, int NW, int WMAX,
          int OCC = LdsOcc<Planes<KIND, V, KIND == 2 ? 3 : 2>::R, NB, NW, WMAX>::value, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64, OCC)
k_cast_rays(const DevModel* __restrict__ models, RayArgs a) {
  using PL = Planes<KIND, V, KIND == 2 ? 3 : 2>;
  constexpr int R = PL::R;
  constexpr int PT = KIND == 2 ? 1 : PL::RB;   // first hit-test point plane
  constexpr int NT = NW * 64;
  __shared__ float X[R * NB * WMAX];
  __shared__ float Rd[KIND == 2 ? NB : 1];
  // the refill's 64-bit base is passed through the first 8 bytes of the planes
  long long* s_base = reinterpret_cast<long long*>(X);
  const int lane = threadIdx.x & 63;
  const bool slot_lane = lane < NB;

  int rid = -1, cnt = 0, iters = 0;
  float t = 0.f, step = 0.f;
  float ro[3] = {0.f, 0.f, 0.f}, di[3] = {0.f, 0.f, 0.f};
  int gi = blockIdx.x % a.n_groups, seen = 0;
  RayGroup G = a.groups[gi];
  bool exhausted = false;   // this group's rays are all claimed

  while (true) {
    // ---- refill dead slots (identical in every wave)
    const unsigned long long dead = __ballot(slot_lane && rid < 0);
    const int n_dead = __popcll(dead);
    __syncthreads();
    if (threadIdx.x == 0)   // -1: no claim (no dead slot, or nothing left to claim)
      *s_base = (n_dead && !exhausted) ? (long long)atomicAdd(G.next_ray, (unsigned long long)n_dead) : -1LL;
    __syncthreads();
    const long long base = *s_base;
    if (base >= 0 && base + n_dead >= G.n) exhausted = true;
    if (base >= 0 && slot_lane && rid < 0) {
      const int rank = __popcll(dead & ((1ull << lane) - 1ull));
      const long long r = base + rank;
      if (r < G.n) {
        rid = (int)r;
        t = 0.f; step = G.init_size; cnt = 0; iters = 0;
        for (int k = 0; k < 3; ++k) { ro[k] = G.roots[3 * r + k]; di[k] = G.dirs[3 * r + k]; }
      }
    }
    const bool any_live = __ballot(slot_lane && rid >= 0) != 0ull;
    if (!any_live) {   // uniform: every wave holds the same slot state
      // every slot was dead and the claim found no ray left (or none was
      // made): the group is drained here, on to the next group
      if (++seen >= a.n_groups) break;
      gi = gi + 1 == a.n_groups ? 0 : gi + 1;
      G = a.groups[gi];
      exhausted = false;
      continue;
    }

    // ---- one iteration = n_substeps substeps (src/queries.py:94-111)
    bool is_hit = false;
    int hit_id = 0, nsteps = 0;
    for (int sub = 0; sub < G.n_substeps; ++sub) {
      bool can_step = !is_hit;
      nsteps += is_hit ? 0 : 1;
      for (int f = 0; f < G.n_models; ++f) {
        __syncthreads();   // previous readers of X are done
        for (int e = threadIdx.x; e < R * NB * 16; e += NT) {
          const int k = e & 15;
          if (k >= 3) X[xoff<NB, WMAX>(e / (NB * 16), (e / 16) % NB, k)] = 0.f;
        }
        if (threadIdx.x < NB) {
          const int b = threadIdx.x;
          const bool live = rid >= 0;
          const float hs = 0.5f * step;
          float hh = 0.f;
          for (int k = 0; k < 3; ++k) {
            const float ps = ro[k] + t * di[k];
            const float hv = hs * di[k];
            const float pm = ps + hv;
            const float pe = ro[k] + (t + G.hit_eps) * di[k];
            float vals[R];
            for (int p = 0; p < R; ++p) vals[p] = 0.f;
            vals[0] = pm;
            if constexpr (KIND == 0 && V == 0) vals[1] = fabsf(hv);
            else if constexpr (KIND != 2) vals[1] = hv;
            vals[PT] = ps;
            vals[PT + 1] = pe;
            hh = __fadd_rn(hh, __fmul_rn(hv, hv));
#pragma unroll
            for (int p = 0; p < R; ++p) X[xoff<NB, WMAX>(p, b, k)] = live ? vals[p] : 0.f;
          }
          if constexpr (KIND == 2) {
            const float nv = sqrtf(hh);
            Rd[b] = sqrtf(__fmul_rn(nv, nv));
          }
        }
        mlp_forward<PL, KIND, V, NB, NW, WMAX, AS>(models[G.model0 + f], X, [&](const float* v, int b) {
          int code;
          if constexpr (KIND == 2) {
            code = sdf_code(v[0], Rd[b], G.lipschitz, 0.f);
          } else {
            float lo, hi, scale;
            final_bounds<KIND, V>(v, lo, hi, scale);
            code = sign_code(lo, hi, 0.f);
          }
          X[4 * b + 0] = __int_as_float(code);
          X[4 * b + 1] = v[PT];
          X[4 * b + 2] = v[PT + 1];
        });
        __syncthreads();
        if (slot_lane) {
          const int code = __float_as_int(X[4 * lane]);
          const float v0 = X[4 * lane + 1], v1 = X[4 * lane + 2];
          can_step = can_step && (code == CODE_POSITIVE || code == CODE_NEGATIVE);
          const bool this_hit = sign_f(v0) != sign_f(v1);
          if (this_hit) hit_id = f + 1;
          is_hit = is_hit || this_hit;
        }
      }
      const float this_step = can_step ? step : G.hit_eps;
      if (!is_hit) t = t + this_step * G.safety;
      step = can_step ? step * G.grow : step * G.shrink;
      step = fmaxf(step, G.hit_eps);
    }

    // ---- termination (src/queries.py:113-129)
    if (slot_lane && rid >= 0) {
      cnt += nsteps;
      ++iters;
      const bool term = is_hit || (t > G.max_dist) || (cnt >= G.n_max_step);
      if (term) {
        if (threadIdx.x < 64) {   // wave 0 writes
          G.out_t[rid] = t;
          G.out_hit[rid] = hit_id;
          G.out_count[rid] = cnt;
          if (G.iter_hist) atomicAdd(&G.iter_hist[min(iters, G.n_max_step + 1)], 1);
        }
        rid = -1;
      }
    }
  }
}
