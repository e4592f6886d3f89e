#pragma once
// Precision mode, layer by layer: every hidden layer of a chunk of boxes is
// one batched contraction over the whole chunk, with the split bf16 planes of
// the activations streamed through HBM between layers.
//
// The fused precision-mode kernel keeps a tile's activations in LDS across
// all layers, which caps it at 16 boxes per workgroup at width 512 (10 bf16
// planes x 16 boxes x 512 features = 160 KB) and makes it re-read every
// weight once per 16 boxes.  Here a workgroup owns 128 output features x 64
// boxes of one layer, so a weight is read once per 64 boxes (and split into
// bf16 terms once per workgroup), at the price of writing and re-reading the
// activation planes (10 planes x 2 B per feature per box) between layers.
// The arithmetic, its order and the rigorous widening are those of
// mlp_forward_bf16: the same products, smallest first, fresh accumulator per
// 32-wide K chunk, the same e-plane and elementwise rules.
//
// HBM activation layout, per chunk of boxes: [box tile of 16][K chunk of 32]
// [plane][box][k] bf16, so that one (box tile, K chunk) block is contiguous
// (NPL * 1 KB) and stages into LDS with 16-byte loads.
#include "niq_bf16.h"

namespace lw {

constexpr int kBoxes = 64;     // boxes per workgroup (4 box tiles)
constexpr int kFeats = 128;    // output features per workgroup (8 feature tiles)
constexpr int kWaves = 8;      // 2 feature groups x 4 box tiles
constexpr int kThreads = kWaves * 64;

template <int V> struct Cfg {
  static constexpr int NPL = PlanesBf16<V>::NPL;
  static constexpr int BLK = NPL * 16 * 32;          // bf16 elements per (box tile, K chunk)
  static constexpr int XCH = 4 * NPL * 64;           // 16-byte chunks per staged K chunk (4 box tiles)
  static constexpr int XPT = (XCH + kThreads - 1) / kThreads;
};

// Bytes of activation scratch for chunks of `nch` boxes (two ping-pong
// buffers of the widest layer).
template <int V>
inline size_t scratch_bytes(long long nch, int wmax) {
  return 2 * (size_t)nch * Cfg<V>::NPL * (size_t)wmax * 2;
}

__device__ __forceinline__ size_t blk_off(long long bt, int kb, int nkb, int npl) {
  return ((size_t)bt * nkb + kb) * (size_t)npl * 512;
}

// Swizzled 16-byte chunk within a [16 rows][4 chunks] plane block: chunk c of
// row r sits at r * 4 + (c ^ ((r >> 1) & 3)), so the 16 lanes of a quarter
// wave reading one chunk of 16 consecutive rows touch distinct banks.
__device__ __forceinline__ int swz(int r, int c) { return r * 4 + (c ^ ((r >> 1) & 3)); }

// Layer-0 input planes (K = 32, features >= d zero) for boxes [box0, box0 + nb).
template <int V>
__global__ void __launch_bounds__(256)
k_lw_prologue(const DevModelBf16* __restrict__ Mp, ClassifyArgs a, long long box0, long long nb,
              unsigned short* __restrict__ X0) {
  using C = Cfg<V>;
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  const float cb0 = Mp->L[0].c_base, ca0 = Mp->L[0].c_aff;
  const long long total = nb * 32;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const long long bl = e >> 5;
    const int k = (int)(e & 31);
    const long long i = box0 + bl;
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
    unsigned short* B = X0 + blk_off(bl >> 4, 0, 1, C::NPL) + (bl & 15) * 32 + k;
    const unsigned short x1 = bf16_bits_rn(s[0]);
    const float r1 = s[0] - bf16_to_f(x1);
    const unsigned short x2 = bf16_bits_rn(r1);
    B[0 * 512] = x1;
    B[1 * 512] = x2;
    B[2 * 512] = bf16_bits_rn(r1 - bf16_to_f(x2));
    float amag = 0.f;
#pragma unroll
    for (int p = 0; p < V; ++p) {
      const unsigned short h = bf16_bits_rn(s[1 + p]);
      B[(3 + 2 * p) * 512] = h;
      B[(4 + 2 * p) * 512] = bf16_bits_rn(s[1 + p] - bf16_to_f(h));
      amag = amag + fabsf(s[1 + p]);
    }
    const float ev = up1(up1(err + cb0 * fabsf(s[0]) + up1(ca0 * up1(amag * (1.f + 4.f * kU24))) + kTinyE));
    B[(3 + 2 * V) * 512] = bf16_bits_up_nonneg(ev);
  }
}

// One hidden layer l (out_pad a multiple of 128) for boxes [box0, box0 + nb):
// reads Xin (K = in_pad) and writes Xout (K = out_pad) in the block layout.
template <int V>
__global__ void __launch_bounds__(kThreads)
k_lw_layer(const DevModelBf16* __restrict__ Mp, int l, ClassifyArgs a, long long box0, long long nb,
           const unsigned short* __restrict__ Xin, unsigned short* __restrict__ Xout) {
  using C = Cfg<V>;
  constexpr int NPL = C::NPL;
  constexpr int NS = V + 1;
  constexpr int R = NS + 1;
  __shared__ uint4 Xl[2][4][NPL][64];       // [buf][box tile][plane][swizzled 16 rows x 4 chunks]
  __shared__ uint4 Wl[2][3][8][64];         // [buf][w1 w2 w3][feature tile][lane]

  const DevLayerBf16 Ly = Mp->L[l];
  const int nkb = Ly.in_pad >> 5;
  const int nkb_out = Ly.out_pad >> 5;
  const int nfg = Ly.out_pad / kFeats;
  // XCD-aware tile order: the feature groups of one box tile share g % 8, so
  // they run on the same XCD and share its L2 copy of the input block
  const long long g = blockIdx.x;
  const int fgi = (int)((g >> 3) % nfg);
  const long long tile = (g / (8LL * nfg)) * 8 + (g & 7);
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  const long long ntile = (nb + kBoxes - 1) / kBoxes;
  if (tile >= ntile || box0 + tile * kBoxes >= n) return;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ln = lane & 15;
  const int lg = lane >> 4;
  const int fg = wave & 1;       // feature group of this wave: feature tiles 4 fg .. 4 fg + 3
  const int bt = wave >> 1;      // box tile of this wave
  const long long bt0 = tile * 4;
  const floatx4 z4 = {0.f, 0.f, 0.f, 0.f};

  // staging registers: this thread's share of one K chunk
  uint4 xs[C::XPT];
  floatx4 ws[2];
  const int wft = threadIdx.x >> 6;           // weight fragment (feature tile) this thread splits
  auto load_stage = [&](int kb) {
#pragma unroll
    for (int q = 0; q < C::XPT; ++q) {
      const int c = threadIdx.x + q * kThreads;
      if (c < C::XCH) {
        const int btl = c / (NPL * 64), w = c % (NPL * 64);
        const uint4* src = reinterpret_cast<const uint4*>(Xin + blk_off(bt0 + btl, kb, nkb, NPL)) + w;
        xs[q] = *src;
      }
    }
    const int ftg = fgi * 8 + wft;
    const floatx4* P = Ly.W32 + 2 * (size_t)(ftg * nkb + kb) * 64 + lane;
    ws[0] = P[0];
    ws[1] = P[64];
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < C::XPT; ++q) {
      const int c = threadIdx.x + q * kThreads;
      if (c < C::XCH) {
        const int btl = c / (NPL * 64), w = c % (NPL * 64);
        const int p = w >> 6, r = (w >> 2) & 15, ch = w & 3;
        Xl[buf][btl][p][swz(r, ch)] = xs[q];
      }
    }
    bf16x8 w1, w2, w3, a1, a2;
    split_w8(ws[0], ws[1], w1, w2, w3, a1, a2);
    Wl[buf][0][wft][lane] = __builtin_bit_cast(uint4, w1);
    Wl[buf][1][wft][lane] = __builtin_bit_cast(uint4, w2);
    Wl[buf][2][wft][lane] = __builtin_bit_cast(uint4, w3);
  };

  floatx4 acc[4][R];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int p = 0; p < R; ++p) acc[f][p] = z4;

  load_stage(0);
  store_stage(0);
  __syncthreads();
  for (int kb = 0; kb < nkb; ++kb) {
    const int cur = kb & 1;
    if (kb + 1 < nkb) load_stage(kb + 1);
    bf16x8 xp[NPL];
#pragma unroll
    for (int p = 0; p < NPL; ++p) xp[p] = __builtin_bit_cast(bf16x8, Xl[cur][bt][p][swz(ln, lg)]);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int ftl = fg * 4 + f;
      const uint4 u1 = Wl[cur][0][ftl][lane], u2 = Wl[cur][1][ftl][lane], u3 = Wl[cur][2][ftl][lane];
      const bf16x8 w1 = __builtin_bit_cast(bf16x8, u1);
      const bf16x8 w2 = __builtin_bit_cast(bf16x8, u2);
      const bf16x8 w3 = __builtin_bit_cast(bf16x8, u3);
      const uint4 m = {0x7fff7fffu, 0x7fff7fffu, 0x7fff7fffu, 0x7fff7fffu};
      const bf16x8 wa1 = __builtin_bit_cast(bf16x8, uint4{u1.x & m.x, u1.y & m.y, u1.z & m.z, u1.w & m.w});
      const bf16x8 wa2 = __builtin_bit_cast(bf16x8, uint4{u2.x & m.x, u2.y & m.y, u2.z & m.z, u2.w & m.w});
      {
        floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[2], z4, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, xp[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3, xp[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, xp[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[0], c, 0, 0, 0);
        acc[f][0] += c;
      }
#pragma unroll
      for (int i = 0; i < V; ++i) {
        floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[4 + 2 * i], z4, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, xp[3 + 2 * i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[3 + 2 * i], c, 0, 0, 0);
        acc[f][1 + i] += c;
      }
      acc[f][NS] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa2, xp[NPL - 1], acc[f][NS], 0, 0, 0);
      acc[f][NS] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, xp[NPL - 1], acc[f][NS], 0, 0, 0);
    }
    if (kb + 1 < nkb) store_stage(cur ^ 1);
    __syncthreads();
  }

  // epilogue: rigorous rule, split, e plane; write the next layer's planes
  const float cb_next = Mp->L[l + 1].c_base;
  const float ca_next = Mp->L[l + 1].c_aff;
  const long long btg = bt0 + bt;
  const long long box = box0 + btg * 16 + ln;
  const bool live = (btg * 16 + ln) < nb && box < n;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int ftg = fgi * 8 + fg * 4 + f;       // global feature tile
    const int feat0 = ftg * 16 + lg * 4;
    const floatx4 bias4 = *reinterpret_cast<const floatx4*>(&Ly.bias[feat0]);
    float v[4][V + 2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j][0] = acc[f][0][j] + bias4[j];
#pragma unroll
      for (int i = 0; i < V; ++i) v[j][1 + i] = acc[f][1 + i][j];
      v[j][V + 1] = up1(acc[f][NS][j] * Ly.e_fix32);
      affine_rule_rigorous<V>(Ly.act, v[j]);
    }
    OutPlanes<V> o;
    split_out4<V>(v, cb_next, ca_next, o);
    if (!live) continue;
    const int kbo = feat0 >> 5, ko = feat0 & 31;
    unsigned short* B = Xout + blk_off(btg, kbo, nkb_out, NPL) + ln * 32 + ko;
#pragma unroll
    for (int q = 0; q < 3; ++q) *reinterpret_cast<uint2*>(B + q * 512) = uint2{o.x[q][0], o.x[q][1]};
#pragma unroll
    for (int i = 0; i < V; ++i) {
      *reinterpret_cast<uint2*>(B + (3 + 2 * i) * 512) = uint2{o.a[i][0][0], o.a[i][0][1]};
      *reinterpret_cast<uint2*>(B + (4 + 2 * i) * 512) = uint2{o.a[i][1][0], o.a[i][1][1]};
    }
    *reinterpret_cast<uint2*>(B + (NPL - 1) * 512) = uint2{o.e[0], o.e[1]};
  }
}

// The output layer (one output feature, out_pad 16): one wave per box tile,
// operands straight from HBM and L1, then the FP64 final test.
template <int V>
__global__ void __launch_bounds__(256)
k_lw_last(const DevModelBf16* __restrict__ Mp, int l, ClassifyArgs a, long long box0, long long nb,
          const unsigned short* __restrict__ Xin) {
  using C = Cfg<V>;
  constexpr int NPL = C::NPL;
  constexpr int NS = V + 1;
  const DevLayerBf16 Ly = Mp->L[l];
  const int nkb = Ly.in_pad >> 5;
  const int lane = threadIdx.x & 63;
  const int ln = lane & 15;
  const int lg = lane >> 4;
  const long long btg = blockIdx.x * 4LL + (threadIdx.x >> 6);
  const long long n = a.n_dev ? (long long)(*a.n_dev) : a.n;
  if (btg * 16 >= nb || box0 + btg * 16 >= n) return;
  const floatx4 z4 = {0.f, 0.f, 0.f, 0.f};
  floatx4 acc[NS + 1];
#pragma unroll
  for (int p = 0; p <= NS; ++p) acc[p] = z4;
  for (int kb = 0; kb < nkb; ++kb) {
    const floatx4* P = Ly.W32 + 2 * (size_t)kb * 64 + lane;
    bf16x8 w1, w2, w3, wa1, wa2;
    split_w8(P[0], P[64], w1, w2, w3, wa1, wa2);
    const unsigned short* B = Xin + blk_off(btg, kb, nkb, NPL) + ln * 32 + lg * 8;
    bf16x8 xp[NPL];
#pragma unroll
    for (int p = 0; p < NPL; ++p) xp[p] = *reinterpret_cast<const bf16x8*>(B + p * 512);
    {
      floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[2], z4, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, xp[1], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3, xp[0], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[1], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, xp[0], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[0], c, 0, 0, 0);
      acc[0] += c;
    }
#pragma unroll
    for (int i = 0; i < V; ++i) {
      floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[4 + 2 * i], z4, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2, xp[3 + 2 * i], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xp[3 + 2 * i], c, 0, 0, 0);
      acc[1 + i] += c;
    }
    acc[NS] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa2, xp[NPL - 1], acc[NS], 0, 0, 0);
    acc[NS] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, xp[NPL - 1], acc[NS], 0, 0, 0);
  }
  // output feature 0 sits in lanes 0..15 (lg == 0), slot j == 0
  if (lg != 0) return;
  const long long i = box0 + btg * 16 + ln;
  if (btg * 16 + ln >= nb || i >= n) return;
  float v[V + 2];
  v[0] = acc[0][0] + Ly.bias[0];
#pragma unroll
  for (int q = 0; q < V; ++q) v[1 + q] = acc[1 + q][0];
  v[V + 1] = up1(acc[NS][0] * Ly.e_fix32);
  double lo, hi, scale;
  final_bounds_fp64<V>(v, lo, hi, scale);
  a.codes[i] = sign_code<double>(lo, hi, (double)a.offset);
  if (a.bounds) {
    float flo = (float)lo, fhi = (float)hi;
    if ((double)flo > lo) flo = nextafterf(flo, -INFINITY);
    if ((double)fhi < hi) fhi = nextafterf(fhi, INFINITY);
    a.bounds[2 * i] = flo;
    a.bounds[2 * i + 1] = fhi;
  }
}

}  // namespace lw
