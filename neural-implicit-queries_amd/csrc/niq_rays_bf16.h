#pragma once
// Persistent certified-step ray caster in precision mode: the step's segment
// box is bounded by the rigorous BF16-MFMA range form (FP64 final test), and
// the two hit-test points ride beside it as plain rows of the same split
// products.  Slot handling, groups, stepping and termination are those of
// k_cast_rays (src/queries.py:39-175).
#include "niq_bf16.h"
#include "niq_kernels.h"

// Point rows split three ways when the LDS holds them, else two ways.
template <int V, int NB, int WMAX>
struct RayPlanesBf16 {
  static constexpr int NPL = PlanesBf16<V>::NPL;
  static constexpr int PS = (size_t)(NPL + 6) * NB * WMAX * 2 <= (size_t)kLdsLimit ? 3 : 2;
  static constexpr int NPT = NPL + 2 * PS;
  static constexpr bool fits = (size_t)NPT * NB * WMAX * 2 <= (size_t)kLdsLimit;
};

template <int V, int NB, int NW, int WMAX, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64)
k_cast_rays_bf16(const DevModelBf16* __restrict__ models, RayArgs a) {
  using RP = RayPlanesBf16<V, NB, WMAX>;
  constexpr int NPL = RP::NPL;
  constexpr int PS = RP::PS;
  constexpr int NT = NW * 64;
  __shared__ __attribute__((aligned(16))) unsigned short Xs[RP::NPT * NB * WMAX];
  float* Xf = reinterpret_cast<float*>(Xs);
  // the refill's 64-bit base is passed through the first 8 bytes of the planes
  long long* s_base = reinterpret_cast<long long*>(Xs);
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

    bool is_hit = false;
    int hit_id = 0, nsteps = 0;
    for (int sub = 0; sub < G.n_substeps; ++sub) {
      bool can_step = !is_hit;
      nsteps += is_hit ? 0 : 1;
      for (int f = 0; f < G.n_models; ++f) {
        const DevModelBf16& M = models[G.model0 + f];
        __syncthreads();   // previous readers of Xs are done
        // layer-0 inputs: features 0..2 from the slot, 3..31 zero
        for (int e = threadIdx.x; e < RP::NPT * NB * 32; e += NT) {
          const int k = e & 31;
          if (k >= 3) Xs[boff16<NB, WMAX>(e / (NB * 32), (e / 32) % NB, k)] = 0;
        }
        if (threadIdx.x < NB) {
          const int b = threadIdx.x;
          const bool live = rid >= 0;
          const float hs = 0.5f * step;
          for (int k = 0; k < 3; ++k) {
            const float ps = ro[k] + t * di[k];
            const float hv = hs * di[k];
            const float pm = ps + hv;
            const float pe = ro[k] + (t + G.hit_eps) * di[k];
            float s[V + 1];
            float err = 0.f;
            s[0] = live ? pm : 0.f;
            if constexpr (V == 0) err = live ? fabsf(hv) : 0.f;
            else s[1] = live ? hv : 0.f;
            store_split<V, NB, WMAX>(Xs, b, k, s, err, M.L[0].c_base, M.L[0].c_aff);
            const float pts[2] = {live ? ps : 0.f, live ? pe : 0.f};
#pragma unroll
            for (int r = 0; r < 2; ++r) {
              const unsigned short h1 = bf16_bits_rn(pts[r]);
              const float r1 = pts[r] - bf16_to_f(h1);
              const unsigned short h2 = bf16_bits_rn(r1);
              Xs[boff16<NB, WMAX>(NPL + r * PS, b, k)] = h1;
              Xs[boff16<NB, WMAX>(NPL + r * PS + 1, b, k)] = h2;
              if constexpr (PS == 3) Xs[boff16<NB, WMAX>(NPL + r * PS + 2, b, k)] = bf16_bits_rn(r1 - bf16_to_f(h2));
            }
          }
        }
        mlp_forward_bf16<V, NB, NW, WMAX, WSRC_FP32, 2, PS, AS>(M, Xs, [&](const float* v, int b) {
          double lo, hi, scale;
          final_bounds_fp64<V>(v, lo, hi, scale);
          Xf[4 * b + 0] = __int_as_float(sign_code<double>(lo, hi, 0.0));
          Xf[4 * b + 1] = v[V + 2];
          Xf[4 * b + 2] = v[V + 3];
        });
        __syncthreads();
        if (slot_lane) {
          const int code = __float_as_int(Xf[4 * lane]);
          const float v0 = Xf[4 * lane + 1], v1 = Xf[4 * lane + 2];
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

    if (slot_lane && rid >= 0) {
      cnt += nsteps;
      ++iters;
      const bool term = is_hit || (t > G.max_dist) || (cnt >= G.n_max_step);
      if (term) {
        if (threadIdx.x < 64) {
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
