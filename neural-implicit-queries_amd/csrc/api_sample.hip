// Level-set sampling and bulk-property sums (src/kd_tree.py:220-336, 804-863).
#include "niq_internal.h"
#include "niq_sample.h"

namespace {

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

struct SampleWs {
  unsigned char* flags;
  int* bcnt;
  long long* bofs;
  SampCounters* C;
};

bool carve(void* ws, size_t bytes, long long n_per_round, SampleWs& w) {
  const long long nb = (n_per_round + kSampBlock - 1) / kSampBlock;
  char* p = static_cast<char*>(ws);
  size_t off = 0;
  w.flags = reinterpret_cast<unsigned char*>(p + off); off += align256((size_t)n_per_round);
  w.bcnt = reinterpret_cast<int*>(p + off); off += align256(sizeof(int) * nb);
  w.bofs = reinterpret_cast<long long*>(p + off); off += align256(sizeof(long long) * nb);
  w.C = reinterpret_cast<SampCounters*>(p + off); off += align256(sizeof(SampCounters));
  return off <= bytes;
}

// AS: the activation set of the kernel (ReLU models take the ReLU-only one).
template <int W, bool MASS, int AS = ACTS_ALL>
niq_status launch_sample_eval(const niq_model* m, const SampleSrc& S, long long n, float width, unsigned char* flags,
                              Sum4* partial, int max_blocks, int* blocks_used, hipStream_t s) {
  if constexpr (AS == ACTS_ALL) {
    if (m->relu_only)
      return launch_sample_eval<W, MASS, ACTS_RELU>(m, S, n, width, flags, partial, max_blocks, blocks_used, s);
  }
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_sample_eval<NB, NW, W, MASS, AS>;
  const long long tiles = (n + NB - 1) / NB;
  unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  if (max_blocks > 0) g = std::min<unsigned>(g, (unsigned)max_blocks);
  if (blocks_used) *blocks_used = (int)g;
  if (tiles == 0) return NIQ_OK;
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, S, n, width, flags, partial);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

template <bool MASS>
niq_status sample_eval(const niq_model* m, const SampleSrc& S, long long n, float width, unsigned char* flags,
                       Sum4* partial, int max_blocks, int* blocks_used, hipStream_t s) {
  switch (m->wt) {
    case 64: return launch_sample_eval<64, MASS>(m, S, n, width, flags, partial, max_blocks, blocks_used, s);
    case 128: return launch_sample_eval<128, MASS>(m, S, n, width, flags, partial, max_blocks, blocks_used, s);
    case 256: return launch_sample_eval<256, MASS>(m, S, n, width, flags, partial, max_blocks, blocks_used, s);
    case 512: return launch_sample_eval<512, MASS>(m, S, n, width, flags, partial, max_blocks, blocks_used, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

SampleSrc make_src(const float* lower, const float* upper, long long n_nodes, uint64_t seed, unsigned tag) {
  SampleSrc S{};
  S.lower = lower; S.upper = upper; S.n_nodes = n_nodes;
  S.key0 = (unsigned)(seed & 0xffffffffull);
  S.key1 = (unsigned)(seed >> 32);
  S.tag = tag;
  S.j0 = 0;
  return S;
}

constexpr int kMassMaxBlocks = 4096;   // partials of the mass pass
constexpr int kBoxSumBlocks = 256;

}  // namespace

extern "C" size_t niq_sample_workspace_bytes(int64_t n_per_round) {
  if (n_per_round < 1) n_per_round = 1;
  const long long nb = (n_per_round + kSampBlock - 1) / kSampBlock;
  return align256((size_t)n_per_round) + align256(sizeof(int) * nb) + align256(sizeof(long long) * nb) +
         align256(sizeof(SampCounters));
}

extern "C" niq_status niq_sample_surface_nodes(const niq_model* m, const float* node_lower, const float* node_upper,
                                               int64_t n_nodes, float width, uint64_t seed,
                                               int64_t first_candidate, int64_t n_per_round, int64_t max_rounds,
                                               float* out_points, int64_t n_samples, int64_t* n_found,
                                               int64_t* n_rounds, void* workspace, size_t workspace_bytes,
                                               void* stream) {
  if (niq_status st = check_model(m, 3)) return st;
  if (!n_found || !n_rounds) return fail(NIQ_ERR_INVALID, "null count argument");
  *n_rounds = 0;
  if (n_samples < 0 || *n_found < 0 || *n_found > n_samples)
    return fail(NIQ_ERR_INVALID, "need 0 <= n_found <= n_samples");
  if (*n_found == n_samples) return NIQ_OK;
  if (n_nodes < 1 || n_nodes >= (1LL << 31)) return fail(NIQ_ERR_INVALID, "need 1 <= n_nodes < 2^31");
  if (n_per_round < 1 || n_per_round > (1LL << 30)) return fail(NIQ_ERR_INVALID, "need 1 <= n_per_round <= 2^30");
  if (max_rounds < 1) return fail(NIQ_ERR_INVALID, "max_rounds must be >= 1");
  if (first_candidate < 0) return fail(NIQ_ERR_INVALID, "first_candidate must be >= 0");
  if (!node_lower || !node_upper || !out_points || !workspace) return fail(NIQ_ERR_INVALID, "null buffer");
  if (!(width == width)) return fail(NIQ_ERR_INVALID, "width is NaN");
  SampleWs w{};
  if (!carve(workspace, workspace_bytes, n_per_round, w)) return fail(NIQ_ERR_INVALID, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  SampCounters h{*n_found, *n_found, n_samples};
  HIPCHK(hipMemcpyAsync(w.C, &h, sizeof(h), hipMemcpyHostToDevice, s));
  SampleSrc S = make_src(node_lower, node_upper, n_nodes, seed, kTagSurface);
  const int nb = (int)((n_per_round + kSampBlock - 1) / kSampBlock);
  for (int64_t r = 0; r < max_rounds; ++r) {
    S.j0 = (unsigned long long)first_candidate + (unsigned long long)r * (unsigned long long)n_per_round;
    if (niq_status st = sample_eval<false>(m, S, n_per_round, width, w.flags, nullptr, 0, nullptr, s)) return st;
    hipLaunchKernelGGL(k_samp_count, dim3(nb), dim3(kSampBlock), 0, s, w.flags, (long long)n_per_round, w.bcnt);
    hipLaunchKernelGGL(k_samp_scan, dim3(1), dim3(kSampBlock), 0, s, w.bcnt, nb, w.bofs, w.C);
    hipLaunchKernelGGL(k_samp_scatter, dim3(nb), dim3(kSampBlock), 0, s, w.flags, (long long)n_per_round, w.bofs,
                       w.C, S, out_points);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&h, w.C, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *n_rounds = r + 1;
    *n_found = h.found;
    if (h.found >= n_samples) break;
  }
  return NIQ_OK;
}

extern "C" niq_status niq_bulk_sums_nodes(const niq_model* m, const float* unknown_lower, const float* unknown_upper,
                                          int64_t n_unknown, const float* interior_lower,
                                          const float* interior_upper, int64_t n_interior, int64_t n_samples,
                                          uint64_t seed, niq_bulk_sums* out, void* stream) {
  if (niq_status st = check_model(m, 3)) return st;
  if (!out) return fail(NIQ_ERR_INVALID, "null output");
  if (n_unknown < 0 || n_unknown >= (1LL << 31) || n_interior < 0)
    return fail(NIQ_ERR_INVALID, "need 0 <= n_unknown < 2^31 and n_interior >= 0");
  if (n_samples < 0) return fail(NIQ_ERR_INVALID, "n_samples must be >= 0");
  if ((n_unknown > 0 && (!unknown_lower || !unknown_upper)) || (n_interior > 0 && (!interior_lower || !interior_upper)))
    return fail(NIQ_ERR_INVALID, "null node list");
  hipStream_t s = (hipStream_t)stream;
  // partials: interior boxes, unknown boxes, mass pass; then the three sums
  const size_t n_part = 2 * kBoxSumBlocks + kMassMaxBlocks + 3;
  Sum4* part = nullptr;
  HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&part), sizeof(Sum4) * n_part, s));
  Sum4* p_int = part;
  Sum4* p_unk = part + kBoxSumBlocks;
  Sum4* p_mass = part + 2 * kBoxSumBlocks;
  Sum4* sums = part + 2 * kBoxSumBlocks + kMassMaxBlocks;
  niq_status st = NIQ_OK;
  hipLaunchKernelGGL(k_box_sums, dim3(kBoxSumBlocks), dim3(256), 0, s, interior_lower, interior_upper,
                     (long long)n_interior, p_int);
  hipLaunchKernelGGL(k_box_sums, dim3(kBoxSumBlocks), dim3(256), 0, s, unknown_lower, unknown_upper,
                     (long long)n_unknown, p_unk);
  int g = 0;
  const long long n_cand = n_unknown > 0 ? n_samples : 0;
  if (n_cand > 0) {
    SampleSrc S = make_src(unknown_lower, unknown_upper, n_unknown, seed, kTagMass);
    st = sample_eval<true>(m, S, n_cand, 0.f, nullptr, p_mass, kMassMaxBlocks, &g, s);
  }
  Sum4 h[3] = {};
  if (st == NIQ_OK) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, p_int, kBoxSumBlocks, sums);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, p_unk, kBoxSumBlocks, sums + 1);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, p_mass, g, sums + 2);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess)
      st = fail(NIQ_ERR_HIP, "bulk sums failed");
  }
  (void)hipFreeAsync(part, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, "bulk sums stream failed");
  if (st != NIQ_OK) return st;
  out->interior_volume = h[0].c;
  out->interior_moment[0] = h[0].x; out->interior_moment[1] = h[0].y; out->interior_moment[2] = h[0].z;
  out->unknown_volume = h[1].c;
  out->n_inside = h[2].c;
  out->inside_moment[0] = h[2].x; out->inside_moment[1] = h[2].y; out->inside_moment[2] = h[2].z;
  return NIQ_OK;
}
