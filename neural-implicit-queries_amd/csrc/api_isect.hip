// Search for any point inside two implicit functions at once
// (find_any_intersection, src/kd_tree.py:402-655): a breadth-first kd-tree
// walk in which each level classifies its nodes under both functions, samples
// 7 points around each centre, and keeps the nodes that could hold a point
// inside both; the first node of a level (in the reference's node order) that
// certifies one ends the search.
#include "niq_internal.h"

namespace {

constexpr int kIsBlock = 1024;   // nodes per compaction block
const int kNoneFound = INT_MAX;  // "no node found a point" marker for the level

struct IsectArgs {
  const float* lower;    // (n, 3)
  const float* upper;
  const int* codeA;
  const int* codeB;
  const float* valsA;    // (n, 7)
  const float* valsB;
  const float* pts;      // (n, 7, 3)
  float eps_w;
  int* found_first;      // min node index that found a point
  float* loc;            // (n, 3)
  int* subdiv;           // (n,)
  int* split_dim;        // (n,)
};

// The 7 samples of each node: centre + eps_w * {0, e_i, -e_i}.
__global__ void k_isect_points(const float* lower, const float* upper, long long n, float eps_w, float* pts) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = 0.5f * (lower[3 * i + k] + upper[3 * i + k]);
    for (int s = 0; s < 7; ++s)
      for (int k = 0; k < 3; ++k) {
        const float off = (s == 0) ? 0.f : ((s - 1) % 3 == k ? (s <= 3 ? 1.f : -1.f) : 0.f);
        pts[(i * 7 + s) * 3 + k] = c[k] + eps_w * off;
      }
  }
}

// process_node (src/kd_tree.py:445-516) for every node.
__global__ void k_isect_flags(IsectArgs a, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float w[3], width = -INFINITY;
    int sd = 0;
    for (int k = 0; k < 3; ++k) {
      w[k] = a.upper[3 * i + k] - a.lower[3 * i + k];
      if (w[k] > width) { width = w[k]; sd = k; }
    }
    const bool small = width < a.eps_w;
    const float* vA = a.valsA + 7 * i;
    const float* vB = a.valsB + 7 * i;
    bool allnA = true, allpA = true, allnB = true, allpB = true;
    int iA = -1, iB = -1, ib = -1;
    for (int s = 0; s < 7; ++s) {
      allnA = allnA && vA[s] < 0.f; allpA = allpA && vA[s] > 0.f;
      allnB = allnB && vB[s] < 0.f; allpB = allpB && vB[s] > 0.f;
      if (iA < 0 && vA[s] < 0.f) iA = s;
      if (iB < 0 && vB[s] < 0.f) iB = s;
      if (ib < 0 && vA[s] < 0.f && vB[s] < 0.f) ib = s;
    }
    const bool nearA = small && !(allnA || allpA);
    const bool nearB = small && !(allnB || allpB);
    bool found = false;
    float l[3] = {-777.f, -777.f, -777.f};
    const float* p = a.pts + 21 * i;
    if (small && iA >= 0 && iB >= 0) {
      found = true;
      for (int k = 0; k < 3; ++k) l[k] = 0.5f * (p[3 * iA + k] + p[3 * iB + k]);
    }
    if (ib >= 0) {
      found = true;
      for (int k = 0; k < 3; ++k) l[k] = p[3 * ib + k];
    }
    for (int k = 0; k < 3; ++k) a.loc[3 * i + k] = l[k];
    if (found) atomicMin(a.found_first, (int)i);
    const int cA = a.codeA[i], cB = a.codeB[i];
    const bool inA = cA == CODE_NEGATIVE || (cA == CODE_UNKNOWN && !nearA);
    const bool inB = cB == CODE_NEGATIVE || (cB == CODE_UNKNOWN && !nearB);
    a.subdiv[i] = (inA && inB) ? 1 : 0;
    a.split_dim[i] = sd;
  }
}

__global__ void k_isect_count(const int* subdiv, long long n, int* bcnt) {
  __shared__ int red[256];
  const long long b0 = (long long)blockIdx.x * kIsBlock;
  int c = 0;
  for (int j = threadIdx.x; j < kIsBlock; j += 256) {
    const long long i = b0 + j;
    if (i < n) c += subdiv[i];
  }
  red[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bcnt[blockIdx.x] = red[0];
}

// Exclusive scan of the block counts (one workgroup); total at boff[nb].
__global__ void k_isect_scan(const int* bcnt, int nb, int* boff) {
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    __shared__ int s[1024];
    const int j = base + threadIdx.x;
    const int v = j < nb ? bcnt[j] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (j < nb) boff[j] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) boff[nb] = carry;
}

// Stable compaction of the kept nodes and their children, interleaved A0, B0,
// A1, B1, ... (src/kd_tree.py:530-551).
__global__ void k_isect_scatter(const float* lower, const float* upper, const int* subdiv, const int* split_dim,
                                long long n, const int* boff, float* nlower, float* nupper) {
  __shared__ int s[kIsBlock];
  const long long b0 = (long long)blockIdx.x * kIsBlock;
  for (int j = threadIdx.x; j < kIsBlock; j += 1024) {
    const long long i = b0 + j;
    s[j] = (i < n) ? subdiv[i] : 0;
  }
  __syncthreads();
  for (int o = 1; o < kIsBlock; o <<= 1) {
    const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  const long long i = b0 + threadIdx.x;
  if (i >= n || !subdiv[i]) return;
  const long long r = (long long)boff[blockIdx.x] + s[threadIdx.x] - 1;
  const int sd = split_dim[i];
  for (int k = 0; k < 3; ++k) {
    const float lo = lower[3 * i + k], up = upper[3 * i + k];
    const float mid = 0.5f * (lo + up);
    nlower[(2 * r) * 3 + k] = lo;
    nupper[(2 * r) * 3 + k] = (k == sd) ? mid : up;
    nlower[(2 * r + 1) * 3 + k] = (k == sd) ? mid : lo;
    nupper[(2 * r + 1) * 3 + k] = up;
  }
}

}  // namespace

extern "C" niq_status niq_find_any_intersection(const niq_model* mA, const niq_eval_cfg* cfgA, const niq_model* mB,
                                                const niq_eval_cfg* cfgB, const float lower[3], const float upper[3],
                                                float eps, int32_t* found, float loc[3], int64_t* n_processed,
                                                void* stream) {
  if (niq_status st = check_model(mA, 3)) return st;
  if (niq_status st = check_model(mB, 3)) return st;
  if (!lower || !upper || !found || !loc) return fail(NIQ_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const float eps_w = eps / std::sqrt(3.0f);
  std::vector<void*> allocs;
  niq_status st = NIQ_OK;
  auto al = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (st == NIQ_OK && hipMallocAsync(&p, (bytes + 255) / 256 * 256, s) != hipSuccess)
      st = fail(NIQ_ERR_HIP, "intersection scratch allocation failed");
    if (p) allocs.push_back(p);
    return p;
  };
  // per-level buffers, grown on demand
  long long cap = 0;
  float *lo = nullptr, *up = nullptr, *nlo = nullptr, *nup = nullptr, *pts = nullptr, *vA = nullptr, *vB = nullptr;
  float* lc = nullptr;
  int *cA = nullptr, *cB = nullptr, *sub = nullptr, *sd = nullptr, *bcnt = nullptr, *boff = nullptr;
  int* first = static_cast<int*>(al(sizeof(int) * 64));
  auto grow = [&](long long need, bool keep_current) {
    if (need <= cap) return;
    long long nc = std::max<long long>(need, 2 * cap);
    float* olo = lo;
    float* oup = up;
    const long long ocap = cap;
    lo = static_cast<float*>(al(12 * nc)); up = static_cast<float*>(al(12 * nc));
    nlo = static_cast<float*>(al(12 * nc)); nup = static_cast<float*>(al(12 * nc));
    pts = static_cast<float*>(al(84 * nc));
    vA = static_cast<float*>(al(28 * nc)); vB = static_cast<float*>(al(28 * nc));
    lc = static_cast<float*>(al(12 * nc));
    cA = static_cast<int*>(al(4 * nc)); cB = static_cast<int*>(al(4 * nc));
    sub = static_cast<int*>(al(4 * nc)); sd = static_cast<int*>(al(4 * nc));
    const long long nbm = (nc + kIsBlock - 1) / kIsBlock;
    bcnt = static_cast<int*>(al(4 * nbm)); boff = static_cast<int*>(al(4 * (nbm + 1)));
    if (keep_current && st == NIQ_OK && ocap > 0) {
      if (hipMemcpyAsync(lo, olo, 12 * ocap, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(up, oup, 12 * ocap, hipMemcpyDeviceToDevice, s) != hipSuccess)
        st = fail(NIQ_ERR_HIP, "intersection buffer growth failed");
    }
    cap = nc;
  };
  grow(1024, false);
  if (st == NIQ_OK && (hipMemcpyAsync(lo, lower, 12, hipMemcpyHostToDevice, s) != hipSuccess ||
                       hipMemcpyAsync(up, upper, 12, hipMemcpyHostToDevice, s) != hipSuccess))
    st = fail(NIQ_ERR_HIP, "intersection set-up failed");
  long long n = 1, processed = 0;
  *found = 0;
  for (int k = 0; k < 3; ++k) loc[k] = -777.f;
  while (st == NIQ_OK) {
    processed += n;
    const int big = kNoneFound;
    if (hipMemcpyAsync(first, &kNoneFound, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
      st = fail(NIQ_ERR_HIP, "intersection counter reset failed");
      break;
    }
    ClassifyArgs a{};
    a.lower = lo; a.upper = up; a.d = 3; a.n = n; a.offset = 0.f;
    a.codes = cA;
    if ((st = classify_dispatch(mA, cfgA, 3, a, n, s))) break;
    a.codes = cB;
    if ((st = classify_dispatch(mB, cfgB, 3, a, n, s))) break;
    const unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_isect_points, dim3(g), dim3(256), 0, s, lo, up, n, eps_w, pts);
    if ((st = points_dev(mA, pts, nullptr, 1, 7 * n, vA, s))) break;
    if ((st = points_dev(mB, pts, nullptr, 1, 7 * n, vB, s))) break;
    IsectArgs ia{lo, up, cA, cB, vA, vB, pts, eps_w, first, lc, sub, sd};
    hipLaunchKernelGGL(k_isect_flags, dim3(g), dim3(256), 0, s, ia, n);
    const int nb = (int)((n + kIsBlock - 1) / kIsBlock);
    hipLaunchKernelGGL(k_isect_count, dim3(nb), dim3(256), 0, s, sub, n, bcnt);
    hipLaunchKernelGGL(k_isect_scan, dim3(1), dim3(1024), 0, s, bcnt, nb, boff);
    int hf = big, hn = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&hf, first, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&hn, boff + nb, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      st = fail(NIQ_ERR_HIP, "intersection level failed");
      break;
    }
    if (hf != big) {   // the first node of the level that found a point
      float hl[3];
      if (hipMemcpyAsync(hl, lc + 3 * (long long)hf, 12, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        st = fail(NIQ_ERR_HIP, "intersection location read failed");
        break;
      }
      *found = 1;
      for (int k = 0; k < 3; ++k) loc[k] = hl[k];
      break;
    }
    if (hn == 0) break;
    if (2LL * hn > (1LL << 31) / 21) { st = fail(NIQ_ERR_CAPACITY, "intersection search exceeds 100M nodes"); break; }
    // children of level l go to nlo/nup; grow first if they do not fit
    if (2LL * hn > cap) {
      // keep this level's nodes and flags: grow into fresh buffers, copying lo/up/sub/sd
      float* olo = lo;
      float* oup = up;
      int* osub = sub;
      int* osd = sd;
      int* oboff = boff;
      const long long on = n;
      grow(2LL * hn, true);
      if (st) break;
      if (hipMemcpyAsync(sub, osub, 4 * on, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(sd, osd, 4 * on, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(boff, oboff, 4 * (nb + 1), hipMemcpyDeviceToDevice, s) != hipSuccess) {
        st = fail(NIQ_ERR_HIP, "intersection buffer growth failed");
        break;
      }
      (void)olo; (void)oup;
    }
    hipLaunchKernelGGL(k_isect_scatter, dim3(nb), dim3(kIsBlock), 0, s, lo, up, sub, sd, n, boff, nlo, nup);
    std::swap(lo, nlo);
    std::swap(up, nup);
    n = 2LL * hn;
  }
  for (void* p : allocs) (void)hipFreeAsync(p, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, "intersection stream failed");
  if (n_processed) *n_processed = processed;
  return st;
}
