// Closest-point search and marching-cubes extraction entry points.
#include "niq_internal.h"
#include "niq_cp_mc.h"

// ---------------------------------------------------------- closest point
namespace {

template <int W, int AS = ACTS_ALL>
niq_status launch_points_dev_w(const niq_model* m, const float* x, const int* n_dev, int mult, long long max_n,
                               float* out, hipStream_t s);

// A closest-point round as two launches: k_cp_forward (box and point tiles of
// 16, grid sized for a full batch) and the one-workgroup update.
template <int KIND, int V, int W, int AS = ACTS_ALL>
niq_status launch_cp_round_w(const niq_model* m, CpState* S, const CpArgs& a, hipStream_t s) {
  if constexpr (!Fits<Planes<KIND, V, 0>::R, 16, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "closest_point round: box planes above LDS");
  } else {
    constexpr int NW = TileCfgWide<W>::NW;
    const long long g_box = (a.B + 15) / 16, g_pts = (7LL * a.B + 15) / 16;
    hipLaunchKernelGGL((k_cp_forward<KIND, V, NW, W, AS>), dim3((unsigned)(g_box + g_pts)), dim3(NW * 64), 0, s,
                       m->dev, S, a, (int)g_box);
    hipLaunchKernelGGL(k_cp_update, dim3(1), dim3(1024), 0, s, S, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int KIND, int V>
niq_status launch_cp_round_kv(const niq_model* m, CpState* S, const CpArgs& a, hipStream_t s) {
  if constexpr (KIND == 0) {
    // ReLU or TanH models: the round compiled for that activation alone
    // (bit-identical; fewer registers, so more workgroups per CU)
    if (m->relu_only) {
      switch (m->wt) {
        case 64: return launch_cp_round_w<KIND, V, 64, ACTS_RELU>(m, S, a, s);
        case 128: return launch_cp_round_w<KIND, V, 128, ACTS_RELU>(m, S, a, s);
        case 256: return launch_cp_round_w<KIND, V, 256, ACTS_RELU>(m, S, a, s);
        case 512: return launch_cp_round_w<KIND, V, 512, ACTS_RELU>(m, S, a, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
    if (m->tanh_only) {
      switch (m->wt) {
        case 64: return launch_cp_round_w<KIND, V, 64, ACTS_TANH>(m, S, a, s);
        case 128: return launch_cp_round_w<KIND, V, 128, ACTS_TANH>(m, S, a, s);
        case 256: return launch_cp_round_w<KIND, V, 256, ACTS_TANH>(m, S, a, s);
        case 512: return launch_cp_round_w<KIND, V, 512, ACTS_TANH>(m, S, a, s);
      }
      return fail(NIQ_ERR_UNSUPPORTED, "width");
    }
  }
  switch (m->wt) {
    case 64: return launch_cp_round_w<KIND, V, 64>(m, S, a, s);
    case 128: return launch_cp_round_w<KIND, V, 128>(m, S, a, s);
    case 256: return launch_cp_round_w<KIND, V, 256>(m, S, a, s);
    case 512: return launch_cp_round_w<KIND, V, 512>(m, S, a, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// The (KIND, V) form of the fused forward for this model and config, or -1:
// it covers the FP32 interval, affine_fixed and slope_interval classifiers at
// their default tile (variant 0) whose box planes fit LDS; other configs run
// the round as gather, classify, points and update launches.
int cp_round_form(const niq_model* m, const niq_eval_cfg* cfg) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  if (cfg && (cfg->precision != NIQ_PREC_FP32 || cfg->variant != 0)) return -1;
  const size_t plane = 16 * (size_t)m->wt * 4;
  if (mode == NIQ_MODE_INTERVAL) return 0;
  if (mode == NIQ_MODE_AFFINE_FIXED) return (5 * plane <= (size_t)kLdsLimit) ? 1 : -1;
  if (mode == NIQ_MODE_SLOPE_INTERVAL) return (7 * plane <= (size_t)kLdsLimit) ? 2 : -1;
  return -1;
}

niq_status launch_cp_round(int form, const niq_model* m, CpState* S, const CpArgs& a, hipStream_t s) {
  switch (form) {
    case 0: return launch_cp_round_kv<0, 0>(m, S, a, s);
    case 1: return launch_cp_round_kv<0, 3>(m, S, a, s);
    case 2: return launch_cp_round_kv<1, 3>(m, S, a, s);
  }
  return fail(NIQ_ERR_INVALID, "closest_point round form");
}

}  // namespace

extern "C" niq_status niq_closest_point(const niq_model* m, const niq_eval_cfg* cfg, const float lower[3],
                                        const float upper[3], const float* query_points, int64_t Q, float eps,
                                        int batch_process_size, float* out_dist, float* out_loc,
                                        int64_t* n_processed, void* stream) {
  if (niq_status st = check_model(m, 3)) return st;
  if (Q < 0 || Q > (1LL << 30)) return fail(NIQ_ERR_INVALID, "query count out of range");
  if (Q > 0 && (!query_points || !out_dist || !out_loc)) return fail(NIQ_ERR_INVALID, "bad buffers");
  const int B = batch_process_size;
  if (B < 1 || B > (1 << 20)) return fail(NIQ_ERR_INVALID, "batch_process_size out of range");
  if (Q == 0) return NIQ_OK;
  hipStream_t s = (hipStream_t)stream;
  const int K = cp_rounds_per_check(B);
  long long cap = cp_initial_stack(Q, B);
  // per-call device scratch (stream ordered); every allocation made is freed
  // on every path, the stack arrays (which may be replaced as they grow) at the
  // end through `a`
  niq_status st = NIQ_OK;
  std::vector<void*> owned;
  auto al = [&](size_t bytes, void** p) {
    *p = nullptr;
    const hipError_t e = hipMallocAsync(p, (bytes + 255) / 256 * 256, s);
    if (e != hipSuccess) *p = nullptr;
    return e;
  };
  auto own = [&](size_t bytes, void** p) {
    if (st != NIQ_OK) { *p = nullptr; return; }
    const hipError_t e = al(bytes, p);
    if (e != hipSuccess) st = fail(NIQ_ERR_HIP, std::string("closest_point: hipMallocAsync: ") + hipGetErrorString(e));
    else owned.push_back(*p);
  };
  auto hipfail = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  void *pS, *pbq, *pblo, *pbup, *ppts, *pcodes, *pvals, *pwin, *pidist, *pineeds;
  own(sizeof(CpState), &pS);
  own(sizeof(int) * B, &pbq);
  own(sizeof(float) * 3 * B, &pblo);
  own(sizeof(float) * 3 * B, &pbup);
  own(sizeof(float) * 21 * B, &ppts);
  own(sizeof(int) * B, &pcodes);
  own(sizeof(float) * 7 * B, &pvals);
  own(sizeof(int) * Q, &pwin);
  own(sizeof(float) * B, &pidist);
  own(sizeof(int) * B, &pineeds);
  CpArgs a{};
  void *pwq = nullptr, *pwlo = nullptr, *pwup = nullptr;
  if (st == NIQ_OK) hipfail(al(sizeof(int) * cap, &pwq), "closest_point: hipMallocAsync");
  if (st == NIQ_OK) hipfail(al(sizeof(float) * 3 * cap, &pwlo), "closest_point: hipMallocAsync");
  if (st == NIQ_OK) hipfail(al(sizeof(float) * 3 * cap, &pwup), "closest_point: hipMallocAsync");
  a.wq = (int*)pwq; a.wlo = (float*)pwlo; a.wup = (float*)pwup;
  auto release = [&]() {
    for (void* p : owned) (void)hipFreeAsync(p, s);
    for (void* p : {(void*)a.wq, (void*)a.wlo, (void*)a.wup})
      if (p) (void)hipFreeAsync(p, s);
  };
  if (st != NIQ_OK) {
    release();
    return st;
  }
  a.B = B; a.cap = cap;
  a.bq = (int*)pbq; a.blo = (float*)pblo; a.bup = (float*)pbup; a.pts = (float*)ppts;
  a.codes = (int*)pcodes; a.vals = (float*)pvals; a.q = query_points; a.Q = Q;
  a.qmin = out_dist; a.qloc = out_loc; a.win = (int*)pwin;
  a.idist = (float*)pidist; a.ineeds = (int*)pineeds;
  // eps / sqrt(d) in float32 (src/kd_tree.py:688)
  a.eps_cube_width = eps / sqrtf(3.0f);
  CpState* S = (CpState*)pS;
  {
    CpState h{};
    h.top = (int)Q;
    std::vector<int> ids(Q);
    for (long long i = 0; i < Q; ++i) ids[i] = (int)i;
    std::vector<float> blo(3 * Q), bup(3 * Q), inf(Q, INFINITY), loc(3 * Q, -777.f);
    for (long long i = 0; i < Q; ++i)
      for (int k = 0; k < 3; ++k) { blo[3 * i + k] = lower[k]; bup[3 * i + k] = upper[k]; }
    hipfail(hipMemcpyAsync(S, &h, sizeof(h), hipMemcpyHostToDevice, s), "closest_point: set-up copy");
    if (st == NIQ_OK) hipfail(hipMemcpyAsync(a.wq, ids.data(), sizeof(int) * Q, hipMemcpyHostToDevice, s), "closest_point: set-up copy");
    if (st == NIQ_OK) hipfail(hipMemcpyAsync(a.wlo, blo.data(), sizeof(float) * 3 * Q, hipMemcpyHostToDevice, s), "closest_point: set-up copy");
    if (st == NIQ_OK) hipfail(hipMemcpyAsync(a.wup, bup.data(), sizeof(float) * 3 * Q, hipMemcpyHostToDevice, s), "closest_point: set-up copy");
    if (st == NIQ_OK) hipfail(hipMemcpyAsync(out_dist, inf.data(), sizeof(float) * Q, hipMemcpyHostToDevice, s), "closest_point: set-up copy");
    if (st == NIQ_OK) hipfail(hipMemcpyAsync(out_loc, loc.data(), sizeof(float) * 3 * Q, hipMemcpyHostToDevice, s), "closest_point: set-up copy");
    if (st == NIQ_OK) hipfail(hipMemsetAsync(a.win, 0xff, sizeof(int) * Q, s), "closest_point: set-up memset");
    // the host vectors and h go out of scope here: wait for their copies on
    // every path
    hipfail(hipStreamSynchronize(s), "closest_point: hipStreamSynchronize");
  }
  if (st != NIQ_OK) {
    release();
    return st;
  }
  ClassifyArgs ca{};
  ca.lower = a.blo; ca.upper = a.bup; ca.d = 3; ca.n_dev = &S->n; ca.offset = 0.f; ca.codes = a.codes;
  const unsigned gg = (unsigned)std::min<long long>((B + 255) / 256, 1024);
  const int form = cp_round_form(m, cfg);
  while (true) {
    for (int r = 0; r < K && st == NIQ_OK; ++r) {
      if (form >= 0) {
        st = launch_cp_round(form, m, S, a, s);
        continue;
      }
      hipLaunchKernelGGL(k_cp_gather, dim3(gg), dim3(256), 0, s, S, a);
      if ((st = classify_dispatch(m, cfg, 3, ca, B, s))) break;
      switch (m->wt) {
        case 64: st = launch_points_dev_w<64>(m, a.pts, &S->n, 7, 7LL * B, a.vals, s); break;
        case 128: st = launch_points_dev_w<128>(m, a.pts, &S->n, 7, 7LL * B, a.vals, s); break;
        case 256: st = launch_points_dev_w<256>(m, a.pts, &S->n, 7, 7LL * B, a.vals, s); break;
        default: st = launch_points_dev_w<512>(m, a.pts, &S->n, 7, 7LL * B, a.vals, s); break;
      }
      if (st) break;
      hipLaunchKernelGGL(k_cp_update, dim3(1), dim3(1024), 0, s, S, a);
    }
    if (st) break;
    CpState h;
    if (hipMemcpyAsync(&h, S, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      st = fail(NIQ_ERR_HIP, "closest_point: device state read failed");
      break;
    }
    if (h.overflow) { st = fail(NIQ_ERR_CAPACITY, "closest_point stack overflow"); break; }
    if (h.top == 0) {
      if (n_processed) *n_processed = h.n_processed;
      break;
    }
    if (h.top + 2LL * K * B > cap) {   // grow the stack before the next rounds
      const long long ncap = std::max(2 * cap, (long long)h.top + 4LL * K * B);
      void *nq = nullptr, *nlo = nullptr, *nup = nullptr;
      if (al(sizeof(int) * ncap, &nq) || al(sizeof(float) * 3 * ncap, &nlo) || al(sizeof(float) * 3 * ncap, &nup)) {
        for (void* p : {nq, nlo, nup})
          if (p) (void)hipFreeAsync(p, s);
        st = fail(NIQ_ERR_HIP, "closest_point: stack growth failed");
        break;
      }
      (void)hipMemcpyAsync(nq, a.wq, sizeof(int) * h.top, hipMemcpyDeviceToDevice, s);
      (void)hipMemcpyAsync(nlo, a.wlo, sizeof(float) * 3 * h.top, hipMemcpyDeviceToDevice, s);
      (void)hipMemcpyAsync(nup, a.wup, sizeof(float) * 3 * h.top, hipMemcpyDeviceToDevice, s);
      (void)hipFreeAsync(a.wq, s); (void)hipFreeAsync(a.wlo, s); (void)hipFreeAsync(a.wup, s);
      a.wq = (int*)nq; a.wlo = (float*)nlo; a.wup = (float*)nup; a.cap = ncap; cap = ncap;
    }
  }
  release();
  if (st == NIQ_OK) HIPCHK(hipGetLastError());
  return st;
}

// -------------------------------------------------------- marching cubes
namespace {

// Bourke's marching-cubes triangle table in packed form: per case, one hex
// digit per edge id, three per triangle.
const char* const kTriHex[256] = {
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98",
    "3b2", "0b28b0", "19023b", "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b",
    "478", "430734", "019847", "419471731", "12a847", "34730412a", "92a902847", "2a9297273794",
    "8473b2", "b47b24204", "90184723b", "47b94b9b2921", "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba",
    "954", "954083", "054150", "854835315", "12a954", "30812a495", "52a542402", "2a5325354348",
    "95423b", "0b208b495", "05401523b", "21525828b485", "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b",
    "978579", "930953573", "078017157", "153357", "978957a12", "a12950530573", "802825857a52", "2a5253357",
    "7957893b2", "95797292027b", "23b018178157", "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5",
    "a65", "0835a6", "9015a6", "1831985a6", "165261", "165126308", "965906026", "598582526328",
    "23ba65", "b08b20a65", "01923b5a6", "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98",
    "5a6478", "43047365a", "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269",
    "3b2784a65", "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847", "65969b4797b9",
    "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264", "024426", "832824426",
    "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63", "8b1810b61914641", "3b6360064", "648b68",
    "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173", "126168189867", "269291679093739", "780706602", "732672",
    "23ba68a89867", "20727b09767a9a7", "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6",
    "76b", "308b76", "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98",
    "723627", "708760620", "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9",
    "684b86", "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93", "a946a4",
    "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495", "76b54a42a402", "348354325a52b76",
    "723762549", "954086062687", "362376150540", "628687218485158", "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a",
    "6956b9b89", "36b063056095", "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a", "a56",
    "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7", "75272b592328982",
    "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175", "903935537", "987597",
    "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458", "04b0b345b2b151b", "0250592b5458b85", "9452b3",
    "25a352345384", "5a2524420", "3a235a385458019", "5a2524192942", "845853351", "045105", "845853905035", "945",
    "4b749b9ab", "0834979b79ab", "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743", "487",
    "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", "",
};

McTable* device_mc_table(hipStream_t s, niq_status* st) {
  static std::mutex mu;
  static std::unordered_map<int, McTable*> tables;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto it = tables.find(dev);
  if (it != tables.end()) return it->second;
  McTable h;
  std::memset(&h, -1, sizeof(h));
  for (int c = 0; c < 256; ++c) {
    const char* p = kTriHex[c];
    for (int i = 0; p[i] && i < 15; ++i) h.tri[c][i] = (signed char)(p[i] <= '9' ? p[i] - '0' : p[i] - 'a' + 10);
  }
  McTable* d = nullptr;
  if (hipMalloc(&d, sizeof(McTable)) != hipSuccess || hipMemcpy(d, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) {
    *st = fail(NIQ_ERR_HIP, "marching-cubes table upload failed");
    return nullptr;
  }
  tables[dev] = d;
  (void)s;
  return d;
}

}  // namespace

extern "C" int niq_mc_table(signed char* out) {
  if (!out) return NIQ_ERR_INVALID;
  std::memset(out, -1, 256 * 16);
  for (int c = 0; c < 256; ++c) {
    const char* p = kTriHex[c];
    for (int i = 0; p[i] && i < 15; ++i) out[c * 16 + i] = (signed char)(p[i] <= '9' ? p[i] - '0' : p[i] - 'a' + 10);
  }
  return NIQ_OK;
}

extern "C" niq_status niq_mc_extract(const niq_model* m, const float* leaf_lower, const float* leaf_upper,
                                     int64_t n_leaves, int n_subcell_depth, float* tri_out, int64_t tri_capacity,
                                     int64_t* n_tris, void* stream) {
  if (niq_status st = check_model(m, 3)) return st;
  if (n_subcell_depth < 0 || n_subcell_depth > 6) return fail(NIQ_ERR_INVALID, "n_subcell_depth must be in [0, 6]");
  if (n_leaves < 0 || (n_leaves > 0 && (!leaf_lower || !leaf_upper))) return fail(NIQ_ERR_INVALID, "bad buffers");
  if (!n_tris) return fail(NIQ_ERR_INVALID, "null n_tris");
  hipStream_t s = (hipStream_t)stream;
  *n_tris = 0;
  if (n_leaves == 0) return NIQ_OK;
  niq_status st = NIQ_OK;
  McTable* T = device_mc_table(s, &st);
  if (!T) return st;
  const int sd = n_subcell_depth;
  const long long np3 = (long long)((1 << sd) + 1) * ((1 << sd) + 1) * ((1 << sd) + 1);
  const long long chunk = std::max<long long>(1, std::min<long long>(n_leaves, (64LL << 20) / np3));
  void *ppts = nullptr, *pvals = nullptr, *pcnt = nullptr, *poff = nullptr, *ptot = nullptr;
  const std::pair<void**, size_t> want[] = {{&ppts, sizeof(float) * 3 * np3 * chunk},
                                            {&pvals, sizeof(float) * np3 * chunk},
                                            {&pcnt, sizeof(long long) * chunk},
                                            {&poff, sizeof(long long) * chunk},
                                            {&ptot, sizeof(long long)}};
  for (const auto& w : want)
    if (st == NIQ_OK && hipMallocAsync(w.first, w.second, s) != hipSuccess)
      st = fail(NIQ_ERR_HIP, "mc extract: scratch allocation failed");
  long long written = 0;
  for (long long l0 = 0; l0 < n_leaves && st == NIQ_OK; l0 += chunk) {
    const long long nl = std::min(chunk, n_leaves - l0);
    const long long npts = nl * np3;
    hipLaunchKernelGGL(k_mc_grid, dim3((unsigned)std::min<long long>((npts + 255) / 256, 65535)), dim3(256), 0, s,
                       leaf_lower, leaf_upper, l0, nl, sd, (float*)ppts);
    if ((st = niq_eval_points(m, (const float*)ppts, npts, 3, (float*)pvals, s))) break;
    const unsigned gl = (unsigned)std::min<long long>(nl, 65535);
    hipLaunchKernelGGL(k_mc_count, dim3(gl), dim3(256), 0, s, T, (const float*)pvals, nl, sd, (long long*)pcnt);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, s, (const long long*)pcnt, nl, written,
                       (long long*)poff, (long long*)ptot);
    hipLaunchKernelGGL(k_mc_emit, dim3(gl), dim3(256), 0, s, T, (const float*)pvals, leaf_lower, leaf_upper, l0, nl,
                       sd, (const long long*)poff, tri_out, (long long)tri_capacity);
    long long tot = 0;
    if (hipMemcpyAsync(&tot, ptot, sizeof(long long), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      st = fail(NIQ_ERR_HIP, "mc extract failed");
      break;
    }
    written = tot;
  }
  for (void* p : {ppts, pvals, pcnt, poff, ptot})
    if (p) (void)hipFreeAsync(p, s);
  if (st) return st;
  HIPCHK(hipGetLastError());
  *n_tris = written;
  if (written > tri_capacity) return fail(NIQ_ERR_CAPACITY, "triangle buffer too small: need " + std::to_string(written));
  return NIQ_OK;
}

namespace {
// AS: the activation set of the kernel (ReLU models take the ReLU-only one).
template <int W, int AS>
niq_status launch_points_dev_w(const niq_model* m, const float* x, const int* n_dev, int mult, long long max_n,
                               float* out, hipStream_t s) {
  if constexpr (AS == ACTS_ALL) {
    if (m->relu_only) return launch_points_dev_w<W, ACTS_RELU>(m, x, n_dev, mult, max_n, out, s);
  }
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_points_dev<NB, NW, W, AS>;
  const long long tiles = (max_n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, x, n_dev, mult, max_n, out);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}
}  // namespace

niq_status niqi::points_dev(const niq_model* m, const float* x, const int* n_dev, int mult, long long max_n,
                            float* out, hipStream_t s) {
  switch (m->wt) {
    case 64: return launch_points_dev_w<64>(m, x, n_dev, mult, max_n, out, s);
    case 128: return launch_points_dev_w<128>(m, x, n_dev, mult, max_n, out, s);
    case 256: return launch_points_dev_w<256>(m, x, n_dev, mult, max_n, out, s);
    case 512: return launch_points_dev_w<512>(m, x, n_dev, mult, max_n, out, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

