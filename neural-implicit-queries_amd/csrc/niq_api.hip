// C ABI of the gfx950 range-analysis hot path (declarations in include/niq.h).
#include "niq.h"
#include "niq_kernels.h"
#include "niq_bf16.h"
#include "niq_cp_mc.h"
#include "niq_rows.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

niq_status fail(niq_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(NIQ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

inline int round16(int x) { return (x + 15) / 16 * 16; }
inline int round32(int x) { return (x + 31) / 32 * 32; }

// host bf16 conversions (same bit rules as the device)
inline unsigned short h_bf16_rn(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
inline float h_bf16_f(unsigned short h) {
  unsigned u = (unsigned)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline unsigned short h_bf16_up_nonneg(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  unsigned short h = (unsigned short)(u >> 16);
  if ((u & 0xffffu) && (u & 0x7fffffffu) < 0x7f800000u) h = (unsigned short)(h + 1);
  return h;
}
// gamma_n(u) = n u / (1 - n u), rounded up generously
inline double gamma_n(double n, double u) { return n * u / (1.0 - n * u) * (1.0 + 1e-6); }

// Template width that holds every padded layer width of a model.
inline int template_width(int wmax) {
  if (wmax <= 64) return 64;
  if (wmax <= 128) return 128;
  if (wmax <= 256) return 256;
  if (wmax <= 512) return 512;
  return -1;
}

int num_cus() {
  static int cus = -1;
  if (cus < 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
  }
  return cus;
}

// Persistent grid: resident workgroups of kernel `f` over the whole device.
template <class F>
long long persistent_blocks(F f, int threads) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find((const void*)f);
  int occ = 0;
  if (it == cache.end()) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, threads, 0) != hipSuccess || occ < 1) occ = 1;
    cache[(const void*)f] = occ;
  } else {
    occ = it->second;
  }
  return (long long)occ * num_cus();
}

inline unsigned grid_for(long long work_blocks, long long persist) {
  long long g = std::min(work_blocks, persist);
  return (unsigned)std::max(1LL, g);
}

}  // namespace

struct niq_model {
  int n_layers = 0;
  int in_dim = 0;
  int wt = 0;                  // template width
  int wt16 = 0;                // template width of the precision-mode copy
  DevModelBf16 host16;         // precision-mode copy (bf16 hi/lo/|W| weights)
  DevModelBf16* dev16 = nullptr;
  void* weights16 = nullptr;
  std::vector<int> dims, acts;
  DevModel host;               // with device pointers
  DevModel* dev = nullptr;     // device copy of `host`
  void* weights = nullptr;     // packed weights and biases
  size_t bytes = 0;
};

extern "C" {

const char* niq_last_error(void) { return g_err.c_str(); }
int niq_version(void) { return 1; }

niq_status niq_model_create(int n_layers, const int* dims, const int* acts, const float* const* W,
                            const float* const* b, niq_model** out) {
  if (!out || !dims || !acts || !W) return fail(NIQ_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_layers < 1 || n_layers > NIQ_MAX_LAYERS)
    return fail(NIQ_ERR_INVALID, "n_layers must be in [1, " + std::to_string(NIQ_MAX_LAYERS) + "]");
  if (dims[0] < 1 || dims[0] > 16) return fail(NIQ_ERR_INVALID, "input dimension must be in [1, 16]");
  if (dims[n_layers] != 1) return fail(NIQ_ERR_INVALID, "the last layer must have one output (squeeze_last)");
  int wmax = 16;
  for (int i = 0; i <= n_layers; ++i) {
    if (dims[i] < 1) return fail(NIQ_ERR_INVALID, "layer widths must be positive");
    wmax = std::max(wmax, round16(dims[i]));
  }
  for (int i = 0; i < n_layers; ++i)
    if (acts[i] < NIQ_ACT_NONE || acts[i] > NIQ_ACT_TANH) return fail(NIQ_ERR_INVALID, "unknown activation");
  const int wt = template_width(wmax);
  if (wt < 0) return fail(NIQ_ERR_UNSUPPORTED, "layer widths above 512 are not supported");

  // Pack each layer's W (in, out) into MFMA A-fragments: fragment (ft, kb) is
  // 64 lanes x 4 floats, lane l / slot j holding W[16 kb + 4 (l / 16) + j][16 ft + l % 16].
  std::vector<size_t> woff(n_layers), poff(n_layers), boff(n_layers);
  size_t total = 0;
  for (int i = 0; i < n_layers; ++i) {
    const size_t ip = round16(dims[i]), op = round16(dims[i + 1]);
    woff[i] = total;
    total += ip * op * sizeof(float);
    poff[i] = total;
    total += ip * op * sizeof(float);
    boff[i] = total;
    total += op * sizeof(float);
    total = (total + 255) / 256 * 256;
  }
  std::vector<unsigned char> host(total, 0);
  for (int i = 0; i < n_layers; ++i) {
    const int in = dims[i], outd = dims[i + 1];
    const int ip = round16(in), op = round16(outd);
    const int nkb = ip / 16;
    float* wp = reinterpret_cast<float*>(host.data() + woff[i]);
    for (int ft = 0; ft < op / 16; ++ft)
      for (int kb = 0; kb < nkb; ++kb)
        for (int l = 0; l < 64; ++l)
          for (int j = 0; j < 4; ++j) {
            const int k = 16 * kb + 4 * (l / 16) + j;
            const int f = 16 * ft + (l % 16);
            const float v = (k < in && f < outd) ? W[i][(size_t)k * outd + f] : 0.f;
            wp[((size_t)(ft * nkb + kb) * 64 + l) * 4 + j] = v;
          }
    float* pw = reinterpret_cast<float*>(host.data() + poff[i]);
    for (int k = 0; k < in; ++k)
      for (int f = 0; f < outd; ++f) pw[(size_t)k * op + f] = W[i][(size_t)k * outd + f];
    float* bp = reinterpret_cast<float*>(host.data() + boff[i]);
    for (int f = 0; f < outd; ++f) bp[f] = (b && b[i]) ? b[i][f] : 0.f;
  }

  niq_model* m = new niq_model();
  m->n_layers = n_layers;
  m->in_dim = dims[0];
  m->wt = wt;
  m->dims.assign(dims, dims + n_layers + 1);
  m->acts.assign(acts, acts + n_layers);
  hipError_t e = hipMalloc(&m->weights, total);
  if (e == hipSuccess) e = hipMemcpy(m->weights, host.data(), total, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&m->dev), sizeof(DevModel));
  if (e != hipSuccess) {
    niq_model_destroy(m);
    return fail(NIQ_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
  }
  std::memset(&m->host, 0, sizeof(DevModel));
  m->host.n_layers = n_layers;
  m->host.wmax = wmax;
  m->host.in_dim = dims[0];
  for (int i = 0; i < n_layers; ++i) {
    DevLayer& L = m->host.L[i];
    L.in_pad = round16(dims[i]);
    L.out_pad = round16(dims[i + 1]);
    L.act = (i + 1 < n_layers) ? acts[i] : ACT_NONE;
    L.Wp = reinterpret_cast<const floatx4*>(static_cast<unsigned char*>(m->weights) + woff[i]);
    L.bias = reinterpret_cast<const float*>(static_cast<unsigned char*>(m->weights) + boff[i]);
    L.W = reinterpret_cast<const float*>(static_cast<unsigned char*>(m->weights) + poff[i]);
  }
  e = hipMemcpy(m->dev, &m->host, sizeof(DevModel), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    niq_model_destroy(m);
    return fail(NIQ_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
  }
  m->bytes = total + sizeof(DevModel);

  // Precision-mode copy: K padded to 32, hidden outputs padded to 32.
  {
    std::vector<size_t> o16(n_layers), ob16(n_layers);
    size_t t16 = 0;
    for (int i = 0; i < n_layers; ++i) {
      const size_t ip = round32(dims[i]);
      const size_t op = (i + 1 < n_layers) ? round32(dims[i + 1]) : round16(dims[i + 1]);
      o16[i] = t16;
      t16 += 4 * ip * op * sizeof(unsigned short);
      t16 = (t16 + 255) / 256 * 256;
      ob16[i] = t16;
      t16 += op * sizeof(float);
      t16 = (t16 + 255) / 256 * 256;
    }
    std::vector<unsigned char> h16(t16, 0);
    int wmax16 = 32;
    std::memset(&m->host16, 0, sizeof(DevModelBf16));
    m->host16.n_layers = n_layers;
    for (int i = 0; i < n_layers; ++i) {
      const int in = dims[i], outd = dims[i + 1];
      const int ip = round32(in);
      const int op = (i + 1 < n_layers) ? round32(outd) : round16(outd);
      wmax16 = std::max(wmax16, std::max(ip, op));
      const int nkb = ip / 32;
      const size_t nfrag = (size_t)(op / 16) * nkb * 64 * 8;
      unsigned short* w1p = reinterpret_cast<unsigned short*>(h16.data() + o16[i]);
      unsigned short* w2p = w1p + nfrag;
      unsigned short* w3p = w2p + nfrag;
      unsigned short* wap = w3p + nfrag;
      for (int ft = 0; ft < op / 16; ++ft)
        for (int kb = 0; kb < nkb; ++kb)
          for (int l = 0; l < 64; ++l)
            for (int sl = 0; sl < 8; ++sl) {
              const int k = 32 * kb + 8 * (l / 16) + sl;
              const int f = 16 * ft + (l % 16);
              const float w = (k < in && f < outd) ? W[i][(size_t)k * outd + f] : 0.f;
              const size_t idx = ((size_t)(ft * nkb + kb) * 64 + l) * 8 + sl;
              const unsigned short b1 = h_bf16_rn(w);
              const float r1 = w - h_bf16_f(b1);
              const unsigned short b2 = h_bf16_rn(r1);
              w1p[idx] = b1;
              w2p[idx] = b2;
              w3p[idx] = h_bf16_rn(r1 - h_bf16_f(b2));
              wap[idx] = h_bf16_up_nonneg(std::fabs(w));
            }
      float* bp = reinterpret_cast<float*>(h16.data() + ob16[i]);
      for (int f = 0; f < outd; ++f) bp[f] = (b && b[i]) ? b[i][f] : 0.f;
      DevLayerBf16& L = m->host16.L[i];
      L.in_pad = ip;
      L.out_pad = op;
      L.act = (i + 1 < n_layers) ? acts[i] : ACT_NONE;
      // Contraction bounds (relative to |X|.|W|): the split residual (base:
      // three-way splits drop terms <= ~4.1 * 2^-24; aff: two-way splits drop
      // <= 3.1 * 2^-16), each 32-wide chunk's MFMA chain with its smallest
      // terms first (1.05 gamma_33 at unit 2^-23, any FP32 rounding mode), and
      // the FP32 sum of the nkb chunk results (gamma_nkb at 2^-24).
      const double g_chunk = 1.05 * gamma_n(33.0, std::ldexp(1.0, -23));
      const double g_sum = gamma_n((double)nkb, std::ldexp(1.0, -24));
      L.c_base = (float)((4.1 * std::ldexp(1.0, -24) + g_chunk + g_sum) * 1.5);
      L.c_aff = (float)((3.1 * std::ldexp(1.0, -16) + g_chunk + g_sum) * 1.5);
      L.e_fix = (float)(1.0 + 1.5 * (gamma_n(33.0, std::ldexp(1.0, -23)) + g_sum));
    }
    m->host16.wmax = wmax16;
    m->wt16 = template_width(wmax16);
    e = hipMalloc(&m->weights16, t16);
    if (e == hipSuccess) e = hipMemcpy(m->weights16, h16.data(), t16, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&m->dev16), sizeof(DevModelBf16));
    if (e != hipSuccess) {
      niq_model_destroy(m);
      return fail(NIQ_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < n_layers; ++i) {
      DevLayerBf16& L = m->host16.L[i];
      const size_t nfrag = (size_t)(L.out_pad / 16) * (L.in_pad / 32) * 64;
      unsigned char* wb = static_cast<unsigned char*>(m->weights16) + o16[i];
      L.W1 = reinterpret_cast<const bf16x8*>(wb);
      L.W2 = L.W1 + nfrag;
      L.W3 = L.W2 + nfrag;
      L.Wa = L.W3 + nfrag;
      L.bias = reinterpret_cast<const float*>(static_cast<unsigned char*>(m->weights16) + ob16[i]);
    }
    e = hipMemcpy(m->dev16, &m->host16, sizeof(DevModelBf16), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      niq_model_destroy(m);
      return fail(NIQ_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
    }
    m->bytes += t16 + sizeof(DevModelBf16);
  }
  *out = m;
  return NIQ_OK;
}

void niq_model_destroy(niq_model* m) {
  if (!m) return;
  if (m->weights) (void)hipFree(m->weights);
  if (m->dev) (void)hipFree(m->dev);
  if (m->weights16) (void)hipFree(m->weights16);
  if (m->dev16) (void)hipFree(m->dev16);
  delete m;
}

size_t niq_model_device_bytes(const niq_model* m) { return m ? m->bytes : 0; }

}  // extern "C"

// ------------------------------------------------------------- dispatch
namespace {

// Launch k_classify for (KIND, V) at template width W.
template <int KIND, int V, int W>
niq_status launch_classify_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  constexpr int R = Planes<KIND, V, 0>::R;
  if constexpr (!Fits<R, NB, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "range form does not fit in LDS at this width");
  } else {
    auto k = k_classify<KIND, V, NB, NW, W>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, a);
    HIPCHK(hipGetLastError());
    if (a.recheck_tol > 0.f) {
      auto k64 = k_recheck_fp64<KIND, V, W>;
      const unsigned g64 = grid_for(std::min<long long>(max_n, a.recheck_cap), persistent_blocks(k64, 256));
      hipLaunchKernelGGL(k64, dim3(g64), dim3(256), 0, s, m->dev, a);
      HIPCHK(hipGetLastError());
    }
    return NIQ_OK;
  }
}

template <int KIND, int V>
niq_status launch_classify(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  switch (m->wt) {
    case 64: return launch_classify_w<KIND, V, 64>(m, a, max_n, s);
    case 128: return launch_classify_w<KIND, V, 128>(m, a, max_n, s);
    case 256: return launch_classify_w<KIND, V, 256>(m, a, max_n, s);
    case 512: return launch_classify_w<KIND, V, 512>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

template <int V, int W>
niq_status launch_bf16_w(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  if constexpr ((size_t)PlanesBf16<V>::NPL * NB * W * 2 > (size_t)kLdsLimit) {
    return fail(NIQ_ERR_UNSUPPORTED, "precision-mode range form does not fit in LDS at this width");
  } else {
    auto k = k_classify_bf16<V, NB, NW, W>;
    const long long tiles = (max_n + NB - 1) / NB;
    if (tiles == 0) return NIQ_OK;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev16, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int V>
niq_status launch_bf16(const niq_model* m, const ClassifyArgs& a, long long max_n, hipStream_t s) {
  switch (m->wt16) {
    case 64: return launch_bf16_w<V, 64>(m, a, max_n, s);
    case 128: return launch_bf16_w<V, 128>(m, a, max_n, s);
    case 256: return launch_bf16_w<V, 256>(m, a, max_n, s);
    case 512: return launch_bf16_w<V, 512>(m, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// affine_all / affine_truncate / affine_append: one box per workgroup, error
// terms streamed through a private HBM scratch.
template <int W, int MODE>
niq_status launch_rows_w(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                         hipStream_t s) {
  constexpr int NW = RowsCfg<W>::NW;
  auto k = k_classify_rows<W, MODE>;
  const long long d0 = std::max(a.d, a.lower ? 0 : a.vin);
  const int n_keep = cfg ? cfg->truncate_count : 0;
  const int n_append = cfg ? cfg->n_append : 0;
  long long max_rows = d0;
  int n_act = 0;
  for (int i = 0; i + 1 < m->n_layers; ++i)
    if (m->acts[i] != NIQ_ACT_NONE) { max_rows += m->host.L[i].out_pad; ++n_act; }
  if (MODE == ROWS_APPEND) max_rows = d0 + (long long)n_act * n_append;
  if (MODE == ROWS_TRUNCATE) max_rows = d0 + n_keep + 2LL * W;
  max_rows = std::max<long long>(max_rows, 1);
  const size_t per_wg = sizeof(float) * 2 * (size_t)max_rows * W;
  if (max_n == 0) return NIQ_OK;
  long long g = std::min<long long>(max_n, persistent_blocks(k, NW * 64));
  g = std::max<long long>(1, std::min<long long>(g, (long long)((16ull << 30) / per_wg)));
  void* scratch = nullptr;
  HIPCHK(hipMallocAsync(&scratch, per_wg * g, s));
  RowsArgs ra{};
  ra.c = a;
  ra.scratch = static_cast<float*>(scratch);
  ra.max_rows = max_rows;
  ra.n_keep = n_keep;
  ra.n_append = n_append;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(NW * 64), 0, s, m->dev, ra);
  const hipError_t e = hipGetLastError();
  HIPCHK(hipFreeAsync(scratch, s));
  if (e != hipSuccess) return fail(NIQ_ERR_HIP, std::string("k_classify_rows: ") + hipGetErrorString(e));
  return NIQ_OK;
}

template <int MODE>
niq_status launch_rows(const niq_model* m, const niq_eval_cfg* cfg, const ClassifyArgs& a, long long max_n,
                       hipStream_t s) {
  switch (m->wt) {
    case 64: return launch_rows_w<64, MODE>(m, cfg, a, max_n, s);
    case 128: return launch_rows_w<128, MODE>(m, cfg, a, max_n, s);
    case 256: return launch_rows_w<256, MODE>(m, cfg, a, max_n, s);
    case 512: return launch_rows_w<512, MODE>(m, cfg, a, max_n, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// Route a classification by (mode, precision, number of box vectors).
niq_status classify_dispatch(const niq_model* m, const niq_eval_cfg* cfg, int nvec,
                             const ClassifyArgs& a, long long max_n, hipStream_t s) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  const int prec = cfg ? cfg->precision : NIQ_PREC_FP32;
  if (prec == NIQ_PREC_BF16) {
    ClassifyArgs b = a;
    b.recheck_tol = 0.f;   // the precision mode is rigorous by construction
    if (mode == NIQ_MODE_INTERVAL) return launch_bf16<0>(m, b, max_n, s);
    if (mode == NIQ_MODE_AFFINE_FIXED) {
      switch (nvec) {
        case 1: return launch_bf16<1>(m, b, max_n, s);
        case 2: return launch_bf16<2>(m, b, max_n, s);
        case 3: return launch_bf16<3>(m, b, max_n, s);
      }
    }
    return fail(NIQ_ERR_UNSUPPORTED, "the precision mode supports interval and affine_fixed with 1..3 box vectors");
  }
  if (prec != NIQ_PREC_FP32) return fail(NIQ_ERR_INVALID, "unknown precision");
  if (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND) {
    if (a.recheck_tol > 0.f) return fail(NIQ_ERR_UNSUPPORTED, "the growing-row modes have no FP64 re-test");
    if (!a.lower && a.vin > 64) return fail(NIQ_ERR_UNSUPPORTED, "at most 64 box vectors");
    if (mode == NIQ_MODE_AFFINE_ALL) return launch_rows<ROWS_ALL>(m, cfg, a, max_n, s);
    if (mode == NIQ_MODE_AFFINE_TRUNCATE) {
      if (!cfg || cfg->truncate_count < 1 || cfg->truncate_count > 64)
        return fail(NIQ_ERR_INVALID, "affine_truncate needs 1 <= truncate_count <= 64");
      return launch_rows<ROWS_TRUNCATE>(m, cfg, a, max_n, s);
    }
    if (!cfg || cfg->n_append < 0 || cfg->n_append > 512)
      return fail(NIQ_ERR_INVALID, "affine_append needs 0 <= n_append <= 512");
    return launch_rows<ROWS_APPEND>(m, cfg, a, max_n, s);
  }
  if (mode == NIQ_MODE_INTERVAL) return launch_classify<0, 0>(m, a, max_n, s);
  if (mode == NIQ_MODE_AFFINE_FIXED) {
    switch (nvec) {
      case 1: return launch_classify<0, 1>(m, a, max_n, s);
      case 2: return launch_classify<0, 2>(m, a, max_n, s);
      case 3: return launch_classify<0, 3>(m, a, max_n, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "affine_fixed supports 1..3 box vectors");
  }
  if (mode == NIQ_MODE_SLOPE_INTERVAL) {
    switch (nvec) {
      case 1: return launch_classify<1, 1>(m, a, max_n, s);
      case 2: return launch_classify<1, 2>(m, a, max_n, s);
      case 3: return launch_classify<1, 3>(m, a, max_n, s);
    }
    return fail(NIQ_ERR_UNSUPPORTED, "slope_interval supports 1..3 box vectors");
  }
  return fail(NIQ_ERR_UNSUPPORTED, "mode not available in the fused kernel");
}

template <int W>
niq_status launch_points_w(const niq_model* m, const float* x, int d, long long n, float* out, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_points<NB, NW, W>;
  const long long tiles = (n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, x, d, n, out);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

// Classification with the optional FP64 re-test list (stream-ordered scratch).
niq_status classify_with_recheck(const niq_model* m, const niq_eval_cfg* cfg, int nvec, ClassifyArgs a,
                                 long long n, hipStream_t s);

niq_status check_model(const niq_model* m, int d) {
  if (!m) return fail(NIQ_ERR_INVALID, "null model");
  if (d != m->in_dim) return fail(NIQ_ERR_INVALID, "point dimension does not match the model input");
  return NIQ_OK;
}

}  // namespace

namespace {
niq_status classify_with_recheck(const niq_model* m, const niq_eval_cfg* cfg, int nvec, ClassifyArgs a,
                                 long long n, hipStream_t s) {
  const float tol = cfg ? cfg->recheck_tol : 0.f;
  if (!(tol > 0.f) || n == 0) return classify_dispatch(m, cfg, nvec, a, n, s);
  if (n > INT_MAX) return fail(NIQ_ERR_UNSUPPORTED, "FP64 re-test supports at most 2^31-1 boxes per call");
  void* scratch = nullptr;
  HIPCHK(hipMallocAsync(&scratch, sizeof(int) * (size_t)(n + 64), s));
  HIPCHK(hipMemsetAsync(scratch, 0, sizeof(int), s));
  a.recheck_tol = tol;
  a.recheck_cnt = static_cast<int*>(scratch);
  a.recheck_idx = static_cast<int*>(scratch) + 64;
  a.recheck_cap = (int)n;
  niq_status st = classify_dispatch(m, cfg, nvec, a, n, s);
  HIPCHK(hipFreeAsync(scratch, s));
  return st;
}
}  // namespace

extern "C" {

niq_status niq_classify_boxes(const niq_model* m, const niq_eval_cfg* cfg, const float* lower,
                              const float* upper, int64_t n, int d, float offset, int32_t* codes,
                              float* bounds, void* stream) {
  if (niq_status st = check_model(m, d)) return st;
  if (n < 0 || (n > 0 && (!lower || !upper || !codes))) return fail(NIQ_ERR_INVALID, "bad buffers");
  ClassifyArgs a{};
  a.lower = lower; a.upper = upper; a.d = d; a.n = n; a.offset = offset;
  a.codes = codes; a.bounds = bounds;
  return classify_with_recheck(m, cfg, d, a, n, (hipStream_t)stream);
}

niq_status niq_classify_general_boxes(const niq_model* m, const niq_eval_cfg* cfg, const float* center,
                                      const float* vecs, int64_t n, int v, int d, float offset,
                                      int32_t* codes, float* bounds, void* stream) {
  if (niq_status st = check_model(m, d)) return st;
  if (v < 1) return fail(NIQ_ERR_INVALID, "need at least one box vector");
  if (n < 0 || (n > 0 && (!center || !vecs || !codes))) return fail(NIQ_ERR_INVALID, "bad buffers");
  ClassifyArgs a{};
  a.center = center; a.vecs = vecs; a.vin = v; a.d = d; a.n = n; a.offset = offset;
  a.codes = codes; a.bounds = bounds;
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  return classify_with_recheck(m, cfg, mode == NIQ_MODE_INTERVAL ? 0 : v, a, n, (hipStream_t)stream);
}

niq_status niq_eval_points(const niq_model* m, const float* x, int64_t n, int d, float* out, void* stream) {
  if (niq_status st = check_model(m, d)) return st;
  if (n < 0 || (n > 0 && (!x || !out))) return fail(NIQ_ERR_INVALID, "bad buffers");
  hipStream_t s = (hipStream_t)stream;
  switch (m->wt) {
    case 64: return launch_points_w<64>(m, x, d, n, out, s);
    case 128: return launch_points_w<128>(m, x, d, n, out, s);
    case 256: return launch_points_w<256>(m, x, d, n, out, s);
    case 512: return launch_points_w<512>(m, x, d, n, out, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

// ----------------------------------------------------------- tree build
namespace {

struct TreeLayout {
  long long cap;      // max nodes in any level list
  long long nbmax;    // max batches per level
  size_t off_counters, off_codes, off_bcnt, off_boff, off_lvl[2], off_recheck, total;
};

niq_status tree_layout(const niq_tree_params* p, TreeLayout* L) {
  if (!p) return fail(NIQ_ERR_INVALID, "null params");
  const int B = p->batch_process_size;
  if (B < 1 || (B & (B - 1)) || B > (1 << 24))
    return fail(NIQ_ERR_INVALID, "batch_process_size must be a power of 2 (a factor of the bucket sizes)");
  if (p->split_depth < 0 && p->node_terminate_thresh <= 0)
    return fail(NIQ_ERR_INVALID, "must specify at least one of node_terminate_thresh or split_depth");
  long long cap = LLONG_MAX;
  if (p->split_depth >= 0) {
    if (p->split_depth > 30) return fail(NIQ_ERR_UNSUPPORTED, "split_depth above 30");
    cap = 1LL << p->split_depth;
  }
  if (p->node_terminate_thresh > 0) cap = std::min(cap, 2LL * p->node_terminate_thresh);
  cap = std::max(cap, 1LL);
  if (cap > (1LL << 30)) return fail(NIQ_ERR_UNSUPPORTED, "tree level capacity above 2^30 nodes");
  L->cap = cap;
  L->nbmax = (cap + B - 1) / B;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 255) / 256 * 256; return r; };
  L->off_counters = take(sizeof(TreeCounters));
  L->off_codes = take(sizeof(int) * cap);
  L->off_bcnt = take(sizeof(int) * 4 * L->nbmax);
  L->off_boff = take(sizeof(int) * 4 * L->nbmax);
  L->off_lvl[0] = take(sizeof(float) * 6 * cap);
  L->off_lvl[1] = take(sizeof(float) * 6 * cap);
  L->off_recheck = take(sizeof(int) * cap);
  L->total = o;
  return NIQ_OK;
}

__global__ void k_tree_init(TreeCounters* C, float* lo, float* up, float l0, float l1, float l2, float u0,
                            float u1, float u2) {
  if (threadIdx.x == 0) {
    C->n[0] = 1; C->n[1] = 0; C->n_int = 0; C->n_ext = 0; C->n_classified = 0; C->overflow = 0;
    C->recheck_cnt = 0; C->n_recheck = 0;
    lo[0] = l0; lo[1] = l1; lo[2] = l2;
    up[0] = u0; up[1] = u1; up[2] = u2;
  }
  for (int i = threadIdx.x; i < 64; i += blockDim.x) { C->level_n[i] = 0; C->level_int[i] = 0; C->level_ext[i] = 0; }
}

}  // namespace

niq_status niq_tree_workspace_bytes(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p,
                                    size_t* bytes) {
  (void)m; (void)cfg;
  TreeLayout L;
  if (niq_status st = tree_layout(p, &L)) return st;
  if (bytes) *bytes = L.total;
  return NIQ_OK;
}

niq_status niq_tree_build(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, void* ws,
                          size_t ws_bytes, niq_tree_result* res, void* stream) {
  if (niq_status st = check_model(m, 3)) return st;
  if (!res) return fail(NIQ_ERR_INVALID, "null result");
  TreeLayout L;
  if (niq_status st = tree_layout(p, &L)) return st;
  if (!ws || ws_bytes < L.total)
    return fail(NIQ_ERR_CAPACITY, "tree workspace too small: need " + std::to_string(L.total) + " bytes");
  if (!res->unknown_lower || !res->unknown_upper || res->unknown_capacity < 0)
    return fail(NIQ_ERR_INVALID, "unknown-node output buffers are required");
  if (p->with_interior_nodes && (!res->interior_lower || !res->interior_upper))
    return fail(NIQ_ERR_INVALID, "interior output buffers are required");
  if (p->with_exterior_nodes && (!res->exterior_lower || !res->exterior_upper))
    return fail(NIQ_ERR_INVALID, "exterior output buffers are required");
  hipStream_t s = (hipStream_t)stream;
  unsigned char* base = static_cast<unsigned char*>(ws);
  TreeCounters* C = reinterpret_cast<TreeCounters*>(base + L.off_counters);
  int* codes = reinterpret_cast<int*>(base + L.off_codes);
  int* bcnt = reinterpret_cast<int*>(base + L.off_bcnt);
  int* boff = reinterpret_cast<int*>(base + L.off_boff);
  float* lvl[2] = {reinterpret_cast<float*>(base + L.off_lvl[0]), reinterpret_cast<float*>(base + L.off_lvl[1])};
  const long long cap = L.cap;
  const int B = p->batch_process_size;

  hipLaunchKernelGGL(k_tree_init, dim3(1), dim3(64), 0, s, C, lvl[0], lvl[0] + 3 * cap, p->lower[0], p->lower[1],
                     p->lower[2], p->upper[0], p->upper[1], p->upper[2]);
  HIPCHK(hipGetLastError());

  const int n_levels = p->split_depth >= 0 ? p->split_depth + 1 : 64;
  int level = 0;
  for (; level < std::min(n_levels, 62); ++level) {
    bool quit_next = (level + 1 == n_levels);
    if (p->node_terminate_thresh > 0) {
      int n_host = 0;
      HIPCHK(hipMemcpyAsync(&n_host, &C->n[level & 1], sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      quit_next = quit_next || (n_host >= p->node_terminate_thresh);
    }
    const int split = quit_next ? 0 : 1;
    const long long max_n = level < 40 ? std::min(cap, 1LL << level) : cap;
    float* cur_lo = lvl[level & 1];
    float* cur_up = cur_lo + 3 * cap;

    ClassifyArgs a{};
    a.lower = cur_lo; a.upper = cur_up; a.d = 3; a.n_dev = &C->n[level & 1]; a.offset = p->offset;
    a.codes = codes;
    if (cfg && cfg->recheck_tol > 0.f) {
      a.recheck_tol = cfg->recheck_tol;
      a.recheck_cnt = &C->recheck_cnt;
      a.recheck_idx = reinterpret_cast<int*>(base + L.off_recheck);
      a.recheck_cap = (int)std::min<long long>(cap, INT_MAX);
    }
    if (niq_status st = classify_dispatch(m, cfg, 3, a, max_n, s)) return st;

    const unsigned nb = (unsigned)((max_n + B - 1) / B);
    hipLaunchKernelGGL(k_tree_count, dim3(nb), dim3(256), 0, s, codes, C, level, B, bcnt);
    HIPCHK(hipGetLastError());

    TreeScatterArgs sa{};
    sa.codes = codes; sa.lower = cur_lo; sa.upper = cur_up;
    if (split) {
      sa.next_lower = lvl[(level + 1) & 1];
      sa.next_upper = sa.next_lower + 3 * cap;
      sa.cap_next = cap;
    } else {
      sa.next_lower = res->unknown_lower;
      sa.next_upper = res->unknown_upper;
      sa.cap_next = res->unknown_capacity;
    }
    if (p->with_interior_nodes) { sa.int_lower = res->interior_lower; sa.int_upper = res->interior_upper; sa.cap_int = res->interior_capacity; }
    if (p->with_exterior_nodes) { sa.ext_lower = res->exterior_lower; sa.ext_upper = res->exterior_upper; sa.cap_ext = res->exterior_capacity; }
    const long long cap_int = p->with_interior_nodes ? res->interior_capacity : LLONG_MAX;
    const long long cap_ext = p->with_exterior_nodes ? res->exterior_capacity : LLONG_MAX;
    hipLaunchKernelGGL(k_tree_scan, dim3(1), dim3(1024), 0, s, bcnt, boff, C, level, B, split, sa.cap_next, cap_int,
                       cap_ext);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_tree_scatter, dim3(nb), dim3(256), 0, s, sa, bcnt, boff, C, level, B, split);
    HIPCHK(hipGetLastError());
    if (quit_next) break;
  }
  TreeCounters hc;
  HIPCHK(hipMemcpyAsync(&hc, C, sizeof(TreeCounters), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  res->n_levels = level + 1;
  res->n_unknown = hc.n[(level + 1) & 1];
  res->n_interior = hc.n_int;
  res->n_exterior = hc.n_ext;
  res->n_classified = hc.n_classified;
  res->n_recheck = hc.n_recheck;
  for (int i = 0; i < 64; ++i) {
    res->level_n[i] = hc.level_n[i];
    res->level_interior[i] = hc.level_int[i];
    res->level_exterior[i] = hc.level_ext[i];
  }
  if (hc.overflow) return fail(NIQ_ERR_CAPACITY, "a tree output buffer was too small");
  return NIQ_OK;
}

// ---------------------------------------------------------- ray casting
size_t niq_cast_workspace_bytes(const niq_cast_opts* opts) {
  const int n_max = opts ? std::max(opts->n_max_step, 1) : 512;
  size_t b = 256;                                  // work counter
  b += ((size_t)(n_max + 2) * sizeof(int) + 255) / 256 * 256;
  b += 8 * sizeof(DevModel);                       // model table
  return b;
}

}  // extern "C"

namespace {

template <int KIND, int V, int W>
niq_status launch_rays_w(const DevModel* models, const RayArgs& a, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  constexpr int R = Planes<KIND, V, 2>::R;
  if constexpr (!Fits<R, NB, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "ray range form does not fit in LDS at this width");
  } else {
    auto k = k_cast_rays<KIND, V, NB, NW, W>;
    const long long tiles = (a.n + NB - 1) / NB;
    const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, models, a);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

template <int KIND, int V>
niq_status launch_rays(int wt, const DevModel* models, const RayArgs& a, hipStream_t s) {
  switch (wt) {
    case 64: return launch_rays_w<KIND, V, 64>(models, a, s);
    case 128: return launch_rays_w<KIND, V, 128>(models, a, s);
    case 256: return launch_rays_w<KIND, V, 256>(models, a, s);
    case 512: return launch_rays_w<KIND, V, 512>(models, a, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "width");
}

long long next_bucket(long long s) {
  for (int e = 7; e <= 30; ++e)
    if (s <= (1LL << e)) return 1LL << e;
  return -1;
}

}  // namespace

extern "C" niq_status niq_cast_rays(const niq_model* const* models, int n_models, const niq_eval_cfg* cfg,
                                    const float* roots, const float* dirs, int64_t n, const niq_cast_opts* o,
                                    float* out_t, int32_t* out_hit_id, int32_t* out_count, void* ws,
                                    size_t ws_bytes, int64_t* n_evals, void* stream) {
  if (!models || n_models < 1 || n_models > 8) return fail(NIQ_ERR_INVALID, "need 1..8 models");
  if (!o) return fail(NIQ_ERR_INVALID, "null opts");
  if (o->n_substeps < 1 || o->n_max_step < 1) return fail(NIQ_ERR_INVALID, "n_substeps and n_max_step must be >= 1");
  if (n < 0 || n > (1LL << 31) - 1) return fail(NIQ_ERR_INVALID, "ray count out of range");
  if (n > 0 && (!roots || !dirs || !out_t || !out_hit_id || !out_count)) return fail(NIQ_ERR_INVALID, "bad buffers");
  const size_t need = niq_cast_workspace_bytes(o);
  if (!ws || ws_bytes < need) return fail(NIQ_ERR_CAPACITY, "cast workspace too small: need " + std::to_string(need));
  int wt = 0;
  for (int i = 0; i < n_models; ++i) {
    if (niq_status st = check_model(models[i], 3)) return st;
    wt = std::max(wt, models[i]->wt);
  }
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  if (cfg && cfg->precision != NIQ_PREC_FP32) return fail(NIQ_ERR_UNSUPPORTED, "cast_rays runs in FP32");
  hipStream_t s = (hipStream_t)stream;
  unsigned char* base = static_cast<unsigned char*>(ws);
  int* counter = reinterpret_cast<int*>(base);
  int* hist = reinterpret_cast<int*>(base + 256);
  const size_t hist_bytes = ((size_t)(o->n_max_step + 2) * sizeof(int) + 255) / 256 * 256;
  DevModel* table = reinterpret_cast<DevModel*>(base + 256 + hist_bytes);
  std::vector<DevModel> host_table(n_models);
  for (int i = 0; i < n_models; ++i) host_table[i] = models[i]->host;
  HIPCHK(hipMemsetAsync(base, 0, 256 + hist_bytes, s));
  HIPCHK(hipMemcpyAsync(table, host_table.data(), sizeof(DevModel) * n_models, hipMemcpyHostToDevice, s));
  if (n > 0) {
    RayArgs a{};
    a.roots = roots; a.dirs = dirs; a.n = n; a.n_models = n_models;
    a.hit_eps = o->hit_eps; a.max_dist = o->max_dist; a.safety = o->safety_factor;
    a.grow = o->interval_grow_fac; a.shrink = o->interval_shrink_fac; a.init_size = o->init_step;
    a.n_max_step = o->n_max_step; a.n_substeps = o->n_substeps;
    a.out_t = out_t; a.out_hit = out_hit_id; a.out_count = out_count;
    a.next_ray = counter; a.iter_hist = hist;
    niq_status st;
    if (mode == NIQ_MODE_AFFINE_FIXED) st = launch_rays<0, 1>(wt, table, a, s);
    else if (mode == NIQ_MODE_INTERVAL) st = launch_rays<0, 0>(wt, table, a, s);
    else if (mode == NIQ_MODE_SLOPE_INTERVAL) st = launch_rays<1, 1>(wt, table, a, s);
    else return fail(NIQ_ERR_UNSUPPORTED, "cast_rays supports interval, affine_fixed and slope_interval");
    if (st) return st;
  }
  std::vector<int> h(o->n_max_step + 2, 0);
  HIPCHK(hipMemcpyAsync(h.data(), hist, sizeof(int) * h.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // Reproduce the reference's N_evals: each lock-step iteration evaluates the
  // whole (bucketed) working array n_substeps times (src/queries.py:164,170-173).
  if (n_evals) {
    long long evals = 0, bucket = n, live = n;
    for (int it = 1; it <= o->n_max_step + 1 && live > 0; ++it) {
      evals += bucket * o->n_substeps;
      live -= h[it];
      if (live > 0) {
        const long long nbkt = next_bucket(live);
        if (nbkt > 0 && nbkt < bucket) bucket = nbkt;
      }
    }
    *n_evals = evals;
  }
  return NIQ_OK;
}

// ---------------------------------------------------------- closest point
namespace {

template <int W>
niq_status launch_points_dev_w(const niq_model* m, const float* x, const int* n_dev, int mult, long long max_n,
                               float* out, hipStream_t s);

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
  const int K = 8;                                  // rounds between host checks
  long long cap = std::max<long long>(2 * Q, Q + 2LL * K * B);
  // per-call device scratch (stream ordered)
  auto al = [&](size_t bytes, void** p) { return hipMallocAsync(p, (bytes + 255) / 256 * 256, s); };
  void *pS, *pbq, *pblo, *pbup, *ppts, *pcodes, *pvals, *pwin, *pidist, *pineeds, *pwq, *pwlo, *pwup;
  HIPCHK(al(sizeof(CpState), &pS));
  HIPCHK(al(sizeof(int) * B, &pbq));
  HIPCHK(al(sizeof(float) * 3 * B, &pblo));
  HIPCHK(al(sizeof(float) * 3 * B, &pbup));
  HIPCHK(al(sizeof(float) * 21 * B, &ppts));
  HIPCHK(al(sizeof(int) * B, &pcodes));
  HIPCHK(al(sizeof(float) * 7 * B, &pvals));
  HIPCHK(al(sizeof(int) * Q, &pwin));
  HIPCHK(al(sizeof(float) * B, &pidist));
  HIPCHK(al(sizeof(int) * B, &pineeds));
  HIPCHK(al(sizeof(int) * cap, &pwq));
  HIPCHK(al(sizeof(float) * 3 * cap, &pwlo));
  HIPCHK(al(sizeof(float) * 3 * cap, &pwup));
  CpArgs a{};
  a.B = B; a.cap = cap;
  a.wq = (int*)pwq; a.wlo = (float*)pwlo; a.wup = (float*)pwup;
  a.bq = (int*)pbq; a.blo = (float*)pblo; a.bup = (float*)pbup; a.pts = (float*)ppts;
  a.codes = (int*)pcodes; a.vals = (float*)pvals; a.q = query_points;
  a.qmin = out_dist; a.qloc = out_loc; a.win = (int*)pwin;
  a.idist = (float*)pidist; a.ineeds = (int*)pineeds;
  // eps / sqrt(d) in float32 (src/kd_tree.py:688)
  a.eps_cube_width = eps / sqrtf(3.0f);
  CpState* S = (CpState*)pS;
  {
    CpState h{};
    h.top = (int)Q;
    HIPCHK(hipMemcpyAsync(S, &h, sizeof(h), hipMemcpyHostToDevice, s));
    std::vector<int> ids(Q);
    for (long long i = 0; i < Q; ++i) ids[i] = (int)i;
    std::vector<float> blo(3 * Q), bup(3 * Q), inf(Q, INFINITY), loc(3 * Q, -777.f);
    for (long long i = 0; i < Q; ++i)
      for (int k = 0; k < 3; ++k) { blo[3 * i + k] = lower[k]; bup[3 * i + k] = upper[k]; }
    HIPCHK(hipMemcpyAsync(a.wq, ids.data(), sizeof(int) * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(a.wlo, blo.data(), sizeof(float) * 3 * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(a.wup, bup.data(), sizeof(float) * 3 * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(out_dist, inf.data(), sizeof(float) * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(out_loc, loc.data(), sizeof(float) * 3 * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(a.win, 0xff, sizeof(int) * Q, s));
    HIPCHK(hipStreamSynchronize(s));   // host vectors go out of scope
  }
  ClassifyArgs ca{};
  ca.lower = a.blo; ca.upper = a.bup; ca.d = 3; ca.n_dev = &S->n; ca.offset = 0.f; ca.codes = a.codes;
  const unsigned gg = (unsigned)std::min<long long>((B + 255) / 256, 1024);
  niq_status st = NIQ_OK;
  while (true) {
    for (int r = 0; r < K && st == NIQ_OK; ++r) {
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
      void *nq, *nlo, *nup;
      if (al(sizeof(int) * ncap, &nq) || al(sizeof(float) * 3 * ncap, &nlo) || al(sizeof(float) * 3 * ncap, &nup)) {
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
  for (void* p : {pS, pbq, pblo, pbup, ppts, pcodes, pvals, pwin, pidist, pineeds}) (void)hipFreeAsync(p, s);
  (void)hipFreeAsync(a.wq, s); (void)hipFreeAsync(a.wlo, s); (void)hipFreeAsync(a.wup, s);
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
  void *ppts, *pvals, *pcnt, *poff, *ptot;
  HIPCHK(hipMallocAsync(&ppts, sizeof(float) * 3 * np3 * chunk, s));
  HIPCHK(hipMallocAsync(&pvals, sizeof(float) * np3 * chunk, s));
  HIPCHK(hipMallocAsync(&pcnt, sizeof(long long) * chunk, s));
  HIPCHK(hipMallocAsync(&poff, sizeof(long long) * chunk, s));
  HIPCHK(hipMallocAsync(&ptot, sizeof(long long), s));
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
  for (void* p : {ppts, pvals, pcnt, poff, ptot}) (void)hipFreeAsync(p, s);
  if (st) return st;
  HIPCHK(hipGetLastError());
  *n_tris = written;
  if (written > tri_capacity) return fail(NIQ_ERR_CAPACITY, "triangle buffer too small: need " + std::to_string(written));
  return NIQ_OK;
}

namespace {
template <int W>
niq_status launch_points_dev_w(const niq_model* m, const float* x, const int* n_dev, int mult, long long max_n,
                               float* out, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_points_dev<NB, NW, W>;
  const long long tiles = (max_n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, x, n_dev, mult, out);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}
}  // namespace
