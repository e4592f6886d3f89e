// Model handles: lowering upload, FP32 fragments and the precision-mode copy.
#include "niq_internal.h"

#include <atomic>

namespace niqi {
thread_local std::string g_err;
}  // namespace niqi

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

  static std::atomic<uint64_t> next_serial{1};
  niq_model* m = new niq_model();
  m->serial = next_serial.fetch_add(1);
  m->n_layers = n_layers;
  m->in_dim = dims[0];
  m->wt = wt;
  m->dims.assign(dims, dims + n_layers + 1);
  m->acts.assign(acts, acts + n_layers);
  // hidden layers ReLU or the identity: the ReLU-only classify kernels apply
  m->relu_only = true;
  m->tanh_only = true;
  for (int i = 0; i + 1 < n_layers; ++i) {
    if (acts[i] != NIQ_ACT_RELU && acts[i] != NIQ_ACT_NONE) m->relu_only = false;
    if (acts[i] != NIQ_ACT_TANH && acts[i] != NIQ_ACT_NONE) m->tanh_only = false;
  }
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
      t16 += 4 * ip * op * sizeof(unsigned short) + ip * op * sizeof(float);
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
      float* w32p = reinterpret_cast<float*>(wap + nfrag);
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
              // FP32 copy: [fragment][half][lane][4]
              const size_t i32 = (((size_t)(ft * nkb + kb) * 2 + sl / 4) * 64 + l) * 4 + sl % 4;
              w32p[i32] = w;
              // bf16 terms of FP32 subnormals are not guaranteed by the
              // hardware conversions: the precision mode refuses such weights
              if (w != 0.f && std::fabs(w) < 1e-30f) m->bf16_ok = false;
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
      // e |w1| + e |w2| summed in one accumulator over the whole K (2 in_pad
      // non-negative terms), and |W| <= (|w1| + |w2|) / (1 - 2^-16)
      L.e_fix32 = (float)((1.0 + 1.5 * gamma_n(2.0 * ip + 1.0, std::ldexp(1.0, -23))) *
                          (1.0 + std::ldexp(1.0, -15)));
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
      L.W32 = reinterpret_cast<const floatx4*>(L.Wa + nfrag);
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

