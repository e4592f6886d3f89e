// Normals and shading of ray-cast hits (src/render.py:53-224).
#include "niq_internal.h"
#include "niq_render.h"

namespace {

static __global__ void __launch_bounds__(256)
k_hit_positions(const float* __restrict__ roots, const float* __restrict__ dirs, const float* __restrict__ t,
                long long n, float* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < 3 * n; i += (long long)gridDim.x * 256)
    out[i] = __fadd_rn(roots[i], __fmul_rn(t[i / 3], dirs[i]));
}

unsigned elem_grid(long long n) { return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384)); }

// AS: the activation set of the kernel (ReLU models take the ReLU-only one).
template <int W, int AS = ACTS_ALL>
niq_status launch_normal_eval(const niq_model* m, const int* list, const int* count, long long max_n,
                              const float* hit_pos, float eps, float* vals, hipStream_t s) {
  if constexpr (AS == ACTS_ALL) {
    if (m->relu_only) return launch_normal_eval<W, ACTS_RELU>(m, list, count, max_n, hit_pos, eps, vals, s);
  }
  constexpr int NB = TileCfg<W>::NBP, NW = TileCfg<W>::NW;
  auto k = k_normal_eval<NB, NW, W, AS>;
  const long long tiles = (4 * max_n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, list, count, hit_pos, eps, vals);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

template <int W>
niq_status launch_grad_normals(const niq_model* m, const int* list, const int* count, long long max_n,
                               const float* hit_pos, float* normals, hipStream_t s) {
  constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
  auto k = k_grad_normals<NB, NW, W>;
  const long long tiles = (max_n + NB - 1) / NB;
  if (tiles == 0) return NIQ_OK;
  const unsigned g = grid_for(tiles, persistent_blocks(k, NW * 64));
  hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, m->dev, list, count, hit_pos, normals);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

}  // namespace

namespace {

// One pinhole camera ray per pixel (src/render.py:26-50), in the reference's
// float32 operation order: x, y from linspace(-1, 1, res) (formed in double,
// the last sample exactly 1, then rounded), pixel p = row * res + col ('xy'
// meshgrid), pos = (look + left (x t)) + up (y t), dir = pos / |pos| with
// |pos| = sqrt((x0^2 + x1^2) + x2^2) correctly rounded.
__global__ void k_camera_rays(float3 eye, float3 look, float3 up, float3 left, int res, float tan_half,
                              float* __restrict__ roots, float* __restrict__ dirs) {
  const long long n = (long long)res * res;
  const double step = res > 1 ? 2.0 / (double)(res - 1) : 0.0;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(p / res), col = (int)(p % res);
    const float cx = col == res - 1 && res > 1 ? 1.f : (float)((double)col * step + -1.0);
    const float cy = row == res - 1 && res > 1 ? 1.f : (float)((double)row * step + -1.0);
    const float sx = __fmul_rn(cx, tan_half), sy = __fmul_rn(cy, tan_half);
    float q[3];
    const float lk[3] = {look.x, look.y, look.z}, lf[3] = {left.x, left.y, left.z}, u[3] = {up.x, up.y, up.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = __fadd_rn(__fadd_rn(lk[k], __fmul_rn(lf[k], sx)), __fmul_rn(u[k], sy));
    // sqrtf: the correctly rounded expansion (HIP compiles it so by default;
    // __fsqrt_rn is a bare v_sqrt_f32, within 1 ulp)
    const float nrm = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(q[0], q[0]), __fmul_rn(q[1], q[1])), __fmul_rn(q[2], q[2])));
    roots[3 * p] = eye.x; roots[3 * p + 1] = eye.y; roots[3 * p + 2] = eye.z;
#pragma unroll
    for (int k = 0; k < 3; ++k) dirs[3 * p + k] = __fdiv_rn(q[k], nrm);
  }
}

}  // namespace

extern "C" niq_status niq_camera_rays(const float eye[3], const float look[3], const float up[3], const float left[3],
                                      int res, float tan_half, float* out_roots, float* out_dirs, void* stream) {
  if (!eye || !look || !up || !left) return fail(NIQ_ERR_INVALID, "null camera vector");
  if (res < 1 || res > 65536) return fail(NIQ_ERR_INVALID, "resolution out of range");
  if (!out_roots || !out_dirs) return fail(NIQ_ERR_INVALID, "bad buffers");
  const long long n = (long long)res * res;
  const unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_camera_rays, dim3(g), dim3(256), 0, (hipStream_t)stream, make_float3(eye[0], eye[1], eye[2]),
                     make_float3(look[0], look[1], look[2]), make_float3(up[0], up[1], up[2]),
                     make_float3(left[0], left[1], left[2]), res, tan_half, out_roots, out_dirs);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

extern "C" niq_status niq_hit_positions(const float* roots, const float* dirs, const float* t, int64_t n,
                                        float* out, void* stream) {
  if (n < 0 || (n > 0 && (!roots || !dirs || !t || !out))) return fail(NIQ_ERR_INVALID, "bad buffers");
  if (n == 0) return NIQ_OK;
  hipLaunchKernelGGL(k_hit_positions, dim3(elem_grid(3 * n)), dim3(256), 0, (hipStream_t)stream, roots, dirs, t,
                     (long long)n, out);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}

extern "C" niq_status niq_outward_normals(const niq_model* const* models, int n_models, const float* hit_pos,
                                          const int32_t* hit_ids, int64_t n, float eps, int method,
                                          float* out_normals, void* stream) {
  if (!models || n_models < 1 || n_models > 8) return fail(NIQ_ERR_INVALID, "need 1..8 models");
  if (method != NIQ_NORMALS_FINITE_DIFFERENCES && method != NIQ_NORMALS_AUTODIFF)
    return fail(NIQ_ERR_INVALID, "unrecognized method");
  for (int i = 0; i < n_models; ++i)
    if (niq_status st = check_model(models[i], 3)) return st;
  if (n < 0 || n >= (1LL << 31)) return fail(NIQ_ERR_INVALID, "need 0 <= n < 2^31");
  if (n > 0 && (!hit_pos || !hit_ids || !out_normals)) return fail(NIQ_ERR_INVALID, "bad buffers");
  if (n == 0) return NIQ_OK;
  hipStream_t s = (hipStream_t)stream;
  // scratch: per-function pixel lists and counts, then the sample values
  const size_t list_bytes = (sizeof(int) * (size_t)n * n_models + 255) / 256 * 256;
  const size_t cnt_bytes = 256;
  const size_t val_bytes = sizeof(float) * 4 * (size_t)n;
  char* scratch = nullptr;
  HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&scratch), list_bytes + cnt_bytes + val_bytes, s));
  int* lists = reinterpret_cast<int*>(scratch);
  int* counts = reinterpret_cast<int*>(scratch + list_bytes);
  float* vals = reinterpret_cast<float*>(scratch + list_bytes + cnt_bytes);
  niq_status st = NIQ_OK;
  if (hipMemsetAsync(counts, 0, cnt_bytes, s) != hipSuccess) st = fail(NIQ_ERR_HIP, "normal scratch clear failed");
  if (st == NIQ_OK)
    hipLaunchKernelGGL(k_hit_lists, dim3(elem_grid(n)), dim3(256), 0, s, hit_ids, (long long)n, n_models,
                       (long long)n, lists, counts, out_normals);
  for (int f = 0; f < n_models && st == NIQ_OK && method == NIQ_NORMALS_AUTODIFF; ++f) {
    const int* list = lists + (size_t)f * n;
    switch (models[f]->wt) {
      case 64: st = launch_grad_normals<64>(models[f], list, counts + f, n, hit_pos, out_normals, s); break;
      case 128: st = launch_grad_normals<128>(models[f], list, counts + f, n, hit_pos, out_normals, s); break;
      case 256: st = launch_grad_normals<256>(models[f], list, counts + f, n, hit_pos, out_normals, s); break;
      case 512: st = launch_grad_normals<512>(models[f], list, counts + f, n, hit_pos, out_normals, s); break;
      default: st = fail(NIQ_ERR_UNSUPPORTED, "width");
    }
  }
  for (int f = 0; f < n_models && st == NIQ_OK && method == NIQ_NORMALS_FINITE_DIFFERENCES; ++f) {
    const int* list = lists + (size_t)f * n;
    switch (models[f]->wt) {
      case 64: st = launch_normal_eval<64>(models[f], list, counts + f, n, hit_pos, eps, vals, s); break;
      case 128: st = launch_normal_eval<128>(models[f], list, counts + f, n, hit_pos, eps, vals, s); break;
      case 256: st = launch_normal_eval<256>(models[f], list, counts + f, n, hit_pos, eps, vals, s); break;
      case 512: st = launch_normal_eval<512>(models[f], list, counts + f, n, hit_pos, eps, vals, s); break;
      default: st = fail(NIQ_ERR_UNSUPPORTED, "width");
    }
    if (st == NIQ_OK)
      hipLaunchKernelGGL(k_normal_finish, dim3(elem_grid(n)), dim3(256), 0, s, list, counts + f, vals, eps,
                         out_normals);
  }
  if (st == NIQ_OK && hipGetLastError() != hipSuccess) st = fail(NIQ_ERR_HIP, "normal kernels failed");
  (void)hipFreeAsync(scratch, s);
  return st;
}

extern "C" niq_status niq_shade_image(const niq_shade_opts* o, const float* ray_dirs, const float* normals,
                                      const int32_t* hit_ids, int64_t n, float* out_rgb, void* stream) {
  if (!o) return fail(NIQ_ERR_INVALID, "null options");
  if (n < 0 || (n > 0 && (!normals || !hit_ids || !out_rgb))) return fail(NIQ_ERR_INVALID, "bad buffers");
  if (o->shading != NIQ_SHADE_NORMAL && o->shading != NIQ_SHADE_MATCAP)
    return fail(NIQ_ERR_INVALID, "Unrecognized shading parameter");
  if (o->shading == NIQ_SHADE_MATCAP) {
    if (!ray_dirs) return fail(NIQ_ERR_INVALID, "matcap shading needs ray directions");
    for (int i = 0; i < 4; ++i)
      if (!o->matcaps[i]) return fail(NIQ_ERR_INVALID, "matcap shading needs the r, g, b, k matcaps");
    if (o->matcap_h < 1 || o->matcap_w < 1) return fail(NIQ_ERR_INVALID, "bad matcap size");
    if (o->n_colors < 0 || o->n_colors > 8) return fail(NIQ_ERR_INVALID, "need 0..8 colours");
  }
  if (o->tonemap && !(o->gamma > 0.f && o->white_level > 0.f)) return fail(NIQ_ERR_INVALID, "bad tonemap parameters");
  if (n == 0) return NIQ_OK;
  ShadeParams P{};
  P.shading = o->shading;
  for (int k = 0; k < 3; ++k) P.up[k] = o->up_dir[k];
  for (int i = 0; i < 4; ++i) P.mat[i] = o->matcaps[i];
  P.mh = o->matcap_h; P.mw = o->matcap_w;
  std::memcpy(P.colors, o->colors, sizeof(P.colors));
  P.n_colors = o->n_colors;
  P.color_per_hit = o->color_per_hit;
  P.composite = o->composite;
  P.tonemap = o->tonemap;
  P.gamma_inv = (float)(1.0 / (double)o->gamma);
  P.white2 = (float)((double)o->white_level * (double)o->white_level);
  P.exposure = o->exposure;
  hipLaunchKernelGGL(k_shade, dim3(elem_grid(n)), dim3(256), 0, (hipStream_t)stream, P, ray_dirs, normals, hit_ids,
                     (long long)n, out_rgb);
  HIPCHK(hipGetLastError());
  return NIQ_OK;
}
