// Device-resident level-set tree build.
#include "niq_internal.h"

extern "C" {

// ----------------------------------------------------------- tree build
namespace {

struct TreeLayout {
  long long cap;      // max nodes in any level list
  long long nbmax;    // max batches per level
  int top_levels;     // levels 0 .. top_levels - 1 are classified in one launch (0: none)
  long long heap_cap; // heap-index entries per level list
  size_t off_counters, off_codes, off_bcnt, off_boff, off_lvl[2], off_recheck, off_lw, lw_bytes;
  size_t off_top_box, off_top_codes, off_heap[2], total;
};

// The top levels of a tree hold at most 2^k nodes each and would each cost a
// whole launch of one-tile latency; instead every node of levels
// 0 .. kTopLevels - 1 (8191 boxes, two tiles per CU) is classified in one
// launch, and those levels take their codes from it by heap index.
constexpr int kTopLevels = 13;

int top_levels_for(const niq_eval_cfg* cfg, const niq_tree_params* p) {
  if (cfg && cfg->recheck_tol > 0.f) return 0;   // the FP64 re-test indexes the level list
  const int n_levels = p->split_depth >= 0 ? p->split_depth + 1 : 64;
  const int t = std::min(n_levels, kTopLevels);
  return t >= 3 ? t : 0;
}

// Boxes per chunk of the layer-by-layer precision mode inside a tree build.
constexpr long long kTreeLwChunk = 32768;

niq_status tree_layout(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, TreeLayout* L) {
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
  // activation scratch of the layer-by-layer precision mode (variant 3):
  // two ping-pong buffers of 10 bf16 planes at the widest layer
  L->lw_bytes = 0;
  if (m && cfg && cfg->precision == NIQ_PREC_BF16 && cfg->variant == 3) {
    const long long nch = std::min<long long>((cap + 63) / 64 * 64, kTreeLwChunk);
    L->lw_bytes = 2 * (size_t)nch * 10 * (size_t)std::max(m->host16.wmax, 32) * 2;
  }
  L->off_lw = take(L->lw_bytes);
  L->top_levels = top_levels_for(cfg, p);
  const long long n_top = L->top_levels ? (1LL << L->top_levels) - 1 : 0;
  L->heap_cap = L->top_levels ? std::min(cap, 1LL << (L->top_levels - 1)) : 0;
  L->off_top_box = take(sizeof(float) * 6 * n_top);
  L->off_top_codes = take(sizeof(int) * n_top);
  L->off_heap[0] = take(sizeof(int) * L->heap_cap);
  L->off_heap[1] = take(sizeof(int) * L->heap_cap);
  L->total = o;
  return NIQ_OK;
}

// Box of heap node h (root 1; children 2h = lower half, 2h + 1 = upper half
// along the widest axis, first on ties): the splits of the level loop
// (k_tree_scatter), replayed from the root in the same float32 operations.
__global__ void k_top_boxes(float l0, float l1, float l2, float u0, float u1, float u2, const float* root,
                            int n_top, float* __restrict__ lo, float* __restrict__ up) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (h > n_top) return;
  float l[3] = {l0, l1, l2}, u[3] = {u0, u1, u2};
  if (root) {
    for (int k = 0; k < 3; ++k) { l[k] = root[k]; u[k] = root[3 + k]; }
  }
  const int depth = 31 - __clz(h);
  for (int d = depth - 1; d >= 0; --d) {
    const float e0 = u[0] - l[0], e1 = u[1] - l[1], e2 = u[2] - l[2];
    int dim = 0; float em = e0;
    if (e1 > em) { dim = 1; em = e1; }
    if (e2 > em) { dim = 2; }
    const float m = 0.5f * (l[dim] + u[dim]);
    if ((h >> d) & 1) l[dim] = m; else u[dim] = m;
  }
  for (int k = 0; k < 3; ++k) { lo[3 * (h - 1) + k] = l[k]; up[3 * (h - 1) + k] = u[k]; }
}

// Codes of a top level's list from the one launch over the top of the tree.
__global__ void k_top_gather(const int* __restrict__ heap, const TreeCounters* __restrict__ C, int level,
                             const int* __restrict__ top_codes, int* __restrict__ codes) {
  const int n = C->n[level & 1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    codes[i] = top_codes[heap[i] - 1];
}

// Root box from the arguments, or from `root` (6 floats on the device:
// lower then upper) when it is given, so a captured build can be replayed for
// any root box.
__global__ void k_tree_init(TreeCounters* C, float* lo, float* up, float l0, float l1, float l2, float u0,
                            float u1, float u2, const float* root, int* heap0) {
  if (threadIdx.x == 0) {
    C->n[0] = 1; C->n[1] = 0; C->n_int = 0; C->n_ext = 0; C->n_classified = 0; C->overflow = 0;
    C->recheck_cnt = 0; C->n_recheck = 0;
    if (heap0) heap0[0] = 1;
    if (root) {
      l0 = root[0]; l1 = root[1]; l2 = root[2];
      u0 = root[3]; u1 = root[4]; u2 = root[5];
    }
    lo[0] = l0; lo[1] = l1; lo[2] = l2;
    up[0] = u0; up[1] = u1; up[2] = u2;
  }
  for (int i = threadIdx.x; i < 64; i += blockDim.x) { C->level_n[i] = 0; C->level_int[i] = 0; C->level_ext[i] = 0; }
}

}  // namespace

niq_status niq_tree_workspace_bytes(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p,
                                    size_t* bytes) {
  TreeLayout L;
  if (niq_status st = tree_layout(m, cfg, p, &L)) return st;
  if (bytes) *bytes = L.total;
  return NIQ_OK;
}

}  // extern "C"

namespace {

// Decode the device counters of a finished build into `res`.
void decode_counters(const TreeCounters& hc, int n_levels, niq_tree_result* res) {
  res->n_levels = n_levels;
  res->n_unknown = hc.n[n_levels & 1];
  res->n_interior = hc.n_int;
  res->n_exterior = hc.n_ext;
  res->n_classified = hc.n_classified;
  res->n_recheck = hc.n_recheck;
  for (int i = 0; i < 64; ++i) {
    res->level_n[i] = hc.level_n[i];
    res->level_interior[i] = hc.level_int[i];
    res->level_exterior[i] = hc.level_ext[i];
  }
}

// Enqueue every level of a build on `stream`.  With sync_allowed == false the
// call neither synchronises nor allocates (so it can be captured in a
// hipGraph); it then requires split_depth and a fused-kernel mode.
niq_status tree_enqueue(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, void* ws,
                        size_t ws_bytes, niq_tree_result* res, void* stream, bool sync_allowed, int* levels_out,
                        size_t* counters_offset, const float* root_dev) {
  if (niq_status st = check_model(m, 3)) return st;
  if (!res) return fail(NIQ_ERR_INVALID, "null result");
  TreeLayout L;
  if (niq_status st = tree_layout(m, cfg, p, &L)) return st;
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

  const int T = L.top_levels;
  int* heap[2] = {reinterpret_cast<int*>(base + L.off_heap[0]), reinterpret_cast<int*>(base + L.off_heap[1])};
  int* top_codes = reinterpret_cast<int*>(base + L.off_top_codes);
  hipLaunchKernelGGL(k_tree_init, dim3(1), dim3(64), 0, s, C, lvl[0], lvl[0] + 3 * cap, p->lower[0], p->lower[1],
                     p->lower[2], p->upper[0], p->upper[1], p->upper[2], root_dev, T ? heap[0] : nullptr);
  HIPCHK(hipGetLastError());
  if (T) {
    const int n_top = (1 << T) - 1;
    float* top_lo = reinterpret_cast<float*>(base + L.off_top_box);
    float* top_up = top_lo + 3 * (size_t)n_top;
    hipLaunchKernelGGL(k_top_boxes, dim3((n_top + 255) / 256), dim3(256), 0, s, p->lower[0], p->lower[1],
                       p->lower[2], p->upper[0], p->upper[1], p->upper[2], root_dev, n_top, top_lo, top_up);
    HIPCHK(hipGetLastError());
    ClassifyArgs a{};
    a.lower = top_lo; a.upper = top_up; a.d = 3; a.n = n_top; a.offset = p->offset;
    a.codes = top_codes;
    if (L.lw_bytes) { a.lw_buf = base + L.off_lw; a.lw_bytes = L.lw_bytes; }
    if (niq_status st = classify_dispatch(m, cfg, 3, a, n_top, s)) return st;
  }

  const int n_levels = p->split_depth >= 0 ? p->split_depth + 1 : 64;
  int level = 0;
  for (; level < std::min(n_levels, 62); ++level) {
    bool quit_next = (level + 1 == n_levels);
    if (p->node_terminate_thresh > 0 && sync_allowed) {
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
    if (L.lw_bytes) { a.lw_buf = base + L.off_lw; a.lw_bytes = L.lw_bytes; }
    if (cfg && cfg->recheck_tol > 0.f) {
      a.recheck_tol = cfg->recheck_tol;
      a.recheck_cnt = &C->recheck_cnt;
      a.recheck_idx = reinterpret_cast<int*>(base + L.off_recheck);
      a.recheck_cap = (int)std::min<long long>(cap, INT_MAX);
    }
    if (level < T) {
      hipLaunchKernelGGL(k_top_gather, dim3((unsigned)((max_n + 255) / 256)), dim3(256), 0, s, heap[level & 1], C,
                         level, top_codes, codes);
      HIPCHK(hipGetLastError());
    } else if (niq_status st = classify_dispatch(m, cfg, 3, a, max_n, s)) {
      return st;
    }

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
    if (split && level + 1 < T) {
      sa.heap = heap[level & 1];
      sa.next_heap = heap[(level + 1) & 1];
      sa.cap_heap = L.heap_cap;
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
  if (levels_out) *levels_out = level + 1;
  if (counters_offset) *counters_offset = L.off_counters;
  return NIQ_OK;
}

}  // namespace

extern "C" {

niq_status niq_tree_build(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, void* ws,
                          size_t ws_bytes, niq_tree_result* res, void* stream) {
  int levels = 0;
  size_t off = 0;
  if (niq_status st = tree_enqueue(m, cfg, p, ws, ws_bytes, res, stream, true, &levels, &off, nullptr)) return st;
  hipStream_t s = (hipStream_t)stream;
  TreeCounters hc;
  HIPCHK(hipMemcpyAsync(&hc, static_cast<unsigned char*>(ws) + off, sizeof(TreeCounters), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  decode_counters(hc, levels, res);
  if (hc.overflow) return fail(NIQ_ERR_CAPACITY, "a tree output buffer was too small");
  return NIQ_OK;
}

niq_status niq_tree_enqueue(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, void* ws,
                            size_t ws_bytes, niq_tree_result* res, size_t* counters_offset, void* stream) {
  return niqi::tree_enqueue_async(m, cfg, p, ws, ws_bytes, res, counters_offset, nullptr, stream);
}

size_t niq_tree_counters_bytes(void) { return sizeof(TreeCounters); }

niq_status niq_tree_decode(const void* host_counters, const niq_tree_params* p, niq_tree_result* res) {
  if (!host_counters || !p || !res || p->split_depth < 0) return fail(NIQ_ERR_INVALID, "bad arguments");
  TreeCounters hc;
  std::memcpy(&hc, host_counters, sizeof(hc));
  decode_counters(hc, p->split_depth + 1, res);
  if (hc.overflow) return fail(NIQ_ERR_CAPACITY, "a tree output buffer was too small");
  return NIQ_OK;
}

}  // extern "C"

namespace niqi {

niq_status tree_enqueue_async(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, void* ws,
                              size_t ws_bytes, niq_tree_result* res, size_t* counters_offset, const float* root_dev,
                              void* stream) {
  if (!p || p->split_depth < 0 || p->node_terminate_thresh > 0)
    return fail(NIQ_ERR_INVALID, "niq_tree_enqueue needs split_depth and no node_terminate_thresh");
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  if (mode == NIQ_MODE_AFFINE_ALL || mode == NIQ_MODE_AFFINE_TRUNCATE || mode == NIQ_MODE_AFFINE_APPEND)
    return fail(NIQ_ERR_UNSUPPORTED, "niq_tree_enqueue runs the fused-kernel modes only");
  if (m && classify_needs_scratch(m, cfg, 3))
    return fail(NIQ_ERR_UNSUPPORTED, "niq_tree_enqueue: this form needs per-call scratch (planes larger than LDS)");
  return tree_enqueue(m, cfg, p, ws, ws_bytes, res, stream, false, nullptr, counters_offset, root_dev);
}

}  // namespace niqi
