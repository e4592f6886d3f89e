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
__device__ __forceinline__ void top_box_body(const float* root6, int n_top, float* __restrict__ lo,
                                             float* __restrict__ up, int h) {
  if (h > n_top) return;
  float l[3] = {root6[0], root6[1], root6[2]}, u[3] = {root6[3], root6[4], root6[5]};
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

__global__ void k_top_boxes(float l0, float l1, float l2, float u0, float u1, float u2, const float* root,
                            int n_top, float* __restrict__ lo, float* __restrict__ up) {
  float r[6] = {l0, l1, l2, u0, u1, u2};
  if (root)
    for (int k = 0; k < 6; ++k) r[k] = root[k];
  top_box_body(r, n_top, lo, up, blockIdx.x * blockDim.x + threadIdx.x + 1);
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
__device__ __forceinline__ void k_tree_init_body(TreeCounters* C, float* lo, float* up, const float* root6,
                                                 int* heap0) {
  if (threadIdx.x == 0) {
    C->n[0] = 1; C->n[1] = 0; C->n_int = 0; C->n_ext = 0; C->n_classified = 0; C->overflow = 0;
    C->recheck_cnt = 0; C->n_recheck = 0;
    if (heap0) heap0[0] = 1;
    lo[0] = root6[0]; lo[1] = root6[1]; lo[2] = root6[2];
    up[0] = root6[3]; up[1] = root6[4]; up[2] = root6[5];
  }
  for (int i = threadIdx.x; i < 64; i += blockDim.x) { C->level_n[i] = 0; C->level_int[i] = 0; C->level_ext[i] = 0; }
}

__global__ void k_tree_init(TreeCounters* C, float* lo, float* up, float l0, float l1, float l2, float u0,
                            float u1, float u2, const float* root, int* heap0) {
  float r[6] = {l0, l1, l2, u0, u1, u2};
  if (root)
    for (int k = 0; k < 6; ++k) r[k] = root[k];
  k_tree_init_body(C, lo, up, r, heap0);
}

// ------------------------------------------------- batched builds
extern "C++" {   // templates below; the file's outer block has C linkage
// One tree of a batched build: its model, its workspace (the layout every
// tree of the batch shares, at its own base), root box and output lists.
struct TreeItem {
  const DevModel* m;
  unsigned char* ws;
  float root[6];
  float* unk_lo; float* unk_up; long long cap_unk;
  float* int_lo; float* int_up; long long cap_int;
  float* ext_lo; float* ext_up; long long cap_ext;
};

// Offsets of the shared layout, for the device.
struct TreeOffs {
  size_t counters, codes, bcnt, boff, lvl[2], top_box, top_codes, heap[2];
  long long cap, heap_cap;
  int n_top, top_levels, with_int, with_ext;
};

__device__ __forceinline__ TreeCounters* item_counters(const TreeItem& it, const TreeOffs& o) {
  return reinterpret_cast<TreeCounters*>(it.ws + o.counters);
}
__device__ __forceinline__ float* item_lvl(const TreeItem& it, const TreeOffs& o, int level) {
  return reinterpret_cast<float*>(it.ws + o.lvl[level & 1]);
}

__global__ void k_tree_init_b(const TreeItem* __restrict__ items, TreeOffs o) {
  const TreeItem& it = items[blockIdx.x];
  TreeCounters* C = item_counters(it, o);
  float* lo = item_lvl(it, o, 0);
  k_tree_init_body(C, lo, lo + 3 * o.cap, it.root, o.top_levels ? reinterpret_cast<int*>(it.ws + o.heap[0]) : nullptr);
}

__global__ void k_top_boxes_b(const TreeItem* __restrict__ items, TreeOffs o) {
  const TreeItem& it = items[blockIdx.y];
  float* lo = reinterpret_cast<float*>(it.ws + o.top_box);
  top_box_body(it.root, o.n_top, lo, lo + 3 * (size_t)o.n_top, blockIdx.x * blockDim.x + threadIdx.x + 1);
}

__global__ void k_top_gather_b(const TreeItem* __restrict__ items, TreeOffs o, int level) {
  const TreeItem& it = items[blockIdx.y];
  const TreeCounters* C = item_counters(it, o);
  const int* heap = reinterpret_cast<const int*>(it.ws + o.heap[level & 1]);
  const int* top_codes = reinterpret_cast<const int*>(it.ws + o.top_codes);
  int* codes = reinterpret_cast<int*>(it.ws + o.codes);
  const int n = C->n[level & 1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    codes[i] = top_codes[heap[i] - 1];
}

__global__ void __launch_bounds__(256) k_tree_count_b(const TreeItem* __restrict__ items, TreeOffs o, int level,
                                                      int B) {
  const TreeItem& it = items[blockIdx.y];
  tree_count_body(reinterpret_cast<const int*>(it.ws + o.codes), item_counters(it, o), level, B,
                  reinterpret_cast<int*>(it.ws + o.bcnt), blockIdx.x);
}

__global__ void __launch_bounds__(1024) k_tree_scan_b(const TreeItem* __restrict__ items, TreeOffs o, int level,
                                                      int B, int split) {
  const TreeItem& it = items[blockIdx.x];
  tree_scan_body(reinterpret_cast<const int*>(it.ws + o.bcnt), reinterpret_cast<int*>(it.ws + o.boff),
                 item_counters(it, o), level, B, split, split ? o.cap : it.cap_unk,
                 o.with_int ? it.cap_int : LLONG_MAX, o.with_ext ? it.cap_ext : LLONG_MAX);
}

__global__ void __launch_bounds__(256) k_tree_scatter_b(const TreeItem* __restrict__ items, TreeOffs o, int level,
                                                        int B, int split) {
  const TreeItem& it = items[blockIdx.y];
  TreeScatterArgs sa{};
  sa.codes = reinterpret_cast<const int*>(it.ws + o.codes);
  sa.lower = item_lvl(it, o, level);
  sa.upper = sa.lower + 3 * o.cap;
  if (split) {
    sa.next_lower = item_lvl(it, o, level + 1);
    sa.next_upper = sa.next_lower + 3 * o.cap;
    sa.cap_next = o.cap;
  } else {
    sa.next_lower = it.unk_lo; sa.next_upper = it.unk_up; sa.cap_next = it.cap_unk;
  }
  if (split && level + 1 < o.top_levels) {
    sa.heap = reinterpret_cast<const int*>(it.ws + o.heap[level & 1]);
    sa.next_heap = reinterpret_cast<int*>(it.ws + o.heap[(level + 1) & 1]);
    sa.cap_heap = o.heap_cap;
  }
  if (o.with_int) { sa.int_lower = it.int_lo; sa.int_upper = it.int_up; sa.cap_int = it.cap_int; }
  if (o.with_ext) { sa.ext_lower = it.ext_lo; sa.ext_upper = it.ext_up; sa.cap_ext = it.cap_ext; }
  tree_scatter_body(sa, reinterpret_cast<const int*>(it.ws + o.bcnt), reinterpret_cast<const int*>(it.ws + o.boff),
                    item_counters(it, o), level, B, split, blockIdx.x);
}

// Tile offsets of one level over the trees of a batch: pref[g] = tiles (of nb
// boxes) of the trees before g, pref[ng] = all.  level -1: the top launch.
__global__ void k_items_prefix(const TreeItem* __restrict__ items, int ng, TreeOffs o, int level, int nb,
                               int* __restrict__ pref) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  for (int g = 0; g < ng; ++g) {
    pref[g] = acc;
    const int n = level < 0 ? o.n_top : item_counters(items[g], o)->n[level & 1];
    acc += (n + nb - 1) / nb;
  }
  pref[ng] = acc;
}

// The classify pass of one level of every tree of a batch in one launch: the
// trees' tiles of NB boxes are numbered one tree after another (pref) and
// workgroups walk them; each tile runs its own tree's model.  level -1 is the
// launch over the top of the trees.  Every (box, feature) sums as in
// k_classify: the codes are a lone build's bit for bit.
template <int KIND, int V, int NB, int NW, int WMAX>
__global__ void __launch_bounds__(NW * 64, (LdsOcc<Planes<KIND, V, 0>::R, NB, NW, WMAX>::value))
k_classify_items(const TreeItem* __restrict__ items, int ng, TreeOffs o, int level, float offset,
                 const int* __restrict__ pref) {
  using PL = Planes<KIND, V, 0>;
  constexpr int NT = NW * 64;
  __shared__ float X[PL::R * NB * WMAX];
  const int total = pref[ng];
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    int g = 0, hi = ng - 1;   // the last tree whose first tile is at or before t
    while (g < hi) {
      const int mid = (g + hi + 1) >> 1;
      if (pref[mid] <= t) g = mid; else hi = mid - 1;
    }
    const int first = pref[g];
    const TreeItem& it = items[g];
    ClassifyArgs a{};
    int* codes;
    if (level < 0) {
      a.lower = reinterpret_cast<const float*>(it.ws + o.top_box);
      a.upper = a.lower + 3 * (size_t)o.n_top;
      a.n = o.n_top;
      codes = reinterpret_cast<int*>(it.ws + o.top_codes);
    } else {
      a.lower = item_lvl(it, o, level);
      a.upper = a.lower + 3 * o.cap;
      a.n = item_counters(it, o)->n[level & 1];
      codes = reinterpret_cast<int*>(it.ws + o.codes);
    }
    a.d = 3;
    const long long i0 = (long long)(t - first) * NB, n = a.n;
    fill_box_inputs<KIND, V, NB, NT, WMAX>(X, a, i0, n);
    mlp_forward<PL, KIND, V, NB, NW, WMAX>(*it.m, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i >= n) return;
      float lo, hi2, scale;
      final_bounds<KIND, V>(v, lo, hi2, scale);
      codes[i] = sign_code(lo, hi2, offset);
    });
  }
}

template <int KIND, int V, int W>
niq_status launch_classify_items_w(const TreeItem* items, int ng, const TreeOffs& o, int level, float offset,
                                   long long max_n, int* pref, hipStream_t s) {
  if constexpr (!Fits<Planes<KIND, V, 0>::R, TileCfg<W>::NB, W>::value) {
    return fail(NIQ_ERR_UNSUPPORTED, "batched tree build: planes above LDS");
  } else {
    constexpr int NB = TileCfg<W>::NB, NW = TileCfg<W>::NW;
    hipLaunchKernelGGL(k_items_prefix, dim3(1), dim3(64), 0, s, items, ng, o, level, NB, pref);
    auto k = k_classify_items<KIND, V, NB, NW, W>;
    const long long max_tiles = (long long)ng * ((max_n + NB - 1) / NB);
    const unsigned g = tile_grid(max_tiles, persistent_blocks(k, NW * 64));
    hipLaunchKernelGGL(k, dim3(g), dim3(NW * 64), 0, s, items, ng, o, level, offset, pref);
    HIPCHK(hipGetLastError());
    return NIQ_OK;
  }
}

// (KIND, V) of the fused classify of a config: 0 interval, 1 affine_fixed
// over 3 box vectors, 2 slope_interval; -1 none.
int batch_form(const niq_eval_cfg* cfg) {
  const int mode = cfg ? cfg->mode : NIQ_MODE_AFFINE_FIXED;
  if (cfg && (cfg->precision != NIQ_PREC_FP32 || cfg->recheck_tol > 0.f)) return -1;
  if (mode == NIQ_MODE_INTERVAL) return 0;
  if (mode == NIQ_MODE_AFFINE_FIXED) return 1;
  if (mode == NIQ_MODE_SLOPE_INTERVAL) return 2;
  return -1;
}

template <int W>
niq_status launch_classify_items(int form, const TreeItem* items, int ng, const TreeOffs& o, int level,
                                 float offset, long long max_n, int* pref, hipStream_t s) {
  switch (form) {
    case 0: return launch_classify_items_w<0, 0, W>(items, ng, o, level, offset, max_n, pref, s);
    case 1: return launch_classify_items_w<0, 3, W>(items, ng, o, level, offset, max_n, pref, s);
    case 2: return launch_classify_items_w<1, 3, W>(items, ng, o, level, offset, max_n, pref, s);
  }
  return fail(NIQ_ERR_UNSUPPORTED, "batched tree build: mode");
}
}  // extern "C++"

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


// The table of trees and the tile offsets ahead of the trees' workspaces.
static size_t batch_head_bytes(int ng) {
  return (sizeof(TreeItem) * ng + 255) / 256 * 256 + (sizeof(int) * (ng + 1) + 255) / 256 * 256;
}

bool tree_batchable(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p) {
  return p && p->split_depth >= 0 && p->node_terminate_thresh <= 0 && batch_form(cfg) >= 0 &&
         !classify_needs_scratch(m, cfg, 3);
}

niq_status tree_batch_bytes(const niq_model* m, const niq_eval_cfg* cfg, const niq_tree_params* p, int ng,
                            size_t* bytes) {
  if (ng < 1 || ng > kMaxTreeBatch) return fail(NIQ_ERR_INVALID, "tree batch size out of range");
  TreeLayout L;
  if (niq_status st = tree_layout(m, cfg, p, &L)) return st;
  *bytes = batch_head_bytes(ng) + L.total * (size_t)ng;
  return NIQ_OK;
}

niq_status tree_batch_build(const niq_model* const* models, int ng, const niq_eval_cfg* cfg,
                            const niq_tree_params* p, const float* roots, void* ws, size_t ws_bytes,
                            niq_tree_result* const* res, niq_status* tree_status, hipStream_t s) {
  if (ng < 1 || ng > kMaxTreeBatch) return fail(NIQ_ERR_INVALID, "tree batch size out of range");
  const int wt = models[0]->wt;
  for (int g = 0; g < ng; ++g) {
    if (niq_status st = check_model(models[g], 3)) return st;
    if (models[g]->wt != wt) return fail(NIQ_ERR_INVALID, "a tree batch needs one template width");
    if (!tree_batchable(models[g], cfg, p)) return fail(NIQ_ERR_UNSUPPORTED, "tree not batchable");
    const niq_tree_result* r = res[g];
    if (!r->unknown_lower || !r->unknown_upper || r->unknown_capacity < 0)
      return fail(NIQ_ERR_INVALID, "unknown-node output buffers are required");
    if (p->with_interior_nodes && (!r->interior_lower || !r->interior_upper))
      return fail(NIQ_ERR_INVALID, "interior output buffers are required");
    if (p->with_exterior_nodes && (!r->exterior_lower || !r->exterior_upper))
      return fail(NIQ_ERR_INVALID, "exterior output buffers are required");
  }
  const int form = batch_form(cfg);
  TreeLayout L;
  if (niq_status st = tree_layout(models[0], cfg, p, &L)) return st;
  size_t need = 0;
  if (niq_status st = tree_batch_bytes(models[0], cfg, p, ng, &need)) return st;
  if (!ws || ws_bytes < need) return fail(NIQ_ERR_CAPACITY, "tree batch workspace too small");
  unsigned char* base = static_cast<unsigned char*>(ws);
  TreeItem* d_items = reinterpret_cast<TreeItem*>(base);
  int* d_pref = reinterpret_cast<int*>(base + (sizeof(TreeItem) * ng + 255) / 256 * 256);
  unsigned char* trees = base + batch_head_bytes(ng);
  std::vector<TreeItem> items(ng);
  for (int g = 0; g < ng; ++g) {
    TreeItem& it = items[g];
    const niq_tree_result* r = res[g];
    it.m = models[g]->dev;
    it.ws = trees + L.total * (size_t)g;
    for (int k = 0; k < 6; ++k) it.root[k] = roots[6 * g + k];
    it.unk_lo = static_cast<float*>(r->unknown_lower); it.unk_up = static_cast<float*>(r->unknown_upper);
    it.cap_unk = r->unknown_capacity;
    it.int_lo = static_cast<float*>(r->interior_lower); it.int_up = static_cast<float*>(r->interior_upper);
    it.cap_int = p->with_interior_nodes ? r->interior_capacity : 0;
    it.ext_lo = static_cast<float*>(r->exterior_lower); it.ext_up = static_cast<float*>(r->exterior_upper);
    it.cap_ext = p->with_exterior_nodes ? r->exterior_capacity : 0;
  }
  TreeOffs o{};
  o.counters = L.off_counters; o.codes = L.off_codes; o.bcnt = L.off_bcnt; o.boff = L.off_boff;
  o.lvl[0] = L.off_lvl[0]; o.lvl[1] = L.off_lvl[1]; o.top_box = L.off_top_box; o.top_codes = L.off_top_codes;
  o.heap[0] = L.off_heap[0]; o.heap[1] = L.off_heap[1];
  o.cap = L.cap; o.heap_cap = L.heap_cap;
  o.top_levels = L.top_levels;
  o.n_top = L.top_levels ? (1 << L.top_levels) - 1 : 0;
  o.with_int = p->with_interior_nodes ? 1 : 0; o.with_ext = p->with_exterior_nodes ? 1 : 0;
  const int B = p->batch_process_size;
  const int n_levels = p->split_depth + 1;
  std::vector<TreeCounters> hc(ng);
  niq_status st = NIQ_OK;
  auto hipfail = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  auto classify = [&](int level, long long max_n) {
    if (st != NIQ_OK) return;
    switch (wt) {
      case 64: st = launch_classify_items<64>(form, d_items, ng, o, level, p->offset, max_n, d_pref, s); break;
      case 128: st = launch_classify_items<128>(form, d_items, ng, o, level, p->offset, max_n, d_pref, s); break;
      case 256: st = launch_classify_items<256>(form, d_items, ng, o, level, p->offset, max_n, d_pref, s); break;
      case 512: st = launch_classify_items<512>(form, d_items, ng, o, level, p->offset, max_n, d_pref, s); break;
      default: st = fail(NIQ_ERR_UNSUPPORTED, "width");
    }
  };
  hipfail(hipMemcpyAsync(d_items, items.data(), sizeof(TreeItem) * ng, hipMemcpyHostToDevice, s), "tree batch table");
  if (st == NIQ_OK) {
    hipLaunchKernelGGL(k_tree_init_b, dim3(ng), dim3(64), 0, s, d_items, o);
    hipfail(hipGetLastError(), "k_tree_init_b");
  }
  if (st == NIQ_OK && o.top_levels) {
    hipLaunchKernelGGL(k_top_boxes_b, dim3((o.n_top + 255) / 256, ng), dim3(256), 0, s, d_items, o);
    hipfail(hipGetLastError(), "k_top_boxes_b");
    classify(-1, o.n_top);
  }
  for (int level = 0; level < n_levels && st == NIQ_OK; ++level) {
    const int split = level + 1 < n_levels ? 1 : 0;
    const long long max_n = level < 40 ? std::min(L.cap, 1LL << level) : L.cap;
    if (level < o.top_levels) {
      hipLaunchKernelGGL(k_top_gather_b, dim3((unsigned)((max_n + 255) / 256), ng), dim3(256), 0, s, d_items, o,
                         level);
      hipfail(hipGetLastError(), "k_top_gather_b");
    } else {
      classify(level, max_n);
    }
    if (st != NIQ_OK) break;
    const unsigned nb = (unsigned)((max_n + B - 1) / B);
    hipLaunchKernelGGL(k_tree_count_b, dim3(nb, ng), dim3(256), 0, s, d_items, o, level, B);
    hipLaunchKernelGGL(k_tree_scan_b, dim3(ng), dim3(1024), 0, s, d_items, o, level, B, split);
    hipLaunchKernelGGL(k_tree_scatter_b, dim3(nb, ng), dim3(256), 0, s, d_items, o, level, B, split);
    hipfail(hipGetLastError(), "tree batch step");
  }
  for (int g = 0; g < ng && st == NIQ_OK; ++g)
    hipfail(hipMemcpyAsync(&hc[g], items[g].ws + L.off_counters, sizeof(TreeCounters), hipMemcpyDeviceToHost, s),
            "tree batch counters");
  // the table and counters live in host vectors: wait for their copies on
  // every path
  const hipError_t e = hipStreamSynchronize(s);
  hipfail(e, "hipStreamSynchronize");
  if (st != NIQ_OK) return st;
  bool overflow = false;
  for (int g = 0; g < ng; ++g) {
    decode_counters(hc[g], n_levels, res[g]);
    overflow = overflow || hc[g].overflow;
    // each tree fails as its lone build would: only when its own lists overflow
    if (tree_status) tree_status[g] = hc[g].overflow ? NIQ_ERR_CAPACITY : NIQ_OK;
  }
  if (overflow) return fail(NIQ_ERR_CAPACITY, "a tree output buffer was too small");
  return NIQ_OK;
}

}  // namespace niqi
