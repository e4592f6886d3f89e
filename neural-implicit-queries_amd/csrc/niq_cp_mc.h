// Closest-point search and marching-cubes extraction kernels.
//
// closest_point follows the reference's batch-stack semantics exactly
// (src/kd_tree.py:659-802): each round pops the top B work items, classifies
// them and samples 7 points per node (with the fused MFMA forward), then one
// workgroup applies the scatter-min, the location update and the stable
// compaction + interleaved push of the children.  The stack, the query minima
// and every count stay on the device.
//
// Marching cubes follows extract_triangles_from_subcells / _from_cell
// (src/extract_cell.py:314-421): the (2^s + 1)^3 sample grid of every leaf is
// evaluated by the point kernel, then one workgroup per leaf forms the cube
// cases of its 8^s subcells and writes their triangles in the reference's
// (leaf, subcell, slot) order at offsets from a device prefix sum.
#pragma once
#include <rocprim/block/block_scan.hpp>

#include "niq_kernels.h"

// ------------------------------------------------------------ closest point
struct CpState {
  int top;              // stack top
  int n;                // items of the current batch
  int pop;              // position the batch was popped from
  int overflow;
  long long n_processed;
  long long rounds;
};

struct CpArgs {
  int B;                        // batch_process_size
  long long cap;                // stack capacity (rows)
  int* wq; float* wlo; float* wup;          // stack: query id, node lower/upper (cap, 3)
  int* bq; float* blo; float* bup;          // popped batch (B), (B, 3)
  float* pts;                   // 7 sample points per item (7B, 3)
  int* codes;                   // (B,)
  float* vals;                  // (7B,)
  const float* q;               // query points (Q, 3)
  long long Q;                  // query count
  float* qmin;                  // (Q,) running minimum distance
  float* qloc;                  // (Q, 3)
  int* win;                     // (Q,) scratch, -1 between rounds
  float* idist; int* ineeds;    // per-item scratch (B,)
  float eps_cube_width;
};

// Pop the top B items into the batch buffers and form their sample points
// (src/kd_tree.py:679-686, 703-704).
static __global__ void __launch_bounds__(256) k_cp_gather(CpState* S, CpArgs a) {
  const int top = S->top;
  const int pop = max(top - a.B, 0);
  const int n = top - pop;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const long long s = pop + j;
    a.bq[j] = a.wq[s];
    float lo[3], up[3], c[3], e[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = a.wlo[3 * s + k]; up[k] = a.wup[3 * s + k];
      a.blo[3 * j + k] = lo[k]; a.bup[3 * j + k] = up[k];
      c[k] = 0.5f * (lo[k] + up[k]);
      e[k] = up[k] - lo[k];
    }
    // offsets [0; I; -I] in this order
    for (int t = 0; t < 7; ++t) {
      for (int k = 0; k < 3; ++k) {
        const float o = (t == 0) ? 0.f : ((t - 1) % 3 == k ? (t <= 3 ? 1.f : -1.f) : 0.f);
        a.pts[(7 * j + t) * 3 + k] = c[k] + e[k] * o;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->n = n;
    S->pop = pop;
    S->n_processed += n;
    if (n > 0) S->rounds += 1;
  }
}

__device__ __forceinline__ float norm3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }

// A value another thread of this round changed with an atomic, read past the
// caches (device scope).
template <class T>
__device__ __forceinline__ T load_fresh(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Push the two children of a subdivided item at rank r of the stable
// compaction: split along the longest axis (first on ties), interleaved
// A0 B0 A1 B1 ... at the pop position (src/kd_tree.py:735-757).
__device__ __forceinline__ void cp_push_children(const CpArgs& a, int pop, int r, int qid, const float* lo,
                                                 const float* up) {
  const float e0 = up[0] - lo[0], e1 = up[1] - lo[1], e2 = up[2] - lo[2];
  int dim = 0; float em = e0;
  if (e1 > em) { dim = 1; em = e1; }
  if (e2 > em) dim = 2;
  const float mid = 0.5f * (lo[dim] + up[dim]);
  const long long pa = (long long)pop + 2LL * r, pb = pa + 1;
  if (pb < a.cap) {
    for (int k = 0; k < 3; ++k) {
      a.wlo[3 * pa + k] = lo[k];
      a.wup[3 * pa + k] = (k == dim) ? mid : up[k];
      a.wlo[3 * pb + k] = (k == dim) ? mid : lo[k];
      a.wup[3 * pb + k] = up[k];
    }
    a.wq[pa] = qid;
    a.wq[pb] = qid;
  }
}

// The per-item part of the update: the distance the item offers its query
// (its centre distance plus the node diagonal when the samples change sign)
// and whether it is subdivided (src/kd_tree.py:692-716).
__device__ __forceinline__ void cp_item(const CpArgs& a, int j, int qid, const float* lo, const float* up,
                                        float qmin_pre, float& idist, int& needs) {
  float c[3], e[3];
  for (int k = 0; k < 3; ++k) { c[k] = 0.5f * (lo[k] + up[k]); e[k] = up[k] - lo[k]; }
  const float width = fmaxf(fmaxf(e[0], e[1]), e[2]);
  const float diag = norm3(e[0], e[1], e[2]);
  const float dc = norm3(a.q[3 * qid] - c[0], a.q[3 * qid + 1] - c[1], a.q[3 * qid + 2] - c[2]);
  const bool is_small = width < a.eps_cube_width;
  const int code = a.codes[j];
  const bool outside = code == CODE_NEGATIVE || code == CODE_POSITIVE;
  bool all_neg = true, all_pos = true;
  for (int t = 0; t < 7; ++t) {
    const float v = a.vals[7 * j + t];
    all_neg = all_neg && (v < 0.f);
    all_pos = all_pos && (v > 0.f);
  }
  idist = !(all_neg || all_pos) ? dc + diag : INFINITY;
  needs = (!outside && !is_small && dc < qmin_pre) ? 1 : 0;
}

// Queries whose running minimum and winning item the register-resident update
// keeps in LDS for the round (one coalesced load and store of the minima
// instead of scattered L2 atomics and reads per item).
constexpr int kCpLdsQ = 4096;

// The update for n <= 2 NT items (the default batch of 2048 on 1024
// threads): item j = threadIdx + k NT stays in thread registers through all
// phases, and the stable compaction scans slice k = 0, 1 in turn, so ranks
// follow the item order as in cp_update_body.  With at most kCpLdsQ queries
// the minima and winners live in LDS for the round: the same atomics (integer
// min over non-negative float bits, max item index) in the same phases, so
// the results are those of the global form.
template <int NT>
__device__ __forceinline__ void cp_update_regs(CpState* S, const CpArgs& a, int n, int pop) {
  using Scan = rocprim::block_scan<int, NT>;
  __shared__ typename Scan::storage_type scan_storage;
  __shared__ float s_qmin[kCpLdsQ];
  __shared__ int s_win[kCpLdsQ];
  const bool lds = a.Q <= kCpLdsQ;   // uniform
  const int nq = (int)a.Q;
  if (lds) {
    for (int i = threadIdx.x; i < nq; i += NT) { s_qmin[i] = a.qmin[i]; s_win[i] = -1; }
    __syncthreads();
  }
  constexpr int KI = 2;
  int qid[KI], needs[KI];
  float lo[KI][3], up[KI][3], idist[KI];
#pragma unroll
  for (int k = 0; k < KI; ++k) {
    const int j = threadIdx.x + k * NT;
    needs[k] = 0; idist[k] = INFINITY; qid[k] = 0;
    if (j < n) {
      qid[k] = a.bq[j];
      for (int d = 0; d < 3; ++d) { lo[k][d] = a.blo[3 * j + d]; up[k][d] = a.bup[3 * j + d]; }
      const float pre = lds ? s_qmin[qid[k]] : load_fresh(&a.qmin[qid[k]]);
      cp_item(a, j, qid[k], lo[k], up[k], pre, idist[k], needs[k]);
    }
  }
  __syncthreads();   // every pre-round minimum is read before the scatter
#pragma unroll
  for (int k = 0; k < KI; ++k)
    if (threadIdx.x + k * NT < n && idist[k] < INFINITY) {
      int* m = lds ? reinterpret_cast<int*>(&s_qmin[qid[k]]) : reinterpret_cast<int*>(&a.qmin[qid[k]]);
      atomicMin(m, __float_as_int(idist[k]));
    }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KI; ++k) {
    const int j = threadIdx.x + k * NT;
    if (j < n && idist[k] == (lds ? s_qmin[qid[k]] : load_fresh(&a.qmin[qid[k]])))
      atomicMax(lds ? &s_win[qid[k]] : &a.win[qid[k]], j);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KI; ++k) {
    const int j = threadIdx.x + k * NT;
    if (j < n && (lds ? s_win[qid[k]] : load_fresh(&a.win[qid[k]])) == j)
      for (int d = 0; d < 3; ++d) a.qloc[3 * qid[k] + d] = 0.5f * (lo[k][d] + up[k][d]);
  }
  __syncthreads();
  if (lds) {
    for (int i = threadIdx.x; i < nq; i += NT) a.qmin[i] = s_qmin[i];
  } else {
#pragma unroll
    for (int k = 0; k < KI; ++k)
      if (threadIdx.x + k * NT < n) a.win[qid[k]] = -1;
  }
  int base = 0;
#pragma unroll
  for (int k = 0; k < KI; ++k) {
    int r = 0, tot = 0;
    Scan().exclusive_scan(needs[k], r, 0, tot, scan_storage);
    if (needs[k]) cp_push_children(a, pop, base + r, qid[k], lo[k], up[k]);
    base += tot;
    __syncthreads();   // the scan storage is reused
  }
  if (threadIdx.x == 0) {
    long long nt = (long long)pop + 2LL * base;
    if (nt > a.cap) { S->overflow = 1; nt = a.cap; }
    S->top = (int)nt;
  }
}

// The round's update over the n popped items, by the NT threads of one
// workgroup (src/kd_tree.py:692-760): distances, scatter-min, location, and
// the stable compaction with the interleaved push at the pop position.
template <int NT>
__device__ __forceinline__ void cp_update_body(CpState* S, const CpArgs& a, int n, int pop) {
  if (n <= 2 * NT) {
    cp_update_regs<NT>(S, a, n, pop);
    return;
  }
  using Scan = rocprim::block_scan<int, NT>;
  __shared__ typename Scan::storage_type scan_storage;
  // phase 1: per-item distance and subdivision flag (pruning uses the minimum
  // from before this round's scatter)
  for (int j = threadIdx.x; j < n; j += NT) {
    float lo[3], up[3];
    for (int k = 0; k < 3; ++k) { lo[k] = a.blo[3 * j + k]; up[k] = a.bup[3 * j + k]; }
    float idist;
    int needs;
    cp_item(a, j, a.bq[j], lo, up, load_fresh(&a.qmin[a.bq[j]]), idist, needs);
    a.idist[j] = idist;
    a.ineeds[j] = needs;
  }
  __syncthreads();
  // phase 2: scatter-min (non-negative floats order as their bit patterns)
  for (int j = threadIdx.x; j < n; j += NT) {
    const float d = a.idist[j];
    if (d < INFINITY) atomicMin(reinterpret_cast<int*>(&a.qmin[a.bq[j]]), __float_as_int(d));
  }
  __syncthreads();
  // phase 3: items equal to the new minimum claim the location; the last
  // such batch entry wins
  for (int j = threadIdx.x; j < n; j += NT) {
    const int qid = a.bq[j];
    if (a.idist[j] == load_fresh(&a.qmin[qid])) atomicMax(&a.win[qid], j);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += NT) {
    const int qid = a.bq[j];
    if (load_fresh(&a.win[qid]) == j) {
      for (int k = 0; k < 3; ++k) a.qloc[3 * qid + k] = 0.5f * (a.blo[3 * j + k] + a.bup[3 * j + k]);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += NT) a.win[a.bq[j]] = -1;
  // phase 4: stable compaction of the subdivided items, children pushed
  // interleaved A0 B0 A1 B1 ... at the pop position
  const int per = (n + NT - 1) / NT;
  const int j0 = threadIdx.x * per, j1 = min(j0 + per, n);
  int cnt = 0;
  for (int j = j0; j < j1; ++j) cnt += a.ineeds[j];
  int r = 0, total = 0;
  Scan().exclusive_scan(cnt, r, 0, total, scan_storage);
  for (int j = j0; j < j1; ++j) {
    if (!a.ineeds[j]) continue;
    float lo[3], up[3];
    for (int k = 0; k < 3; ++k) { lo[k] = a.blo[3 * j + k]; up[k] = a.bup[3 * j + k]; }
    cp_push_children(a, pop, r, a.bq[j], lo, up);
    ++r;
  }
  if (threadIdx.x == 0) {
    long long nt = (long long)pop + 2LL * total;
    if (nt > a.cap) { S->overflow = 1; nt = a.cap; }
    S->top = (int)nt;
  }
}

// One workgroup applies the update of the round k_cp_gather popped.
static __global__ void __launch_bounds__(1024) k_cp_update(CpState* S, CpArgs a) {
  const int n = S->n;
  if (n == 0) return;
  cp_update_body<1024>(S, a, n, S->pop);
}

// The sample point t of a node: its centre plus (upper - lower) times the
// offset row t of [0; I; -I] (src/kd_tree.py:703-704).
__device__ __forceinline__ float cp_sample(const float* lo, const float* up, int t, int k) {
  const float c = 0.5f * (lo[k] + up[k]);
  const float e = up[k] - lo[k];
  const float o = (t == 0) ? 0.f : ((t - 1) % 3 == k ? (t <= 3 ? 1.f : -1.f) : 0.f);
  return c + e * o;
}

// The forward half of a closest-point round in one launch (the update is
// k_cp_update, the next launch on the stream).  Workgroups [0, g_box) take 16
// popped nodes each through the box forward (Planes<KIND, V>, the forward of
// k_classify) and copy them to the batch buffers; the others take 16 of the
// 7n sample points each through the point forward (the forward of
// k_points_dev), so box and point tiles fill the chip together.  Every
// (node, feature) and (point, feature) sums in the order of those kernels:
// codes and values are theirs bit for bit.
template <int KIND, int V, int NW, int WMAX, int AS = ACTS_ALL>
__global__ void __launch_bounds__(NW * 64, (LdsOcc<Planes<KIND, V, 0>::R, 16, NW, WMAX>::value))
k_cp_forward(const DevModel* __restrict__ Mp, CpState* S, CpArgs a, int g_box) {
  using PB = Planes<KIND, V, 0>;
  using PP = Planes<2, 0, 1>;
  constexpr int NB = 16, NT = NW * 64;
  __shared__ float X[PB::R * NB * WMAX];
  const int top = S->top;
  const int pop = max(top - a.B, 0);
  const int n = top - pop;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->n = n;
    S->pop = pop;
    S->n_processed += n;
    if (n > 0) S->rounds += 1;
  }
  if ((int)blockIdx.x < g_box) {
    const long long i0 = (long long)blockIdx.x * NB;
    if (i0 >= n) return;
    for (int b = threadIdx.x; b < NB; b += NT) {
      const long long i = i0 + b;
      if (i < n) {
        const long long s = pop + i;
        a.bq[i] = a.wq[s];
        for (int k = 0; k < 3; ++k) { a.blo[3 * i + k] = a.wlo[3 * s + k]; a.bup[3 * i + k] = a.wup[3 * s + k]; }
      }
    }
    ClassifyArgs ca{};
    ca.lower = a.wlo + 3LL * pop;
    ca.upper = a.wup + 3LL * pop;
    ca.d = 3;
    fill_box_inputs<KIND, V, NB, NT, WMAX>(X, ca, i0, n);
    mlp_forward<PB, KIND, V, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long i = i0 + b;
      if (i >= n) return;
      float lo, hi, scale;
      final_bounds<KIND, V>(v, lo, hi, scale);
      a.codes[i] = sign_code(lo, hi, 0.f);
    });
  } else {
    const long long p0 = (long long)(blockIdx.x - g_box) * NB;
    const long long np = 7LL * n;
    if (p0 >= np) return;
    for (int e = threadIdx.x; e < NB * 16; e += NT) {
      const int b = e >> 4, k = e & 15;
      const long long q = p0 + b;
      float val = 0.f;
      if (q < np && k < 3) {
        const long long s = pop + q / 7;
        val = cp_sample(a.wlo + 3 * s, a.wup + 3 * s, (int)(q % 7), k);
      }
      X[xoff<NB, WMAX>(0, b, k)] = val;
    }
    mlp_forward<PP, 2, 0, NB, NW, WMAX, AS>(*Mp, X, [&](const float* v, int b) {
      const long long q = p0 + b;
      if (q < np) a.vals[q] = v[0];
    });
  }
}

// -------------------------------------------------------- marching cubes
// Bourke's public triangle table, one case per entry: up to 5 triangles of 3
// edge ids, -1 padded.  Filled by the host from its packed form.
struct McTable { signed char tri[256][16]; };

// Grid points of leaves [l0, l0 + nl): (2^s + 1)^3 per leaf, 'ij' order,
// coordinates lower + i * (upper - lower) / 2^s, the last one = upper.
static __global__ void __launch_bounds__(256) k_mc_grid(const float* __restrict__ lo, const float* __restrict__ up,
                                               long long l0, long long nl, int s, float* __restrict__ pts) {
  const int np = (1 << s) + 1;
  const long long total = nl * np * np * np;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const long long leaf = l0 + g / (np * np * np);
    const int rem = (int)(g % (np * np * np));
    const int ii[3] = {rem / (np * np), (rem / np) % np, rem % np};
    for (int k = 0; k < 3; ++k) {
      const float l = lo[3 * leaf + k], u = up[3 * leaf + k];
      const float step = (u - l) / (float)(1 << s);
      pts[3 * g + k] = (ii[k] == np - 1) ? u : l + (float)ii[k] * step;
    }
  }
}

// vertex v at logical corner (x, y, z) bits (src/extract_cell.py:293-304)
__constant__ int c_vert_logical[8][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0},
                                         {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};
// edge -> vertex pair (src/extract_cell.py:274-289)
__constant__ int c_edge_verts[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                        {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

__device__ __forceinline__ int mc_case(const float* vals, long long leaf_rel, int s, int sub, float* vv) {
  const int n = 1 << s, np = n + 1;
  const int si = sub / (n * n), sj = (sub / n) % n, sk = sub % n;
  int cid = 0;
  for (int v = 0; v < 8; ++v) {
    const int gi = si + c_vert_logical[v][0], gj = sj + c_vert_logical[v][1], gk = sk + c_vert_logical[v][2];
    vv[v] = vals[leaf_rel * np * np * np + (gi * np + gj) * np + gk];
    cid |= (vv[v] < 0.f) ? (1 << v) : 0;
  }
  return cid;
}

// Per-leaf triangle counts (one workgroup per leaf, threads over subcells).
static __global__ void __launch_bounds__(256) k_mc_count(const McTable* __restrict__ T, const float* __restrict__ vals,
                                                long long nl, int s, long long* __restrict__ cnt) {
  const int nsub = 1 << (3 * s);
  __shared__ int red[256];
  for (long long leaf = blockIdx.x; leaf < nl; leaf += gridDim.x) {
    int c = 0;
    for (int sub = threadIdx.x; sub < nsub; sub += blockDim.x) {
      float vv[8];
      const int cid = mc_case(vals, leaf, s, sub, vv);
      for (int t = 0; t < 5; ++t) c += (T->tri[cid][3 * t] >= 0);
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) cnt[leaf] = red[0];
    __syncthreads();
  }
}

// Exclusive prefix sum of n int64 counts in one workgroup (chunked).
static __global__ void __launch_bounds__(1024) k_scan_i64(const long long* __restrict__ in, long long n,
                                                 long long base, long long* __restrict__ out,
                                                 long long* __restrict__ total) {
  __shared__ long long s[1024];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = base;
  __syncthreads();
  for (long long b0 = 0; b0 < n; b0 += 1024) {
    const long long i = b0 + threadIdx.x;
    const long long v = i < n ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1) {
      const long long a = (int)threadIdx.x >= w ? s[threadIdx.x - w] : 0;
      __syncthreads();
      s[threadIdx.x] += a;
      __syncthreads();
    }
    if (i < n) out[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Emit the triangles of each leaf at its offset, in (subcell, slot) order.
static __global__ void __launch_bounds__(256) k_mc_emit(const McTable* __restrict__ T, const float* __restrict__ vals,
                                               const float* __restrict__ lo, const float* __restrict__ up,
                                               long long l0, long long nl, int s,
                                               const long long* __restrict__ off, float* __restrict__ tri,
                                               long long cap) {
  const int n = 1 << s;
  const int nsub = n * n * n;
  __shared__ int sc[256];
  for (long long leaf = blockIdx.x; leaf < nl; leaf += gridDim.x) {
    const long long L = l0 + leaf;
    float clo[3], d[3];
    for (int k = 0; k < 3; ++k) {
      clo[k] = lo[3 * L + k];
      d[k] = (up[3 * L + k] - clo[k]) / (float)n;
    }
    long long base = off[leaf];
    for (int s0 = 0; s0 < nsub; s0 += blockDim.x) {
      const int sub = s0 + threadIdx.x;
      float vv[8];
      int cid = 0, c = 0;
      if (sub < nsub) {
        cid = mc_case(vals, leaf, s, sub, vv);
        for (int t = 0; t < 5; ++t) c += (T->tri[cid][3 * t] >= 0);
      }
      sc[threadIdx.x] = c;
      __syncthreads();
      for (int w = 1; w < 256; w <<= 1) {
        const int a = (int)threadIdx.x >= w ? sc[threadIdx.x - w] : 0;
        __syncthreads();
        sc[threadIdx.x] += a;
        __syncthreads();
      }
      const long long my = base + sc[threadIdx.x] - c;
      const int chunk_total = sc[255];
      if (sub < nsub && c > 0) {
        const int si = sub / (n * n), sj = (sub / n) % n, sk = sub % n;
        float slo[3], sup[3];
        const int ii[3] = {si, sj, sk};
        for (int k = 0; k < 3; ++k) {
          slo[k] = clo[k] + (float)ii[k] * d[k];
          sup[k] = slo[k] + d[k];
        }
        for (int t = 0; t < c; ++t) {
          const long long o = my + t;
          if (o >= cap) break;
          for (int corner = 0; corner < 3; ++corner) {
            const int e = T->tri[cid][3 * t + corner];
            const int va = c_edge_verts[e][0], vb = c_edge_verts[e][1];
            const float fa = vv[va], fb = vv[vb];
            float tc = -fa / (fb - fa);
            if (tc != tc) tc = 0.f;
            else if (isinf(tc)) tc = tc > 0.f ? 3.4028234663852886e38f : -3.4028234663852886e38f;
            tc = fminf(fmaxf(tc, 0.f), 1.f);
            for (int k = 0; k < 3; ++k) {
              const float pa = c_vert_logical[va][k] ? sup[k] : slo[k];
              const float pb = c_vert_logical[vb][k] ? sup[k] : slo[k];
              tri[(o * 3 + corner) * 3 + k] = (1.f - tc) * pa + tc * pb;
            }
          }
        }
      }
      base += chunk_total;
      __syncthreads();
    }
  }
}
