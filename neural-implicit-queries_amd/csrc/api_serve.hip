// Serving loop: worker threads with one HIP stream each, admission by device
// memory, ray requests merged into one persistent cast (one ray group each)
// and replayed tree-build graphs.
#include "niq_internal.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <memory>
#include <thread>

namespace {

using Clock = std::chrono::steady_clock;

enum ReqKind { REQ_RAYS, REQ_TREE, REQ_CP };

niq_eval_cfg default_cfg() {
  niq_eval_cfg c{};
  c.mode = NIQ_MODE_AFFINE_FIXED;
  c.precision = NIQ_PREC_FP32;
  return c;
}

struct Request {
  uint64_t id = 0;
  ReqKind kind = REQ_RAYS;
  Clock::time_point t_submit, t_start, t_done;
  int state = 0;
  niq_status status = NIQ_OK;
  std::string err;
  size_t need = 0;   // device bytes reserved while the request runs
  niq_eval_cfg cfg{};
  // rays
  std::vector<const niq_model*> models;
  const float* roots = nullptr;
  const float* dirs = nullptr;
  int64_t n = 0;
  niq_cast_opts opts{};
  float* out_t = nullptr;
  int32_t* out_hit = nullptr;
  int32_t* out_count = nullptr;
  int64_t n_evals = 0;
  int n_batched = 0;
  // tree
  const niq_model* model = nullptr;
  niq_tree_params tp{};
  niq_tree_result tres{};
  int graph_replay = 0;
  // closest point
  float lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0};
  const float* q = nullptr;
  int64_t Q = 0;
  float eps = 0.f;
  int B = 0;
  float* out_dist = nullptr;
  float* out_loc = nullptr;
  int64_t n_processed = 0;
};

// Template widths of a ray request's functions (FP32 and precision-mode
// copies): requests merge into one cast only at equal widths, so no function
// runs in a wider tile than its own cast would give it.
int req_wt(const Request& r, bool w16) {
  int w = 0;
  for (const niq_model* m : r.models) w = std::max(w, w16 ? m->wt16 : m->wt);
  return w;
}

// Ray requests that can share one persistent cast: the same eval config
// (mode, precision, variants) and template widths.  Models, options and
// sizes may differ: each request is a group of the cast with its own
// functions, options, inputs and outputs.
bool same_cast(const Request& a, const Request& b) {
  return a.kind == REQ_RAYS && b.kind == REQ_RAYS && std::memcmp(&a.cfg, &b.cfg, sizeof(niq_eval_cfg)) == 0 &&
         req_wt(a, false) == req_wt(b, false) && req_wt(a, true) == req_wt(b, true);
}

// Tree requests that can share one batched build: the same eval config (its
// tile variants aside: they change no result, and a batch runs the default
// tile) and build params (root boxes aside), one template width, a batchable
// form.
bool same_tree_batch(const Request& a, const Request& b) {
  if (a.kind != REQ_TREE || b.kind != REQ_TREE) return false;
  niq_eval_cfg ca = a.cfg, cb = b.cfg;
  ca.variant = cb.variant = 0;
  ca.ray_variant = cb.ray_variant = 0;
  if (std::memcmp(&ca, &cb, sizeof(niq_eval_cfg)) != 0 || a.model->wt != b.model->wt) return false;
  const niq_tree_params &p = a.tp, &q = b.tp;
  return p.split_depth == q.split_depth && p.node_terminate_thresh == q.node_terminate_thresh &&
         p.batch_process_size == q.batch_process_size && p.with_interior_nodes == q.with_interior_nodes &&
         p.with_exterior_nodes == q.with_exterior_nodes && p.offset == q.offset &&
         tree_batchable(b.model, &b.cfg, &b.tp);
}

// One captured tree build: its workspace, internal outputs and root box stay
// allocated while the graph is cached.
struct TreeGraph {
  // The model is matched by its serial number, not its address: a model
  // destroyed and another created can share an address, and a graph's
  // captured arguments hold the old model's device weights.
  uint64_t model_serial = 0;
  niq_eval_cfg cfg{};
  int split_depth = 0, B = 0, with_int = 0, with_ext = 0;
  float offset = 0.f;
  int64_t cap_unk = 0, cap_int = 0, cap_ext = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0, counters_off = 0, bytes = 0;
  float* unk = nullptr;    // [2][cap_unk][3]
  float* fin_i = nullptr;  // [2][cap_int][3]
  float* fin_e = nullptr;
  float* root = nullptr;   // 6 floats
  hipGraphExec_t exec = nullptr;
  bool busy = false;       // checked out by a worker replaying it

  bool matches(const Request& r) const {
    return model_serial == r.model->serial && std::memcmp(&cfg, &r.cfg, sizeof(cfg)) == 0 && split_depth == r.tp.split_depth &&
           B == r.tp.batch_process_size && with_int == r.tp.with_interior_nodes &&
           with_ext == r.tp.with_exterior_nodes && offset == r.tp.offset && cap_unk == r.tres.unknown_capacity &&
           cap_int == (r.tp.with_interior_nodes ? r.tres.interior_capacity : 0) &&
           cap_ext == (r.tp.with_exterior_nodes ? r.tres.exterior_capacity : 0);
  }
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (void* p : {ws, (void*)unk, (void*)fin_i, (void*)fin_e, (void*)root})
      if (p) (void)hipFree(p);
    exec = nullptr; ws = nullptr; unk = fin_i = fin_e = root = nullptr;
  }
};

// Tree graphs the server keeps (besides the byte cap of 1/8 of its memory cap).
constexpr int kMaxGraphs = 64;

}  // namespace

struct niq_server {
  int device = 0;
  size_t cap = 0;
  size_t graph_cap = 0;
  int64_t max_merge = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  size_t in_use = 0;        // reserved by running requests
  size_t graph_bytes = 0;   // held by cached graphs
  // Captured tree builds shared by all workers, most recently used first.  A
  // worker checks one out (busy) for a replay, so one instance never runs on
  // two streams at once; a key in use everywhere gets another instance.
  std::list<TreeGraph> graphs;
  int running = 0;
  bool stop = false;
  uint64_t next_id = 1;
  std::vector<std::thread> workers;

  void worker_main();
  niq_status submit(std::unique_ptr<Request> r, uint64_t* id);
};

namespace {

struct Worker {
  niq_server* srv = nullptr;
  hipStream_t s = nullptr;
  void* pinned = nullptr;   // root box (256 B) + tree counters

  void finish(Request* r, niq_status st) {
    r->status = st;
    if (st != NIQ_OK) r->err = niq_last_error();
  }

  void run_rays(std::vector<Request*>& group);
  void run_tree(Request* r);
  void run_tree_batch(std::vector<Request*>& group);
  void run_cp(Request* r);
  TreeGraph* graph_for(Request* r, bool* reused);
  void return_graph(TreeGraph* g);
};

void Worker::run_rays(std::vector<Request*>& group) {
  // one persistent cast over every merged request, each its own ray group
  const int ng = (int)group.size();
  std::vector<CastGroupReq> reqs(ng);
  int wt = 0, wt16 = 0;
  for (int g = 0; g < ng; ++g) {
    Request* r = group[g];
    CastGroupReq& q = reqs[g];
    q.models = r->models.data(); q.n_models = (int)r->models.size();
    q.roots = r->roots; q.dirs = r->dirs; q.n = r->n; q.opts = r->opts;
    q.out_t = r->out_t; q.out_hit = r->out_hit; q.out_count = r->out_count;
    wt = std::max(wt, req_wt(*r, false));
    wt16 = std::max(wt16, req_wt(*r, true));
  }
  const niq_eval_cfg& cfg = group[0]->cfg;
  const size_t ws_bytes = cast_layout(reqs.data(), ng, cfg.precision == NIQ_PREC_BF16).total;
  niq_status st = NIQ_OK;
  void* ws = nullptr;
  auto hipfail = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipfail(hipMallocAsync(&ws, ws_bytes, s), "hipMallocAsync");
  if (st == NIQ_OK) st = cast_run_groups(reqs.data(), ng, &cfg, static_cast<unsigned char*>(ws), wt, wt16, s);
  if (ws) hipfail(hipFreeAsync(ws, s), "hipFreeAsync");
  hipfail(hipStreamSynchronize(s), "hipStreamSynchronize");
  for (int g = 0; g < ng; ++g) {
    group[g]->n_evals = reqs[g].n_evals;
    group[g]->n_batched = ng;
    finish(group[g], st);
  }
}

TreeGraph* Worker::graph_for(Request* r, bool* reused) {
  *reused = false;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    for (auto it = srv->graphs.begin(); it != srv->graphs.end(); ++it) {
      if (!it->busy && it->matches(*r)) {
        srv->graphs.splice(srv->graphs.begin(), srv->graphs, it);
        srv->graphs.front().busy = true;
        *reused = true;
        return &srv->graphs.front();
      }
    }
  }
  size_t ws_bytes = 0;
  if (niq_tree_workspace_bytes(r->model, &r->cfg, &r->tp, &ws_bytes) != NIQ_OK) return nullptr;
  TreeGraph g;
  g.model_serial = r->model->serial; g.cfg = r->cfg; g.split_depth = r->tp.split_depth; g.B = r->tp.batch_process_size;
  g.with_int = r->tp.with_interior_nodes; g.with_ext = r->tp.with_exterior_nodes; g.offset = r->tp.offset;
  g.cap_unk = r->tres.unknown_capacity;
  g.cap_int = g.with_int ? r->tres.interior_capacity : 0;
  g.cap_ext = g.with_ext ? r->tres.exterior_capacity : 0;
  g.ws_bytes = ws_bytes;
  g.bytes = ws_bytes + 24 * (size_t)(g.cap_unk + g.cap_int + g.cap_ext) + 256;
  std::list<TreeGraph> evicted;   // released after the lock is dropped
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    // least recently used idle graphs make room (count and bytes); a busy
    // graph is held by a worker and stays
    for (auto it = srv->graphs.end(); it != srv->graphs.begin();) {
      if (srv->graph_bytes + g.bytes <= srv->graph_cap && (int)srv->graphs.size() < kMaxGraphs) break;
      --it;
      if (it->busy) continue;
      srv->graph_bytes -= it->bytes;
      auto victim = it++;   // it: the graph after the victim (or end)
      evicted.splice(evicted.end(), srv->graphs, victim);
    }
    if (srv->graph_bytes + g.bytes > srv->graph_cap) {
      for (TreeGraph& e : evicted) e.release();
      return nullptr;
    }
    srv->graph_bytes += g.bytes;
  }
  for (TreeGraph& e : evicted) e.release();
  bool ok = hipMalloc(&g.ws, ws_bytes) == hipSuccess &&
            hipMalloc((void**)&g.unk, 24 * (size_t)std::max<int64_t>(g.cap_unk, 1)) == hipSuccess &&
            hipMalloc((void**)&g.root, 256) == hipSuccess;
  if (ok && g.with_int) ok = hipMalloc((void**)&g.fin_i, 24 * (size_t)std::max<int64_t>(g.cap_int, 1)) == hipSuccess;
  if (ok && g.with_ext) ok = hipMalloc((void**)&g.fin_e, 24 * (size_t)std::max<int64_t>(g.cap_ext, 1)) == hipSuccess;
  hipGraph_t graph = nullptr;
  if (ok) {
    niq_tree_result inner{};
    inner.unknown_lower = g.unk;
    inner.unknown_upper = g.unk + 3 * g.cap_unk;
    inner.unknown_capacity = g.cap_unk;
    if (g.with_int) {
      inner.interior_lower = g.fin_i; inner.interior_upper = g.fin_i + 3 * g.cap_int; inner.interior_capacity = g.cap_int;
    }
    if (g.with_ext) {
      inner.exterior_lower = g.fin_e; inner.exterior_upper = g.fin_e + 3 * g.cap_ext; inner.exterior_capacity = g.cap_ext;
    }
    ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      const niq_status st =
          tree_enqueue_async(r->model, &r->cfg, &r->tp, g.ws, g.ws_bytes, &inner, &g.counters_off, g.root, s);
      const hipError_t e = hipStreamEndCapture(s, &graph);
      ok = st == NIQ_OK && e == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
  }
  if (!ok) {
    (void)hipGetLastError();
    g.release();
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->graph_bytes -= g.bytes;
    return nullptr;
  }
  g.busy = true;
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->graphs.push_front(g);
  return &srv->graphs.front();
}

void Worker::return_graph(TreeGraph* g) {
  std::lock_guard<std::mutex> lk(srv->mu);
  g->busy = false;
}

void Worker::run_tree(Request* r) {
  r->n_batched = 1;
  const int mode = r->cfg.mode;
  const bool capturable = r->tp.split_depth >= 0 && r->tp.node_terminate_thresh <= 0 &&
                          mode != NIQ_MODE_AFFINE_ALL && mode != NIQ_MODE_AFFINE_TRUNCATE &&
                          mode != NIQ_MODE_AFFINE_APPEND && !classify_needs_scratch(r->model, &r->cfg, 3);
  bool reused = false;
  TreeGraph* g = capturable ? graph_for(r, &reused) : nullptr;
  niq_status st = NIQ_OK;
  auto hipfail = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  if (g) {
    float* box = static_cast<float*>(pinned);
    for (int k = 0; k < 3; ++k) { box[k] = r->tp.lower[k]; box[3 + k] = r->tp.upper[k]; }
    unsigned char* counters = static_cast<unsigned char*>(pinned) + 256;
    hipfail(hipMemcpyAsync(g->root, box, 24, hipMemcpyHostToDevice, s), "root box");
    hipfail(hipGraphLaunch(g->exec, s), "hipGraphLaunch");
    hipfail(hipMemcpyAsync(counters, static_cast<unsigned char*>(g->ws) + g->counters_off, niq_tree_counters_bytes(),
                           hipMemcpyDeviceToHost, s), "counters");
    hipfail(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (st == NIQ_OK) st = niq_tree_decode(counters, &r->tp, &r->tres);
    if (st == NIQ_OK) {
      auto rows = [&](float* dlo, float* dup, const float* src, int64_t cap, int64_t n) {
        if (n <= 0) return;
        hipfail(hipMemcpyAsync(dlo, src, 12 * n, hipMemcpyDeviceToDevice, s), "rows");
        hipfail(hipMemcpyAsync(dup, src + 3 * cap, 12 * n, hipMemcpyDeviceToDevice, s), "rows");
      };
      rows(static_cast<float*>(r->tres.unknown_lower), static_cast<float*>(r->tres.unknown_upper), g->unk, g->cap_unk,
           r->tres.n_unknown);
      if (g->with_int)
        rows(static_cast<float*>(r->tres.interior_lower), static_cast<float*>(r->tres.interior_upper), g->fin_i,
             g->cap_int, r->tres.n_interior);
      if (g->with_ext)
        rows(static_cast<float*>(r->tres.exterior_lower), static_cast<float*>(r->tres.exterior_upper), g->fin_e,
             g->cap_ext, r->tres.n_exterior);
      hipfail(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    r->graph_replay = reused ? 2 : 1;
    return_graph(g);
  } else {
    size_t ws_bytes = 0;
    void* ws = nullptr;
    st = niq_tree_workspace_bytes(r->model, &r->cfg, &r->tp, &ws_bytes);
    if (st == NIQ_OK) hipfail(hipMallocAsync(&ws, ws_bytes, s), "hipMallocAsync");
    if (st == NIQ_OK) st = niq_tree_build(r->model, &r->cfg, &r->tp, ws, ws_bytes, &r->tres, s);
    if (ws) hipfail(hipFreeAsync(ws, s), "hipFreeAsync");
    hipfail(hipStreamSynchronize(s), "hipStreamSynchronize");
  }
  finish(r, st);
}

// Several tree builds merged: every level of all of them in shared launches.
void Worker::run_tree_batch(std::vector<Request*>& group) {
  const int ng = (int)group.size();
  std::vector<const niq_model*> models(ng);
  std::vector<float> roots(6 * ng);
  std::vector<niq_tree_result*> res(ng);
  for (int g = 0; g < ng; ++g) {
    Request* r = group[g];
    models[g] = r->model;
    for (int k = 0; k < 3; ++k) { roots[6 * g + k] = r->tp.lower[k]; roots[6 * g + 3 + k] = r->tp.upper[k]; }
    res[g] = &r->tres;
  }
  const Request& r0 = *group[0];
  niq_status st = NIQ_OK;
  size_t bytes = 0;
  void* ws = nullptr;
  st = tree_batch_bytes(r0.model, &r0.cfg, &r0.tp, ng, &bytes);
  if (st == NIQ_OK && hipMallocAsync(&ws, bytes, s) != hipSuccess) st = fail(NIQ_ERR_HIP, "tree batch workspace");
  std::vector<niq_status> tree_st(ng, NIQ_ERR_INVALID);
  bool ran = false;
  if (st == NIQ_OK) {
    st = tree_batch_build(models.data(), ng, &r0.cfg, &r0.tp, roots.data(), ws, bytes, res.data(), tree_st.data(), s);
    ran = st == NIQ_OK || st == NIQ_ERR_CAPACITY;
  }
  if (ws) (void)hipFreeAsync(ws, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, "hipStreamSynchronize failed");
  for (int g = 0; g < ng; ++g) {
    group[g]->n_batched = ng;
    // a build that ran reports each tree's own outcome; otherwise every tree
    // gets the batch's error
    finish(group[g], ran && (st == NIQ_OK || st == NIQ_ERR_CAPACITY) ? tree_st[g] : st);
  }
}

void Worker::run_cp(Request* r) {
  r->n_batched = 1;
  niq_status st = niq_closest_point(r->model, &r->cfg, r->lower, r->upper, r->q, r->Q, r->eps, r->B, r->out_dist,
                                    r->out_loc, &r->n_processed, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == NIQ_OK) st = fail(NIQ_ERR_HIP, "hipStreamSynchronize failed");
  finish(r, st);
}

}  // namespace

void niq_server::worker_main() {
  Worker w;
  w.srv = this;
  const bool ready = hipSetDevice(device) == hipSuccess &&
                     hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking) == hipSuccess &&
                     hipHostMalloc(&w.pinned, 4096, hipHostMallocDefault) == hipSuccess;
  while (true) {
    std::vector<Request*> group;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        if (queue.empty()) return stop;
        return running == 0 || in_use + queue.front()->need <= cap;
      });
      if (queue.empty()) break;   // stopping and drained
      Request* r = queue.front();
      queue.pop_front();
      group.push_back(r);
      size_t need = r->need;
      if (r->kind == REQ_RAYS) {
        // continuous batching: fold in every waiting cast of the same models,
        // config and options that still fits
        int64_t rays = r->n;
        for (auto it = queue.begin(); it != queue.end() && (int)group.size() < kMaxRayGroups;) {
          Request* o = *it;
          if (same_cast(*r, *o) && rays + o->n <= max_merge && in_use + need + o->need <= cap) {
            group.push_back(o);
            rays += o->n;
            need += o->need;
            it = queue.erase(it);
          } else {
            ++it;
          }
        }
      } else if (r->kind == REQ_TREE && tree_batchable(r->model, &r->cfg, &r->tp)) {
        // builds of the same config and depth run their levels together; the
        // waiting ones are shared out over the workers, so that several
        // batches overlap on the device (a batch's small levels are latency)
        int compatible = 1;
        for (const Request* o : queue) compatible += same_tree_batch(*r, *o) ? 1 : 0;
        const int n_workers = std::max<int>(1, (int)workers.size());
        const int target = std::min(kMaxTreeBatch, (compatible + n_workers - 1) / n_workers);
        for (auto it = queue.begin(); it != queue.end() && (int)group.size() < target;) {
          Request* o = *it;
          if (same_tree_batch(*r, *o) && in_use + need + o->need <= cap) {
            group.push_back(o);
            need += o->need;
            it = queue.erase(it);
          } else {
            ++it;
          }
        }
      }
      in_use += need;
      ++running;
      const auto now = Clock::now();
      for (Request* g : group) { g->state = 1; g->t_start = now; }
    }
    if (!ready) {
      for (Request* r : group) { r->status = NIQ_ERR_HIP; r->err = "serving worker could not set up its HIP stream"; }
    } else if (group[0]->kind == REQ_RAYS) {
      w.run_rays(group);
    } else if (group[0]->kind == REQ_TREE) {
      if (group.size() > 1) w.run_tree_batch(group);
      else w.run_tree(group[0]);
    } else {
      w.run_cp(group[0]);
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      size_t need = 0;
      const auto now = Clock::now();
      for (Request* r : group) { need += r->need; r->state = 2; r->t_done = now; }
      in_use -= need;
      --running;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
  if (w.pinned) (void)hipHostFree(w.pinned);
  if (w.s) (void)hipStreamDestroy(w.s);
}

niq_status niq_server::submit(std::unique_ptr<Request> r, uint64_t* id) {
  if (r->need > cap) return fail(NIQ_ERR_CAPACITY, "request needs " + std::to_string(r->need) +
                                                     " device bytes, above the server's memory cap");
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(mu);
    if (stop) return fail(NIQ_ERR_INVALID, "server is stopping");
    r->id = next_id++;
    if (id) *id = r->id;
    queue.push_back(r.get());
    reqs.emplace(r->id, std::move(r));
  }
  cv_work.notify_one();
  return NIQ_OK;
}

extern "C" {

niq_status niq_serve_create(int n_streams, size_t mem_cap_bytes, int64_t max_merge_rays, niq_server** out) {
  if (!out) return fail(NIQ_ERR_INVALID, "null output");
  if (n_streams < 1 || n_streams > 16) return fail(NIQ_ERR_INVALID, "n_streams must be 1..16");
  if (mem_cap_bytes < (1u << 20)) return fail(NIQ_ERR_INVALID, "memory cap below 1 MiB");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  auto* srv = new niq_server();
  srv->device = dev;
  srv->cap = mem_cap_bytes;
  srv->graph_cap = mem_cap_bytes / 8;
  srv->max_merge = max_merge_rays > 0 ? std::min<int64_t>(max_merge_rays, INT_MAX) : (1LL << 24);
  for (int i = 0; i < n_streams; ++i) srv->workers.emplace_back([srv] { srv->worker_main(); });
  *out = srv;
  return NIQ_OK;
}

niq_status niq_serve_destroy(niq_server* srv) {
  if (!srv) return NIQ_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (auto& t : srv->workers) t.join();
  // no worker runs now: release the shared graphs on the server's device
  int prev = 0;
  const bool have_prev = hipGetDevice(&prev) == hipSuccess;
  if (!srv->graphs.empty() && hipSetDevice(srv->device) == hipSuccess) {
    for (TreeGraph& g : srv->graphs) g.release();
    if (have_prev) (void)hipSetDevice(prev);
  }
  delete srv;
  return NIQ_OK;
}

niq_status niq_serve_submit_rays(niq_server* srv, const niq_model* const* models, int n_models,
                                 const niq_eval_cfg* cfg, const float* roots, const float* dirs, int64_t n,
                                 const niq_cast_opts* opts, float* out_t, int32_t* out_hit_id, int32_t* out_count,
                                 uint64_t* request_id) {
  if (!srv) return fail(NIQ_ERR_INVALID, "null server");
  if (!models || n_models < 1 || n_models > 8) return fail(NIQ_ERR_INVALID, "need 1..8 models");
  if (!opts) return fail(NIQ_ERR_INVALID, "null opts");
  if (n < 0 || n > srv->max_merge) return fail(NIQ_ERR_INVALID, "ray count out of range");
  int wt = 0, wt16 = 0;
  if (niq_status st = cast_check(models, n_models, cfg, roots, dirs, n, opts, out_t, out_hit_id, out_count, &wt, &wt16))
    return st;
  auto r = std::make_unique<Request>();
  r->kind = REQ_RAYS;
  r->models.assign(models, models + n_models);
  r->cfg = cfg ? *cfg : default_cfg();
  r->roots = roots; r->dirs = dirs; r->n = n; r->opts = *opts;
  r->out_t = out_t; r->out_hit = out_hit_id; r->out_count = out_count;
  r->need = niq_cast_workspace_bytes(opts);
  return srv->submit(std::move(r), request_id);
}

niq_status niq_serve_submit_tree(niq_server* srv, const niq_model* m, const niq_eval_cfg* cfg,
                                 const niq_tree_params* p, const niq_tree_result* res, uint64_t* request_id) {
  if (!srv) return fail(NIQ_ERR_INVALID, "null server");
  if (niq_status st = check_model(m, 3)) return st;
  if (!p || !res) return fail(NIQ_ERR_INVALID, "null params or result");
  if (!res->unknown_lower || !res->unknown_upper || res->unknown_capacity < 0)
    return fail(NIQ_ERR_INVALID, "unknown-node output buffers are required");
  if (p->with_interior_nodes && (!res->interior_lower || !res->interior_upper))
    return fail(NIQ_ERR_INVALID, "interior output buffers are required");
  if (p->with_exterior_nodes && (!res->exterior_lower || !res->exterior_upper))
    return fail(NIQ_ERR_INVALID, "exterior output buffers are required");
  auto r = std::make_unique<Request>();
  r->kind = REQ_TREE;
  r->model = m;
  r->cfg = cfg ? *cfg : default_cfg();
  r->tp = *p;
  r->tres = *res;
  size_t ws = 0;
  if (niq_status st = niq_tree_workspace_bytes(m, &r->cfg, p, &ws)) return st;
  r->need = ws + 8192;   // and its share of a merged build's table of trees
  return srv->submit(std::move(r), request_id);
}

niq_status niq_serve_submit_closest_point(niq_server* srv, const niq_model* m, const niq_eval_cfg* cfg,
                                          const float lower[3], const float upper[3], const float* query_points,
                                          int64_t Q, float eps, int batch_process_size, float* out_dist,
                                          float* out_loc, uint64_t* request_id) {
  if (!srv) return fail(NIQ_ERR_INVALID, "null server");
  if (niq_status st = check_model(m, 3)) return st;
  if (!lower || !upper) return fail(NIQ_ERR_INVALID, "null box");
  if (Q < 0 || Q > (1LL << 30)) return fail(NIQ_ERR_INVALID, "query count out of range");
  if (Q > 0 && (!query_points || !out_dist || !out_loc)) return fail(NIQ_ERR_INVALID, "bad buffers");
  const int B = batch_process_size;
  if (B < 1 || B > (1 << 20)) return fail(NIQ_ERR_INVALID, "batch_process_size out of range");
  auto r = std::make_unique<Request>();
  r->kind = REQ_CP;
  r->model = m;
  r->cfg = cfg ? *cfg : default_cfg();
  for (int k = 0; k < 3; ++k) { r->lower[k] = lower[k]; r->upper[k] = upper[k]; }
  r->q = query_points; r->Q = Q; r->eps = eps; r->B = B;
  r->out_dist = out_dist; r->out_loc = out_loc;
  // per-round buffers, the query's winner ids and a node stack with room to
  // grow twice (28 bytes an entry)
  const size_t stack = (size_t)cp_initial_stack(Q, B);
  r->need = 152 * (size_t)B + 4 * (size_t)Q + 28 * 4 * stack + 4096;
  return srv->submit(std::move(r), request_id);
}

niq_status niq_serve_poll(niq_server* srv, uint64_t request_id, int wait, niq_serve_info* info,
                          niq_tree_result* tree_res) {
  if (!srv) return fail(NIQ_ERR_INVALID, "null server");
  std::unique_ptr<Request> done;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(request_id);
    if (it == srv->reqs.end()) return fail(NIQ_ERR_INVALID, "unknown request id " + std::to_string(request_id));
    if (wait) {
      // look the id up again on every wake: another poller may have taken the
      // request (and freed it) while this one waited
      srv->cv_done.wait(lk, [&] {
        auto f = srv->reqs.find(request_id);
        return f == srv->reqs.end() || f->second->state == 2;
      });
      it = srv->reqs.find(request_id);
      if (it == srv->reqs.end())
        return fail(NIQ_ERR_INVALID, "request id " + std::to_string(request_id) + " was polled done by another caller");
    }
    Request* r = it->second.get();
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->state = r->state;
      info->status = r->state == 2 ? r->status : NIQ_OK;
      using ms = std::chrono::duration<double, std::milli>;
      if (r->state >= 1) info->queue_ms = ms(r->t_start - r->t_submit).count();
      if (r->state == 2) info->latency_ms = ms(r->t_done - r->t_submit).count();
      info->n_evals = r->n_evals;
      info->n_batched = r->n_batched;
      info->graph_replay = r->graph_replay;
      info->n_processed = r->n_processed;
    }
    if (r->state != 2) return NIQ_OK;
    done = std::move(it->second);
    srv->reqs.erase(it);
  }
  if (tree_res && done->kind == REQ_TREE) {
    const niq_tree_result& s = done->tres;
    tree_res->n_unknown = s.n_unknown; tree_res->n_interior = s.n_interior; tree_res->n_exterior = s.n_exterior;
    tree_res->n_classified = s.n_classified; tree_res->n_recheck = s.n_recheck; tree_res->n_levels = s.n_levels;
    for (int i = 0; i < 64; ++i) {
      tree_res->level_n[i] = s.level_n[i];
      tree_res->level_interior[i] = s.level_interior[i];
      tree_res->level_exterior[i] = s.level_exterior[i];
    }
  }
  if (done->status != NIQ_OK) g_err = done->err;
  return NIQ_OK;
}

size_t niq_serve_bytes_in_use(niq_server* srv) {
  if (!srv) return 0;
  std::lock_guard<std::mutex> lk(srv->mu);
  return srv->in_use + srv->graph_bytes;
}

}  // extern "C"
