"""Serving loop over the resident models (niq_serve_* in the C ABI).

The reference answers one query at a time from one Python thread
(src/queries.py:134, src/kd_tree.py:103,766).  A ``Server`` instead keeps
several models resident and runs requests as they arrive, on worker threads
inside the library, each with its own HIP stream:

* admission by device memory: a request starts once the bytes it needs fit
  under ``mem_cap_bytes`` beside the work in flight (arrival order is kept);
* continuous batching: ray requests waiting with the same models, config and
  cast options are merged into the cast of the request taken;
* tree builds with a ``split_depth`` in the fused-kernel modes that wait in the
  queue together (same config, depth and template width, any models) run as
  one batched build, every level of all of them in shared launches; a lone
  one replays a hipGraph captured once per (model, config, depth,
  capacities), shared by the workers;
* every request reports its queueing time and latency.

Each result equals the result of the matching standalone call
(``queries.cast_rays``, ``kd_tree.construct_uniform_unknown_levelset_tree``,
``kd_tree.closest_point``): the same kernels run on the same inputs, and rays
are independent of one another, and a batched tree build runs each tree's
levels exactly as its lone build does.  ``info['n_batched']`` is the number
of requests that shared the launches.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import mlp as _mlp
from .implicit_function import as_device_f32
from .kd_tree import TreeBuffers, tree_request
from .queries import cast_opts_struct


class Request:
    """One submitted query; ``result()`` waits and returns what the standalone
    call returns."""

    def __init__(self, server, rid, kind, keep, finish):
        self._server = server
        self.id = rid
        self.kind = kind
        self._keep = keep          # tensors the device reads or writes
        self._finish = finish
        self._tree_res = keep.get("tree_res")
        self.info = None
        self._result = None
        self._error = None

    def _poll(self, wait):
        if self.info is not None:
            return True
        lib = _lib.load_library()
        info = _lib.ServeInfo()
        res = ctypes.byref(self._tree_res) if self._tree_res is not None else None
        _lib.check(lib.niq_serve_poll(self._server._h, self.id, int(wait), ctypes.byref(info), res))
        if info.state != 2:
            return False
        self.info = {"status": info.status, "queue_ms": info.queue_ms, "latency_ms": info.latency_ms,
                     "n_evals": int(info.n_evals), "n_batched": int(info.n_batched),
                     "graph_replay": bool(info.graph_replay), "graph_reused": info.graph_replay == 2,
                     "n_processed": int(info.n_processed)}
        try:
            _lib.check(info.status)
            self._result = self._finish(self)
        except Exception as e:   # the request failed: keep its error for result()
            self._error = e
        return True

    def done(self):
        return self._poll(False)

    def result(self):
        self._poll(True)
        if self._error is not None:
            raise self._error
        return self._result

    def __del__(self):
        # outputs must not be freed while a worker may still write them
        try:
            if self.info is None and self._server._h is not None:
                self._poll(True)
        except Exception:
            pass


class Server:
    """Continuous-batching query server over one GPU."""

    def __init__(self, n_streams=4, mem_cap_bytes=200 * 10 ** 9, max_merge_rays=1 << 24):
        _lib.require_gpu()
        lib = _lib.load_library()
        h = ctypes.c_void_p()
        _lib.check(lib.niq_serve_create(int(n_streams), int(mem_cap_bytes), int(max_merge_rays), ctypes.byref(h)))
        self._h = h
        self._models = []     # keep model handles alive while requests may use them

    # -- lifecycle
    def close(self):
        if self._h is not None:
            _lib.load_library().niq_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def bytes_in_use(self):
        return int(_lib.load_library().niq_serve_bytes_in_use(self._h))

    @staticmethod
    def _inputs_ready():
        # inputs are read by the server's own streams: finish the producer's work
        torch.cuda.current_stream().synchronize()

    def _model(self, params):
        m = _mlp.device_model(params)
        self._models.append(m)
        return m

    # -- requests
    def submit_rays(self, funcs_tuple, params_tuple, roots, dirs, opts):
        """Queue a cast; ``result()`` gives (t, hit_id, count, N_evals) as
        ``queries.cast_rays`` (src/queries.py:134-175)."""
        if not isinstance(funcs_tuple, (tuple, list)):
            funcs_tuple = (funcs_tuple,)
        if not isinstance(params_tuple, (tuple, list)):
            params_tuple = (params_tuple,)
        if len(funcs_tuple) != len(params_tuple):
            raise ValueError("funcs_tuple and params_tuple lengths differ")
        if len({f.cfg.mode for f in funcs_tuple}) != 1:
            raise ValueError("all functions of one cast_rays call must use the same range-analysis mode")
        roots = as_device_f32(roots)
        dirs = as_device_f32(dirs)
        N = roots.shape[0]
        models = [self._model(p) for p in params_tuple]
        handles = (ctypes.c_void_p * len(models))(*[m.handle.value for m in models])
        o = cast_opts_struct(opts)
        dev = roots.device
        out_t = torch.empty(N, dtype=torch.float32, device=dev)
        out_hit = torch.empty(N, dtype=torch.int32, device=dev)
        out_count = torch.empty(N, dtype=torch.int32, device=dev)
        self._inputs_ready()
        rid = ctypes.c_uint64()
        lib = _lib.load_library()
        _lib.check(lib.niq_serve_submit_rays(self._h, handles, len(models), ctypes.byref(funcs_tuple[0].cfg),
                                             _lib.ptr(roots), _lib.ptr(dirs), N, ctypes.byref(o),
                                             _lib.ptr(out_t), _lib.ptr(out_hit), _lib.ptr(out_count),
                                             ctypes.byref(rid)))
        keep = {"roots": roots, "dirs": dirs, "t": out_t, "hit": out_hit, "count": out_count,
                "cfg": funcs_tuple[0].cfg, "handles": handles}
        return Request(self, rid.value, "rays", keep,
                       lambda r: (r._keep["t"], r._keep["hit"], r._keep["count"], r.info["n_evals"]))

    def submit_tree(self, func, params, lower, upper, node_terminate_thresh=None, split_depth=None,
                    with_interior_nodes=False, with_exterior_nodes=False, offset=0., batch_process_size=2048):
        """Queue a level-set tree build; ``result()`` gives the dict of
        ``construct_uniform_unknown_levelset_tree`` (src/kd_tree.py:103-218)."""
        m, p, cap, cap_fin = tree_request(func, params, lower, upper, node_terminate_thresh, split_depth,
                                          with_interior_nodes, with_exterior_nodes, offset, batch_process_size)
        self._models.append(m)
        bufs = TreeBuffers(p, cap, cap_fin, torch.device("cuda"))
        self._inputs_ready()
        rid = ctypes.c_uint64()
        lib = _lib.load_library()
        _lib.check(lib.niq_serve_submit_tree(self._h, m.handle, ctypes.byref(func.cfg), ctypes.byref(p),
                                             ctypes.byref(bufs.r), ctypes.byref(rid)))
        keep = {"bufs": bufs, "tree_res": bufs.r, "cfg": func.cfg, "p": p}
        return Request(self, rid.value, "tree", keep, lambda r: r._keep["bufs"].outputs()[0])

    def submit_closest_point(self, func, params, lower, upper, query_points, eps=0.001, batch_process_size=2048):
        """Queue a closest-point batch; ``result()`` gives (dist, loc) as
        ``kd_tree.closest_point`` (src/kd_tree.py:766-802)."""
        m = self._model(params)
        q = torch.as_tensor(query_points).to(device="cuda", dtype=torch.float32).contiguous()
        Q = q.shape[0]
        lo = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(lower, dtype=np.float64).reshape(-1)])
        up = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(upper, dtype=np.float64).reshape(-1)])
        dist = torch.empty(Q, dtype=torch.float32, device=q.device)
        loc = torch.empty((Q, 3), dtype=torch.float32, device=q.device)
        self._inputs_ready()
        rid = ctypes.c_uint64()
        lib = _lib.load_library()
        _lib.check(lib.niq_serve_submit_closest_point(self._h, m.handle, ctypes.byref(func.cfg), lo, up,
                                                      _lib.ptr(q), Q, float(eps), int(batch_process_size),
                                                      _lib.ptr(dist), _lib.ptr(loc), ctypes.byref(rid)))
        keep = {"q": q, "dist": dist, "loc": loc, "cfg": func.cfg}
        return Request(self, rid.value, "closest_point", keep, lambda r: (r._keep["dist"], r._keep["loc"]))


def latency_summary(requests):
    """p50/p99 latency (ms) of finished requests, overall and per kind."""
    out = {}
    kinds = sorted({r.kind for r in requests})
    for k in [None] + kinds:
        lat = np.array([r.info["latency_ms"] for r in requests if r.info is not None and (k is None or r.kind == k)])
        if lat.size == 0:
            continue
        out["all" if k is None else k] = {"n": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
                                          "p99_ms": float(np.percentile(lat, 99))}
    return out
