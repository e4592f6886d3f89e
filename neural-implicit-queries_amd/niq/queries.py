"""Ray casting against neural implicits (src/queries.py:23-175).

``cast_rays`` keeps the reference's signature and return values.  All rays
march inside one persistent kernel (niq_cast_rays): a workgroup holds a tile
of ray slots, bounds the next segment of every live ray with the fused
range-analysis forward pass, and refills finished slots from a device counter.
The host synchronises once per call.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import mlp as _mlp
from .implicit_function import as_device_f32


def get_default_cast_opts():
    """src/queries.py:23-36."""
    d = {
        'hit_eps': 0.001,
        'max_dist': 10.,
        'n_max_step': 512,
        'n_substeps': 1,
        'safety_factor': 0.98,
        'interval_grow_fac': 1.5,
        'interval_shrink_fac': 0.5,
        'interval_init_size': 0.1,  # relative, as a factor of max_dist
        'refine_width_fac': 2.,
        'n_side_init': 16,
    }
    return d


def cast_opts_struct(opts):
    o = _lib.CastOpts()
    o.hit_eps = float(opts['hit_eps'])
    o.max_dist = float(opts['max_dist'])
    o.n_max_step = int(opts['n_max_step'])
    o.n_substeps = int(opts['n_substeps'])
    o.safety_factor = float(opts['safety_factor'])
    o.interval_grow_fac = float(opts['interval_grow_fac'])
    o.interval_shrink_fac = float(opts['interval_shrink_fac'])
    # jnp.ones(N) * (init_size * max_dist): the product is formed in Python
    # (double) and rounded once to float32 (src/queries.py:148)
    o.init_step = float(np.float32(opts['interval_init_size'] * opts['max_dist']))
    return o


def cast_rays(funcs_tuple, params_tuple, roots, dirs, opts, stream=None):
    """Returns (t (N,) f32, hit_id (N,) int32, count (N,) int32, N_evals)."""
    if not isinstance(funcs_tuple, (tuple, list)):
        funcs_tuple = (funcs_tuple,)
    if not isinstance(params_tuple, (tuple, list)):
        params_tuple = (params_tuple,)
    if len(funcs_tuple) != len(params_tuple):
        raise ValueError("funcs_tuple and params_tuple lengths differ")
    if len({bytes(f.cfg) for f in funcs_tuple}) != 1:
        raise ValueError("all functions of one cast_rays call must use the same range-analysis mode and options")
    _lib.require_gpu()
    lib = _lib.load_library()
    roots = as_device_f32(roots)
    dirs = as_device_f32(dirs)
    N = roots.shape[0]
    models = [_mlp.device_model(p) for p in params_tuple]
    handles = (ctypes.c_void_p * len(models))(*[m.handle.value for m in models])
    o = cast_opts_struct(opts)
    dev = roots.device
    out_t = torch.empty(N, dtype=torch.float32, device=dev)
    out_hit = torch.empty(N, dtype=torch.int32, device=dev)
    out_count = torch.empty(N, dtype=torch.int32, device=dev)
    ws_bytes = lib.niq_cast_workspace_bytes(ctypes.byref(o))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    n_evals = ctypes.c_int64(0)
    _lib.check(lib.niq_cast_rays(handles, len(models), ctypes.byref(funcs_tuple[0].cfg),
                                 _lib.ptr(roots), _lib.ptr(dirs), N, ctypes.byref(o),
                                 _lib.ptr(out_t), _lib.ptr(out_hit), _lib.ptr(out_count),
                                 _lib.ptr(ws), ws_bytes, ctypes.byref(n_evals),
                                 _lib.stream_handle(stream)))
    return out_t, out_hit, out_count, int(n_evals.value)


def cast_rays_frustum(funcs_tuple, params_tuple, cam_params, opts, stream=None):
    """src/queries.py:454-587: (t, hit_id, count) of shape (res_x, res_y),
    indexed [x, y], and N_evals.  ``cam_params`` is the reference's
    (root_pos, look_dir, up_dir, left_dir, fov_x, fov_y, res_x, res_y)."""
    if not isinstance(funcs_tuple, (tuple, list)):
        funcs_tuple = (funcs_tuple,)
    if not isinstance(params_tuple, (tuple, list)):
        params_tuple = (params_tuple,)
    if len(funcs_tuple) != len(params_tuple):
        raise ValueError("funcs_tuple and params_tuple lengths differ")
    if len({bytes(f.cfg) for f in funcs_tuple}) != 1:
        raise ValueError("all functions of one cast must use the same range-analysis mode and options")
    _lib.require_gpu()
    lib = _lib.load_library()
    root_pos, look_dir, up_dir, left_dir, fov_x, fov_y, res_x, res_y = cam_params
    cam = _lib.FrustumCam()
    for k in range(3):
        cam.root[k] = float(np.asarray(root_pos, np.float32).reshape(-1)[k])
        cam.look[k] = float(np.asarray(look_dir, np.float32).reshape(-1)[k])
        cam.up[k] = float(np.asarray(up_dir, np.float32).reshape(-1)[k])
        cam.left[k] = float(np.asarray(left_dir, np.float32).reshape(-1)[k])
    cam.fov_x, cam.fov_y = float(np.float32(fov_x)), float(np.float32(fov_y))
    cam.res_x, cam.res_y = int(res_x), int(res_y)
    fo = _lib.FrustumOpts()
    fo.cast = cast_opts_struct(opts)
    fo.refine_width_fac = float(opts['refine_width_fac'])
    fo.n_side_init = int(opts['n_side_init'])
    models = [_mlp.device_model(p) for p in params_tuple]
    handles = (ctypes.c_void_p * len(models))(*[m.handle.value for m in models])
    dev = torch.device("cuda")
    out_t = torch.empty((cam.res_x, cam.res_y), dtype=torch.float32, device=dev)
    out_hit = torch.empty((cam.res_x, cam.res_y), dtype=torch.int32, device=dev)
    out_count = torch.empty((cam.res_x, cam.res_y), dtype=torch.int32, device=dev)
    n_evals = ctypes.c_int64(0)
    _lib.check(lib.niq_cast_rays_frustum(handles, len(models), ctypes.byref(funcs_tuple[0].cfg), ctypes.byref(cam),
                                         ctypes.byref(fo), _lib.ptr(out_t), _lib.ptr(out_hit), _lib.ptr(out_count),
                                         ctypes.byref(n_evals), _lib.stream_handle(stream)))
    return out_t, out_hit, out_count, int(n_evals.value)
