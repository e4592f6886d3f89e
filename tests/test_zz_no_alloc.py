"""niq_cast_rays_noalloc against niq_cast_rays, through the C ABI.

The no-allocation entry point runs the same persistent ray kernel as
niq_cast_rays with its workspace in a static device array: its outputs
(t, hit_id, count, N_evals) must equal the workspace entry point's bit for
bit.  Checked on two calls in a row, the second with fewer rays and a smaller
n_max_step, so that the state the first call leaves in the shared array
(work counter, histogram) cannot leak into the second; and the size limit is
refused with NIQ_ERR_CAPACITY before anything runs.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

from niq import _lib, implicit_mlp_utils, models, render, queries

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc_fn():
    lib = _lib.load_library()
    fn = lib.niq_cast_rays_noalloc
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    fn.restype = I
    fn.argtypes = [P, I, P, P, P, I64, P, P, P, P, ctypes.POINTER(I64), P]
    return lib, fn


def cast_both(func, params, roots, dirs, opts):
    from niq import mlp as nmlp
    lib, fn = noalloc_fn()
    m = nmlp.device_model(params)
    handles = (ctypes.c_void_p * 1)(m.handle.value)
    o = queries.cast_opts_struct(opts)
    n = roots.shape[0]
    outs = []
    for which in ("workspace", "noalloc"):
        t = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        hit = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        cnt = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        evals = ctypes.c_int64(-1)
        stream = _lib.stream_handle()
        torch.cuda.synchronize()
        if which == "workspace":
            ws_bytes = lib.niq_cast_workspace_bytes(ctypes.byref(o))
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
            _lib.check(lib.niq_cast_rays(handles, 1, ctypes.byref(func.cfg), _lib.ptr(roots), _lib.ptr(dirs), n,
                                         ctypes.byref(o), _lib.ptr(t), _lib.ptr(hit), _lib.ptr(cnt), _lib.ptr(ws),
                                         ws_bytes, ctypes.byref(evals), stream))
        else:
            _lib.check(fn(handles, 1, ctypes.byref(func.cfg), _lib.ptr(roots), _lib.ptr(dirs), n, ctypes.byref(o),
                          _lib.ptr(t), _lib.ptr(hit), _lib.ptr(cnt), ctypes.byref(evals), stream))
        torch.cuda.synchronize()
        outs.append((t, hit, cnt, evals.value))
    return outs


def camera(res):
    eye = np.array([0., 0., 3.], np.float32)
    look, up, _ = render.look_at(eye)
    return render.generate_camera_rays(eye, look, up, res=res, fov_deg=60.0, device=None)


def test_noalloc_cast_equals_workspace_cast_twice():
    p = models.config_model("c1_relu_3x32")
    f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    roots, dirs = camera(96)
    roots = torch.as_tensor(np.asarray(roots, np.float32), device="cuda").contiguous()
    dirs = torch.as_tensor(np.asarray(dirs, np.float32), device="cuda").contiguous()
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    for k, (n, n_max_step) in enumerate([(96 * 96, 512), (2500, 64)]):
        note(f"[noalloc] call {k + 1}: {n} rays, n_max_step {n_max_step}")
        opts["n_max_step"] = n_max_step
        (t0, h0, c0, e0), (t1, h1, c1, e1) = cast_both(f, p, roots[:n].contiguous(), dirs[:n].contiguous(), opts)
        assert torch.equal(t0, t1) and torch.equal(h0, h1) and torch.equal(c0, c1)
        assert e0 == e1 and e0 > 0
        assert int((c1 >= 0).sum()) == n
        note(f"  hits {int((h1 > 0).sum())}, N_evals {e1}")


def test_noalloc_refuses_above_its_scratch():
    lib, fn = noalloc_fn()
    p = models.config_model("c1_relu_3x32")
    f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    from niq import mlp as nmlp
    m = nmlp.device_model(p)
    handles = (ctypes.c_void_p * 1)(m.handle.value)
    opts = queries.get_default_cast_opts()
    opts["n_max_step"] = 20000
    o = queries.cast_opts_struct(opts)
    roots = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    out = torch.zeros(4, dtype=torch.float32, device="cuda")
    outi = torch.zeros(4, dtype=torch.int32, device="cuda")
    evals = ctypes.c_int64(0)
    st = fn(handles, 1, ctypes.byref(f.cfg), _lib.ptr(roots), _lib.ptr(roots), 4, ctypes.byref(o), _lib.ptr(out),
            _lib.ptr(outi), _lib.ptr(outi), ctypes.byref(evals), _lib.stream_handle())
    assert st == _lib.NIQ_ERR_CAPACITY
