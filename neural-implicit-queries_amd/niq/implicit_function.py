"""The ``ImplicitFunction`` plugin interface, backed by the gfx950 kernels.

Mirrors src/implicit_function.py:15-70: sign codes, ``classify_box`` (centre
and diag(half-widths), src/implicit_function.py:28-37),
``classify_general_box`` and ``__call__``.  Where the reference evaluates one
box per call under ``jax.vmap``, these methods take a batch (leading axis) or
a single box/point and return a ``torch`` tensor on the GPU.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import mlp as _mlp

SIGN_UNKNOWN = 0    # could be anything
SIGN_POSITIVE = 1   # definitely positive throughout
SIGN_NEGATIVE = 2   # definitely negative throughout


def as_device_f32(x):
    """float32, contiguous, on the current GPU."""
    if isinstance(x, torch.Tensor):
        t = x.detach()
    else:
        t = torch.as_tensor(np.asarray(x, dtype=np.float32))
    return t.to(device="cuda", dtype=torch.float32).contiguous()


class ImplicitFunction:
    """Base class: subclasses set ``self.cfg`` (an ``_lib.EvalCfg``)."""

    def __init__(self, style):
        if style not in ["classify-only", "classify-and-distance"]:
            raise ValueError("unrecognized style")
        self.style = style
        self.cfg = _lib.eval_cfg()

    # -- plain evaluation -------------------------------------------------
    def __call__(self, params, x):
        _lib.require_gpu()
        x = as_device_f32(x)
        single = x.dim() == 1
        xb = x.reshape(1, -1) if single else x
        out = torch.empty(xb.shape[0], dtype=torch.float32, device=xb.device)
        m = _mlp.device_model(params)
        _lib.check(_lib.load_library().niq_eval_points(
            m.handle, _lib.ptr(xb), xb.shape[0], xb.shape[1], _lib.ptr(out), _lib.stream_handle()))
        return out[0] if single else out

    # -- classification -----------------------------------------------------
    def classify_box(self, params, box_lower, box_upper, offset=0.0, return_bounds=False):
        _lib.require_gpu()
        lo = as_device_f32(box_lower)
        up = as_device_f32(box_upper)
        single = lo.dim() == 1
        lo2 = lo.reshape(1, -1) if single else lo
        up2 = up.reshape(1, -1) if single else up
        if lo2.shape != up2.shape:
            raise ValueError("box_lower and box_upper shapes differ")
        n, d = lo2.shape
        codes = torch.empty(n, dtype=torch.int32, device=lo2.device)
        bounds = torch.empty((n, 2), dtype=torch.float32, device=lo2.device) if return_bounds else None
        m = _mlp.device_model(params)
        _lib.check(_lib.load_library().niq_classify_boxes(
            m.handle, ctypes.byref(self.cfg), _lib.ptr(lo2), _lib.ptr(up2), n, d, float(offset),
            _lib.ptr(codes), _lib.ptr(bounds), _lib.stream_handle()))
        if single:
            codes = codes[0]
            bounds = bounds[0] if bounds is not None else None
        return (codes, bounds) if return_bounds else codes

    def classify_general_box(self, params, box_center, box_vecs, offset=0.0, return_bounds=False):
        _lib.require_gpu()
        c = as_device_f32(box_center)
        v = as_device_f32(box_vecs)
        single = c.dim() == 1
        c2 = c.reshape(1, -1) if single else c
        v2 = v.reshape(1, *v.shape) if single else v
        n, d = c2.shape
        if v2.dim() != 3 or v2.shape[0] != n or v2.shape[2] != d:
            raise ValueError("bad box_vecs shape")
        codes = torch.empty(n, dtype=torch.int32, device=c2.device)
        bounds = torch.empty((n, 2), dtype=torch.float32, device=c2.device) if return_bounds else None
        m = _mlp.device_model(params)
        _lib.check(_lib.load_library().niq_classify_general_boxes(
            m.handle, ctypes.byref(self.cfg), _lib.ptr(c2), _lib.ptr(v2), n, v2.shape[1], d,
            float(offset), _lib.ptr(codes), _lib.ptr(bounds), _lib.stream_handle()))
        if single:
            codes = codes[0]
            bounds = bounds[0] if bounds is not None else None
        return (codes, bounds) if return_bounds else codes

    def min_distance_to_zero(self, params, box_center, box_axis_vec):
        raise RuntimeError("min_distance_to_zero is not part of the range-analysis hot path")

    def min_distance_to_zero_in_direction(self, params, source_point, bound_vec, source_range=None,
                                          return_source_value=False):
        raise RuntimeError("min_distance_to_zero_in_direction is not part of the range-analysis hot path")
