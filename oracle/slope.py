"""Oracle: slope-interval range analysis, batched over boxes.

Test infrastructure only (see ``oracle/__init__.py``).

A slope-interval value is ``(primal (N,w), slope_center (N,v,w),
slope_width (N,v,w))`` (src/slope_interval.py:164-173).  Restated from:
  * coordinates_in_general_box                    -- src/slope_interval.py:192-197
  * slope_bounds / primal_may_contain_bounds      -- src/slope_interval.py:200-210
  * classify_general_box                          -- src/slope_interval.py:29-51
  * dense / relu / elu / sin / encode / squeeze   -- src/slope_interval_layers.py:11-143
  * biggest_magnitude                             -- src/utils.py:205-206
TanH's derivative-bound rule is this project's addition.
"""
import numpy as np

from . import mlp as omlp
from .affine import cos_bound, sign_code, box_to_general


def slope_bounds(c, w):
    return c - w, c + w


def primal_bounds(primal, sl, su):
    """src/slope_interval.py:205-210; row sum sequential."""
    rad = np.zeros_like(primal)
    for r in range(sl.shape[1]):
        rad = rad + np.maximum(su[:, r, :], -sl[:, r, :])
    return primal - rad, primal + rad


def _mul_nonneg(sl, su, dfl, dfu):
    """src/slope_interval_layers.py:38-56 (df >= 0 case)."""
    nl = np.minimum(sl * dfl[:, None, :], sl * dfu[:, None, :])
    nu = np.maximum(su * dfl[:, None, :], su * dfu[:, None, :])
    return nl, nu


def activation(name, primal, c, w):
    t = primal.dtype.type
    sl, su = slope_bounds(c, w)
    pl, pu = primal_bounds(primal, sl, su)
    if name == "relu":            # src/slope_interval_layers.py:36-58
        dfl = np.where(pl > 0, t(1), t(0))
        dfu = np.where(pu < 0, t(0), t(1))
        nl, nu = _mul_nonneg(sl, su, dfl, dfu)
    elif name == "elu":           # src/slope_interval_layers.py:60-82
        with np.errstate(over="ignore"):
            dfl = np.minimum(np.exp(pl), t(1))
            dfu = np.minimum(np.exp(pu), t(1))
        nl, nu = _mul_nonneg(sl, su, dfl, dfu)
    elif name == "sin":           # src/slope_interval_layers.py:84-109
        dfl, dfu = cos_bound(pl, pu)
        vals = [sl * dfl[:, None, :], sl * dfu[:, None, :],
                su * dfl[:, None, :], su * dfu[:, None, :]]
        nl = np.min(np.stack(vals, 0), axis=0)
        nu = np.max(np.stack(vals, 0), axis=0)
    elif name == "tanh":          # addition: tanh' = 1 - tanh^2 on [pl, pu]
        dl = t(1) - np.tanh(pl) ** 2
        du = t(1) - np.tanh(pu) ** 2
        dfl = np.minimum(dl, du)
        dfu = np.where((pl <= 0) & (pu >= 0), t(1), np.maximum(dl, du))
        nl, nu = _mul_nonneg(sl, su, dfl, dfu)
    else:
        raise ValueError(name)
    new_c = t(0.5) * (nl + nu)
    new_w = nu - new_c
    return omlp.act_point(name, primal), new_c, new_w


def forward(params, primal, c, w, dtype=np.float32):
    for name, args in omlp.op_list(params):
        if name == "dense":       # src/slope_interval_layers.py:11-34
            A = np.asarray(args["A"], dtype=dtype)
            primal = primal @ A
            if "b" in args:
                primal = primal + np.asarray(args["b"], dtype=dtype)
            c = c @ A
            w = w @ np.abs(A)
        elif name in omlp.ACTIVATIONS:
            primal, c, w = activation(name, primal, c, w)
        elif name == "squeeze_last":
            primal, c, w = primal[:, 0], c[:, :, 0], w[:, :, 0]
        elif name == "pow2_frequency_encode":
            coefs = np.asarray(args["coefs"], dtype=dtype)
            N = primal.shape[0]
            p2 = primal[:, :, None] * coefs
            if "shift" in args:
                p2 = p2 + np.asarray(args["shift"], dtype=dtype)
            primal = p2.reshape(N, -1)
            c = (c[:, :, :, None] * coefs).reshape(N, c.shape[1], -1)
            w = (w[:, :, :, None] * coefs).reshape(N, w.shape[1], -1)
        elif name == "spatial_transformation":
            R_inv, t_inv = omlp.spatial_inverse(args["R"], args["t"], dtype)
            primal = primal @ R_inv + t_inv
            c = c @ R_inv
            w = w @ np.abs(R_inv)
        else:
            raise ValueError(f"unsupported op {name}")
    return primal, c, w


def bound_general_boxes(params, center, vecs, dtype=np.float32):
    center = np.asarray(center, dtype=dtype)
    vecs = np.asarray(vecs, dtype=dtype)
    c = vecs.copy()                   # src/slope_interval.py:192-197
    w = np.zeros_like(c)
    primal, c, w = forward(params, center, c, w, dtype)
    sl, su = slope_bounds(c[:, :, None], w[:, :, None])
    lo, hi = primal_bounds(primal[:, None], sl, su)
    return lo[:, 0], hi[:, 0]


def classify_general_boxes(params, center, vecs, offset=0.0, dtype=np.float32,
                           return_bounds=False):
    lo, hi = bound_general_boxes(params, center, vecs, dtype)
    code = sign_code(lo, hi, offset)
    if return_bounds:
        return code, lo, hi
    return code


def classify_boxes(params, lower, upper, offset=0.0, dtype=np.float32, return_bounds=False):
    center, vecs = box_to_general(lower, upper, dtype)
    return classify_general_boxes(params, center, vecs, offset, dtype, return_bounds)
