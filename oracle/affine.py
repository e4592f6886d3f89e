"""Oracle: affine-arithmetic range analysis, batched over boxes.

Test infrastructure only (see ``oracle/__init__.py``).

An affine value is the triple ``(base (N,w), aff (N,v,w), err (N,w))``
(src/affine.py:59-60), batched over N boxes.  Rules restated from:
  * AffineContext modes                         -- src/affine.py:62-76
  * radius / may_contain_bounds                 -- src/affine.py:85-91, 119-125
  * coordinates_in_general_box                  -- src/affine.py:109-117
  * truncate_affine                             -- src/affine.py:127-162
  * apply_linear_approx                         -- src/affine.py:164-193
  * dense / relu / elu / sin / encode / squeeze -- src/affine_layers.py:11-180
  * classify_general_box sign test              -- src/affine.py:34-55
  * classify_box (centre + diag(half))          -- src/implicit_function.py:28-37
TanH is not in the reference; ``tanh_rule`` is this project's addition (the
same extrema-enumeration structure as the reference's sin rule).
"""
from dataclasses import dataclass

import numpy as np

from . import mlp as omlp

SIGN_UNKNOWN = 0
SIGN_POSITIVE = 1
SIGN_NEGATIVE = 2

MODES = ("interval", "affine_fixed", "affine_truncate", "affine_append", "affine_all")


@dataclass(frozen=True)
class AffineContext:
    """src/affine.py:62-76."""
    mode: str = "affine_fixed"
    truncate_count: int = -777
    truncate_policy: str = "absolute"
    n_append: int = 0

    def __post_init__(self):
        if self.mode not in MODES:
            raise ValueError("invalid mode")
        if self.mode == "affine_truncate" and self.truncate_count is None:
            raise ValueError("must specify truncate count")


def radius(aff, err):
    """src/affine.py:85-91: sum over error-term rows of |aff|, plus err.

    Sequential in row order (the canonical order of this restatement)."""
    rad = np.zeros_like(err)
    for r in range(aff.shape[1]):
        rad = rad + np.abs(aff[:, r, :])
    return rad + err


def may_contain_bounds(base, aff, err):
    """src/affine.py:119-125."""
    rad = radius(aff, err)
    return base - rad, base + rad


def coordinates_in_general_box(ctx, center, vecs):
    """src/affine.py:109-117.  center (N,d), vecs (N,v,d)."""
    base = center
    if ctx.mode == "interval":
        aff = np.zeros((center.shape[0], 0, center.shape[1]), dtype=center.dtype)
        err = np.zeros_like(center)
        for r in range(vecs.shape[1]):
            err = err + np.abs(vecs[:, r, :])
    else:
        aff = vecs
        err = np.zeros_like(center)
    return base, aff, err


# Selection-gap log: when set to a list, truncate_affine and affine_append
# append, per call, each box's relative gap between the last row kept and the
# first row dropped (the margin of that selection; inf when nothing is
# dropped), so a test can tell cells whose row choice is a near tie.
_gap_log = None


def log_selection_gaps(log):
    """Start (a list) or stop (None) recording selection gaps."""
    global _gap_log
    _gap_log = log


def _record_gap(vals_desc, k):
    """vals_desc (N, m) sorted decreasing; gap between entries k-1 and k."""
    if _gap_log is None:
        return
    if k <= 0 or k >= vals_desc.shape[1]:
        _gap_log.append(np.full(vals_desc.shape[0], np.inf))
        return
    a = vals_desc[:, k - 1].astype(np.float64)
    b = vals_desc[:, k].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(a > 0, (a - b) / a, np.inf)
    _gap_log.append(g)


def truncate_affine(ctx, base, aff, err):
    """src/affine.py:127-162 ('absolute' policy).

    The reference's 'relative' policy divides a (v,) row-magnitude vector by a
    (w,) |base| vector (src/affine.py:146), which only broadcasts when v == w;
    it is therefore rejected here as in every call that could reach it."""
    if ctx.mode != "affine_truncate":
        return base, aff, err
    n_keep = ctx.truncate_count
    if aff.shape[1] <= n_keep:
        return base, aff, err
    if ctx.truncate_policy != "absolute":
        raise ValueError("only the 'absolute' truncate policy is supported")
    mags = np.sum(np.abs(aff), axis=-1)                      # (N, v)
    order = np.argsort(-mags, axis=-1, kind="stable")        # decreasing
    _record_gap(np.take_along_axis(mags, order, axis=1), n_keep)
    aff = np.take_along_axis(aff, order[:, :, None], axis=1)
    aff_keep = aff[:, :n_keep, :]
    aff_drop = aff[:, n_keep:, :]
    drop = np.zeros_like(err)
    for r in range(aff_drop.shape[1]):
        drop = drop + np.abs(aff_drop[:, r, :])
    return base, aff_keep, err + drop


def apply_linear_approx(ctx, base, aff, err, alpha, beta, delta):
    """src/affine.py:164-193."""
    base = alpha * base + beta
    aff = alpha[:, None, :] * aff
    delta = np.abs(delta)
    if ctx.mode in ("interval", "affine_fixed"):
        err = alpha * err + delta
    elif ctx.mode in ("affine_truncate", "affine_all"):
        err = alpha * err
        N, w = delta.shape
        new_aff = np.zeros((N, w, w), dtype=delta.dtype)
        idx = np.arange(w)
        new_aff[:, idx, idx] = delta
        aff = np.concatenate((aff, new_aff), axis=1)
        base, aff, err = truncate_affine(ctx, base, aff, err)
    elif ctx.mode == "affine_append":
        err = alpha * err
        n = ctx.n_append
        # lax.top_k: largest first, lower index first on ties
        full = np.argsort(-delta, axis=-1, kind="stable")
        _record_gap(np.take_along_axis(delta, full, axis=1), n)
        order = full[:, :n]
        keep_vals = np.take_along_axis(delta, order, axis=-1)
        N, w = delta.shape
        new_aff = np.zeros((N, n, w), dtype=delta.dtype)
        rows = np.arange(n)
        for i in range(N):
            new_aff[i, rows, order[i]] = keep_vals[i]
        aff = np.concatenate((aff, new_aff), axis=1)
        dropped = np.sum(delta, axis=-1) - np.sum(keep_vals, axis=-1)
        err = err + dropped[:, None]
    return base, aff, err


# ------------------------------------------------------------- linearisations
def nan_to_num(x):
    """jnp.nan_to_num(x, nan=0.0): NaN -> 0, +-inf -> +-max finite."""
    return np.nan_to_num(x, nan=0.0)


def relu_rule(lower, upper):
    """Chebyshev ReLU, src/affine_layers.py:34-56.  Returns alpha, beta, delta."""
    t = lower.dtype.type
    relu = lambda x: np.maximum(x, t(0))
    with np.errstate(divide="ignore", invalid="ignore"):
        alpha = (relu(upper) - relu(lower)) / (upper - lower)
    alpha = np.where(lower >= 0, t(1), alpha)
    alpha = np.where(upper < 0, t(0), alpha)
    alpha = nan_to_num(alpha)
    alpha = np.clip(alpha, t(0), t(1))
    beta = (relu(lower) - alpha * lower) / t(2)
    delta = beta
    return alpha, beta, delta


def elu_rule(lower, upper):
    """src/affine_layers.py:59-98."""
    t = lower.dtype.type
    lowerF = omlp.elu(lower)
    upperF = omlp.elu(upper)
    with np.errstate(over="ignore"):
        lowerS = np.minimum(np.exp(lower), t(1))
        upperS = np.minimum(np.exp(upper), t(1))
    with np.errstate(divide="ignore", invalid="ignore"):
        alpha = (upperF - lowerF) / (upper - lower)
    alpha = np.where(lower >= 0, t(1), alpha)
    alpha = nan_to_num(alpha)
    alpha = np.minimum(np.maximum(alpha, lowerS), upperS)   # jnp.clip
    r_upper = lowerF - alpha * lower
    with np.errstate(divide="ignore"):
        x_lower = np.minimum(np.maximum(np.log(alpha), lower), upper)
    r_lower = (alpha - t(1)) - alpha * x_lower
    beta = t(0.5) * (r_upper + r_lower)
    delta = t(0.5) * np.abs(r_upper - r_lower)
    pos = lower >= 0
    alpha = np.where(pos, t(1), alpha)
    beta = np.where(pos, t(0), beta)
    delta = np.where(pos, t(0), delta)
    return alpha, beta, delta


def sin_bound(lower, upper):
    """src/utils.py:209-228."""
    t = lower.dtype.type
    f_lower = np.sin(lower)
    f_upper = np.sin(upper)
    two_pi = t(2.0 * np.pi)
    lower = lower / two_pi
    upper = upper / two_pi
    contains_min = np.ceil(lower - t(0.75)) < (upper - t(0.75))
    contains_max = np.ceil(lower - t(0.25)) < (upper - t(0.25))
    out_lower = np.where(contains_min, t(-1), np.minimum(f_lower, f_upper))
    out_upper = np.where(contains_max, t(1), np.maximum(f_lower, f_upper))
    return out_lower, out_upper


def cos_bound(lower, upper):
    """src/utils.py:230-231."""
    t = lower.dtype.type
    h = t(np.pi / 2)
    return sin_bound(lower + h, upper + h)


def sin_rule(lower, upper):
    """src/affine_layers.py:100-138."""
    t = lower.dtype.type
    two_pi = t(2.0 * np.pi)
    slope_lower, slope_upper = cos_bound(lower, upper)
    alpha = t(0.5) * (slope_lower + slope_upper)
    alpha = np.clip(alpha, t(-1), t(1))
    intA = np.arccos(alpha)
    intB = -intA

    def first(x):
        return two_pi * np.ceil((lower + x) / two_pi) - x

    def last(x):
        return two_pi * np.floor((upper - x) / two_pi) + x

    locs = [lower, upper, first(intA), last(intA), first(intB), last(intB)]
    locs = [np.minimum(np.maximum(x, lower), upper) for x in locs]
    vals = [np.sin(x) - alpha * x for x in locs]
    r_lower = np.min(np.stack(vals, 0), axis=0)
    r_upper = np.max(np.stack(vals, 0), axis=0)
    beta = t(0.5) * (r_upper + r_lower)
    delta = r_upper - beta
    return alpha, beta, delta


def tanh_rule(lower, upper):
    """TanH linearisation (no reference counterpart; this project's rule).

    alpha is the secant slope (tanh'(lower) when lower == upper), clipped to
    [0, 1].  g(x) = tanh(x) - alpha*x has critical points where
    1 - tanh(x)^2 = alpha, i.e. x = +-atanh(sqrt(1 - alpha)); its extrema on
    [lower, upper] lie at the endpoints or at those points clipped into the
    interval.  beta, delta centre and half-width the range of g, so
    |tanh(x) - (alpha*x + beta)| <= delta on [lower, upper]."""
    t = lower.dtype.type
    tl = np.tanh(lower)
    tu = np.tanh(upper)
    with np.errstate(divide="ignore", invalid="ignore"):
        alpha = (tu - tl) / (upper - lower)
    dl = t(1) - tl * tl
    alpha = np.where(upper > lower, alpha, dl)
    alpha = nan_to_num(alpha)
    alpha = np.clip(alpha, t(0), t(1))
    with np.errstate(divide="ignore"):
        xs = np.arctanh(np.sqrt(t(1) - alpha))
    locs = [lower, upper, xs, -xs]
    locs = [np.minimum(np.maximum(x, lower), upper) for x in locs]
    vals = [np.tanh(x) - alpha * x for x in locs]
    r_lower = np.min(np.stack(vals, 0), axis=0)
    r_upper = np.max(np.stack(vals, 0), axis=0)
    beta = t(0.5) * (r_upper + r_lower)
    delta = t(0.5) * (r_upper - r_lower)
    return alpha, beta, delta


RULES = {"relu": relu_rule, "elu": elu_rule, "sin": sin_rule, "tanh": tanh_rule}


# --------------------------------------------------------------- layer rules
def dense(base, aff, err, A, b):
    """src/affine_layers.py:11-32: base.A+b, aff rows .A, err.|A|."""
    base = base @ A
    aff = aff @ A
    err = err @ np.abs(A)
    if b is not None:
        base = base + b
    return base, aff, err


def activation(ctx, name, base, aff, err):
    lower, upper = may_contain_bounds(base, aff, err)
    alpha, beta, delta = RULES[name](lower, upper)
    return apply_linear_approx(ctx, base, aff, err, alpha, beta, delta)


def forward(ctx, params, base, aff, err, dtype=np.float32):
    """Walk the op list (src/mlp.py:96-113) with the affine rule table."""
    for name, args in omlp.op_list(params):
        if name == "dense":
            A = np.asarray(args["A"], dtype=dtype)
            b = np.asarray(args["b"], dtype=dtype) if "b" in args else None
            base, aff, err = dense(base, aff, err, A, b)
        elif name in RULES:
            base, aff, err = activation(ctx, name, base, aff, err)
        elif name == "squeeze_last":
            # src/affine_layers.py:164-173
            base, aff, err = base[:, 0], aff[:, :, 0], err[:, 0]
        elif name == "pow2_frequency_encode":
            # src/affine_layers.py:140-162
            coefs = np.asarray(args["coefs"], dtype=dtype)
            shift = np.asarray(args["shift"], dtype=dtype) if "shift" in args else None
            N, d = base.shape
            b2 = base[:, :, None] * coefs[None, None, :]
            if shift is not None:
                b2 = b2 + shift
            base = b2.reshape(N, -1)
            aff = (aff[:, :, :, None] * coefs).reshape(N, aff.shape[1], -1)
            err = (err[:, :, None] * coefs).reshape(N, -1)
        elif name == "spatial_transformation":
            # src/affine_layers.py:175-180
            R_inv, t_inv = omlp.spatial_inverse(args["R"], args["t"], dtype)
            base, aff, err = dense(base, aff, err, R_inv, t_inv)
        else:
            raise ValueError(f"unsupported op {name}")
    return base, aff, err


def sign_code(lower, upper, offset):
    """src/affine.py:51-53 (strict inequalities)."""
    t = lower.dtype.type
    code = np.full(lower.shape, SIGN_UNKNOWN, dtype=np.int32)
    code = np.where(lower > t(offset), SIGN_POSITIVE, code)
    code = np.where(upper < -t(offset), SIGN_NEGATIVE, code)
    return code


def bound_general_boxes(ctx, params, center, vecs, dtype=np.float32):
    """Output affine form of a batch of general boxes; returns base, rad."""
    center = np.asarray(center, dtype=dtype)
    vecs = np.asarray(vecs, dtype=dtype)
    base, aff, err = coordinates_in_general_box(ctx, center, vecs)
    base, aff, err = forward(ctx, params, base, aff, err, dtype)
    return base, radius(aff[:, :, None], err[:, None])[:, 0]


def classify_general_boxes(ctx, params, center, vecs, offset=0.0, dtype=np.float32,
                           return_bounds=False):
    """src/affine.py:34-55 batched: codes (N,) int32 [, lower, upper]."""
    base, rad = bound_general_boxes(ctx, params, center, vecs, dtype)
    lower, upper = base - rad, base + rad
    code = sign_code(lower, upper, offset)
    if return_bounds:
        return code, lower, upper
    return code


def box_to_general(lower, upper, dtype=np.float32):
    """src/implicit_function.py:28-37: centre and diag(upper - centre)."""
    lower = np.asarray(lower, dtype=dtype)
    upper = np.asarray(upper, dtype=dtype)
    t = np.dtype(dtype).type
    center = t(0.5) * (lower + upper)
    pos_vec = upper - center
    N, d = lower.shape
    vecs = np.zeros((N, d, d), dtype=dtype)
    idx = np.arange(d)
    vecs[:, idx, idx] = pos_vec
    return center, vecs


def classify_boxes(ctx, params, lower, upper, offset=0.0, dtype=np.float32,
                   return_bounds=False):
    center, vecs = box_to_general(lower, upper, dtype)
    return classify_general_boxes(ctx, params, center, vecs, offset, dtype, return_bounds)
