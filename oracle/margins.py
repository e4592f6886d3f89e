"""Oracle: near-threshold accounting for decisions of the queries.

Test infrastructure only (see ``oracle/__init__.py``).

A GPU and a CPU FP32 evaluation of the same bound or value differ in summation
order, so their decisions (the sign test of a classified box,
src/affine.py:47-53; the sign of a sample, src/queries.py:67-70,
src/kd_tree.py:706-708) can differ only where the quantity tested lies close
to its threshold.  ``TwinFunc`` evaluates every decision in FP32 (the result
the queries use, exactly as ``OracleFunc``) and again in FP64 from the same
FP32 inputs, and marks a decision *near* when the FP64 value lies within

    NEAR_REL * scale + NOISE_FAC * |x32 - x64|

of the threshold: ``scale`` is |lower| + |upper| for a bound, and for a point
value the magnitude of the final sum that forms it (``out_scale``);
``|x32 - x64|`` is the FP32 rounding noise of that decision.  In
affine_truncate and affine_append a classification is also near when its row
selection is a near tie (``TIE_REL``).  The query oracles (rays, frustum, closest point) accumulate these
marks per ray, frustum pixel or query when given TwinFunc objects, so a test
can require every result that differs from the oracle to carry one.
"""
import numpy as np

from . import affine as oaff
from . import funcs as ofuncs
from . import mlp as omlp

NEAR_REL = 1e-6
NOISE_FAC = 4.0
# affine_truncate / affine_append choose rows by comparing FP32 magnitudes: a
# box whose last kept and first dropped magnitudes lie within TIE_REL
# (relative) of each other is a near tie, whose row choice rounding can flip
TIE_REL = 1e-5


def out_scale(params, x):
    """Magnitude of the sum that forms f(x): sum_i |A_i0 h_i| + |b_0| over the
    last dense layer, with h its actual (FP64) input.  An FP32 value of f is
    correctly signed unless f lies within a few rounding errors of this sum
    (plus what earlier layers propagate, which the FP32-FP64 noise term of
    ``near_values`` measures)."""
    h = np.asarray(x, np.float64)
    ops = omlp.op_list(params)
    last = max(i for i, (name, _) in enumerate(ops) if name == "dense")
    for i, (name, args) in enumerate(ops):
        if name == "dense":
            A = np.asarray(args["A"], np.float64)
            b = np.asarray(args["b"], np.float64) if "b" in args else np.zeros(A.shape[1])
            if i == last:
                return np.abs(h) @ np.abs(A[:, :1])[:, 0] + abs(b[0])
            h = h @ A + b
        elif name in omlp.ACTIVATIONS:
            h = omlp.act_point(name, h)
        elif name == "pow2_frequency_encode":
            coefs = np.asarray(args["coefs"], np.float64)
            h = h[:, :, None] * coefs[None, None, :]
            if "shift" in args:
                h = h + np.asarray(args["shift"], np.float64)
            h = h.reshape(h.shape[0], -1)
        elif name == "spatial_transformation":
            R_inv, t_inv = omlp.spatial_inverse(args["R"], args["t"], np.float64)
            h = h @ R_inv + t_inv
        else:
            raise ValueError(f"unsupported op {name}")
    raise ValueError("no dense layer")


def near_bounds(lo32, hi32, lo64, hi64):
    """Sign tests of a bound [lo, hi] against 0 that FP32 rounding could flip."""
    scale = np.abs(lo64) + np.abs(hi64)
    n_lo = np.abs(lo32.astype(np.float64) - lo64)
    n_hi = np.abs(hi32.astype(np.float64) - hi64)
    return ((np.abs(lo64) <= NEAR_REL * scale + NOISE_FAC * n_lo) |
            (np.abs(hi64) <= NEAR_REL * scale + NOISE_FAC * n_hi))


def near_values(v32, v64, scale):
    """Signs of point values that FP32 rounding could flip."""
    return np.abs(v64) <= NEAR_REL * scale + NOISE_FAC * np.abs(v32.astype(np.float64) - v64)


class TwinFunc:
    """``OracleFunc`` in FP32 that also marks near-threshold decisions.

    Every call returns what the FP32 ``OracleFunc`` returns; ``last_near``
    then holds one flag per batch element of that call, and ``n_near`` counts
    all flagged decisions so far."""

    def __init__(self, params, mode="affine_fixed", **kwargs):
        self.f32 = ofuncs.OracleFunc(params, mode, np.float32, **kwargs)
        self.f64 = ofuncs.OracleFunc(params, mode, np.float64, **kwargs)
        self.params = params
        self.mode = mode
        self.last_near = None
        self.n_near = 0

    def _mark(self, near):
        self.last_near = near
        self.n_near += int(np.count_nonzero(near))

    def __call__(self, x):
        x32 = np.atleast_2d(np.asarray(x, np.float32))
        v32 = self.f32(x32)
        v64 = self.f64(x32.astype(np.float64))
        self._mark(near_values(v32, v64, out_scale(self.params, x32)))
        return v32

    def _sdf_near(self, center, vecs):
        c32 = np.asarray(center, np.float32)
        v32 = np.asarray(self.f32(c32), np.float64).reshape(-1)
        v64 = np.asarray(self.f64(c32.astype(np.float64)), np.float64).reshape(-1)
        r64 = ofuncs.sdf_radius(np.asarray(vecs, np.float32).astype(np.float64), np.float64) * self.f32.lipschitz
        r32 = ofuncs.sdf_radius(vecs, np.float32).astype(np.float64) * self.f32.lipschitz
        noise = np.abs(v32 - v64) + np.abs(r32 - r64)
        gap = np.abs(np.abs(v64) - r64)
        return ((gap <= NEAR_REL * (np.abs(v64) + r64) + NOISE_FAC * noise) |
                (np.abs(v64) <= NEAR_REL * out_scale(self.params, c32) + NOISE_FAC * np.abs(v32 - v64)))

    def classify_general_box(self, center, vecs, offset=0.0):
        center = np.asarray(center, np.float32)
        vecs = np.asarray(vecs, np.float32)
        if self.mode == "sdf":
            self._mark(self._sdf_near(center, vecs))
            return self.f32.classify_general_box(center, vecs, offset)
        gaps = []
        if self.mode in ("affine_truncate", "affine_append"):
            oaff.log_selection_gaps(gaps)
        try:
            lo32, hi32 = self.f32.bounds_general(center, vecs)
        finally:
            oaff.log_selection_gaps(None)
        lo64, hi64 = self.f64.bounds_general(center.astype(np.float64), vecs.astype(np.float64))
        near = near_bounds(lo32, hi32, lo64 - offset, hi64 + offset)
        if gaps:
            near = near | (np.min(np.stack(gaps), axis=0) <= TIE_REL)
        self._mark(near)
        return oaff.sign_code(lo32, hi32, offset)

    def classify_box(self, lower, upper, offset=0.0):
        center, vecs = oaff.box_to_general(lower, upper, np.float32)
        return self.classify_general_box(center, vecs, offset)
