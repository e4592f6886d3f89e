"""Oracle: implicit-function objects by mode (batched).

Test infrastructure only (see ``oracle/__init__.py``).

Mirrors ``generate_implicit_from_file`` (src/implicit_mlp_utils.py:12-64) for
the range-analysis modes, and the ``ImplicitFunction`` interface
(src/implicit_function.py:15-70), vectorised over boxes / points.
"""
import numpy as np

from . import affine as oaff
from . import slope as oslope
from . import mlp as omlp

MODES = ("interval", "affine_fixed", "affine_truncate", "affine_append",
         "affine_all", "slope_interval", "sdf")


def sdf_radius(vecs, dtype=np.float32):
    """src/sdf.py:34: sqrt(sum_v |vec_v|^2), each norm and sum in order."""
    vecs = np.asarray(vecs, dtype=dtype)
    t = np.dtype(dtype).type
    nv = np.sqrt(np.sum(vecs * vecs, axis=-1, dtype=dtype)).astype(dtype)     # (N, v)
    r2 = np.zeros(vecs.shape[0], dtype)
    for i in range(vecs.shape[1]):
        r2 = (r2 + nv[:, i] * nv[:, i]).astype(dtype)
    return np.sqrt(r2).astype(dtype)


class OracleFunc:
    def __init__(self, params, mode="affine_fixed", dtype=np.float32, **kwargs):
        if mode not in MODES:
            raise RuntimeError("unrecognized mode")
        self.params = params
        self.mode = mode
        self.dtype = dtype
        self.lipschitz = kwargs.get("sdf_lipschitz", 1.0)
        if mode not in ("slope_interval", "sdf"):
            self.ctx = oaff.AffineContext(
                mode,
                truncate_count=kwargs.get("affine_n_truncate", -777),
                truncate_policy=kwargs.get("affine_truncate_policy", "absolute"),
                n_append=kwargs.get("affine_n_append", 0))

    def __call__(self, x):
        return omlp.eval_points(self.params, np.atleast_2d(x), self.dtype)

    def bounds_general(self, center, vecs):
        """Returns (lower, upper) of the output over each general box (sdf:
        the value at the centre -/+ lipschitz * radius)."""
        if self.mode == "sdf":
            t = np.dtype(self.dtype).type
            val = np.asarray(self(center), self.dtype).reshape(-1)
            r = (sdf_radius(vecs, self.dtype) * t(self.lipschitz)).astype(self.dtype)
            return val - r, val + r
        if self.mode == "slope_interval":
            return oslope.bound_general_boxes(self.params, center, vecs, self.dtype)
        base, rad = oaff.bound_general_boxes(self.ctx, self.params, center, vecs, self.dtype)
        return base - rad, base + rad

    def classify_general_box(self, center, vecs, offset=0.0):
        if self.mode == "sdf":
            return self.sdf_code(center, vecs, offset)
        lo, hi = self.bounds_general(center, vecs)
        return oaff.sign_code(lo, hi, offset)

    def sdf_code(self, center, vecs, offset=0.0):
        """WeakSDFImplicitFunction.classify_general_box (src/sdf.py:31-50)."""
        t = np.dtype(self.dtype).type
        val = np.asarray(self(center), self.dtype).reshape(-1)
        rad = sdf_radius(vecs, self.dtype)
        can_change = (np.abs(val) - (rad * t(self.lipschitz)).astype(self.dtype)) < 0
        code = np.full(val.shape, oaff.SIGN_UNKNOWN, np.int32)
        code = np.where(~can_change & (val > t(offset)), oaff.SIGN_POSITIVE, code)
        code = np.where(~can_change & (val < -t(offset)), oaff.SIGN_NEGATIVE, code)
        return code

    def bounds_box(self, lower, upper):
        center, vecs = oaff.box_to_general(lower, upper, self.dtype)
        return self.bounds_general(center, vecs)

    def classify_box(self, lower, upper, offset=0.0):
        if self.mode == "sdf":
            center, vecs = oaff.box_to_general(lower, upper, self.dtype)
            return self.sdf_code(center, vecs, offset)
        lo, hi = self.bounds_box(lower, upper)
        return oaff.sign_code(lo, hi, offset)
