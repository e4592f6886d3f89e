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
         "affine_all", "slope_interval")


class OracleFunc:
    def __init__(self, params, mode="affine_fixed", dtype=np.float32, **kwargs):
        if mode not in MODES:
            raise RuntimeError("unrecognized mode")
        self.params = params
        self.mode = mode
        self.dtype = dtype
        if mode != "slope_interval":
            self.ctx = oaff.AffineContext(
                mode,
                truncate_count=kwargs.get("affine_n_truncate", -777),
                truncate_policy=kwargs.get("affine_truncate_policy", "absolute"),
                n_append=kwargs.get("affine_n_append", 0))

    def __call__(self, x):
        return omlp.eval_points(self.params, np.atleast_2d(x), self.dtype)

    def bounds_general(self, center, vecs):
        """Returns (lower, upper) of the output over each general box."""
        if self.mode == "slope_interval":
            return oslope.bound_general_boxes(self.params, center, vecs, self.dtype)
        base, rad = oaff.bound_general_boxes(self.ctx, self.params, center, vecs, self.dtype)
        return base - rad, base + rad

    def classify_general_box(self, center, vecs, offset=0.0):
        lo, hi = self.bounds_general(center, vecs)
        return oaff.sign_code(lo, hi, offset)

    def bounds_box(self, lower, upper):
        center, vecs = oaff.box_to_general(lower, upper, self.dtype)
        return self.bounds_general(center, vecs)

    def classify_box(self, lower, upper, offset=0.0):
        lo, hi = self.bounds_box(lower, upper)
        return oaff.sign_code(lo, hi, offset)
