"""Affine-arithmetic implicit functions (src/affine.py:17-76).

``AffineContext`` keeps the reference's fields and validation; an
``AffineImplicitFunction`` routes every query to the gfx950 kernels with the
context's mode.
"""
import dataclasses
from dataclasses import dataclass

from . import _lib
from .implicit_function import ImplicitFunction


@dataclass(frozen=True)
class AffineContext():
    """src/affine.py:62-76."""
    mode: str = 'affine_fixed'
    truncate_count: int = -777
    truncate_policy: str = 'absolute'
    affine_domain_terms: int = 0
    n_append: int = 0

    def __post_init__(self):
        if self.mode not in ['interval', 'affine_fixed', 'affine_truncate', 'affine_append', 'affine_all']:
            raise ValueError("invalid mode")
        if self.mode == 'affine_truncate':
            if self.truncate_count is None:
                raise ValueError("must specify truncate count")


class AffineImplicitFunction(ImplicitFunction):
    """src/affine.py:17-55.  ``affine_func`` is accepted for signature
    compatibility; the op list comes from ``params`` at call time."""

    def __init__(self, affine_func, ctx, precision="fp32", recheck_tol=0.0):
        super().__init__("classify-only")
        self.affine_func = affine_func
        self.ctx = ctx
        self.mode_dict = {'ctx': self.ctx}
        self.precision = precision
        self.cfg = _lib.eval_cfg(ctx.mode, precision, truncate_count=max(ctx.truncate_count, 0),
                                 n_append=ctx.n_append, recheck_tol=recheck_tol)

    def with_precision(self, precision, recheck_tol=None):
        tol = self.cfg.recheck_tol if recheck_tol is None else recheck_tol
        return AffineImplicitFunction(self.affine_func, self.ctx, precision, tol)
