"""Slope-interval implicit functions (src/slope_interval.py:16-51)."""
from . import _lib
from .implicit_function import ImplicitFunction


class SlopeIntervalImplicitFunction(ImplicitFunction):
    """src/slope_interval.py:16-51; ``slope_interval_func`` is accepted for
    signature compatibility."""

    def __init__(self, slope_interval_func=None, recheck_tol=0.0):
        super().__init__("classify-and-distance")
        self.slope_interval_func = slope_interval_func
        self.cfg = _lib.eval_cfg("slope_interval", recheck_tol=recheck_tol)
