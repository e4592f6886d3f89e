"""Weak-SDF implicit functions (src/sdf.py:17-50)."""
from . import _lib
from .implicit_function import ImplicitFunction


class WeakSDFImplicitFunction(ImplicitFunction):
    """src/sdf.py:17-50: classifies a box by the value at its centre against
    ``lipschitz_bound`` times its radius.  ``sdf_func`` is accepted for
    signature compatibility; the op list comes from ``params`` at call time."""

    def __init__(self, sdf_func=None, lipschitz_bound=1.):
        super().__init__("classify-only")
        self.sdf_func = sdf_func
        self.lipschitz_bound = lipschitz_bound
        self.cfg = _lib.eval_cfg("sdf", sdf_lipschitz=float(lipschitz_bound))
