"""niq: the range-analysis hot path of neural-implicit-queries on MI355X.

Drop-in modules mirroring the reference's ``src/`` layout:
``mlp``, ``implicit_function``, ``affine``, ``slope_interval``,
``implicit_mlp_utils``, ``kd_tree``, ``queries``, ``render``, ``bucketing``.
Every query runs on hand-written gfx950 HIP kernels through the C ABI in
``libniq.so`` (include/niq.h); there is no CPU fallback.
"""
from . import _lib  # noqa: F401
