"""Tile autotuner for the classify pass and the ray caster.

The search space is the kernels' tile shapes, each a compiled variant:

* plain FP32 classify (interval and affine_fixed over 3 box vectors),
  ``EvalCfg.variant``: 1 the default tile (``TileCfg``: 16-32 boxes per
  workgroup, 4-8 waves, the register budget LDS occupancy allows), 2 twice
  the boxes per workgroup where LDS allows (fewer waves at width 512), 3 the
  default tile with a 256-register budget (no spills, fewer resident
  workgroups), 4 16 boxes with the most waves (the narrowest width tile: the
  waves of a tile split the layer's 16-feature tiles among them);
* precision mode (BF16) classify: 1 streams the packed bf16 weight copies,
  2 loads the FP32 weights once and splits them in registers, 3 runs layer
  by layer (one batched launch per hidden layer, activations through HBM);
* the growing-row classify (affine_all, affine_truncate, affine_append; one
  box per workgroup, its error terms streamed in chunks): 1 the default, 2
  half the error-term rows per chunk, 3 a two-waves-per-SIMD register budget
  at widths up to 128;
* the ray caster, ``EvalCfg.ray_variant``: 0 the default tile, 1 the same
  tile with a 256-register budget, 2 16-slot tiles at widths up to 128 and
  32-slot tiles at width 256 (256 registers).

``niq_tile_shape`` reports each variant's (boxes or slots per workgroup, or
error-term rows per chunk; waves; waves per SIMD of the register budget).
Variants compute the same results: the plain FP32 tiles, the growing-row
chunks and the ray tiles sum every (box, feature) in the same order
(bit-identical); the precision-mode weight sources give the
same splits (the e-plane bound differs by at most 2^-15 relative) and
variant 3 repeats variant 2's arithmetic exactly.  ``tune`` times the
classify variants on boxes of the size of the finest tree cells of
``split_depth``; ``tune_rays`` times the ray variants on a camera's rays.
Both keep the fastest in the function's config and cache the choice per
(layer widths, activations, mode, precision).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import mlp as _mlp

VARIANTS = (1, 2, 3, 4)
VARIANTS_ROWS = (1, 2, 3)
VARIANTS_BF16 = (1, 2, 3)
RAY_VARIANTS = (0, 1, 2)
_cache = {}
_ray_cache = {}
ROWS_MODES = (_lib.MODE["affine_truncate"], _lib.MODE["affine_append"], _lib.MODE["affine_all"])


def variants_of(func):
    if int(func.cfg.precision) == 1:
        return VARIANTS_BF16
    return VARIANTS_ROWS if int(func.cfg.mode) in ROWS_MODES else VARIANTS


def _key(func, params):
    m = _mlp.device_model(params)
    acts = tuple(L.act for L in m.layers)
    return (tuple(m.dims), acts, int(func.cfg.mode), int(func.cfg.precision))


def tile_shape(kernel, variant, width):
    """{"boxes_per_workgroup", "waves", "waves_per_simd"} of a variant:
    kernel "classify" (plain FP32), "rays" or "rows" (growing-row classify:
    {"error_rows_per_chunk", "waves", "waves_per_simd"}, one box per
    workgroup)."""
    lib = _lib.load_library()
    nb, nw, occ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.check(lib.niq_tile_shape({"classify": 0, "rays": 1, "rows": 2}[kernel], int(variant), int(width),
                                  ctypes.byref(nb), ctypes.byref(nw), ctypes.byref(occ)))
    first = "error_rows_per_chunk" if kernel == "rows" else "boxes_per_workgroup"
    return {first: nb.value, "waves": nw.value, "waves_per_simd": occ.value}


def _time(fn, reps):
    fn()                                  # warm-up (and first-launch set-up)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(func, params, split_depth=18, n_boxes=1 << 17, reps=3, seed=0):
    """Time every classify variant of ``func`` on ``n_boxes`` random cells of
    half-width 2^-(split_depth/3) in [-1, 1]^3; set the fastest and return
    {"best", "ms": {variant: ms}, "shape"}."""
    key = _key(func, params)
    if key in _cache:
        func.cfg.variant = _cache[key]["best"]
        return _cache[key]
    rng = np.random.default_rng(seed)
    h = 2.0 ** (-(split_depth // 3))
    c = rng.uniform(-1 + h, 1 - h, size=(n_boxes, 3))
    lo = torch.as_tensor((c - h).astype(np.float32), device="cuda")
    up = torch.as_tensor((c + h).astype(np.float32), device="cuda")
    times = {}
    saved = func.cfg.variant
    for v in variants_of(func):
        func.cfg.variant = v
        times[v] = _time(lambda: func.classify_box(params, lo, up), reps)
    func.cfg.variant = saved
    best = min(times, key=times.get)
    func.cfg.variant = best
    res = {"best": best, "ms": times, "n_boxes": n_boxes}
    if int(func.cfg.precision) == 0:
        kernel = "rows" if int(func.cfg.mode) in ROWS_MODES else "classify"
        res["shape"] = tile_shape(kernel, best, max(_mlp.device_model(params).dims))
    _cache[key] = res
    return res


def tune_rays(func, params, roots, dirs, opts, reps=2):
    """Time every ray-caster variant of ``func`` on the given rays; set the
    fastest in ``func.cfg.ray_variant`` and return {"best", "ms", "shape"}."""
    from . import queries
    key = _key(func, params)
    if key in _ray_cache:
        func.cfg.ray_variant = _ray_cache[key]["best"]
        return _ray_cache[key]
    times = {}
    saved = func.cfg.ray_variant
    for v in RAY_VARIANTS:
        func.cfg.ray_variant = v
        times[v] = _time(lambda: queries.cast_rays((func,), (params,), roots, dirs, opts), reps)
    func.cfg.ray_variant = saved
    best = min(times, key=times.get)
    func.cfg.ray_variant = best
    res = {"best": best, "ms": times, "n_rays": int(roots.shape[0]),
           "shape": tile_shape("rays", best, max(_mlp.device_model(params).dims))}
    _ray_cache[key] = res
    return res
