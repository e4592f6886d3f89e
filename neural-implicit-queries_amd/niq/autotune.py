"""Kernel-variant autotuner for the classify pass.

The fused classify kernel comes in variants chosen by ``EvalCfg.variant``:

* plain FP32 (interval and affine_fixed over 3 box vectors): 1 is the
  default tile (``TileCfg``: 16-32 boxes per workgroup), 2 the alternative
  tile (``TileCfgAlt``: twice the boxes per workgroup where LDS allows, fewer
  waves at width 512);
* precision mode (BF16): 1 streams the packed bf16 weight copies, 2 loads the
  FP32 weights once and splits them in registers, 3 runs layer by layer
  (one batched launch per hidden layer over a chunk of boxes, activations
  through HBM).

Variants compute the same bounds: the plain FP32 tiles sum every (box,
feature) in the same order (bit-identical), the precision-mode weight sources
give the same splits (the e-plane bound differs by at most 2^-15 relative)
and variant 3 repeats variant 2's arithmetic exactly.  ``tune`` times each
variant on boxes of the size the finest tree cells of ``split_depth`` have,
keeps the fastest in ``func.cfg.variant`` and caches the choice per (layer
widths, activations, mode, precision).
"""
import numpy as np
import torch

from . import mlp as _mlp

VARIANTS = (1, 2)
VARIANTS_BF16 = (1, 2, 3)
_cache = {}


def variants_of(func):
    return VARIANTS_BF16 if int(func.cfg.precision) == 1 else VARIANTS


def _key(func, params):
    m = _mlp.device_model(params)
    acts = tuple(L.act for L in m.layers)
    return (tuple(m.dims), acts, int(func.cfg.mode), int(func.cfg.precision))


def tune(func, params, split_depth=18, n_boxes=1 << 17, reps=3, seed=0):
    """Time every variant of ``func``'s classify pass on ``n_boxes`` random
    cells of half-width 2^-(split_depth/3) in [-1, 1]^3; set and return the
    fastest as {variant: ms}."""
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
        func.classify_box(params, lo, up)          # warm-up (and first-launch set-up)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            func.classify_box(params, lo, up)
        e1.record()
        torch.cuda.synchronize()
        times[v] = e0.elapsed_time(e1) / reps
    func.cfg.variant = saved
    best = min(times, key=times.get)
    func.cfg.variant = best
    res = {"best": best, "ms": times, "n_boxes": n_boxes}
    _cache[key] = res
    return res
