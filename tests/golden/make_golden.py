"""Writes the committed golden fixtures of the ReLU MLP 3->32x3->1 (seed 0).

  c1_relu_3x32.npz  the model, in the reference's op-list format
                    ("%04d.<op>.<arg>", dense.A shaped (in, out))
  c1_golden.npz     the oracle's outputs on seeded inputs:
    box_lower/box_upper (1024, 3) and, per mode (interval, affine_fixed,
      slope_interval, affine_all): <mode>_bounds32 / <mode>_bounds64 (1024, 2)
      and <mode>_codes (1024,)
    points_x (1024, 3), points_f32 (1024,)
    tree_*: the affine_fixed level-set tree of [-1, 1]^3 at split depth 12
      with interior and exterior lists (the reference's bucketed arrays)
    ray_*: a 64x64 camera at (0, 0, 3) looking at the origin, 60 degree field
      of view, max_dist 6, other options at their defaults: roots, dirs,
      t, hit_id, count, n_evals, and near (the rays one of whose step or hit
      decisions lay within the FP32 rounding band of its threshold,
      oracle/margins.py)
    cp_*: closest_point of 64 seeded queries, eps 0.01, batches of 256 (dist,
      loc, near per query, and the round each query was last processed in
      and the first round with a near decision: queries share the stack);
      cpw_*: the same with one batch for the whole stack (no coupling)

The oracle follows the reference's algorithm (file:line in each oracle
module); these fixtures freeze its outputs so that later changes to the
oracle, the model generator or the kernels are caught against fixed data.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "neural-implicit-queries_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import funcs as ofuncs, tree as otree, rays as orays, closest as ocp, mlp as omlp  # noqa: E402
from oracle import margins  # noqa: E402
from niq import models, mlp as nmlp  # noqa: E402

MODEL = "c1_relu_3x32"
MODES = ("interval", "affine_fixed", "slope_interval", "affine_all")
LO = np.array([-1., -1., -1.], np.float32)
UP = np.array([1., 1., 1.], np.float32)
CP_WIDE_BATCH = 1 << 20   # above the largest closest-point stack of these queries


def golden_inputs():
    """The seeded inputs (shared with the tests)."""
    rng = np.random.default_rng(11)
    c = rng.uniform(-1, 1, size=(1024, 3))
    h = np.exp(rng.uniform(np.log(5e-4), np.log(0.25), size=(1024, 3)))   # log-uniform half-widths
    box_lo, box_up = (c - h).astype(np.float32), (c + h).astype(np.float32)
    x = np.random.default_rng(12).uniform(-1, 1, size=(1024, 3)).astype(np.float32)
    eye = np.array([0., 0., 3.], np.float32)
    look, up, _ = orays.look_at(eye)
    roots, dirs = orays.generate_camera_rays(eye, look, up, res=64, fov_deg=60.0)
    q = np.random.default_rng(13).uniform(-1, 1, size=(64, 3)).astype(np.float32)
    return box_lo, box_up, x, roots, dirs, q


def ray_opts():
    opts = orays.get_default_cast_opts()
    opts["max_dist"] = 6.0
    return opts


def main():
    p = models.config_model(MODEL)
    nmlp.save(os.path.join(HERE, f"{MODEL}.npz"), p)
    box_lo, box_up, x, roots, dirs, q = golden_inputs()
    out = {"box_lower": box_lo, "box_upper": box_up}
    for mode in MODES:
        f32 = ofuncs.OracleFunc(p, mode, np.float32)
        lo32, hi32 = f32.bounds_box(box_lo, box_up)
        lo64, hi64 = ofuncs.OracleFunc(p, mode, np.float64).bounds_box(box_lo, box_up)
        out[f"{mode}_bounds32"] = np.stack([lo32, hi32], 1).astype(np.float32)
        out[f"{mode}_bounds64"] = np.stack([lo64, hi64], 1).astype(np.float64)
        out[f"{mode}_codes"] = f32.classify_box(box_lo, box_up).astype(np.int32)
    out["points_x"] = x
    out["points_f32"] = omlp.eval_points(p, x, np.float32).astype(np.float32)
    tree = otree.construct_uniform_unknown_levelset_tree(ofuncs.OracleFunc(p, "affine_fixed"), LO, UP,
                                                         split_depth=12, with_interior_nodes=True,
                                                         with_exterior_nodes=True)
    for key in ("unknown", "interior", "exterior"):
        for part in ("valid", "lower", "upper"):
            out[f"tree_{key}_{part}"] = np.asarray(tree[f"{key}_node_{part}"])
    out["tree_n_classified"] = np.int64(tree["stats"]["n_classified"])
    t, h, c, ne, near = orays.cast_rays((margins.TwinFunc(p, "affine_fixed"),), roots, dirs, ray_opts(),
                                        track_near=True)
    out.update(ray_roots=roots, ray_dirs=dirs, ray_t=np.asarray(t, np.float32), ray_hit_id=np.asarray(h, np.int32),
               ray_count=np.asarray(c, np.int32), ray_n_evals=np.int64(ne), ray_near=near)
    d, loc, info = ocp.closest_point(margins.TwinFunc(p, "affine_fixed"), LO, UP, q, eps=0.01,
                                     batch_process_size=256, track_near=True)
    out.update(cp_query=q, cp_dist=np.asarray(d, np.float32), cp_loc=np.asarray(loc, np.float32),
               cp_near=info["near"], cp_last_round=info["last_round"],
               cp_near_round=np.int64(-1 if info["near_round"] is None else info["near_round"]))
    d, loc, info = ocp.closest_point(margins.TwinFunc(p, "affine_fixed"), LO, UP, q, eps=0.01,
                                     batch_process_size=CP_WIDE_BATCH, track_near=True)
    out.update(cpw_dist=np.asarray(d, np.float32), cpw_loc=np.asarray(loc, np.float32), cpw_near=info["near"])
    np.savez_compressed(os.path.join(HERE, "c1_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "c1_golden.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
