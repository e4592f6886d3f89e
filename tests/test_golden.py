"""The committed golden fixtures (tests/golden/c1_golden.npz, made by
tests/golden/make_golden.py) against the model generator and the oracle (CPU).

* The model file is the generator's ReLU MLP 3->32x3->1 (seed 0), bit for bit,
  in the reference's op-list format.
* Recomputed by the oracle on this machine, the fixture's outputs hold to the
  parity bars (BLAS may sum in another order on another CPU): bounds within
  1e-5 * scale + 2 |fp32 - fp64| per cell, sign codes identical away from the
  threshold, point values within 1e-5 * scale, tree lists identical, rays
  identical for >= 99.5 % of pixels, closest-point distances identical for
  >= 98 % of queries.
* Size-independent facts of the fixture itself: the tree's three lists
  partition the finest grid, every leaf has the finest width, rays that hit
  stop within max_dist.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import make_golden as mg  # noqa: E402
from oracle import funcs as ofuncs, tree as otree, rays as orays, closest as ocp, mlp as omlp  # noqa: E402
from oracle import margins  # noqa: E402
from niq import models, mlp as nmlp  # noqa: E402

NEAR_REL = 1e-6


@pytest.fixture(scope="module")
def gold():
    with np.load(os.path.join(GOLDEN, "c1_golden.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def params():
    return nmlp.load(os.path.join(GOLDEN, "c1_relu_3x32.npz"))


def test_model_file_is_the_generators(params):
    gen = models.config_model(mg.MODEL)
    assert sorted(gen) == sorted(params)
    for k in gen:
        assert np.asarray(params[k]).dtype == np.asarray(gen[k]).dtype
        np.testing.assert_array_equal(np.asarray(params[k]), np.asarray(gen[k]))


def test_inputs_are_the_seeded_ones(gold):
    box_lo, box_up, x, roots, dirs, q = mg.golden_inputs()
    for a, k in ((box_lo, "box_lower"), (box_up, "box_upper"), (x, "points_x"), (roots, "ray_roots"),
                 (dirs, "ray_dirs"), (q, "cp_query")):
        np.testing.assert_array_equal(a, gold[k])


def bounds_ok(got, b32, b64):
    noise = np.abs(b32 - b64)
    scale = np.max(np.abs(b32))
    return np.abs(got - b32) <= 1e-5 * scale + 2 * noise


def codes_ok(got, ref, b64, b32):
    noise = np.abs(b32 - b64)
    scale = np.abs(b64[:, 0]) + np.abs(b64[:, 1])
    near = (np.abs(b64[:, 0]) <= NEAR_REL * scale + 4 * noise[:, 0]) | \
           (np.abs(b64[:, 1]) <= NEAR_REL * scale + 4 * noise[:, 1])
    return (got == ref) | near


@pytest.mark.parametrize("mode", mg.MODES)
def test_oracle_bounds_match_fixture(gold, params, mode):
    f = ofuncs.OracleFunc(params, mode, np.float32)
    lo, hi = f.bounds_box(gold["box_lower"], gold["box_upper"])
    got = np.stack([lo, hi], 1)
    assert np.all(bounds_ok(got, gold[f"{mode}_bounds32"], gold[f"{mode}_bounds64"]))
    codes = f.classify_box(gold["box_lower"], gold["box_upper"])
    assert np.all(codes_ok(codes, gold[f"{mode}_codes"], gold[f"{mode}_bounds64"], gold[f"{mode}_bounds32"]))
    # the fixture's codes are the signs of its own FP32 bounds
    b = gold[f"{mode}_bounds32"]
    want = np.where(b[:, 0] > 0, 1, np.where(b[:, 1] < 0, 2, 0))
    np.testing.assert_array_equal(gold[f"{mode}_codes"], want)
    assert (gold[f"{mode}_codes"] != 0).sum() > 0


def test_oracle_points_match_fixture(gold, params):
    got = omlp.eval_points(params, gold["points_x"], np.float32)
    ref = gold["points_f32"]
    assert np.all(np.abs(got - ref) <= 1e-5 * np.max(np.abs(ref)))


def test_oracle_tree_matches_fixture(gold, params):
    tree = otree.construct_uniform_unknown_levelset_tree(ofuncs.OracleFunc(params, "affine_fixed"), mg.LO, mg.UP,
                                                         split_depth=12, with_interior_nodes=True,
                                                         with_exterior_nodes=True)
    assert tree["stats"]["n_classified"] == int(gold["tree_n_classified"])
    for key in ("unknown", "interior", "exterior"):
        v = gold[f"tree_{key}_valid"]
        np.testing.assert_array_equal(np.asarray(tree[f"{key}_node_valid"]), v)
        for part in ("lower", "upper"):
            np.testing.assert_array_equal(np.asarray(tree[f"{key}_node_{part}"])[v], gold[f"tree_{key}_{part}"][v])


def test_fixture_tree_partitions_the_grid(gold):
    """Unknown leaves are finest cells; finest unknown cells plus the
    determinate nodes (of any level) cover [-1, 1]^3 exactly once."""
    lo = gold["tree_unknown_lower"][gold["tree_unknown_valid"]]
    up = gold["tree_unknown_upper"][gold["tree_unknown_valid"]]
    np.testing.assert_array_equal(up - lo, np.full_like(lo, 2.0 / 16))
    vol = float(np.prod(up - lo, 1).sum())
    for key in ("interior", "exterior"):
        v = gold[f"tree_{key}_valid"]
        vol += float(np.prod(gold[f"tree_{key}_upper"][v] - gold[f"tree_{key}_lower"][v], 1).sum())
    assert vol == pytest.approx(8.0, abs=1e-9)


def test_oracle_rays_match_fixture(gold, params):
    """The oracle reproduces the fixture's rays; a ray may differ (another
    BLAS's FP32 sums) only where the fixture marks a near-threshold decision,
    and the recomputed near marks agree wherever the rays agree."""
    t, h, c, ne, near = orays.cast_rays((margins.TwinFunc(params, "affine_fixed"),), gold["ray_roots"],
                                        gold["ray_dirs"], mg.ray_opts(), track_near=True)
    same = (t == gold["ray_t"]) & (h == gold["ray_hit_id"]) & (c == gold["ray_count"])
    assert not np.any(~same & ~gold["ray_near"])
    np.testing.assert_array_equal(near[same], gold["ray_near"][same])
    if same.all():
        assert ne == int(gold["ray_n_evals"])
    hit = gold["ray_hit_id"] > 0
    assert 100 < hit.sum() < hit.size
    assert np.all(gold["ray_t"][hit] <= 6.0) and np.all(gold["ray_count"] >= 1)
    assert gold["ray_near"].sum() < 0.05 * hit.sum()   # the band marks a few hits, not all


def test_oracle_closest_point_matches_fixture(gold, params):
    d, loc, info = ocp.closest_point(margins.TwinFunc(params, "affine_fixed"), mg.LO, mg.UP, gold["cp_query"],
                                     eps=0.01, batch_process_size=mg.CP_WIDE_BATCH, track_near=True)
    same = d == gold["cpw_dist"]
    assert not np.any(~same & ~gold["cpw_near"])
    np.testing.assert_array_equal(loc[same], gold["cpw_loc"][same])
    d, loc, info = ocp.closest_point(margins.TwinFunc(params, "affine_fixed"), mg.LO, mg.UP, gold["cp_query"],
                                     eps=0.01, batch_process_size=256, track_near=True)
    same = d == gold["cp_dist"]
    coupled = gold["cp_near"] | (gold["cp_last_round"] >= gold["cp_near_round"])
    assert not np.any(~same & ~coupled)
    np.testing.assert_array_equal(loc[same], gold["cp_loc"][same])
