"""The gfx950 kernels (through the C ABI) against the committed golden
fixtures of the ReLU MLP 3->32x3->1 (tests/golden/c1_golden.npz), to the
parity bars of test_gpu_parity.py: bounds within 1e-5 * scale + 2 |fp32 -
fp64| per cell, sign codes identical away from the threshold, point values
within 1e-5 * scale, tree lists identical (order included), rays identical
except those the fixture marks near-threshold (with the reference's N_evals
when all are), closest-point distances identical except near-threshold
queries, and within 5 eps for all.
"""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from niq import implicit_mlp_utils, kd_tree, queries, mlp as nmlp
from test_golden import bounds_ok, codes_ok
from test_gpu_parity import note, only_near

pytestmark = pytest.mark.gpu

MODES = ("interval", "affine_fixed", "slope_interval", "affine_all")
LO = np.array([-1., -1., -1.], np.float32)
UP = np.array([1., 1., 1.], np.float32)


@pytest.fixture(scope="module")
def gold():
    with np.load(os.path.join(GOLDEN, "c1_golden.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def params():
    return nmlp.load(os.path.join(GOLDEN, "c1_relu_3x32.npz"))


def dev(x):
    return torch.as_tensor(x, device="cuda")


@pytest.mark.parametrize("mode", MODES)
def test_classify_matches_golden(gold, params, mode):
    note(f"[golden classify] {mode}")
    func, _ = implicit_mlp_utils.implicit_from_params(params, mode)
    codes, bounds = func.classify_box(params, dev(gold["box_lower"]), dev(gold["box_upper"]), return_bounds=True)
    codes, bounds = codes.cpu().numpy(), bounds.cpu().numpy()
    assert np.all(bounds_ok(bounds, gold[f"{mode}_bounds32"], gold[f"{mode}_bounds64"]))
    ok = codes_ok(codes, gold[f"{mode}_codes"], gold[f"{mode}_bounds64"], gold[f"{mode}_bounds32"])
    note(f"  {(codes != gold[f'{mode}_codes']).sum()} code differences, all near the threshold")
    assert np.all(ok)


def test_points_match_golden(gold, params):
    func, _ = implicit_mlp_utils.implicit_from_params(params, "affine_fixed")
    got = func(params, dev(gold["points_x"])).cpu().numpy()
    ref = gold["points_f32"]
    assert np.all(np.abs(got - ref) <= 1e-5 * np.max(np.abs(ref)))


def test_tree_matches_golden(gold, params):
    note("[golden tree] depth 12")
    func, _ = implicit_mlp_utils.implicit_from_params(params, "affine_fixed")
    out, st = kd_tree.build_tree(func, params, LO, UP, split_depth=12, with_interior_nodes=True,
                                 with_exterior_nodes=True)
    assert st["n_classified"] == int(gold["tree_n_classified"])
    for key in ("unknown", "interior", "exterior"):
        v = gold[f"tree_{key}_valid"]
        np.testing.assert_array_equal(out[f"{key}_node_valid"].cpu().numpy(), v)
        for part in ("lower", "upper"):
            np.testing.assert_array_equal(out[f"{key}_node_{part}"].cpu().numpy()[v], gold[f"tree_{key}_{part}"][v])


def test_rays_match_golden(gold, params):
    note("[golden rays] 64x64")
    func, _ = implicit_mlp_utils.implicit_from_params(params, "affine_fixed")
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, h, c, ne = queries.cast_rays((func,), (params,), dev(gold["ray_roots"]), dev(gold["ray_dirs"]), opts)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == gold["ray_t"]) & (h == gold["ray_hit_id"]) & (c == gold["ray_count"])
    only_near(same, gold["ray_near"], "rays", f"n_evals {ne} (fixture {int(gold['ray_n_evals'])})")
    if same.all():
        assert ne == int(gold["ray_n_evals"])


def test_closest_point_matches_golden(gold, params):
    """One batch for the whole stack (no query depends on another): distances
    identical except near-threshold queries.  Batches of 256 (the queries
    share the stack, so a near decision of one query reorders the rounds of
    the others): identical except near-threshold queries and queries processed
    after the first near decision; every distance within a few eps."""
    func, _ = implicit_mlp_utils.implicit_from_params(params, "affine_fixed")
    d, loc = kd_tree.closest_point(func, params, LO, UP, dev(gold["cp_query"]), eps=0.01, batch_process_size=1 << 20)
    d, loc = d.cpu().numpy(), loc.cpu().numpy()
    same = d == gold["cpw_dist"]
    only_near(same, gold["cpw_near"], "distances (one batch)")
    np.testing.assert_array_equal(loc[same], gold["cpw_loc"][same])
    d, loc = kd_tree.closest_point(func, params, LO, UP, dev(gold["cp_query"]), eps=0.01, batch_process_size=256)
    d, loc = d.cpu().numpy(), loc.cpu().numpy()
    same = d == gold["cp_dist"]
    diff = np.abs(d - gold["cp_dist"])
    coupled = gold["cp_near"] | (gold["cp_last_round"] >= gold["cp_near_round"])
    only_near(same, coupled, "distances (batches of 256; near or processed after a near decision)",
              f"largest difference {diff.max():.3g}")
    assert diff.max() <= 5 * 0.01
    np.testing.assert_array_equal(loc[same], gold["cp_loc"][same])
