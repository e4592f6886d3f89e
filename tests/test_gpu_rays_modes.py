"""GPU parity of the ray and search queries across range-analysis modes and
activations: cast_rays, cast_rays_frustum, closest_point and
find_any_intersection.

The ray caster bounds each step's segment with the mode's rule (interval,
affine_fixed, slope_interval) and tests the two hit points with a plain
evaluation; every mode and activation is checked against the oracle's
lock-step restatement of src/queries.py:39-175 on the same rays.  Bar (as in
test_gpu_parity.py): (t, hit_id, count) identical for every ray, frustum
pixel and closest-point query except those the oracle marks near-threshold
(one of their decisions had its FP64 value within the FP32 rounding band of
the threshold, oracle/margins.py); both counts are printed.
"""
import numpy as np
import pytest
import torch

from oracle import funcs as ofuncs, rays as orays, margins
from niq import implicit_mlp_utils, kd_tree, queries
from test_gpu_parity import note, model, camera, only_near

pytestmark = pytest.mark.gpu


def cast_both(p, mode, roots, dirs):
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, h, c, ne = queries.cast_rays((func,), (p,), torch.as_tensor(roots, device="cuda"),
                                    torch.as_tensor(dirs, device="cuda"), opts)
    rt, rh, rc, rne, near = orays.cast_rays((margins.TwinFunc(p, mode),), roots, dirs, opts, track_near=True)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == rt) & (h == rh) & (c == rc)
    return same, near, h, rh, ne, rne


@pytest.mark.parametrize("mode", ["interval", "affine_fixed", "slope_interval"])
@pytest.mark.parametrize("name", ["c1_relu_3x32", "small_elu", "small_sin", "small_tanh", "fox"])
def test_rays_match_oracle_per_mode(name, mode):
    note(f"[rays per mode] {name} {mode}")
    p = model(name)
    roots, dirs = camera(32, fov=45.0 if name == "fox" else 60.0)
    roots, dirs = np.asarray(roots, np.float32), np.asarray(dirs, np.float32)
    same, near, h, rh, ne, rne = cast_both(p, mode, roots, dirs)
    only_near(same, near, "rays", f"{int((rh > 0).sum())} oracle hits; n_evals gpu {ne} oracle {rne}")
    if same.all():
        assert ne == rne


def test_rays_match_oracle_c2_window():
    """The 3->256x6->1 net (the bench's ray workload): a 16x16 window at the
    centre of its 1024^2 camera."""
    note("[rays c2 window]")
    p = model("c2_relu_256x6")
    roots, dirs = camera(1024)
    roots, dirs = np.asarray(roots, np.float32), np.asarray(dirs, np.float32)
    rows = np.arange(504, 520)
    pix = (rows[:, None] * 1024 + rows[None, :]).reshape(-1)
    same, near, h, rh, ne, rne = cast_both(p, "affine_fixed", roots[pix], dirs[pix])
    only_near(same, near, "rays", f"{int((rh > 0).sum())} oracle hits")


@pytest.mark.parametrize("mode", ["interval", "slope_interval"])
@pytest.mark.parametrize("name", ["c1_relu_3x32", "small_sin", "fox"])
def test_frustum_matches_oracle_per_mode(name, mode):
    """Frustum casting (src/queries.py:454-587) in the other general-box modes:
    per pixel (t, hit_id, count) identical to the oracle's except at
    near-threshold pixels, N_evals identical when every pixel is."""
    from oracle import frustum as ofr
    from test_gpu_parity import frust_cam
    note(f"[frustum per mode] {name} {mode}")
    p = model(name)
    f, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    opts["n_side_init"] = 8
    cam = frust_cam(48)
    t, h, c, ne = queries.cast_rays_frustum((f,), (p,), cam, opts)
    rt, rh, rc, rne, near = ofr.cast_rays_frustum((margins.TwinFunc(p, mode),), *cam, opts, track_near=True)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == rt) & (h == rh) & (c == rc)
    only_near(same, near, "pixels", f"hits {int((h > 0).sum())} / {int((rh > 0).sum())}; N_evals {ne} / {rne}")
    if same.all():
        assert ne == rne


@pytest.mark.parametrize("mode", ["slope_interval"])
def test_closest_point_matches_oracle_per_mode(mode):
    """closest_point (src/kd_tree.py:659-802) with another classifier: the
    same distances except for near-threshold queries, which stay within a few
    eps.  (The interval classifier is so loose on this net that the oracle's
    search visits ~5e7 nodes; it is left to the GPU-only tests.)"""
    from oracle import closest as ocp
    note(f"[closest point per mode] {mode}")
    p = model("c1_relu_3x32")
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    q = np.random.default_rng(5).uniform(-1, 1, size=(128, 3)).astype(np.float32)
    # one batch for the whole stack: no query depends on another's decisions
    d, _ = kd_tree.closest_point(func, p, lo, up, torch.as_tensor(q, device="cuda"), eps=0.02,
                                 batch_process_size=1 << 20)
    rd, _, info = ocp.closest_point(margins.TwinFunc(p, mode), lo, up, q, eps=0.02, batch_process_size=1 << 20,
                                    track_near=True)
    d = d.cpu().numpy()
    same = d == rd
    only_near(same, info["near"], "distances", f"largest difference {np.abs(d - rd).max():.3g}")
    assert np.abs(d - rd).max() <= 5 * 0.02


@pytest.mark.parametrize("mode", ["interval", "slope_interval"])
@pytest.mark.parametrize("seeds", [(11, 12), (13, 14)])
def test_find_any_intersection_matches_oracle_per_mode(seeds, mode):
    """find_any_intersection (src/kd_tree.py:402-655) with both functions in
    another mode: the same found flag and point as the oracle."""
    from oracle import intersect as oisect
    from niq import models
    note(f"[intersection per mode] seeds {seeds} {mode}")
    a = models.random_mlp([32, 32], seed=seeds[0])
    b = models.random_mlp([48, 24], activation="tanh", seed=seeds[1])
    fa, _ = implicit_mlp_utils.implicit_from_params(a, mode)
    fb, _ = implicit_mlp_utils.implicit_from_params(b, mode)
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    found, _, _, loc = kd_tree.find_any_intersection((fa, fb), (a, b), lo, up, 0.02)
    rf, _, _, rloc, rn = oisect.find_any_intersection(ofuncs.OracleFunc(a, mode), ofuncs.OracleFunc(b, mode),
                                                      lo, up, 0.02)
    note(f"  gpu {found} {loc.tolist()}; oracle {rf} {rloc.tolist()} ({rn} nodes)")
    assert found == rf
    assert np.array_equal(loc.numpy(), rloc)


GROWING = [("affine_all", {}), ("affine_truncate", {"affine_n_truncate": 8, "affine_truncate_policy": "absolute"}),
           ("affine_append", {"affine_n_append": 4})]


@pytest.mark.parametrize("mode,kw", GROWING)
@pytest.mark.parametrize("name", ["c1_relu_3x32", "small_tanh"])
def test_rays_growing_row_modes_match_oracle(name, mode, kw):
    """cast_rays in the growing-row modes (the reference classifies the step's
    segment box with classify_general_box in any mode, src/queries.py:55-58):
    identical to the oracle except near-threshold or near-tie rays."""
    note(f"[rays growing rows] {name} {mode}")
    p = model(name)
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode, **kw)
    roots, dirs = camera(24)
    roots, dirs = np.asarray(roots, np.float32), np.asarray(dirs, np.float32)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, h, c, ne = queries.cast_rays((func,), (p,), torch.as_tensor(roots, device="cuda"),
                                    torch.as_tensor(dirs, device="cuda"), opts)
    rt, rh, rc, rne, near = orays.cast_rays((margins.TwinFunc(p, mode, **kw),), roots, dirs, opts, track_near=True)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == rt) & (h == rh) & (c == rc)
    only_near(same, near, "rays", f"{int((rh > 0).sum())} oracle hits; n_evals gpu {ne} oracle {rne}")
    assert (rh > 0).sum() > 0
    if same.all():
        assert ne == rne


@pytest.mark.parametrize("mode,kw", GROWING)
def test_frustum_growing_row_modes_match_oracle(mode, kw):
    """cast_rays_frustum in the growing-row modes (src/queries.py:218-225
    classifies the frustum's general box in any mode)."""
    from oracle import frustum as ofr
    from test_gpu_parity import frust_cam
    note(f"[frustum growing rows] {mode}")
    p = model("small_tanh")
    f, _ = implicit_mlp_utils.implicit_from_params(p, mode, **kw)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    opts["n_side_init"] = 4
    cam = frust_cam(24)
    t, h, c, ne = queries.cast_rays_frustum((f,), (p,), cam, opts)
    rt, rh, rc, rne, near = ofr.cast_rays_frustum((margins.TwinFunc(p, mode, **kw),), *cam, opts, track_near=True)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == rt) & (h == rh) & (c == rc)
    only_near(same, near, "pixels", f"hits {int((h > 0).sum())} / {int((rh > 0).sum())}; N_evals {ne} / {rne}")
    if same.all():
        assert ne == rne
