"""GPU parity of cast_rays across range-analysis modes and activations.

The ray caster bounds each step's segment with the mode's rule (interval,
affine_fixed, slope_interval) and tests the two hit points with a plain
evaluation; every mode and activation is checked against the oracle's
lock-step restatement of src/queries.py:39-175 on the same rays.  Bar (as in
test_gpu_parity.py): (t, hit_id, count) identical for >= 99.5 % of the rays
and hit_id identical for >= 99.8 % (a near-threshold step decision moves a
ray by one step); the C2 window (256 rays) allows one differing ray.
"""
import numpy as np
import pytest
import torch

from oracle import funcs as ofuncs, rays as orays
from niq import implicit_mlp_utils, queries
from test_gpu_parity import note, model, camera

pytestmark = pytest.mark.gpu


def cast_both(p, mode, roots, dirs):
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, h, c, ne = queries.cast_rays((func,), (p,), torch.as_tensor(roots, device="cuda"),
                                    torch.as_tensor(dirs, device="cuda"), opts)
    rt, rh, rc, rne = orays.cast_rays((ofuncs.OracleFunc(p, mode),), roots, dirs, opts)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == rt) & (h == rh) & (c == rc)
    return same, h, rh, ne, rne


@pytest.mark.parametrize("mode", ["interval", "affine_fixed", "slope_interval"])
@pytest.mark.parametrize("name", ["c1_relu_3x32", "small_elu", "small_sin", "small_tanh", "fox"])
def test_rays_match_oracle_per_mode(name, mode):
    note(f"[rays per mode] {name} {mode}")
    p = model(name)
    roots, dirs = camera(32, fov=45.0 if name == "fox" else 60.0)
    roots, dirs = np.asarray(roots, np.float32), np.asarray(dirs, np.float32)
    same, h, rh, ne, rne = cast_both(p, mode, roots, dirs)
    note(f"  {same.mean() * 100:.2f}% rays identical, {int((rh > 0).sum())} oracle hits; "
         f"n_evals gpu {ne} oracle {rne}")
    assert same.mean() >= 0.995
    assert (h == rh).mean() >= 0.998
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
    same, h, rh, ne, rne = cast_both(p, "affine_fixed", roots[pix], dirs[pix])
    note(f"  {int(same.sum())} of {same.size} rays identical, {int((rh > 0).sum())} oracle hits")
    assert same.size - same.sum() <= 1
