"""GPU edge cases: empty and ragged batches through the C ABI.

* Empty batches return empty outputs from every batched entry point (no
  launch, no error).
* Ragged batches (1, a tile +- 1, sizes that are no multiple of any tile)
  give bit for bit the results the same items get inside a large batch:
  a box's, point's or ray's arithmetic does not depend on its neighbours or
  on the tile shape the launch size selects.  The large batches themselves
  are checked against the oracle in test_gpu_parity.py; here single items
  are checked against the oracle too.
* Trees of split depth 0, 1 and 2 match the oracle's lists exactly.
"""
import numpy as np
import pytest
import torch

from oracle import funcs as ofuncs, tree as otree, mlp as omlp, rays as orays
from niq import implicit_mlp_utils, kd_tree, queries, render
from test_gpu_parity import note, model, boxes, camera, rel_close, LO, UP

pytestmark = pytest.mark.gpu

RAGGED = (1, 2, 15, 17, 63, 65, 1000)


def dev(x):
    return torch.as_tensor(x, device="cuda")


@pytest.mark.parametrize("mode", ["interval", "affine_fixed", "slope_interval", "affine_all", "sdf"])
def test_empty_batches(mode):
    note(f"[empty batches] {mode}")
    p = model("c1_relu_3x32")
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    e = torch.zeros((0, 3), device="cuda")
    assert func(p, e).shape == (0,)
    codes, bounds = func.classify_box(p, e, e, return_bounds=True)
    assert codes.shape == (0,) and bounds.shape == (0, 2)
    assert func.classify_general_box(p, e, torch.zeros((0, 3, 3), device="cuda")).shape == (0,)
    if mode in ("interval", "affine_fixed", "slope_interval", "sdf"):
        t, h, c, ne = queries.cast_rays((func,), (p,), e, e, queries.get_default_cast_opts())
        assert t.shape == h.shape == c.shape == (0,) and ne == 0
    if mode == "affine_fixed":
        d, loc = kd_tree.closest_point(func, p, LO, UP, e)
        assert d.shape == (0,) and loc.shape == (0, 3)
        ids = torch.zeros(0, dtype=torch.int32, device="cuda")
        for method in ("finite_differences", "autodiff"):
            assert render.outward_normals((func,), (p,), e, ids, 1e-3, method=method).shape == (0, 3)
        img = render.shade_image("normal", None, None, e, ids, np.array([0., 1., 0.]), None, ((1., 1., 1.),))
        assert img.shape[0] == 0
        pts, found, rounds = kd_tree.sample_nodes(p, dev(LO[None]), dev(UP[None]), 1, 0, 0.1, 5, 1024)
        assert pts.shape == (0, 3) and found == 0 and rounds == 0


@pytest.mark.parametrize("name,mode", [("c2_relu_256x6", "affine_fixed"), ("c4_relu_512x8", "affine_fixed"),
                                       ("fox", "slope_interval"), ("fox", "interval"),
                                       ("small_tanh", "affine_all"), ("fox", "sdf")])
def test_ragged_classify_matches_large_batch(name, mode):
    note(f"[ragged classify] {name} {mode}")
    p = model(name)
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    lo, up = boxes(4096, seed=7, scale=0.3)
    big_c, big_b = func.classify_box(p, dev(lo), dev(up), return_bounds=True)
    big_c, big_b = big_c.cpu().numpy(), big_b.cpu().numpy()
    start = 0
    for n in RAGGED:
        c, b = func.classify_box(p, dev(lo[start:start + n]), dev(up[start:start + n]), return_bounds=True)
        np.testing.assert_array_equal(c.cpu().numpy(), big_c[start:start + n])
        np.testing.assert_array_equal(b.cpu().numpy(), big_b[start:start + n])
        start += n
    # one box on its own against the oracle (bounds to the parity bar)
    r_lo, r_hi = ofuncs.OracleFunc(p, mode, np.float32).bounds_box(lo[:1], up[:1])
    d_lo, d_hi = ofuncs.OracleFunc(p, mode, np.float64).bounds_box(lo[:1], up[:1])
    c1, b1 = func.classify_box(p, dev(lo[:1]), dev(up[:1]), return_bounds=True)
    b1 = b1.cpu().numpy()[0]
    tol = 1e-5 * max(abs(r_lo[0]), abs(r_hi[0]))
    assert abs(b1[0] - r_lo[0]) <= tol + 2 * abs(r_lo[0] - d_lo[0])
    assert abs(b1[1] - r_hi[0]) <= tol + 2 * abs(r_hi[0] - d_hi[0])


@pytest.mark.parametrize("name", ["c2_relu_256x6", "fox"])
def test_ragged_points_match_large_batch(name):
    note(f"[ragged points] {name}")
    p = model(name)
    func, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    x = np.random.default_rng(8).uniform(-1, 1, size=(4096, 3)).astype(np.float32)
    big = func(p, dev(x)).cpu().numpy()
    start = 0
    for n in RAGGED:
        np.testing.assert_array_equal(func(p, dev(x[start:start + n])).cpu().numpy(), big[start:start + n])
        start += n
    ref = omlp.eval_points(p, x[:64], np.float32)
    assert np.all(rel_close(big[:64], ref))


@pytest.mark.parametrize("mode", ["affine_fixed", "sdf"])
def test_ragged_rays_match_large_batch(mode):
    note(f"[ragged rays] fox {mode}")
    p = model("fox")
    func, _ = implicit_mlp_utils.implicit_from_params(p, mode)
    roots, dirs = camera(48, fov=45.0)
    roots, dirs = np.asarray(roots, np.float32), np.asarray(dirs, np.float32)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, h, c, _ = queries.cast_rays((func,), (p,), dev(roots), dev(dirs), opts)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    assert (h > 0).sum() > 50
    # ragged windows starting at the image centre, where rays hit
    start = 24 * 48
    for n in (1, 2, 15, 17, 65):
        tt, hh, cc, _ = queries.cast_rays((func,), (p,), dev(roots[start:start + n]), dev(dirs[start:start + n]), opts)
        np.testing.assert_array_equal(tt.cpu().numpy(), t[start:start + n])
        np.testing.assert_array_equal(hh.cpu().numpy(), h[start:start + n])
        np.testing.assert_array_equal(cc.cpu().numpy(), c[start:start + n])
        start += n
    # a 17-ray window on its own against the oracle (near-threshold rays may
    # differ, as in the full-image parity tests)
    of = ofuncs.OracleFunc(p, mode, sdf_lipschitz=1.0)
    i = 24 * 48 + 16
    rt, rh, rc, _ = orays.cast_rays((of,), roots[i:i + 17], dirs[i:i + 17], opts)
    same = (rt == t[i:i + 17]) & (rh == h[i:i + 17]) & (rc == c[i:i + 17])
    note(f"  {int(same.sum())} of 17 rays identical to the oracle")
    assert same.sum() >= 16 and (rh > 0).any()


@pytest.mark.parametrize("depth", [0, 1, 2])
def test_shallow_trees_match_oracle(depth):
    note(f"[shallow tree] depth {depth}")
    p = model("fox")
    func, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    out, st = kd_tree.build_tree(func, p, LO, UP, split_depth=depth, with_interior_nodes=True,
                                 with_exterior_nodes=True)
    ref = otree.construct_uniform_unknown_levelset_tree(ofuncs.OracleFunc(p, "affine_fixed"), LO, UP,
                                                        split_depth=depth, with_interior_nodes=True,
                                                        with_exterior_nodes=True)
    assert st["n_classified"] == ref["stats"]["n_classified"]
    for key in ("unknown", "interior", "exterior"):
        v = out[f"{key}_node_valid"].cpu().numpy()
        np.testing.assert_array_equal(v, ref[f"{key}_node_valid"])
        np.testing.assert_array_equal(out[f"{key}_node_lower"].cpu().numpy()[v], ref[f"{key}_node_lower"][v])
        np.testing.assert_array_equal(out[f"{key}_node_upper"].cpu().numpy()[v], ref[f"{key}_node_upper"][v])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_large_cast_writes_every_ray(precision):
    """A cast of far more rays than the persistent grid holds slots (2^20 rays
    of a 1024^2 camera; the grid holds a few tens of thousands): every output
    is written (buffers prefilled with a sentinel) and equals the same rays
    cast in 16 chunks, bit for bit, so no ray is left unclaimed or claimed
    twice whatever the order in which slots free up."""
    import ctypes
    from niq import _lib, mlp as nmlp
    note(f"[large cast] 2^20 rays, {precision}")
    p = model("c1_relu_3x32")
    func, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed", precision=precision)
    roots, dirs = render.generate_camera_rays(np.array([0., 0., 3.], np.float32), *render.look_at((0., 0., 3.))[:2],
                                              res=1024, fov_deg=60.0)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    lib = _lib.load_library()
    m = nmlp.device_model(p)
    handles = (ctypes.c_void_p * 1)(m.handle.value)
    o = queries.cast_opts_struct(opts)
    n = roots.shape[0]
    t = torch.full((n,), -7.0, device="cuda")
    h = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    c = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    ws_bytes = lib.niq_cast_workspace_bytes(ctypes.byref(o))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    ne = ctypes.c_int64(0)
    torch.cuda.synchronize()
    _lib.check(lib.niq_cast_rays(handles, 1, ctypes.byref(func.cfg), _lib.ptr(roots), _lib.ptr(dirs), n,
                                 ctypes.byref(o), _lib.ptr(t), _lib.ptr(h), _lib.ptr(c), _lib.ptr(ws), ws_bytes,
                                 ctypes.byref(ne), _lib.stream_handle()))
    torch.cuda.synchronize()
    unwritten = int(((t == -7.0) | (h == -7) | (c == -7)).sum().item())
    note(f"  unwritten rays {unwritten}; hits {int((h > 0).sum().item())}; N_evals {ne.value}")
    assert unwritten == 0
    assert int((c < 1).sum().item()) == 0
    k = n // 16
    for i in range(16):
        ct, ch, cc, _ = queries.cast_rays((func,), (p,), roots[i * k:(i + 1) * k], dirs[i * k:(i + 1) * k], opts)
        assert torch.equal(ct, t[i * k:(i + 1) * k]) and torch.equal(ch, h[i * k:(i + 1) * k])
        assert torch.equal(cc, c[i * k:(i + 1) * k])
