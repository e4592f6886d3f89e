"""GPU tests at the benchmark configurations' own models and shapes.

* 3->512x8->1 ReLU (He init, seed 2), query B: an 8x8 window at the centre of
  the 2048x2048 camera (60 degree field of view, eye at distance 3, t in
  [0, 6]) against the oracle in FP32, to the near-threshold bar of
  test_gpu_parity.py; and the precision mode on the same rays never reports
  a hit that the FP32 cast does not (same function, t within hit_eps).
* 64 level-set trees over the 16 seeded TanH MLPs (widths 32-256, 3-8 hidden
  layers) submitted at once to the serving loop together with the
  3->128x8->1 TanH tree under affine_all at split depth 21: every tree equals
  its standalone build, bit for bit.
* The serving mix's eight resident MLPs (seeds 10-17, shapes drawn from the
  three single-GPU models) with a 512-px ray tile, a 65,536-point
  closest-point batch and a depth-21 tree among smaller requests: every
  result equals its standalone call, bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import rays as orays, margins
from niq import implicit_mlp_utils, kd_tree, queries, render, models, serve
from test_gpu_parity import note, only_near

pytestmark = pytest.mark.gpu

LO = np.array([-1., -1., -1.], np.float32)
UP = np.array([1., 1., 1.], np.float32)


def camera(res):
    eye = np.array([0., 0., 3.], np.float32)
    look, up, _ = render.look_at(eye)
    return render.generate_camera_rays(eye, look, up, res=res, fov_deg=60.0, device=None)


def cast_opts():
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    return opts


def same_tree(a, b):
    return a.keys() == b.keys() and all(torch.equal(a[k], b[k]) for k in a)


def same_rays(a, b):
    return all(torch.equal(x, y) for x, y in zip(a[:3], b[:3]))


def test_c4_ray_window_matches_oracle():
    note("[c4 rays] 8x8 window of the 2048^2 camera, 3->512x8->1")
    p = models.config_model("c4_relu_512x8")
    roots, dirs = camera(2048)
    roots, dirs = np.asarray(roots, np.float32), np.asarray(dirs, np.float32)
    rows = np.arange(1020, 1028)
    pix = (rows[:, None] * 2048 + rows[None, :]).reshape(-1)
    r, d = roots[pix], dirs[pix]
    opts = cast_opts()
    f32, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    t, h, c, ne = queries.cast_rays((f32,), (p,), torch.as_tensor(r, device="cuda"), torch.as_tensor(d, device="cuda"),
                                    opts)
    rt, rh, rc, rne, near = orays.cast_rays((margins.TwinFunc(p, "affine_fixed"),), r, d, opts, track_near=True)
    t, h, c = t.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()
    same = (t == rt) & (h == rh) & (c == rc)
    only_near(same, near, "rays", f"{int((rh > 0).sum())} oracle hits, {int(rc.sum())} oracle steps")
    assert (rh > 0).sum() > 0
    f16, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed", precision="bf16")
    t16, h16, c16, _ = queries.cast_rays((f16,), (p,), torch.as_tensor(r, device="cuda"),
                                         torch.as_tensor(d, device="cuda"), opts)
    t16, h16, c16 = t16.cpu().numpy(), h16.cpu().numpy(), c16.cpu().numpy()
    hit16 = h16 > 0
    note(f"  precision mode: {int(hit16.sum())} hits (FP32 {int((h > 0).sum())}), {int(c16.sum())} steps")
    assert np.all(h[hit16] == h16[hit16])
    assert np.all(np.abs(t16[hit16] - t[hit16]) <= 1.001e-3)


def test_c3_trees_served_together_equal_standalone():
    note("[c3] 64 trees over 16 MLPs + the affine_all depth-21 tree, served at once")
    fams = []
    for w, nl, p in models.concurrent_tree_models():
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
        fams.append((f, p))
    pc3 = models.config_model("c3_tanh_128x8")
    fc3, _ = implicit_mlp_utils.implicit_from_params(pc3, "affine_all")
    depth = 15
    alone = [kd_tree.construct_uniform_unknown_levelset_tree(*fams[i % 16], LO, UP, split_depth=depth)
             for i in range(64)]
    alone_c3 = kd_tree.construct_uniform_unknown_levelset_tree(fc3, pc3, LO, UP, split_depth=21)
    with serve.Server(n_streams=4) as srv:
        big = srv.submit_tree(fc3, pc3, LO, UP, split_depth=21)
        reqs = [srv.submit_tree(*fams[i % 16], LO, UP, split_depth=depth) for i in range(64)]
        got = [r.result() for r in reqs]
        got_c3 = big.result()
        lat = serve.latency_summary(reqs + [big])
    note(f"  latency {lat}")
    for i, (a, g) in enumerate(zip(alone, got)):
        assert same_tree(a, g), f"tree {i}"
    assert same_tree(alone_c3, got_c3)


def test_c5_mix_equals_standalone():
    note("[c5] eight resident models: 512-px tile, 65,536-point closest-point batch, depth-21 tree")
    resident, _ = models.serving_models()
    funcs = []
    for name, p in resident:
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
        funcs.append((name, f, p))
    c1 = [i for i, (n, _, _) in enumerate(funcs) if n.startswith("c1")]
    c2 = [i for i, (n, _, _) in enumerate(funcs) if n.startswith("c2")]
    c3 = [i for i, (n, _, _) in enumerate(funcs) if n.startswith("c3")]
    opts = cast_opts()
    r512, d512 = camera(512)
    r64, d64 = camera(64)
    rng = np.random.default_rng(0)
    q_big = torch.as_tensor(rng.uniform(-1, 1, size=(65536, 3)).astype(np.float32), device="cuda")
    q_one = torch.as_tensor(rng.uniform(-1, 1, size=(1, 3)).astype(np.float32), device="cuda")
    # (kind, model index, arguments)
    plan = [("rays", c1[0], (r512, d512)), ("rays", c2[0], (r64, d64)), ("rays", c3[0], (r64, d64)),
            ("cp", c1[1], q_big), ("cp", c3[0], q_one), ("tree", c3[0], 21), ("tree", c2[1], 12),
            ("tree", c1[2], 16), ("rays", c1[3], (r64, d64))]

    def alone(kind, k, arg):
        _, f, p = funcs[k]
        if kind == "rays":
            return queries.cast_rays((f,), (p,), arg[0], arg[1], opts)
        if kind == "cp":
            return kd_tree.closest_point(f, p, LO, UP, arg, eps=0.05)
        return kd_tree.construct_uniform_unknown_levelset_tree(f, p, LO, UP, split_depth=arg)

    want = [alone(*x) for x in plan]
    with serve.Server(n_streams=4, mem_cap_bytes=200 * 10 ** 9) as srv:
        reqs = []
        for kind, k, arg in plan:
            _, f, p = funcs[k]
            if kind == "rays":
                reqs.append(srv.submit_rays((f,), (p,), arg[0], arg[1], opts))
            elif kind == "cp":
                reqs.append(srv.submit_closest_point(f, p, LO, UP, arg, eps=0.05))
            else:
                reqs.append(srv.submit_tree(f, p, LO, UP, split_depth=arg))
        got = [r.result() for r in reqs]
        lat = serve.latency_summary(reqs)
    note(f"  latency {lat}")
    for (kind, k, _), w, g in zip(plan, want, got):
        if kind == "rays":
            assert same_rays(w, g), (kind, k)
        elif kind == "cp":
            assert torch.equal(w[0], g[0]) and torch.equal(w[1], g[1]), (kind, k)
        else:
            assert same_tree(w, g), (kind, k)
