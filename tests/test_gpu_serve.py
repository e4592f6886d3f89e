"""Serving loop: each result is the same alone and in the mix.

Bar: identical outputs, bit for bit.  A served request runs the same kernels
on the same inputs as the standalone call: rays are independent of one
another (a merged cast gives every ray what a lone cast gives it, each
request being its own group of rays with its own models and options), a
batched tree build runs each tree's levels as its lone build does, and each
closest-point batch runs as one call of its own.  Checked with requests
merged (one worker kept busy so that ray and tree requests queue), with three
workers overlapping on the device, with tree builds replayed from a captured
hipGraph over different root boxes, with memory admission serialising
requests that do not fit together, and with requests that cannot be replayed
from a graph (affine_all, slope_interval at width 512) or use the sdf mode.
"""
import sys
import time

import numpy as np
import pytest
import torch

from niq import implicit_mlp_utils, kd_tree, queries, render, models, serve
from niq import _lib
from niq import mlp as nmlp

pytestmark = pytest.mark.gpu

LO = np.array([-1., -1., -1.], np.float32)
UP = np.array([1., 1., 1.], np.float32)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def camera(res, eye=(0., 0., 3.)):
    eye = np.array(eye, np.float32)
    look, up, _ = render.look_at(eye)
    return render.generate_camera_rays(eye, look, up, res=res, fov_deg=60.0, device=None)


def setup():
    p1 = models.config_model("c1_relu_3x32")
    p2 = models.random_mlp([48, 40, 24], activation="tanh", seed=3)
    f1, _ = implicit_mlp_utils.implicit_from_params(p1, "affine_fixed")
    f2, _ = implicit_mlp_utils.implicit_from_params(p2, "affine_fixed")
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    return (f1, p1), (f2, p2), opts


def tiles(roots, dirs, sizes):
    out, o = [], 0
    for s in sizes:
        out.append((roots[o:o + s], dirs[o:o + s]))
        o += s
    return out


def same_rays(a, b):
    return all(torch.equal(x, y) for x, y in zip(a[:3], b[:3]))


def same_tree(a, b):
    return a.keys() == b.keys() and all(torch.equal(a[k], b[k]) for k in a)


def test_merged_rays_equal_lone_casts():
    note("[serve] merged ray casts")
    (f1, p1), (f2, p2), opts = setup()
    roots, dirs = camera(80)
    chunks = tiles(roots, dirs, [700, 1900, 64, 2000, 1736])
    alone = [queries.cast_rays((f1,), (p1,), r, d, opts) for r, d in chunks]
    with serve.Server(n_streams=1) as srv:
        # the tree keeps the only worker busy, so the casts queue and merge
        busy = srv.submit_tree(f2, p2, LO, UP, split_depth=16)
        reqs = [srv.submit_rays((f1,), (p1,), r, d, opts) for r, d in chunks]
        busy.result()
        got = [r.result() for r in reqs]
        batched = [r.info["n_batched"] for r in reqs]
    note(f"  merged group sizes {batched}")
    assert max(batched) > 1
    for a, g in zip(alone, got):
        assert same_rays(a, g)
        assert a[3] == g[3]          # each request's own N_evals


def test_rays_of_different_models_and_options_share_a_cast():
    """Ray requests of different models (same width template and config) and
    different cast options are merged into one persistent cast, each as its
    own ray group: every request's (t, hit_id, count, N_evals) equals its lone
    cast's, bit for bit."""
    note("[serve] one cast over requests of different models and options")
    p1 = models.random_mlp([24, 24], activation="relu", seed=31)
    p2 = models.random_mlp([40, 32, 16], activation="tanh", seed=32)
    p3 = models.config_model("c1_relu_3x32")
    fs = []
    for p in (p1, p2, p3):
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
        fs.append((f, p))
    o1 = queries.get_default_cast_opts()
    o1["max_dist"] = 6.0
    o2 = dict(o1, hit_eps=0.002, n_max_step=200, interval_grow_fac=1.3)
    roots, dirs = camera(64)
    chunks = tiles(roots, dirs, [900, 1500, 400, 1296])
    plan = [((fs[0],), o1, chunks[0]), ((fs[1],), o2, chunks[1]), ((fs[2], fs[0]), o1, chunks[2]),
            ((fs[1],), o1, chunks[3])]
    alone = [queries.cast_rays(tuple(f for f, _ in fp), tuple(p for _, p in fp), r, d, o) for fp, o, (r, d) in plan]
    (fb, pb), _, _ = setup()
    with serve.Server(n_streams=1) as srv:
        busy = srv.submit_tree(fb, pb, LO, UP, split_depth=16)
        reqs = [srv.submit_rays(tuple(f for f, _ in fp), tuple(p for _, p in fp), r, d, o) for fp, o, (r, d) in plan]
        busy.result()
        got = [r.result() for r in reqs]
        batched = [r.info["n_batched"] for r in reqs]
    note(f"  merged group sizes {batched}")
    assert max(batched) >= 3
    for i, (a, g) in enumerate(zip(alone, got)):
        assert same_rays(a, g), f"request {i}"
        assert a[3] == g[3], f"request {i}: N_evals {g[3]} alone {a[3]}"


def test_mix_equals_standalone():
    note("[serve] mixed requests on three streams")
    (f1, p1), (f2, p2), opts = setup()
    roots, dirs = camera(64)
    chunks = tiles(roots, dirs, [1024, 3072])
    boxes = [(LO, UP), (np.array([-1, -1, -1], np.float32), np.array([0, 0, 0], np.float32)),
             (np.array([-0.5, -0.25, -1], np.float32), np.array([0.5, 0.75, 0], np.float32))]
    q = np.random.default_rng(5).uniform(-1, 1, size=(48, 3)).astype(np.float32)

    alone_rays = [queries.cast_rays((f, ), (p, ), r, d, opts) for (f, p) in ((f1, p1), (f2, p2)) for r, d in chunks]
    alone_trees = [kd_tree.construct_uniform_unknown_levelset_tree(f1, p1, lo, up, split_depth=10,
                                                                   with_interior_nodes=True,
                                                                   with_exterior_nodes=True)
                   for lo, up in boxes]
    alone_trees.append(kd_tree.construct_uniform_unknown_levelset_tree(f2, p2, LO, UP, split_depth=12))
    alone_cp = kd_tree.closest_point(f2, p2, LO, UP, q, eps=0.05)

    with serve.Server(n_streams=3) as srv:
        t_rays, t_trees = [], []
        # interleaved arrivals; each tree shape is submitted more than once so
        # that later builds replay the captured graph
        for rep in range(2):
            for i, (lo, up) in enumerate(boxes):
                t_trees.append((i, srv.submit_tree(f1, p1, lo, up, split_depth=10, with_interior_nodes=True,
                                                   with_exterior_nodes=True)))
                if i < len(chunks):
                    for k, (f, p) in enumerate(((f1, p1), (f2, p2))):
                        r, d = chunks[i]
                        t_rays.append((k * len(chunks) + i, srv.submit_rays((f,), (p,), r, d, opts)))
            t_trees.append((3, srv.submit_tree(f2, p2, LO, UP, split_depth=12)))
        cp = srv.submit_closest_point(f2, p2, LO, UP, q, eps=0.05)
        for i, r in t_rays:
            assert same_rays(alone_rays[i], r.result()), f"ray request {i}"
        replays, batched = 0, 0
        for i, r in t_trees:
            assert same_tree(alone_trees[i], r.result()), f"tree request {i}"
            # a tree waits in the queue with others of its config (merged
            # into one batched build) or runs alone (a captured graph)
            assert r.info["graph_replay"] > 0 or r.info["n_batched"] > 1, i
            replays += r.info["graph_replay"] > 0
            batched += r.info["n_batched"] > 1
        d, loc = cp.result()
        assert torch.equal(d, alone_cp[0]) and torch.equal(loc, alone_cp[1])
        lat = serve.latency_summary([r for _, r in t_rays] + [r for _, r in t_trees] + [cp])
        if replays:
            assert srv.bytes_in_use() > 0      # cached graphs stay resident
    note(f"  {replays} of {len(t_trees)} trees from graphs, {batched} in batched builds; latency {lat}")


@pytest.mark.parametrize("mode", ["affine_fixed", "interval", "slope_interval"])
def test_tree_requests_merge_into_one_build(mode):
    """Tree requests waiting behind a running one, with different models and
    root boxes but one config, depth and template width, run as one batched
    build; each result equals its lone build bit for bit."""
    note(f"[serve] batched tree builds, {mode}")
    (_, p1), (_, p2), _ = setup()
    f1, _ = implicit_mlp_utils.implicit_from_params(p1, mode)
    f2, _ = implicit_mlp_utils.implicit_from_params(p2, mode)
    assert max(nmlp.device_model(p1).dims) <= 64 and max(nmlp.device_model(p2).dims) <= 64   # one template width
    boxes = [(LO, UP), (np.array([-1, -1, -1], np.float32), np.array([0, 0, 0], np.float32)),
             (np.array([-0.5, -0.25, -1], np.float32), np.array([0.5, 0.75, 0], np.float32))]
    kw = dict(split_depth=11, with_interior_nodes=True, with_exterior_nodes=True)
    plan = [(f, p, lo, up) for (f, p) in ((f1, p1), (f2, p2)) for lo, up in boxes]
    alone = [kd_tree.construct_uniform_unknown_levelset_tree(f, p, lo, up, **kw) for f, p, lo, up in plan]
    with serve.Server(n_streams=1) as srv:
        busy = srv.submit_tree(f1, p1, LO, UP, split_depth=16)
        reqs = [srv.submit_tree(f, p, lo, up, **kw) for f, p, lo, up in plan]
        busy.result()
        got = [r.result() for r in reqs]
        sizes = [r.info["n_batched"] for r in reqs]
    note(f"  batch sizes {sizes}")
    assert max(sizes) >= 3
    for i, (a, g) in enumerate(zip(alone, got)):
        assert same_tree(a, g), f"tree request {i}"


def test_memory_admission_serialises():
    note("[serve] memory admission")
    (f1, p1), _, _ = setup()
    alone = kd_tree.construct_uniform_unknown_levelset_tree(f1, p1, LO, UP, split_depth=18)
    lib = _lib.load_library()
    m, p, cap, cap_fin = kd_tree.tree_request(f1, p1, LO, UP, split_depth=18)
    import ctypes
    need = ctypes.c_size_t()
    _lib.check(lib.niq_tree_workspace_bytes(m.handle, ctypes.byref(f1.cfg), ctypes.byref(p), ctypes.byref(need)))
    # room for one build at a time (graphs take 1/8 of the cap: too little
    # here, and two builds do not fit one merged build either)
    with serve.Server(n_streams=2, mem_cap_bytes=int(need.value * 1.5)) as srv:
        t0 = time.perf_counter()
        a = srv.submit_tree(f1, p1, LO, UP, split_depth=18)
        b = srv.submit_tree(f1, p1, LO, UP, split_depth=18)
        gap_ms = (time.perf_counter() - t0) * 1e3   # at least b's submit time after a's
        ra, rb = a.result(), b.result()
        assert not a.info["graph_replay"] and not b.info["graph_replay"]
        assert a.info["n_batched"] == 1 and b.info["n_batched"] == 1
        run_a = a.info["latency_ms"] - a.info["queue_ms"]
        note(f"  a ran {run_a:.3f} ms; b waited {b.info['queue_ms']:.3f} ms, submitted within {gap_ms:.3f} ms of a")
        # b starts only once a is done: it waits a's run less the submit gap
        assert b.info["queue_ms"] >= run_a - gap_ms - 0.2
        with pytest.raises(_lib.CapacityError):
            srv.submit_tree(f1, p1, LO, UP, split_depth=20)
    assert same_tree(alone, ra) and same_tree(alone, rb)


def test_other_modes_equal_lone_calls():
    """Requests the serving loop cannot replay from a captured graph (growing
    error-term rows, slope_interval planes in global scratch at width 512)
    and the sdf mode: each result equals the standalone call's."""
    import os
    from conftest import GOLDEN
    from niq import mlp as nmlp
    note("[serve] other modes")
    p_tanh = models.random_mlp([48, 40, 24], activation="tanh", seed=3)
    f_all, _ = implicit_mlp_utils.implicit_from_params(p_tanh, "affine_all")
    p_wide = models.random_mlp([512, 512], seed=4)
    f_slope, _ = implicit_mlp_utils.implicit_from_params(p_wide, "slope_interval")
    p_fox = nmlp.load(os.path.join(GOLDEN, "fox.npz"))
    f_sdf, _ = implicit_mlp_utils.implicit_from_params(p_fox, "sdf")
    f_c1, _ = implicit_mlp_utils.implicit_from_params(models.config_model("c1_relu_3x32"), "slope_interval")
    p_c1 = models.config_model("c1_relu_3x32")
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    roots, dirs = camera(32)
    q = torch.as_tensor(np.random.default_rng(2).uniform(-1, 1, size=(40, 3)).astype(np.float32), device="cuda")
    alone_all = kd_tree.construct_uniform_unknown_levelset_tree(f_all, p_tanh, LO, UP, split_depth=9)
    alone_slope = kd_tree.construct_uniform_unknown_levelset_tree(f_slope, p_wide, LO, UP, split_depth=6)
    alone_rays = queries.cast_rays((f_sdf,), (p_fox,), roots, dirs, opts)
    alone_cp = kd_tree.closest_point(f_c1, p_c1, LO, UP, q, eps=0.05)
    with serve.Server(n_streams=3) as srv:
        a = srv.submit_tree(f_all, p_tanh, LO, UP, split_depth=9)
        b = srv.submit_tree(f_slope, p_wide, LO, UP, split_depth=6)
        c = srv.submit_rays((f_sdf,), (p_fox,), roots, dirs, opts)
        d = srv.submit_closest_point(f_c1, p_c1, LO, UP, q, eps=0.05)
        ra, rb, rc, rd = a.result(), b.result(), c.result(), d.result()
        assert not a.info["graph_replay"] and not b.info["graph_replay"]
    assert same_tree(alone_all, ra) and same_tree(alone_slope, rb)
    assert same_rays(alone_rays, rc)
    assert torch.equal(rd[0], alone_cp[0]) and torch.equal(rd[1], alone_cp[1])


def test_graph_not_replayed_for_a_new_model_at_a_freed_address():
    """A cached tree graph holds its model's device weights in its captured
    arguments; a model created after another is destroyed may get the same
    handle address, and must not be served by the old model's graph."""
    import gc
    from niq import mlp as nmlp
    note("[serve] graph cache after a model is destroyed and another created")
    with serve.Server(n_streams=1) as srv:
        pa = models.random_mlp([40, 40], activation="relu", seed=21)
        fa, _ = implicit_mlp_utils.implicit_from_params(pa, "affine_fixed")
        first = srv.submit_tree(fa, pa, LO, UP, split_depth=9)
        first.result()
        again = srv.submit_tree(fa, pa, LO, UP, split_depth=9)
        again.result()
        assert not first.info["graph_reused"] and again.info["graph_reused"]   # the same model reuses its graph
        addr_a = nmlp.device_model(pa).handle.value
        # destroy model a: drop every reference, then create b of the same shape
        srv._models.clear()
        nmlp.forget_device_model(pa)
        del first, again, fa, pa
        gc.collect()
        pb = models.random_mlp([40, 40], activation="relu", seed=22)
        fb, _ = implicit_mlp_utils.implicit_from_params(pb, "affine_fixed")
        addr_b = nmlp.device_model(pb).handle.value
        note(f"  handle of a {addr_a:#x}, of b {addr_b:#x}")
        alone = kd_tree.construct_uniform_unknown_levelset_tree(fb, pb, LO, UP, split_depth=9)
        got = srv.submit_tree(fb, pb, LO, UP, split_depth=9)
        res = got.result()
        assert not got.info["graph_reused"]
    assert same_tree(alone, res)


def test_batched_tree_overflow_fails_only_its_tree():
    """In a batched build, a tree whose output list is too small fails with
    NIQ_ERR_CAPACITY, as its lone build would; the other trees of the batch
    finish and equal their lone builds."""
    import ctypes
    note("[serve] overflow inside a batched build")
    (f1, p1), (f2, p2), _ = setup()
    kw = dict(split_depth=11, with_interior_nodes=True, with_exterior_nodes=True)
    alone = [kd_tree.construct_uniform_unknown_levelset_tree(f, p, LO, UP, **kw) for f, p in ((f1, p1), (f2, p2))]
    lib = _lib.load_library()
    with serve.Server(n_streams=1) as srv:
        busy = srv.submit_tree(f1, p1, LO, UP, split_depth=16)
        good1 = srv.submit_tree(f1, p1, LO, UP, **kw)
        # the second model's build with a one-row unknown-node list
        m, p, cap, cap_fin = kd_tree.tree_request(f2, p2, LO, UP, **kw)
        srv._models.append(m)
        bufs = kd_tree.TreeBuffers(p, 1, cap_fin, torch.device("cuda"))
        srv._inputs_ready()
        rid = ctypes.c_uint64()
        _lib.check(lib.niq_serve_submit_tree(srv._h, m.handle, ctypes.byref(f2.cfg), ctypes.byref(p),
                                             ctypes.byref(bufs.r), ctypes.byref(rid)))
        bad = serve.Request(srv, rid.value, "tree", {"bufs": bufs, "tree_res": bufs.r, "cfg": f2.cfg, "p": p},
                            lambda r: r._keep["bufs"].outputs()[0])
        good2 = srv.submit_tree(f2, p2, LO, UP, **kw)
        busy.result()
        assert same_tree(alone[0], good1.result()) and same_tree(alone[1], good2.result())
        with pytest.raises(_lib.CapacityError):
            bad.result()
    note(f"  batch sizes {good1.info['n_batched']}, {bad.info['n_batched']}, {good2.info['n_batched']}")
    assert good1.info["n_batched"] == 3 and bad.info["n_batched"] == 3 and good2.info["n_batched"] == 3
