"""Benchmark of the range-analysis hot path on one MI355X.

Workload (one "step"): the level-set tree of the ReLU MLP 3->256x6->1
(seeded, median-shifted output bias) under affine_fixed range analysis, built
on [-1,1]^3 to split depth 18 (64^3 finest cells), FP32.  ``value`` is boxes
bounded per second over K timed steps (inputs resident: the model is uploaded
before timing).  The same model's 1024x1024 ray cast (60 deg FOV, eye at
distance 3, t in [0,6]) is timed beside it and reported as rays/s.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build targets one
GPU per process and no multi-GPU figure is produced.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-implicit-queries_amd"))
sys.path.insert(0, ROOT)

import numpy as np

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes of one 262,144-box classify launch of the 3->256x6->1 model:
# rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by what the counters read for a
# plain 1 GiB copy (profiles/k_classify_c2_relu256x6_tree18.md)
C2_CLASSIFY_TRAFFIC_BYTES = 19309870 + 1048562


def progress(msg):
    """A step of the bench finished (stderr; stdout carries the one JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--depth", type=int, default=18)
    ap.add_argument("--ray-res", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-precision", action="store_true", help="skip the BF16 vs FP32 comparison")
    ap.add_argument("--precision-depth", type=int, default=21)
    ap.add_argument("--precision-ray-res", type=int, default=512,
                    help="side of the ray cast timed in FP32 and BF16 (0 skips it)")
    ap.add_argument("--no-concurrent", action="store_true", help="skip the 64 concurrent trees")
    ap.add_argument("--no-next", action="store_true", help="skip frustum, sampling, bulk, render, intersection, sdf")
    ap.add_argument("--no-c3", action="store_true", help="skip the affine_all tree")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-sized tensor (a plain copy kernel of known bytes, "
                         "to scale the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--serve-requests", type=int, default=90,
                    help="requests in the serving trace (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=3.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-max-points", type=int, default=512,
                    help="largest closest-point batch of the serving trace (the configuration's is 65,536)")
    ap.add_argument("--no-cp-cost", action="store_true", help="skip the closest-point cost samples")
    return ap.parse_args()


def cpu_baseline(params, target_s):
    """The oracle (NumPy FP32 restatement of the reference) on host cores:
    the first levels of the same tree, deepening until ~target_s of work."""
    from oracle import funcs as ofuncs, tree as otree
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
    except Exception:
        threads = os.cpu_count() or 1
    f = ofuncs.OracleFunc(params, "affine_fixed", np.float32)
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    depth, boxes, secs = 6, 0, 0.0
    while True:
        t0 = time.perf_counter()
        out = otree.construct_uniform_unknown_levelset_tree(f, lo, up, split_depth=depth)
        dt = time.perf_counter() - t0
        boxes, secs = out["stats"]["n_classified"], dt
        if dt * 2.2 > target_s or depth >= 18:
            break
        depth += 1
    return {"value": boxes / secs, "unit": "boxes/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle tree of the same model to split depth {depth} "
                      f"({boxes} boxes, {secs:.1f} s, NumPy FP32 with {threads} BLAS threads)"}


def precision_compare(depth, ray_res=512):
    """ReLU MLP 3->512x8->1 (He init, seed 2): the level-set tree to `depth`
    with the plain FP32 contraction and with the BF16-MFMA contraction plus
    rigorous widening.  One warm-up and one timed build each."""
    import torch
    from niq import models, implicit_mlp_utils, kd_tree, queries, render
    from niq import mlp as nmlp
    p = models.config_model("c4_relu_512x8")
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    res = {"workload": f"levelset tree, ReLU MLP 3->512x8->1, affine_fixed, split depth {depth}"}
    dims = [3] + [512] * 8 + [1]
    for prec in ("fp32", "bf16"):
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed", precision=prec)
        kd_tree.build_tree(f, p, lo, up, split_depth=min(depth, 12))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st = kd_tree.build_tree(f, p, lo, up, split_depth=depth)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[prec] = {"seconds": dt, "boxes": st["n_classified"], "boxes_per_s": st["n_classified"] / dt,
                     "unknown_leaves": st["n_unknown"],
                     "tflops": models.flops_per_box(dims, 5) * st["n_classified"] / dt / 1e12}
    res["speedup_bf16_over_fp32"] = res["fp32"]["seconds"] / res["bf16"]["seconds"]
    res["bound_quality"] = {name: bf16_bound_quality(name) for name in ("c4_relu_512x8", "fox", "bunny")}
    progress("precision: bound quality " + ", ".join(
        f"{k} median inflation vs fp64 {v['inflation_vs_fp64']['median']:.3f}, false-empty {v['false_empty']}"
        for k, v in res["bound_quality"].items()))
    if ray_res > 0:
        # the 2048^2 cast of this net (query B), sampled: every 4th pixel of the
        # 60 deg camera at distance 3 in both directions (512^2 rays, 1/16 of
        # the image), t in [0, 6], in FP32 and in the precision mode
        eye = np.array([0., 0., 3.], np.float32)
        look, upd, _ = render.look_at(eye)
        roots, dirs = render.generate_camera_rays(eye, look, upd, res=2048, fov_deg=60.)
        sel = torch.arange(0, 2048, 4)
        pix = (sel[:, None] * 2048 + sel[None, :]).reshape(-1).to(roots.device)
        roots, dirs = roots[pix].contiguous(), dirs[pix].contiguous()
        opts = queries.get_default_cast_opts()
        opts["max_dist"] = 6.0
        fl = models.flops_per_box(dims, rows=5)
        for prec in ("fp32", "bf16"):
            f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed", precision=prec)
            queries.cast_rays((f,), (p,), roots[:1024], dirs[:1024], opts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, hid, cnt, _ = queries.cast_rays((f,), (p,), roots, dirs, opts)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            steps = int(cnt.sum().item())
            line = {"sample": "every 4th pixel of the 2048x2048 camera in both directions (262,144 rays, 1/16)",
                    "rays_per_s": roots.shape[0] / dt, "ray_steps_per_s": steps / dt, "seconds": dt,
                    "hits": int((hid > 0).sum().item()), "ray_steps": steps, "flops_per_step": fl,
                    "full_image_seconds_estimate": dt * 16}
            if prec == "fp32":
                line["tflops"] = steps * fl / dt / 1e12
                line["frac_fp32_mfma_peak"] = line["tflops"] / PEAK_FP32_TFLOPS
            res[f"rays_{prec}"] = line
            progress(f"precision: 3->512x8->1 rays {prec}: {roots.shape[0] / dt:.4g} rays/s, "
                     f"{int((hid > 0).sum().item())} hits")
        # cast_rays in both forms.  On this He-init net the precision mode's
        # a-priori widening exceeds the function's scale, so it certifies no
        # step; rays are timed on the reference's trained fox SDF instead
        # (60 deg camera at distance 3, t in [0, 6]).
        p = nmlp.load(os.path.join(ROOT, "tests", "golden", "fox.npz"))
        eye = np.array([0., 0., 3.], np.float32)
        look, upd, _ = render.look_at(eye)
        roots, dirs = render.generate_camera_rays(eye, look, upd, res=ray_res, fov_deg=60.)
        opts = queries.get_default_cast_opts()
        opts["max_dist"] = 6.0
        rays = {"model": "fox (the reference's sample SDF)", "res": ray_res}
        for prec in ("fp32", "bf16"):
            f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed", precision=prec)
            queries.cast_rays((f,), (p,), roots[:1024], dirs[:1024], opts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, hid, cnt, _ = queries.cast_rays((f,), (p,), roots, dirs, opts)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            rays[prec] = {"rays_per_s": ray_res ** 2 / dt, "seconds": dt, "hits": int((hid > 0).sum().item()),
                          "ray_steps": int(cnt.sum().item())}
        rays["speedup_bf16_over_fp32"] = rays["fp32"]["seconds"] / rays["bf16"]["seconds"]
        res["rays"] = rays
    return res


def bf16_bound_quality(name, n_cells=512, seed=4):
    """Precision-mode bounds at the finest cells of a depth-21 tree (half-width
    1/128, random centres): width inflation against the plain FP32 kernel and
    against the FP64 oracle (median, p90, max), and the false-empty count (a
    cell the precision mode calls POSITIVE or NEGATIVE whose FP64 bound does
    not), with the number of cells either form calls determinate."""
    import torch
    from niq import models, implicit_mlp_utils, mlp as nmlp
    from oracle import funcs as ofuncs, affine as oaff
    p = (models.config_model(name) if name in models.CONFIGS
         else nmlp.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz")))
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=(n_cells, 3))
    lo, up = (c - 1 / 128).astype(np.float32), (c + 1 / 128).astype(np.float32)
    tl, tu = torch.as_tensor(lo, device="cuda"), torch.as_tensor(up, device="cuda")
    f16, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed", precision="bf16")
    f32, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    c16, b16 = (x.cpu().numpy() for x in f16.classify_box(p, tl, tu, return_bounds=True))
    c32, b32 = (x.cpu().numpy() for x in f32.classify_box(p, tl, tu, return_bounds=True))
    d_lo, d_hi = ofuncs.OracleFunc(p, "affine_fixed", np.float64).bounds_box(lo, up)
    c64 = oaff.sign_code(d_lo, d_hi, 0.0)
    w16 = b16[:, 1] - b16[:, 0]
    stats = lambda r: {"median": float(np.median(r)), "p90": float(np.percentile(r, 90)), "max": float(np.max(r))}
    return {"cells": n_cells, "half_width": 1 / 128,
            "inflation_vs_fp32": stats(w16 / (b32[:, 1] - b32[:, 0])),
            "inflation_vs_fp64": stats(w16 / (d_hi - d_lo)),
            "false_empty": int(((c16 != 0) & (c16 != c64)).sum()),
            "determinate_bf16": int((c16 != 0).sum()), "determinate_fp64": int((c64 != 0).sum())}


def closest_point_cost(eps=0.05, samples=(("c1", 4096), ("c2", 16), ("c3", 16))):
    """What one closest-point query costs on each serving shape: a batch of
    `n` uniform points in [-1,1]^3 (eps 0.05) on the first resident model of
    the shape, timed once after a warm-up.  Every node popped from the stack
    runs one box forward (5 rows) and 7 point forwards, so a node is
    12 rows through the layers; the FLOP rate is against the FP32 MFMA peak.
    ``s_per_65536_points`` extrapolates the per-point time linearly to the
    largest batch of the serving configuration."""
    import torch
    from niq import models, implicit_mlp_utils, kd_tree
    from niq import mlp as nmlp
    resident, _ = models.serving_models()
    rng = np.random.default_rng(0)
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    out = {}
    for prefix, n in samples:
        name, p = [r for r in resident if r[0].startswith(prefix)][0]
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
        q = torch.as_tensor(rng.uniform(-1, 1, (n, 3)).astype(np.float32), device="cuda")
        kd_tree.closest_point(f, p, lo, up, q[:1], eps=eps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kd_tree.closest_point(f, p, lo, up, q, eps=eps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        nodes = kd_tree.closest_point.last_n_processed
        dims = list(nmlp.device_model(p).dims)
        fl = models.flops_per_box(dims, rows=12)
        out[name] = {"points": n, "nodes": nodes, "ms": dt * 1e3, "nodes_per_point": nodes / n,
                     "nodes_per_s": nodes / dt, "rounds_of_2048": -(-nodes // 2048),
                     "us_per_round": dt * 1e6 / -(-nodes // 2048), "flops_per_node": fl,
                     "tflops": nodes * fl / dt / 1e12, "frac_fp32_mfma": nodes * fl / dt / 1e12 / PEAK_FP32_TFLOPS,
                     "s_per_65536_points": dt / n * 65536}
        progress(f"closest point {name}: {n} points {dt * 1e3:.1f} ms, {nodes} nodes")
    return out


def serving_trace(n_req, rate, seed=0, max_side=512, max_points=512):
    """The serving configuration: eight resident MLPs (seeds 10-17, shapes
    drawn from the three single-GPU models) under affine_fixed; Poisson
    arrivals (seed 0) of 50 % ray-cast tiles (side U{64..max_side} of the
    60-degree camera at distance 3), 30 % closest-point batches
    (U{1..max_points} points in [-1,1]^3, eps 0.05) and 20 % level-set trees
    (split depth U{12..21}); memory cap 200 GB.  n_req requests (the
    configuration asks for 10,000) at `rate` requests/s.  The configuration's
    closest-point batches reach 65,536 points; on the randomly initialised
    C2/C3-shaped models one point pops about 10^5 nodes, so a batch of 65,536
    is minutes of GPU work on its own (``closest_point_cost``) and max_points
    defaults to 512.  Reports p50/p99 latency per kind and overall,
    requests/s and how many ray requests were merged into shared casts."""
    import torch
    from niq import models, implicit_mlp_utils, queries, render, serve
    resident_models, rng = models.serving_models(seed)
    resident = []
    for name, p in resident_models:
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
        resident.append((name, f, p))
    eye = np.array([0., 0., 3.], np.float32)
    look, upd, _ = render.look_at(eye)
    cams = {}
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    plan = []
    for i in range(n_req):
        u = rng.uniform()
        k = int(rng.integers(0, len(resident)))
        if u < 0.5:
            plan.append(("rays", k, int(rng.integers(64, max_side + 1))))
        elif u < 0.8:
            plan.append(("closest_point", k, int(rng.integers(1, max_points + 1))))
        else:
            plan.append(("tree", k, int(rng.integers(12, 22))))
    # inputs built before the clock starts (resident in HBM)
    inputs = []
    for kind, k, arg in plan:
        if kind == "rays":
            if arg not in cams:
                cams[arg] = render.generate_camera_rays(eye, look, upd, res=arg, fov_deg=60.)
            inputs.append(cams[arg])
        elif kind == "closest_point":
            inputs.append(torch.as_tensor(rng.uniform(-1, 1, size=(arg, 3)).astype(np.float32), device="cuda"))
        else:
            inputs.append(None)
    torch.cuda.synchronize()
    reqs = []
    with serve.Server(n_streams=4, mem_cap_bytes=200 * 10 ** 9) as srv:
        t0 = time.perf_counter()
        for (kind, k, arg), inp, at in zip(plan, inputs, arrivals):
            wait = at - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            _, f, p = resident[k]
            if kind == "rays":
                reqs.append(srv.submit_rays((f,), (p,), inp[0], inp[1], opts))
            elif kind == "closest_point":
                reqs.append(srv.submit_closest_point(f, p, lo, up, inp, eps=0.05))
            else:
                reqs.append(srv.submit_tree(f, p, lo, up, split_depth=arg))
            if len(reqs) % 25 == 0:
                progress(f"serving: {len(reqs)}/{len(plan)} requests submitted")
        for i, r in enumerate(reqs):
            r.result()
            if (i + 1) % 25 == 0:
                progress(f"serving: {i + 1}/{len(reqs)} requests done")
        wall = time.perf_counter() - t0
    lat = serve.latency_summary(reqs)
    merged = sum(1 for r in reqs if r.kind == "rays" and r.info["n_batched"] > 1)
    replays = sum(1 for r in reqs if r.kind == "tree" and r.info["graph_replay"])
    reused = sum(1 for r in reqs if r.kind == "tree" and r.info["graph_reused"])
    return {"workload": f"{n_req} requests, Poisson {rate:g}/s (seed {seed}), 8 resident MLPs, 50% ray tiles "
                        f"64-{max_side} px, 30% closest-point 1-{max_points} pts (eps 0.05), 20% trees depth 12-21",
            "requests_per_s": n_req / wall, "wall_s": wall, "latency_ms": lat,
            "rays_merged_requests": merged, "arrival_rate_per_s": rate, "trees_from_graphs": replays,
            "trees_from_cached_graphs": reused, "models": [m[0] for m in resident]}


def next_rows():
    """The queries beyond the tree and the ray cast, on the reference's trained
    fox SDF (tests/golden/fox.npz): frustum against per-ray casting at 1024^2,
    level-set sampling, bulk properties, a matcap render, an intersection
    search and a tree in the sdf mode.  Each is timed once after a warm-up."""
    import torch
    from niq import implicit_mlp_utils, kd_tree, queries, render, models, mlp as nmlp

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    fox = nmlp.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "fox.npz"))
    ff, _ = implicit_mlp_utils.implicit_from_params(fox, "affine_fixed")
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    eye = np.array([0., 0., 3.], np.float32)
    look, upd, left = render.look_at(eye)
    res = 1024
    out = {"model": "fox (the reference's sample SDF)"}
    sf, rf = timed(lambda: queries.cast_rays_frustum((ff,), (fox,), (eye, look, upd, left, 60.0, 60.0, res, res), opts))
    roots, dirs = render.generate_camera_rays(eye, look, upd, res=res, fov_deg=60.)
    sr, rr = timed(lambda: queries.cast_rays((ff,), (fox,), roots, dirs, opts))
    out["frustum_1024"] = {"seconds": sf, "rays_per_s": res * res / sf, "hits": int((rf[1] > 0).sum().item()),
                           "per_ray_seconds": sr, "per_ray_rays_per_s": res * res / sr,
                           "per_ray_hits": int((rr[1] > 0).sum().item()), "speedup_over_per_ray": sr / sf}
    n_samp = 1000000
    ss, _ = timed(lambda: kd_tree.sample_surface(ff, fox, lo, up, n_samp, 0.005, 1))
    out["sample_surface"] = {"samples": n_samp, "width": 0.005, "seconds": ss, "samples_per_s": n_samp / ss,
                             "rounds": kd_tree.sample_surface.last_rounds}
    sb, (mass, _) = timed(lambda: kd_tree.bulk_properties(ff, fox, lo, up, 1))
    out["bulk_properties"] = {"samples": 1000000, "seconds": sb, "mass": float(mass)}
    mats = tuple(np.random.default_rng(0).uniform(0, 1, (256, 256, 3)).astype(np.float32) for _ in range(4))
    e2 = np.array([0., 0.3, 2.5], np.float32)
    l2, u2, f2 = render.look_at(e2)
    sx, _ = timed(lambda: render.render_image((ff,), (fox,), e2, l2, u2, f2, res, 45.0, True, dict(opts, n_side_init=16),
                                              shading="matcap_color", matcaps=mats, tonemap=True))
    out["render_image_frustum_matcap_1024"] = {"seconds": sx, "frames_per_s": 1.0 / sx}
    a = models.random_mlp([32, 32], seed=11)
    b = models.random_mlp([48, 24], activation="tanh", seed=12)
    fa, _ = implicit_mlp_utils.implicit_from_params(a, "affine_fixed")
    fb, _ = implicit_mlp_utils.implicit_from_params(b, "affine_fixed")
    si, _ = timed(lambda: kd_tree.find_any_intersection((fa, fb), (a, b), lo, up, 0.001))
    out["find_any_intersection"] = {"seconds": si, "nodes": kd_tree.find_any_intersection.last_n_processed}
    fs, _ = implicit_mlp_utils.implicit_from_params(fox, "sdf")
    st, (_, stats) = timed(lambda: kd_tree.build_tree(fs, fox, lo, up, split_depth=21))
    out["sdf_tree_21"] = {"seconds": st, "boxes": stats["n_classified"], "boxes_per_s": stats["n_classified"] / st}
    return out


def c3_affine_all(depth=21):
    """TanH MLP 3->128x8->1 (Xavier init, seed 1) under affine_all, whose
    error terms grow to 3 + 1024 rows: the level-set tree to `depth`."""
    import torch
    from niq import models, implicit_mlp_utils, kd_tree, autotune
    p = models.config_model("c3_tanh_128x8")
    f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_all")
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    tuned = autotune.tune(f, p, split_depth=depth, n_boxes=1 << 14, reps=1)
    kd_tree.build_tree(f, p, lo, up, split_depth=9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, st = kd_tree.build_tree(f, p, lo, up, split_depth=depth)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lv = st["level_counts"]
    dims = [3] + [128] * 8 + [1]
    fl = models.flops_per_box_affine_all(dims)
    eb = models.error_bytes_per_box_affine_all(dims)
    tflops = fl * st["n_classified"] / dt / 1e12
    gbs = eb * st["n_classified"] / dt / 1e9
    # dominant launch: the classify of the largest level's box count, timed
    # alone on torch's stream with HIP events (affine_all's work per box does
    # not depend on the box: every box carries 3 + 128 rows per activation)
    n_big = int(max(lv))
    g = torch.Generator(device="cpu").manual_seed(0)
    c = torch.rand((n_big, 3), generator=g) * 2 - 1
    h = torch.full((n_big, 3), 2.0 ** -7)
    blo, bup = (c - h).cuda(), (c + h).cuda()
    f.classify_box(p, blo, bup)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(3):
        f.classify_box(p, blo, bup)
    ev1.record()
    torch.cuda.synchronize()
    launch_ms = ev0.elapsed_time(ev1) / 3
    l_tflops = fl * n_big / (launch_ms * 1e-3) / 1e12
    l_gbs = eb * n_big / (launch_ms * 1e-3) / 1e9
    return {"workload": f"levelset tree, TanH MLP 3->128x8->1, affine_all, split depth {depth}",
            "seconds": dt, "boxes": st["n_classified"], "boxes_per_s": st["n_classified"] / dt,
            "unknown_leaves": st["n_unknown"], "active_boxes_per_level": [min(lv), max(lv)],
            "variant": tuned["best"], "tile": tuned.get("shape"), "variant_ms": tuned["ms"],
            # the whole tree's algorithmic work over its wall time: the
            # error-term contraction (FP32 MFMA) and the error-term matrix
            # streamed through HBM (read and written once per layer)
            "roofline": {"kernel": "k_classify_rows (affine_all, W=128)", "flops_per_box": fl,
                         "achieved_tflops": tflops, "peak_tflops": PEAK_FP32_TFLOPS,
                         "frac_mfma": tflops / PEAK_FP32_TFLOPS, "hbm_bytes_per_box": eb,
                         "achieved_gbs": gbs, "peak_gbs": PEAK_HBM_GBS, "frac_hbm": gbs / PEAK_HBM_GBS,
                         "largest_level_launch": {"boxes": n_big, "launch_ms": launch_ms,
                                                  "achieved_tflops": l_tflops,
                                                  "frac_mfma": l_tflops / PEAK_FP32_TFLOPS,
                                                  "achieved_gbs": l_gbs, "frac_hbm": l_gbs / PEAK_HBM_GBS}}}


def concurrent_trees(n_models=16, n_trees=64, depth=15, seed=1, n_streams=4):
    """64 level-set trees over 16 resident TanH MLPs (Xavier init, seeds
    20-35; widths from {32, 64, 128, 256} and 3-8 hidden layers drawn from
    seed 1), affine_fixed with autotuned kernel variants, built one after
    another and then all submitted at once to the serving loop (4 streams)."""
    import torch
    from niq import models, implicit_mlp_utils, kd_tree, serve, autotune
    fams = []
    tuned = []
    for w, nl, p in models.concurrent_tree_models(n_models, seed):
        f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
        tuned.append(autotune.tune(f, p, split_depth=depth, n_boxes=1 << 14, reps=2))
        fams.append((f, p, w, nl))
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)
    for f, p, _, _ in fams:
        kd_tree.build_tree(f, p, lo, up, split_depth=8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    boxes, flops = 0, 0
    for i in range(n_trees):
        f, p, w, nl = fams[i % n_models]
        _, st = kd_tree.build_tree(f, p, lo, up, split_depth=depth)
        boxes += st["n_classified"]
        flops += st["n_classified"] * models.flops_per_box([3] + [w] * nl + [1], rows=5)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    with serve.Server(n_streams=n_streams) as srv:
        # the first pass captures the graphs; the timed pass replays them
        warm = [srv.submit_tree(fams[i % n_models][0], fams[i % n_models][1], lo, up, split_depth=depth)
                for i in range(n_trees)]
        for r in warm:
            r.result()
        t0 = time.perf_counter()
        reqs = [srv.submit_tree(fams[i % n_models][0], fams[i % n_models][1], lo, up, split_depth=depth)
                for i in range(n_trees)]
        for r in reqs:
            r.result()
        t_srv = time.perf_counter() - t0
    lat = serve.latency_summary(reqs)
    return {"workload": f"{n_trees} trees, split depth {depth}, over {n_models} TanH MLPs "
                        "(widths 32-256, 3-8 hidden layers), affine_fixed",
            "boxes": boxes, "sequential_s": t_seq, "served_s": t_srv,
            "boxes_per_s_sequential": boxes / t_seq, "boxes_per_s_served": boxes / t_srv,
            "speedup_served_over_sequential": t_seq / t_srv,
            # the trees' contraction work (5 rows through every layer per box)
            # over each wall time: how close the builds already are to the
            # FP32 MFMA roof, which bounds what overlapping them can gain
            "tflops_sequential": flops / t_seq / 1e12, "tflops_served": flops / t_srv / 1e12,
            "frac_fp32_mfma_served": flops / t_srv / 1e12 / PEAK_FP32_TFLOPS,
            "latency_ms": lat.get("tree"),
            # requests per shared build (trees of one config, depth and
            # template width waiting together run their levels in shared
            # launches)
            "batch_sizes": sorted({int(r.info["n_batched"]) for r in reqs}),
            "trees_from_graphs": sum(1 for r in reqs if r.info["graph_replay"]),
            "variants": [int(f.cfg.variant) for f, _, _, _ in fams],
            "tiles": [r.get("shape") for r in tuned],
            "shapes": [f"{w}x{nl}" for _, _, w, nl in fams]}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU per process; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from niq import models, implicit_mlp_utils, kd_tree, queries, render

    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        progress(f"copy calibration: 3 copies of {args.copy_calibration} MiB")

    name = "c2_relu_256x6"
    params = models.config_model(name)
    func, params = implicit_mlp_utils.implicit_from_params(params, "affine_fixed")
    dims = [3] + [256] * 6 + [1]
    flops_box = models.flops_per_box(dims, rows=5)
    lo, up = np.full(3, -1, np.float32), np.full(3, 1, np.float32)

    def step():
        return kd_tree.build_tree(func, params, lo, up, split_depth=args.depth)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_boxes = 0
    lat = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        _, st = step()
        lat.append(time.perf_counter() - ts)
        n_boxes += st["n_classified"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    boxes_per_s = n_boxes / elapsed

    # dominant kernel: the fused classify pass, timed alone on torch's stream
    # over the finest level's worth of boxes with HIP events
    n_cls = 1 << 18
    g = torch.Generator(device="cpu").manual_seed(0)
    c = torch.rand((n_cls, 3), generator=g) * 2 - 1
    h = torch.full((n_cls, 3), 1.0 / 64)
    blo, bup = (c - h).cuda(), (c + h).cuda()
    func.classify_box(params, blo, bup)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    ev0.record()
    for _ in range(reps):
        func.classify_box(params, blo, bup)
    ev1.record()
    torch.cuda.synchronize()
    cls_ms = ev0.elapsed_time(ev1) / reps
    achieved_tflops = flops_box * n_cls / (cls_ms * 1e-3) / 1e12

    # rays (same model), one cast of res^2 rays
    eye = np.array([0., 0., 3.], np.float32)
    look, upd, _ = render.look_at(eye)
    roots, dirs = render.generate_camera_rays(eye, look, upd, res=args.ray_res, fov_deg=60.)
    opts = queries.get_default_cast_opts()
    opts["max_dist"] = 6.0
    queries.cast_rays((func,), (params,), roots[:4096], dirs[:4096], opts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, hid, cnt, n_evals = queries.cast_rays((func,), (params,), roots, dirs, opts)
    torch.cuda.synchronize()
    ray_s = time.perf_counter() - t1
    ray_steps = int(cnt.sum().item())
    flops_step = models.flops_per_box(dims, rows=5)


    line = {
        "metric": "boxes bounded/s (level-set tree, affine_fixed FP32)",
        "value": boxes_per_s,
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ReLU MLP, He init, median-shifted output bias)",
        "config": {"workload": f"levelset tree, ReLU MLP 3->256x6->1, affine_fixed, split depth {args.depth}",
                   "boxes_per_step": n_boxes // args.steps, "parallelism": "single GPU"},
        "rays": {"rays_per_s": args.ray_res ** 2 / ray_s, "ray_steps_per_s": ray_steps / ray_s,
                 "res": args.ray_res, "seconds": ray_s, "hits": int((hid > 0).sum().item()),
                 "n_evals": n_evals,
                 # each step bounds the segment (base, one error row, err) and
                 # evaluates its two hit-test points: 5 rows through the MLP
                 "roofline": {"bound": "mfma", "achieved": ray_steps * flops_step / ray_s / 1e12,
                              "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                              "frac": ray_steps * flops_step / ray_s / 1e12 / PEAK_FP32_TFLOPS,
                              "kernel": "k_cast_rays (affine_fixed, W=256)", "flops_per_step": flops_step}},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP32_TFLOPS,
                     "traffic": C2_CLASSIFY_TRAFFIC_BYTES if n_cls == 1 << 18 else None,
                     "kernel": "k_classify (affine_fixed, W=256)", "launch_ms": cls_ms,
                     "flops_per_box": flops_box, "boxes_per_launch": n_cls},
    }
    line["latency_ms"] = {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)}
    progress(f"tree: {boxes_per_s:.4g} boxes/s; rays: {args.ray_res ** 2 / ray_s:.4g} rays/s")
    if not args.no_precision:
        line["precision"] = precision_compare(args.precision_depth, args.precision_ray_res)
        progress(f"precision: bf16 {line['precision']['speedup_bf16_over_fp32']:.3f}x over fp32")
    if not args.no_c3:
        line["c3_affine_all"] = c3_affine_all()
        progress(f"c3 affine_all: {line['c3_affine_all']['boxes_per_s']:.4g} boxes/s")
    if not args.no_next:
        line["next_rows"] = next_rows()
        progress(f"next rows: frustum {line['next_rows']['frustum_1024']['rays_per_s']:.4g} rays/s, "
                 f"sampling {line['next_rows']['sample_surface']['samples_per_s']:.4g} samples/s")
    if not args.no_concurrent:
        line["concurrent_trees"] = concurrent_trees()
        progress(f"concurrent trees: {line['concurrent_trees']['boxes_per_s_served']:.4g} boxes/s served")
    if not args.no_cp_cost:
        line["closest_point"] = closest_point_cost()
    if args.serve_requests > 0:
        line["serving"] = serving_trace(args.serve_requests, args.serve_rate, max_points=args.serve_max_points)
        progress(f"serving: {line['serving']['requests_per_s']:.4g} requests/s")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(params, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
