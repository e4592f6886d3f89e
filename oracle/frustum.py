"""Oracle: frustum ray casting.

Test infrastructure only (see ``oracle/__init__.py``).

Restates:
  * cast_rays_frustum_iter               -- src/queries.py:178-364
  * subdivide_frusta                     -- src/queries.py:371-426
  * frustum_needs_subdiv_to_pixel        -- src/queries.py:428-432
  * write_frust_output                   -- src/queries.py:434-452
  * cast_rays_frustum                    -- src/queries.py:454-587
  * camera_ray                           -- src/render.py:17-24
Every frustum evolves independently of the others (its steps, splits and
termination read only its own state and the global iteration count), so
the live frusta are kept in a compact list here; the reference's padded
array sizes are tracked only to reproduce its ``N_evals``.  Arithmetic is
float32 as in the reference (Python-float options rounded to float32).
"""
import numpy as np

from .affine import SIGN_POSITIVE, SIGN_NEGATIVE
from .tree import get_next_bucket_size

f32 = np.float32


def camera_ray(look_dir, up_dir, left_dir, tan_x, tan_y, theta_x, theta_y):
    """src/render.py:17-24, vectorised over theta (N,); tan_* = tan(fov/2)."""
    pos = (look_dir[None, :] + left_dir[None, :] * (theta_x * tan_x)[:, None]
           + up_dir[None, :] * (theta_y * tan_y)[:, None]).astype(f32)
    return (pos / np.linalg.norm(pos, axis=-1, keepdims=True).astype(f32)).astype(f32)


def frustum_geometry(rng, cam):
    """The per-frustum part of take_several_steps (src/queries.py:275-308):
    corner rays, unit mid ray and the expansion factor."""
    look, up, left, tan_x, tan_y, res_x, res_y = cam
    xl, yl, xu, yu = (rng[:, i] for i in range(4))
    rx1, ry1 = f32(res_x + 1.0), f32(res_y + 1.0)
    xc_lo = (f32(2.0) * xl.astype(f32) / rx1 - f32(1.0)).astype(f32)
    xc_hi = (f32(2.0) * (xu - 1).astype(f32) / rx1 - f32(1.0)).astype(f32)
    yc_lo = (f32(2.0) * yl.astype(f32) / ry1 - f32(1.0)).astype(f32)
    yc_hi = (f32(2.0) * (yu - 1).astype(f32) / ry1 - f32(1.0)).astype(f32)
    r_uu = camera_ray(look, up, left, tan_x, tan_y, xc_hi, yc_hi)
    r_lu = camera_ray(look, up, left, tan_x, tan_y, xc_lo, yc_hi)
    r_ul = camera_ray(look, up, left, tan_x, tan_y, xc_hi, yc_lo)
    r_ll = camera_ray(look, up, left, tan_x, tan_y, xc_lo, yc_lo)
    mid = (f32(0.5) * (r_uu + r_ll)).astype(f32)
    mid_len = np.linalg.norm(mid, axis=-1).astype(f32)
    mid = (mid / mid_len[:, None]).astype(f32)
    expand = (f32(1.0) / mid_len).astype(f32)
    single = (xl + 1 == xu) & (yl + 1 == yu)
    return r_uu, r_lu, r_ul, mid, expand, single


def subdivide(rng):
    """src/queries.py:384-400: split along the wider index axis (x on ties);
    returns the A half (overwrites the entry) and the B half (appended)."""
    x_gap = rng[:, 2] - rng[:, 0]
    y_gap = rng[:, 3] - rng[:, 1]
    sx = x_gap >= y_gap
    x_mid = (rng[:, 0] + rng[:, 2]) // 2
    y_mid = (rng[:, 1] + rng[:, 3]) // 2
    A = rng.copy()
    A[:, 2] = np.where(sx, x_mid, rng[:, 2])
    A[:, 3] = np.where(~sx, y_mid, rng[:, 3])
    B = rng.copy()
    B[:, 0] = np.where(sx, x_mid, rng[:, 0])
    B[:, 1] = np.where(~sx, y_mid, rng[:, 1])
    return A, B


def cast_rays_frustum(funcs, root, look, up, left, fov_x, fov_y, res_x, res_y, opts, track_near=False):
    """src/queries.py:454-587.  Returns (t (res_x, res_y), hit_id, count, N_evals)
    indexed [x, y] like the reference's output arrays.  With ``track_near``
    (the funcs are ``margins.TwinFunc``) a fifth (res_x, res_y) result marks
    the pixels whose frustum, or one of its ancestors, took a decision near
    its threshold."""
    if not isinstance(funcs, (tuple, list)):
        funcs = (funcs,)
    root = np.asarray(root, f32)
    look, up, left = (np.asarray(v, f32) for v in (look, up, left))
    tan_x = f32(np.tan(np.deg2rad(f32(fov_x)) / f32(2)))
    tan_y = f32(np.tan(np.deg2rad(f32(fov_y)) / f32(2)))
    cam = (look, up, left, tan_x, tan_y, int(res_x), int(res_y))
    hit_eps = f32(opts["hit_eps"])
    max_dist = f32(opts["max_dist"])
    safety = f32(opts["safety_factor"])
    grow = f32(opts["interval_grow_fac"])
    shrink = f32(opts["interval_shrink_fac"])
    refine = f32(opts["refine_width_fac"])
    n_sub = int(opts["n_substeps"])
    n_max = int(opts["n_max_step"])
    n_side = int(opts["n_side_init"])
    half_fov_x = (np.deg2rad(f32(fov_x)) / f32(2)).astype(f32)
    half_fov_y = (np.deg2rad(f32(fov_y)) / f32(2)).astype(f32)

    # initial grid of tiles (src/queries.py:497-505)
    xt = np.linspace(0, res_x, n_side + 1).astype(np.int32)
    yt = np.linspace(0, res_y, n_side + 1).astype(np.int32)
    rng = np.stack((np.tile(xt[:-1], n_side), np.repeat(yt[:-1], n_side),
                    np.tile(xt[1:], n_side), np.repeat(yt[1:], n_side)), axis=-1).astype(np.int32)
    n0 = rng.shape[0]
    t = np.zeros(n0, f32)
    step = np.full(n0, f32(opts["interval_init_size"] * opts["max_dist"]), f32)
    cnt = np.zeros(n0, f32)
    near = np.zeros(n0, bool)
    fin = []    # (range, t, hit_id, count, near) of terminated frusta
    bucket, empty = n0, n0
    N_evals = 0
    it = 0
    while True:
        N_evals += bucket
        r_uu, r_lu, r_ul, mid, expand, single = frustum_geometry(rng, cam)
        is_hit = np.zeros(rng.shape[0], bool)
        hit_id = np.zeros(rng.shape[0], np.int32)
        demands = np.zeros(rng.shape[0], bool)
        nsteps = np.zeros(rng.shape[0], np.int32)
        for _ in range(n_sub):                                   # take_step, src/queries.py:202-269
            t_up = (t + step).astype(f32)
            t_adj = (t_up * expand).astype(f32)
            right_front = ((r_uu - r_lu) * t_adj[:, None] / f32(2)).astype(f32)
            up_front = ((r_uu - r_ul) * t_adj[:, None] / f32(2)).astype(f32)
            can_step = ~is_hit
            nsteps += (~is_hit).astype(np.int32)
            center = (root[None, :] + (f32(0.5) * (t + t_adj))[:, None] * mid).astype(f32)
            cvec = ((f32(0.5) * (t_adj - t))[:, None] * mid).astype(f32)
            vecs = np.stack((cvec, right_front, up_front), axis=1)
            for fid, func in enumerate(funcs, start=1):
                code = func.classify_general_box(center, vecs)
                if track_near:
                    near |= func.last_near
                can_step &= (code == SIGN_POSITIVE) | (code == SIGN_NEGATIVE)
                p0 = (root[None, :] + t[:, None] * mid).astype(f32)
                p1 = (root[None, :] + (t + hit_eps)[:, None] * mid).astype(f32)
                v0 = func(p0)
                if track_near:
                    near |= func.last_near
                v1 = func(p1)
                if track_near:
                    near |= func.last_near
                this_hit = np.sign(v0) != np.sign(v1)
                hit_id = np.where(this_hit, fid, hit_id).astype(np.int32)
                is_hit |= this_hit
            this_step = np.where(can_step, step, hit_eps * single.astype(f32)).astype(f32)
            t = np.where(is_hit, t, t + this_step * safety).astype(f32)
            step = np.where(can_step, step * grow, step * shrink).astype(f32)
            demands = demands | (step < hit_eps) | is_hit
            step = np.maximum(step, hit_eps)
        area = ((rng[:, 2] - rng[:, 0]) * (rng[:, 3] - rng[:, 1])).astype(f32)
        cnt = (cnt + nsteps.astype(f32) * (f32(1.0) / area)).astype(f32)
        is_hit = is_hit & (area == 1)
        term = is_hit | (t > max_dist) | (it >= n_max)             # src/queries.py:333-336
        for arr in zip(rng[term], t[term], hit_id[term], cnt[term], near[term]):
            fin.append(arr)
        keep = ~term
        # refinement test on the updated t and step (src/queries.py:348-362)
        wx = (f32(2) * np.sin(half_fov_x * (rng[:, 2] - rng[:, 0]).astype(f32) / f32(res_x)) * t).astype(f32)
        wy = (f32(2) * np.sin(half_fov_y * (rng[:, 3] - rng[:, 1]).astype(f32) / f32(res_y)) * t).astype(f32)
        can_subd = (rng[:, 2] > rng[:, 0] + 1) | (rng[:, 3] > rng[:, 1] + 1)
        needs = ((wx > refine * step) | (wy > refine * step) | demands) & can_subd & keep
        it += n_sub
        rng, t, step, cnt, needs, near = rng[keep], t[keep], step[keep], cnt[keep], needs[keep], near[keep]
        n_valid, n_refine = rng.shape[0], int(needs.sum())
        if n_valid == 0:
            break
        # padded-array bookkeeping of the reference (src/queries.py:557-569)
        new_bucket = get_next_bucket_size(n_valid + n_refine)
        if empty + n_refine > bucket or new_bucket < bucket:
            bucket, empty = new_bucket, n_valid
        empty += n_refine
        A, B = subdivide(rng[needs])
        rng_new = rng.copy()
        rng_new[needs] = A
        rng = np.concatenate((rng_new, B), axis=0)
        t = np.concatenate((t, t[needs]))
        step = np.concatenate((step, step[needs]))
        cnt = np.concatenate((cnt, cnt[needs]))
        near = np.concatenate((near, near[needs]))

    # each finished frustum covers its pixels (src/queries.py:571-585)
    out_t = np.zeros((res_x, res_y), f32)
    out_hit = np.zeros((res_x, res_y), np.int32)
    out_count = np.zeros((res_x, res_y), np.int32)
    out_near = np.zeros((res_x, res_y), bool)
    for r, tv, hv, cv, nv in fin:
        out_t[r[0]:r[2], r[1]:r[3]] = tv
        out_hit[r[0]:r[2], r[1]:r[3]] = hv
        out_count[r[0]:r[2], r[1]:r[3]] = np.int32(cv)
        out_near[r[0]:r[2], r[1]:r[3]] = nv
    if track_near:
        return out_t, out_hit, out_count, N_evals, out_near
    return out_t, out_hit, out_count, N_evals
