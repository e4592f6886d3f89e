"""Oracle: pinhole camera rays and certified-step ray casting.

Test infrastructure only (see ``oracle/__init__.py``).

Restates:
  * get_default_cast_opts                 -- src/queries.py:23-36
  * cast_rays_iter / cast_rays            -- src/queries.py:39-175
  * camera_ray / generate_camera_rays     -- src/render.py:17-50
  * look_at                               -- src/render.py:229-242
  * normalize / orthogonal_dir            -- src/geometry.py:12-19
All arithmetic is float32 with Python-float options rounded to float32 first
(JAX's weak typing of Python scalars).
"""
import numpy as np

from .affine import SIGN_POSITIVE, SIGN_NEGATIVE
from .tree import get_next_bucket_size

f32 = np.float32


def get_default_cast_opts():
    """src/queries.py:23-36."""
    return {
        "hit_eps": 0.001, "max_dist": 10.0, "n_max_step": 512, "n_substeps": 1,
        "safety_factor": 0.98, "interval_grow_fac": 1.5, "interval_shrink_fac": 0.5,
        "interval_init_size": 0.1, "refine_width_fac": 2.0, "n_side_init": 16,
    }


def normalize(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True).astype(x.dtype)


def orthogonal_dir(x, remove_dir):
    remove_dir = normalize(remove_dir)
    x = x - np.dot(x, remove_dir) * remove_dir
    return normalize(x)


def look_at(eye_pos, target=None, up_dir="y"):
    """src/render.py:229-242."""
    eye_pos = np.asarray(eye_pos, f32)
    target = np.zeros(3, f32) if target is None else np.asarray(target, f32)
    up = {"y": np.array((0.0, 1.0, 0.0), f32), "z": np.array((0.0, 0.0, 1.0), f32)}[up_dir]
    look_dir = normalize(target - eye_pos)
    up = orthogonal_dir(up, look_dir)
    left_dir = np.cross(look_dir, up).astype(f32)
    return look_dir, up, left_dir


def generate_camera_rays(eye_pos, look_dir, up_dir, res=1024, fov_deg=30.0):
    """src/render.py:26-50.  Pixel p = row*res + col ('xy' meshgrid)."""
    ax = np.linspace(-1.0, 1.0, res, dtype=f32)
    cam_x, cam_y = np.meshgrid(ax, ax)
    cam_x = cam_x.reshape(-1)
    cam_y = cam_y.reshape(-1)
    look_dir = np.asarray(look_dir, f32)
    up_dir = np.asarray(up_dir, f32)
    up_dir = up_dir - np.dot(look_dir, up_dir) * look_dir
    up_dir = normalize(up_dir)
    left_dir = np.cross(look_dir, up_dir).astype(f32)
    tan_half = f32(np.tan(np.deg2rad(f32(fov_deg)) / f32(2)))
    pos = (look_dir[None, :] + left_dir[None, :] * (cam_x * tan_half)[:, None]
           + up_dir[None, :] * (cam_y * tan_half)[:, None])
    dirs = normalize(pos.astype(f32))
    roots = np.tile(np.asarray(eye_pos, f32), (dirs.shape[0], 1))
    return roots, dirs


def cast_rays(funcs, roots, dirs, opts, track_near=False):
    """src/queries.py:134-175 for a tuple of oracle functions.

    Returns (t (N,), hit_id (N,) int32, count (N,) int32, N_evals).  The
    iteration is lock-step over all live rays exactly as the reference; the
    bucketed array size is tracked only to reproduce its ``N_evals``.  With
    ``track_near`` (the funcs are ``margins.TwinFunc``) a fifth result marks
    the rays one of whose decisions lay near its threshold."""
    if not isinstance(funcs, (tuple, list)):
        funcs = (funcs,)
    roots = np.asarray(roots, f32)
    dirs = np.asarray(dirs, f32)
    N = roots.shape[0]
    hit_eps = f32(opts["hit_eps"])
    max_dist = f32(opts["max_dist"])
    safety = f32(opts["safety_factor"])
    grow = f32(opts["interval_grow_fac"])
    shrink = f32(opts["interval_shrink_fac"])
    n_sub = int(opts["n_substeps"])
    n_max = int(opts["n_max_step"])

    out_t = np.zeros(N, f32)
    out_hit = np.zeros(N, np.int32)
    out_count = np.zeros(N, np.int32)
    idx = np.arange(N)
    t = np.zeros(N, f32)
    step = np.full(N, f32(opts["interval_init_size"] * opts["max_dist"]), f32)
    count = np.zeros(N, np.int32)
    bucket = N
    N_evals = 0
    near = np.zeros(N, bool)
    while idx.shape[0] > 0:
        r, d = roots[idx], dirs[idx]
        is_hit = np.zeros(idx.shape[0], bool)
        hit_id = np.zeros(idx.shape[0], np.int32)
        nsteps = np.zeros(idx.shape[0], np.int32)
        for _ in range(n_sub):                                # src/queries.py:44-92
            can_step = ~is_hit
            nsteps += (~is_hit).astype(np.int32)
            for fid, func in enumerate(funcs, start=1):
                pos_start = r + t[:, None] * d
                half_vec = (f32(0.5) * step)[:, None] * d
                pos_mid = pos_start + half_vec
                code = func.classify_general_box(pos_mid, half_vec[:, None, :])
                if track_near:
                    near[idx] |= func.last_near
                can_step &= (code == SIGN_POSITIVE) | (code == SIGN_NEGATIVE)
                pos_eps = r + (t + hit_eps)[:, None] * d
                v0 = func(pos_start)
                if track_near:
                    near[idx] |= func.last_near
                v1 = func(pos_eps)
                if track_near:
                    near[idx] |= func.last_near
                this_hit = np.sign(v0) != np.sign(v1)
                hit_id = np.where(this_hit, fid, hit_id).astype(np.int32)
                is_hit |= this_hit
            this_step = np.where(can_step, step, hit_eps)
            t = np.where(is_hit, t, t + this_step * safety)
            step = np.where(can_step, step * grow, step * shrink)
            step = np.maximum(step, hit_eps)
        N_evals += bucket * n_sub
        count = count + nsteps
        term = is_hit | (t > max_dist) | (count >= n_max)     # src/queries.py:115-120
        out_t[idx[term]] = t[term]
        out_hit[idx[term]] = hit_id[term]
        out_count[idx[term]] = count[term]
        keep = ~term
        idx, t, step, count = idx[keep], t[keep], step[keep], count[keep]
        n_valid = idx.shape[0]
        if n_valid > 0 and get_next_bucket_size(n_valid) < bucket:   # src/queries.py:170-173
            bucket = get_next_bucket_size(n_valid)
    if track_near:
        return out_t, out_hit, out_count, N_evals, near
    return out_t, out_hit, out_count, N_evals
