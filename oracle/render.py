"""Oracle: normals and shading of ray-cast hits.

Test infrastructure only (see ``oracle/__init__.py``).

Restates (float32, vectorised over pixels):
  * outward_normal / outward_normals (both methods)       -- src/render.py:53-90
  * shade_image ('normal', 'matcap_color')                -- src/render.py:160-224
  * the compositing of render_image                        -- src/render.py:139-149
  * tonemap_image                                          -- src/render.py:152-158
  * geometry.normalize / orthogonal_dir                    -- src/geometry.py:12-19
The matcap lookup is scipy.ndimage.map_coordinates with order=1 and
mode='nearest' (as jax.scipy.ndimage implements it): along each axis the
lower index floor(x) with weight 1 - (x - floor(x)) and the upper index
floor(x) + 1 with weight x - floor(x), indices clamped to the image, terms
summed in (lower, upper) x (lower, upper) order.
"""
import numpy as np

f32 = np.float32


def normalize(x):
    """src/geometry.py:12-13 over the last axis."""
    x = np.asarray(x, f32)
    return (x / np.sqrt(np.sum(x * x, axis=-1, keepdims=True, dtype=f32)).astype(f32)).astype(f32)


def orthogonal_dir(x, remove_dir):
    """src/geometry.py:15-19: x without its remove_dir component, unit length
    (x (3,), remove_dir (N, 3))."""
    r = normalize(remove_dir)
    d = np.sum(x[None, :] * r, axis=-1, keepdims=True, dtype=f32)
    return normalize((x[None, :] - d * r).astype(f32))


def tet_offsets(eps):
    """The tetrahedron of sample offsets (src/render.py:67-72)."""
    e = f32(eps)
    return np.array([[e, -e, -e], [-e, -e, e], [-e, e, -e], [e, e, e]], f32)


def outward_normals(funcs, hit_pos, hit_ids, eps):
    """src/render.py:53-90 (finite differences): per pixel the normalised
    tetrahedron gradient of function hit_id; zeros where no function matches."""
    hit_pos = np.asarray(hit_pos, f32)
    hit_ids = np.asarray(hit_ids)
    off = tet_offsets(eps)
    out = np.zeros_like(hit_pos)
    for fid, func in enumerate(funcs, start=1):
        sel = hit_ids == fid
        if not sel.any():
            continue
        p = hit_pos[sel]
        s = np.stack([np.asarray(func((p + off[q][None, :]).astype(f32)), f32).reshape(-1) for q in range(4)], 1)
        grad = np.zeros_like(p)
        for q in range(4):                    # sum over the 4 samples in order
            grad = (grad + off[q][None, :] * s[:, q:q + 1]).astype(f32)
        out[sel] = normalize(grad)
    return out


def outward_normals_autodiff(funcs, hit_pos, hit_ids):
    """src/render.py:62-64 (method='autodiff'): the normalised forward-mode
    gradient of function hit_id; zeros where no function matches."""
    from .mlp import eval_points_grad
    hit_pos = np.asarray(hit_pos, f32)
    hit_ids = np.asarray(hit_ids)
    out = np.zeros_like(hit_pos)
    for fid, func in enumerate(funcs, start=1):
        sel = hit_ids == fid
        if sel.any():
            _, g = eval_points_grad(func.params, hit_pos[sel], f32)
            out[sel] = normalize(g)
    return out


def map_coordinates_linear_nearest(img, x, y):
    """Bilinear lookup of img (H, W) at (x, y) along axes (0, 1), indices
    clamped; float32."""
    H, W = img.shape
    x = np.asarray(x, f32)
    y = np.asarray(y, f32)
    x0 = np.floor(x)
    y0 = np.floor(y)
    wx1 = (x - x0).astype(f32)
    wy1 = (y - y0).astype(f32)
    wx0 = (f32(1) - wx1).astype(f32)
    wy0 = (f32(1) - wy1).astype(f32)
    xi = [np.clip(x0.astype(np.int64), 0, H - 1), np.clip(x0.astype(np.int64) + 1, 0, H - 1)]
    yi = [np.clip(y0.astype(np.int64), 0, W - 1), np.clip(y0.astype(np.int64) + 1, 0, W - 1)]
    wx = [wx0, wx1]
    wy = [wy0, wy1]
    acc = None
    for a in range(2):
        for b in range(2):
            term = ((wx[a] * wy[b]).astype(f32) * img[xi[a], yi[b]]).astype(f32)
            acc = term if acc is None else (acc + term).astype(f32)
    return acc


def sample_matcap(matcap, x, y):
    """One matcap image (H, W, C) at the coordinates: (N, C)."""
    return np.stack([map_coordinates_linear_nearest(matcap[:, :, c], x, y) for c in range(matcap.shape[2])], -1)


def shade_image(shading, ray_dirs, hit_pos, normals, hit_ids, up_dir, matcaps, colors):
    """src/render.py:160-224 with constant per-function colours."""
    normals = np.asarray(normals, f32)
    if shading == "normal":
        return ((normals + f32(1)) / f32(2)).astype(f32)
    if shading != "matcap_color":
        raise RuntimeError("Unrecognized shading parameter")
    ray_dirs = np.asarray(ray_dirs, f32)
    ray_up = orthogonal_dir(np.asarray(up_dir, f32), ray_dirs)
    ray_left = np.cross(ray_dirs, ray_up).astype(f32)
    u = (np.sum(-ray_left * normals, axis=-1, dtype=f32) * f32(0.98)).astype(f32)
    v = (np.sum(ray_up * normals, axis=-1, dtype=f32) * f32(0.98)).astype(f32)
    H, W = matcaps[0].shape[0], matcaps[0].shape[1]
    mx = ((u + f32(1)) / f32(2) * f32(H)).astype(f32)
    my = ((-v + f32(1)) / f32(2) * f32(W)).astype(f32)
    mats = [sample_matcap(np.asarray(m, f32), mx, my) for m in matcaps]
    col = np.ones((normals.shape[0], 3), f32)
    for fid, c in enumerate(colors, start=1):
        col = np.where((np.asarray(hit_ids) == fid)[:, None], np.asarray(c, f32)[None, :], col)
    cr, cg, cb = col[:, 0:1], col[:, 1:2], col[:, 2:3]
    ck = (f32(1) - (cr + cb + cg)).astype(f32)
    out = (cr * mats[0] + cb * mats[2]).astype(f32)
    out = (out + cg * mats[1]).astype(f32)
    return (out + ck * mats[3]).astype(f32)


def tonemap_image(img, gamma=2.2, white_level=0.75, exposure=1.0):
    """src/render.py:152-158."""
    img = (np.asarray(img, f32) * f32(exposure)).astype(f32)
    num = (img * (f32(1) + img / f32(white_level * white_level))).astype(f32)
    den = (f32(1) + img).astype(f32)
    return np.power((num / den).astype(f32), f32(1.0 / gamma)).astype(f32)


def composite(hit_color, hit_ids, tonemap=False):
    """src/render.py:139-149: white where nothing was hit, then the tonemap."""
    img = np.where((np.asarray(hit_ids) != 0)[:, None], hit_color, f32(1)).astype(f32)
    return tonemap_image(img) if tonemap else img
