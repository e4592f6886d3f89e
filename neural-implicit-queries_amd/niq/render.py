"""Camera rays, normals, shading and images (src/render.py:17-253).

Camera rays are formed on the GPU (niq_camera_rays) in the reference's
float32 operations, or on the host when asked for host arrays.  Hit
positions, finite-difference normals, shading, compositing and the tonemap
run on the gfx950 kernels (niq_hit_positions, niq_outward_normals,
niq_shade_image).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import mlp as _mlp
from . import queries
from .implicit_function import as_device_f32

f32 = np.float32


def _normalize(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True).astype(x.dtype)


def orthogonal_dir(x, remove_dir):
    """src/geometry.py:15-19."""
    remove_dir = _normalize(remove_dir)
    x = x - np.dot(x, remove_dir) * remove_dir
    return _normalize(x)


def look_at(eye_pos, target=None, up_dir='y'):
    """src/render.py:229-242 -> (look_dir, up_dir, left_dir)."""
    eye_pos = np.asarray(eye_pos, f32)
    target = np.zeros(3, f32) if target is None else np.asarray(target, f32)
    if up_dir == 'y':
        up = np.array((0., 1., 0.), f32)
    elif up_dir == 'z':
        up = np.array((0., 0., 1.), f32)
    else:
        up = np.asarray(up_dir, f32)
    look_dir = _normalize(target - eye_pos)
    up = orthogonal_dir(up, look_dir)
    left_dir = np.cross(look_dir, up).astype(f32)
    return look_dir, up, left_dir


def _camera_frame(look_dir, up_dir, fov_deg):
    """The per-camera part of src/render.py:26-50 (float32): look, up made
    orthogonal to look and normalised, left = look x up, tan(fov / 2)."""
    look_dir = np.asarray(look_dir, f32)
    up_dir = np.asarray(up_dir, f32)
    up_dir = up_dir - np.dot(look_dir, up_dir) * look_dir
    up_dir = _normalize(up_dir)
    left_dir = np.cross(look_dir, up_dir).astype(f32)
    tan_half = f32(np.tan(np.deg2rad(f32(fov_deg)) / f32(2)))
    return look_dir, up_dir, left_dir, tan_half


def generate_camera_rays(eye_pos, look_dir, up_dir, res=1024, fov_deg=30., device="cuda"):
    """src/render.py:26-50: (roots, dirs) of shape (res*res, 3); pixel
    p = row * res + col over linspace(-1, 1, res)^2 ('xy' meshgrid).  On a
    device the rays are formed there (niq_camera_rays, the same float32
    operations); with device=None as host arrays."""
    if device is not None:
        _lib.require_gpu()
        lib = _lib.load_library()
        look, up, left, tan_half = _camera_frame(look_dir, up_dir, fov_deg)
        vec = lambda v: (ctypes.c_float * 3)(*[float(x) for x in np.asarray(v, f32).reshape(-1)])
        roots = torch.empty((res * res, 3), dtype=torch.float32, device=device)
        dirs = torch.empty((res * res, 3), dtype=torch.float32, device=device)
        with torch.cuda.device(roots.device):
            _lib.check(lib.niq_camera_rays(vec(eye_pos), vec(look), vec(up), vec(left), int(res), float(tan_half),
                                           _lib.ptr(roots), _lib.ptr(dirs), _lib.stream_handle()))
        return roots, dirs
    ax = np.linspace(-1., 1., res, dtype=f32)
    cam_x, cam_y = np.meshgrid(ax, ax)
    cam_x = cam_x.reshape(-1)
    cam_y = cam_y.reshape(-1)
    look_dir = np.asarray(look_dir, f32)
    up_dir = np.asarray(up_dir, f32)
    up_dir = up_dir - np.dot(look_dir, up_dir) * look_dir
    up_dir = _normalize(up_dir)
    left_dir = np.cross(look_dir, up_dir).astype(f32)
    tan_half = f32(np.tan(np.deg2rad(f32(fov_deg)) / f32(2)))
    pos = (look_dir[None, :] + left_dir[None, :] * (cam_x * tan_half)[:, None]
           + up_dir[None, :] * (cam_y * tan_half)[:, None])
    dirs = _normalize(pos.astype(f32))
    roots = np.tile(np.asarray(eye_pos, f32), (dirs.shape[0], 1))
    return roots, dirs


def camera_ray(look_dir, up_dir, left_dir, fov_deg_x, fov_deg_y, theta_x, theta_y):
    """src/render.py:17-24 (float32, host)."""
    look_dir, up_dir, left_dir = (np.asarray(v, f32) for v in (look_dir, up_dir, left_dir))
    tx = f32(np.tan(np.deg2rad(f32(fov_deg_x)) / f32(2)))
    ty = f32(np.tan(np.deg2rad(f32(fov_deg_y)) / f32(2)))
    pos = look_dir + left_dir * (f32(theta_x) * tx) + up_dir * (f32(theta_y) * ty)
    return _normalize(pos.astype(f32))


def _tuple(x):
    if isinstance(x, list):
        x = tuple(x)
    return x if isinstance(x, tuple) else (x,)


def outward_normals(funcs_tuple, params_tuple, hit_pos, hit_ids, eps, method='finite_differences'):
    """src/render.py:82-90 (vmapped outward_normal, :53-80): (N, 3) unit normals
    of function hit_id at each hit (tetrahedron finite differences, or the
    exact forward-mode gradient for method='autodiff'); zeros where hit_id
    matches none."""
    methods = {'finite_differences': 0, 'autodiff': 1}
    if method not in methods:
        raise ValueError("unrecognized method")
    funcs_tuple, params_tuple = _tuple(funcs_tuple), _tuple(params_tuple)
    if len(funcs_tuple) != len(params_tuple):
        raise ValueError("funcs_tuple and params_tuple lengths differ")
    _lib.require_gpu()
    lib = _lib.load_library()
    pos = as_device_f32(hit_pos).reshape(-1, 3)
    ids = torch.as_tensor(hit_ids).to(device=pos.device, dtype=torch.int32).reshape(-1).contiguous()
    if ids.shape[0] != pos.shape[0]:
        raise ValueError("hit_pos and hit_ids lengths differ")
    models = [_mlp.device_model(p) for p in params_tuple]
    handles = (ctypes.c_void_p * len(models))(*[m.handle.value for m in models])
    out = torch.empty_like(pos)
    _lib.check(lib.niq_outward_normals(handles, len(models), _lib.ptr(pos), _lib.ptr(ids), pos.shape[0],
                                       float(eps), methods[method], _lib.ptr(out), _lib.stream_handle()))
    return out


def outward_normal(funcs_tuple, params_tuple, hit_pos, hit_id, eps, method='finite_differences'):
    """src/render.py:53-80 for one hit."""
    return outward_normals(funcs_tuple, params_tuple, as_device_f32(hit_pos).reshape(1, 3),
                           torch.as_tensor([int(hit_id)]), eps, method)[0]


def _shade(shading, ray_dirs, hit_pos, hit_normals, hit_ids, up_dir, matcaps, shading_color_tuple,
           shading_color_func, composite, tonemap):
    if shading not in ("normal", "matcap_color"):
        raise RuntimeError("Unrecognized shading parameter")
    _lib.require_gpu()
    lib = _lib.load_library()
    nrm = as_device_f32(hit_normals).reshape(-1, 3)
    ids = torch.as_tensor(hit_ids).to(device=nrm.device, dtype=torch.int32).reshape(-1).contiguous()
    n = nrm.shape[0]
    if ids.shape[0] != n:
        raise ValueError("hit_ids and hit_normals lengths differ")
    o = _lib.ShadeOpts()
    o.shading = 0 if shading == "normal" else 1
    keep = []
    dirs = None
    if shading == "matcap_color":
        dirs = as_device_f32(ray_dirs).reshape(-1, 3)
        if dirs.shape[0] != n:
            raise ValueError("ray_dirs and hit_normals lengths differ")
        up = np.asarray(up_dir, f32).reshape(3)
        for k in range(3):
            o.up_dir[k] = float(up[k])
        if matcaps is None or len(matcaps) != 4:
            raise ValueError("matcap shading needs the 4 matcaps (r, g, b, k)")
        mats = [as_device_f32(m) for m in matcaps]
        if any(m.dim() != 3 or m.shape != mats[0].shape or m.shape[2] != 3 for m in mats):
            raise ValueError("matcaps must be 4 equal (h, w, 3) images")
        keep += mats
        for i in range(4):
            o.matcaps[i] = mats[i].data_ptr()
        o.matcap_h, o.matcap_w = int(mats[0].shape[0]), int(mats[0].shape[1])
        if shading_color_func is not None:
            col = as_device_f32(shading_color_func(as_device_f32(hit_pos).reshape(-1, 3))).reshape(-1, 3)
            if col.shape[0] != n:
                raise ValueError("shading_color_func must return one colour per hit")
            keep.append(col)
            o.color_per_hit = col.data_ptr()
        else:
            cols = shading_color_tuple
            if not isinstance(cols[0], (tuple, list)):
                cols = (cols,)
            if len(cols) > 8:
                raise ValueError("at most 8 shading colours")
            for i, c in enumerate(cols):
                for k in range(3):
                    o.colors[i][k] = float(c[k])
            o.n_colors = len(cols)
    o.composite = int(composite)
    o.tonemap = int(tonemap)
    o.gamma, o.white_level, o.exposure = 2.2, 0.75, 1.0
    out = torch.empty((n, 3), dtype=torch.float32, device=nrm.device)
    _lib.check(lib.niq_shade_image(ctypes.byref(o), _lib.ptr(dirs), _lib.ptr(nrm), _lib.ptr(ids), n,
                                   _lib.ptr(out), _lib.stream_handle()))
    # the temporaries in `keep` may be freed now: the caching allocator reuses
    # their memory only for work queued after this kernel on the same stream
    del keep
    return out


def shade_image(shading, ray_dirs, hit_pos, hit_normals, hit_ids, up_dir, matcaps, shading_color_tuple,
                shading_color_func=None):
    """src/render.py:160-224: (N, 3) colours ('normal' or 'matcap_color')."""
    return _shade(shading, ray_dirs, hit_pos, hit_normals, hit_ids, up_dir, matcaps, shading_color_tuple,
                  shading_color_func, composite=False, tonemap=False)


def tonemap_image(img, gamma=2.2, white_level=.75, exposure=1.):
    """src/render.py:152-158 (elementwise, on the tensor's device)."""
    img = as_device_f32(img) * exposure
    num = img * (1.0 + (img / (white_level * white_level)))
    den = (1.0 + img)
    return torch.pow(num / den, 1.0 / gamma)


def render_image(funcs_tuple, params_tuple, eye_pos, look_dir, up_dir, left_dir, res, fov_deg, frustum, opts,
                 shading="normal", shading_color_tuple=((0.157, 0.613, 1.000)), matcaps=None, tonemap=False,
                 shading_color_func=None):
    """src/render.py:94-150: (img (res, res, 3), depth, counts, hit_ids,
    n_eval, -1), with frustum or per-ray casting."""
    funcs_tuple, params_tuple = _tuple(funcs_tuple), _tuple(params_tuple)
    if isinstance(shading_color_tuple, list):
        shading_color_tuple = tuple(shading_color_tuple)
    if not isinstance(shading_color_tuple[0], tuple):
        shading_color_tuple = (shading_color_tuple,)
    L = len(funcs_tuple)
    if len(params_tuple) != L or len(shading_color_tuple) != L:
        raise ValueError("render_image tuple arguments should all be same length")
    _lib.require_gpu()
    lib = _lib.load_library()
    ray_roots, ray_dirs = generate_camera_rays(eye_pos, look_dir, up_dir, res=res, fov_deg=fov_deg)
    if frustum:
        cam_params = eye_pos, look_dir, up_dir, left_dir, fov_deg, fov_deg, res, res
        t_raycast, hit_ids, counts, n_eval = queries.cast_rays_frustum(funcs_tuple, params_tuple, cam_params, opts)
        # outputs are indexed [x, y]; pixels are numbered y * res + x
        t_raycast = t_raycast.t().contiguous().reshape(-1)
        hit_ids = hit_ids.t().contiguous().reshape(-1)
        counts = counts.t().contiguous().reshape(-1)
    else:
        t_raycast, hit_ids, counts, n_eval = queries.cast_rays(funcs_tuple, params_tuple, ray_roots, ray_dirs, opts)
    n = ray_roots.shape[0]
    hit_pos = torch.empty((n, 3), dtype=torch.float32, device=ray_roots.device)
    _lib.check(lib.niq_hit_positions(_lib.ptr(ray_roots), _lib.ptr(ray_dirs), _lib.ptr(t_raycast), n,
                                     _lib.ptr(hit_pos), _lib.stream_handle()))
    hit_normals = outward_normals(funcs_tuple, params_tuple, hit_pos, hit_ids, opts['hit_eps'])
    img = _shade(shading, ray_dirs, hit_pos, hit_normals, hit_ids, up_dir, matcaps, shading_color_tuple,
                 shading_color_func, composite=True, tonemap=tonemap)
    return (img.reshape(res, res, 3), t_raycast.reshape(res, res), counts.reshape(res, res),
            hit_ids.reshape(res, res), n_eval, -1)


def load_matcap(fname_pattern):
    """src/render.py:245-253: the r, g, b, k images, / 256."""
    from PIL import Image
    imgs = []
    for c in ['r', 'g', 'b', 'k']:
        im = np.asarray(Image.open(fname_pattern.format(c)), dtype=np.float32) / f32(256.)
        imgs.append(im.astype(f32))
    return tuple(imgs)
