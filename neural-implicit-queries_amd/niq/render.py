"""Camera rays for ``cast_rays`` (src/render.py:17-50, 229-242).

Only ray generation is kept; shading, normals and image output are outside
the range-analysis path.  Rays are formed in float32 on the host exactly as
the reference forms them, then moved to the GPU.
"""
import numpy as np
import torch

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


def generate_camera_rays(eye_pos, look_dir, up_dir, res=1024, fov_deg=30., device="cuda"):
    """src/render.py:26-50: (roots, dirs) of shape (res*res, 3); pixel
    p = row * res + col over linspace(-1, 1, res)^2 ('xy' meshgrid)."""
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
    if device is None:
        return roots, dirs
    return torch.as_tensor(roots, device=device), torch.as_tensor(dirs, device=device)
