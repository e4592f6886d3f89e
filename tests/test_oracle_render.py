"""Pins for the oracle's normals and shading (CPU).

Normals: a plane's normal exactly; a smooth net's float64 gradient
direction.  Matcap lookup: scipy.ndimage.map_coordinates (order=1,
mode='nearest') on random images and coordinates inside and outside.
Matcap coordinates: the camera-facing normal samples the image centre, the
ray-up normal the top edge.  Colour mix: constant matcaps give
c_r R + c_g G + c_b B + (1 - c_r - c_g - c_b) K.  Tonemap: identity at
gamma = exposure = white_level = 1, a pure power law in gamma.
"""
import numpy as np
from scipy import ndimage

from oracle import funcs as ofuncs, render as orender
from niq import models
from test_oracle_queries import plane_mlp

f32 = np.float32


def test_normals_of_planes():
    n1 = np.array([0.36, -0.48, 0.8])
    n2 = np.array([0.0, 0.6, -0.8])
    f1 = ofuncs.OracleFunc(plane_mlp(n1, 0.1))
    f2 = ofuncs.OracleFunc(plane_mlp(n2, -0.2))
    rng = np.random.default_rng(0)
    pos = rng.uniform(-1, 1, (300, 3)).astype(f32)
    ids = rng.integers(0, 3, 300)
    nr = orender.outward_normals((f1, f2), pos, ids, 1e-3)
    assert np.allclose(nr[ids == 1], n1, atol=1e-3)
    assert np.allclose(nr[ids == 2], n2, atol=1e-3)
    assert np.all(nr[ids == 0] == 0)


def test_normals_follow_the_gradient():
    p = models.random_mlp([32, 32], activation="tanh", seed=6)
    f = ofuncs.OracleFunc(p)
    f64 = ofuncs.OracleFunc(p, dtype=np.float64)
    rng = np.random.default_rng(1)
    pos = rng.uniform(-0.8, 0.8, (200, 3)).astype(f32)
    nr = orender.outward_normals((f,), pos, np.ones(200, int), 1e-3)
    h = 1e-6
    g = np.stack([(f64(pos + h * e) - f64(pos - h * e)).reshape(-1) / (2 * h) for e in np.eye(3)], 1)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    cos = np.sum(g * nr, axis=1)
    assert np.all(cos > 0.999)


def test_matcap_lookup_is_scipy_map_coordinates():
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, (37, 29)).astype(f32)
    x = rng.uniform(-3, 40, 2000).astype(f32)
    y = rng.uniform(-3, 32, 2000).astype(f32)
    got = orender.map_coordinates_linear_nearest(img, x, y)
    want = ndimage.map_coordinates(img.astype(np.float64), np.stack([x, y]).astype(np.float64), order=1,
                                   mode="nearest")
    assert np.allclose(got, want, atol=2e-6)


def test_matcap_coordinates_and_colour_mix():
    H, W = 64, 48
    up = np.array([0.0, 1.0, 0.0], f32)
    dirs = orender.normalize(np.array([[0.1, 0.2, -1.0], [0.0, 0.0, -1.0]], f32))
    # ramp matcaps: channel value = row index / H, so the lookup returns the x coordinate
    ramp = np.repeat((np.arange(H, dtype=f32) / H)[:, None, None], W, 1).repeat(3, 2)
    ramp_y = np.repeat((np.arange(W, dtype=f32) / W)[None, :, None], H, 0).repeat(3, 2)
    facing = -dirs
    cx = orender.shade_image("matcap_color", dirs, None, facing, np.ones(2), up, (ramp,) * 4, ((1.0, 0.0, 0.0),))
    cy = orender.shade_image("matcap_color", dirs, None, facing, np.ones(2), up, (ramp_y,) * 4, ((1.0, 0.0, 0.0),))
    assert np.allclose(cx, (H / 2) / H, atol=1e-5) and np.allclose(cy, (W / 2) / W, atol=1e-5)
    ray_up = orender.orthogonal_dir(up, dirs)
    cy2 = orender.shade_image("matcap_color", dirs, None, ray_up, np.ones(2), up, (ramp_y,) * 4, ((1.0, 0.0, 0.0),))
    assert np.allclose(cy2, (1 - 0.98) / 2, atol=1e-5)
    # constant matcaps: the colour mix
    R, G, B, K = (np.full((H, W, 3), v, f32) for v in (0.9, 0.5, 0.2, 0.05))
    c = (0.157, 0.613, 0.1)
    col = orender.shade_image("matcap_color", dirs, None, facing, np.ones(2), up, (R, G, B, K), (c,))
    want = c[0] * 0.9 + c[1] * 0.5 + c[2] * 0.2 + (1 - sum(c)) * 0.05
    assert np.allclose(col, want, atol=1e-6)
    # a hit_id beyond the colours uses white (c_k = -2)
    col2 = orender.shade_image("matcap_color", dirs, None, facing, np.full(2, 5), up, (R, G, B, K), (c,))
    assert np.allclose(col2, 0.9 + 0.5 + 0.2 - 2 * 0.05, atol=1e-6)


def test_normal_shading_and_tonemap():
    nr = np.array([[1.0, 0.0, -1.0]], f32)
    assert np.allclose(orender.shade_image("normal", None, None, nr, np.ones(1), None, None, None), [[1, 0.5, 0]])
    x = np.linspace(0, 3, 50).astype(f32)
    assert np.allclose(orender.tonemap_image(x, gamma=1.0, white_level=1.0, exposure=1.0), x, rtol=1e-6)
    assert np.allclose(orender.tonemap_image(x, gamma=2.2, white_level=1.0, exposure=1.0), x ** (1 / 2.2),
                       rtol=1e-5)
    img = orender.composite(np.zeros((3, 3), f32), np.array([0, 1, 2]))
    assert np.array_equal(img, [[1, 1, 1], [0, 0, 0], [0, 0, 0]])


def test_autodiff_normals_match_float64_gradient():
    """Forward-mode gradients against float64 central differences (smooth net)
    and a plane's normal exactly; relu's derivative at a kink is jax's 0."""
    from oracle import mlp as omlp
    p = models.random_mlp([32, 32], activation="tanh", seed=6)
    f = ofuncs.OracleFunc(p)
    f64 = ofuncs.OracleFunc(p, dtype=np.float64)
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.8, 0.8, (200, 3)).astype(f32)
    nr = orender.outward_normals_autodiff((f,), pos, np.ones(200, int))
    h = 1e-6
    g = np.stack([(f64(pos + h * e) - f64(pos - h * e)).reshape(-1) / (2 * h) for e in np.eye(3)], 1)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    assert np.all(np.sum(g * nr, axis=1) > 0.99999)
    n1 = np.array([0.36, -0.48, 0.8])
    pl = ofuncs.OracleFunc(plane_mlp(n1, 0.1))
    assert np.allclose(orender.outward_normals_autodiff((pl,), pos, np.ones(200, int)), n1, atol=1e-6)
    # the value part equals plain evaluation
    v, _ = omlp.eval_points_grad(p, pos)
    assert np.array_equal(v, omlp.eval_points(p, pos))
    x = np.zeros(4, f32)
    assert np.array_equal(omlp.act_deriv("relu", x), np.zeros(4, f32))
    assert np.array_equal(omlp.act_deriv("elu", x), np.ones(4, f32))
