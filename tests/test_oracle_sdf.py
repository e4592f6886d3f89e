"""Pins for the oracle's weak-SDF test (src/sdf.py:17-50), CPU.

A unit-normal plane is an exact SDF: a box is determinate exactly when its
circumscribed ball misses the plane (closed form in float64).  On random
ReLU nets, a Lipschitz bound from the product of the layers' spectral norms
makes every determinate box sign-constant under dense sampling.
"""
import numpy as np
import pytest

from oracle import funcs as ofuncs, affine as oaff
from niq import models
from test_oracle_queries import plane_mlp


def test_sdf_plane_closed_form():
    rng = np.random.default_rng(4)
    nrm = np.array([0.36, -0.48, 0.8])
    c0 = 0.1
    f = ofuncs.OracleFunc(plane_mlp(nrm, c0), "sdf")
    c = rng.uniform(-1, 1, (4000, 3)).astype(np.float32)
    h = rng.uniform(0.001, 0.3, (4000, 3)).astype(np.float32)
    lo, up = c - h, c + h
    code = f.classify_box(lo, up)
    cen = 0.5 * (lo.astype(np.float64) + up)
    val = cen @ nrm - c0
    rad = np.linalg.norm(up - cen, axis=1)
    want = np.where(np.abs(val) < rad, oaff.SIGN_UNKNOWN, np.where(val > 0, oaff.SIGN_POSITIVE, oaff.SIGN_NEGATIVE))
    clear = np.abs(np.abs(val) - rad) > 1e-5
    assert np.array_equal(code[clear], want[clear])
    assert {0, 1, 2} <= set(code.tolist())
    # general boxes: radius sqrt(|v1|^2 + |v2|^2); offsets move the sign test only
    vecs = rng.normal(0, 0.1, (4000, 2, 3)).astype(np.float32)
    code = f.classify_general_box(c, vecs, offset=0.05)
    val = c.astype(np.float64) @ nrm - c0
    rad = np.sqrt((np.linalg.norm(vecs.astype(np.float64), axis=2) ** 2).sum(1))
    det = np.abs(val) >= rad
    want = np.where(det & (val > 0.05), 1, np.where(det & (val < -0.05), 2, 0))
    clear = (np.abs(np.abs(val) - rad) > 1e-5) & (np.abs(np.abs(val) - 0.05) > 1e-5)
    assert np.array_equal(code[clear], want[clear])


@pytest.mark.parametrize("seed", [1, 2])
def test_sdf_is_sound_with_a_lipschitz_bound(seed):
    p = models.random_mlp([32, 32], seed=seed)
    lip = 1.0
    for i in range(3):
        A = np.asarray(p[f"{2 * i:04d}.dense.A"], np.float64)
        lip *= np.linalg.norm(A, 2)
    f = ofuncs.OracleFunc(p, "sdf", sdf_lipschitz=lip)
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (600, 3)).astype(np.float32)
    h = rng.uniform(0.0005, 0.01, (600, 3)).astype(np.float32)
    code = f.classify_box(c - h, c + h)
    det = code != oaff.SIGN_UNKNOWN
    assert det.sum() > 50
    u = rng.uniform(-1, 1, (600, 64, 3)).astype(np.float32)
    pts = (c[:, None, :] + u * h[:, None, :]).reshape(-1, 3)
    v = f(pts).reshape(600, 64)
    assert np.all(v[code == oaff.SIGN_POSITIVE] > 0)
    assert np.all(v[code == oaff.SIGN_NEGATIVE] < 0)
