"""Camera rays formed on the GPU (niq_camera_rays) against the host
formation and the oracle's restatement of src/render.py:26-50, bit for bit,
and against the committed golden camera."""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from oracle import rays as orays
from niq import render
from test_gpu_parity import note

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("res,fov,eye,target", [(64, 60.0, (0., 0., 3.), None), (97, 45.0, (0.3, -0.2, 2.5), None),
                                                (1, 30.0, (0., 1., 2.), None), (2, 90.0, (1., 1., 1.), (0.1, 0., 0.)),
                                                (2048, 60.0, (0., 0., 3.), None)])
def test_camera_rays_match_host_and_oracle(res, fov, eye, target):
    note(f"[camera rays] res {res} fov {fov} eye {eye}")
    eye = np.array(eye, np.float32)
    look, up, left = render.look_at(eye, target)
    r_dev, d_dev = render.generate_camera_rays(eye, look, up, res=res, fov_deg=fov)
    r_host, d_host = render.generate_camera_rays(eye, look, up, res=res, fov_deg=fov, device=None)
    olook, oup, _ = orays.look_at(eye, target)
    r_or, d_or = orays.generate_camera_rays(eye, olook, oup, res=res, fov_deg=fov)
    np.testing.assert_array_equal(d_host, d_or)
    np.testing.assert_array_equal(r_host, r_or)
    d_dev, r_dev = d_dev.cpu().numpy(), r_dev.cpu().numpy()
    diff = int((d_dev != d_host).any(axis=1).sum())
    note(f"  {diff} of {res * res} directions differ from the host formation")
    np.testing.assert_array_equal(d_dev, d_host)
    np.testing.assert_array_equal(r_dev, r_host)


def test_camera_rays_match_golden():
    with np.load(os.path.join(GOLDEN, "c1_golden.npz")) as z:
        roots, dirs = z["ray_roots"], z["ray_dirs"]
    eye = np.array([0., 0., 3.], np.float32)
    look, up, _ = render.look_at(eye)
    r, d = render.generate_camera_rays(eye, look, up, res=64, fov_deg=60.0)
    np.testing.assert_array_equal(d.cpu().numpy(), dirs)
    np.testing.assert_array_equal(r.cpu().numpy(), roots)
