"""Pins for the oracle's sampling queries (CPU).

Philox4x32-10: the Random123 known-answer vectors.  Candidate stream: every
candidate lies in its node, nodes are drawn uniformly.  Surface sampling:
the round size does not change the samples (one-at-a-time rejection loop as
the brute force), samples of a plane lie in its |f| < width slab and are
uniform there.  Bulk properties: closed-form volume and centroid of
half-spaces cut from the cube (axis-aligned and oblique, the latter through
the Irwin-Hall distribution).
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import funcs as ofuncs, sampling as osamp
from test_oracle_queries import plane_mlp

LO = np.array([-1., -1., -1.], np.float32)
UP = np.array([1., 1., 1.], np.float32)


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        got = osamp.philox4x32_10([np.array([c], np.uint32) for c in ctr], key)
        assert tuple(int(w[0]) for w in got) == want


def test_candidates_lie_in_uniformly_drawn_nodes():
    rng = np.random.default_rng(3)
    n_nodes = 7
    lo = rng.uniform(-1, 0, (n_nodes, 3)).astype(np.float32)
    up = (lo + rng.uniform(0.01, 1, (n_nodes, 3))).astype(np.float32)
    pos, node = osamp.candidates(1234, osamp.TAG_SURFACE, 5, 70000, lo, up, n_nodes)
    assert np.all(pos >= lo[node]) and np.all(pos <= up[node])
    counts = np.bincount(node, minlength=n_nodes)
    expect = 70000 / n_nodes
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    assert chi2 < 30.0            # 6 dof: p < 1e-4 beyond this
    # uniform within the node: the normalised coordinate has mean 1/2, variance 1/12
    u = (pos - lo[node]) / (up[node] - lo[node])
    assert np.all(np.abs(u.mean(0) - 0.5) < 0.01)
    assert np.all(np.abs(u.var(0) - 1 / 12) < 0.005)
    # a different tag or key is a different stream
    pos2, _ = osamp.candidates(1234, osamp.TAG_MASS, 5, 1000, lo, up, n_nodes)
    pos3, _ = osamp.candidates(1235, osamp.TAG_SURFACE, 5, 1000, lo, up, n_nodes)
    assert not np.array_equal(pos[:1000], pos2) and not np.array_equal(pos[:1000], pos3)
    # candidates are addressed by index: an offset window is a slice
    pos4, _ = osamp.candidates(1234, osamp.TAG_SURFACE, 105, 50, lo, up, n_nodes)
    assert np.array_equal(pos4, pos[100:150])


def one_at_a_time(func, lo, up, n_nodes, n_samples, width, seed):
    """Brute force: test candidates one by one and keep the first n_samples."""
    out = []
    j = 0
    while len(out) < n_samples:
        p, _ = osamp.candidates(seed, osamp.TAG_SURFACE, j, 1, lo, up, n_nodes)
        if abs(float(func(p)[0])) < width:
            out.append(p[0])
        j += 1
    return np.array(out, np.float32)


def test_round_size_does_not_change_samples():
    f = ofuncs.OracleFunc(plane_mlp([0.6, -0.8, 0.0], 0.1))
    lo, up = LO[None, :], UP[None, :]
    want = one_at_a_time(f, lo, up, 1, 40, 0.05, 77)
    for per_round in (7, 64, 1000):
        got, rounds, _ = osamp.sample_nodes(f, lo, up, 1, 40, 0.05, 77, per_round)
        assert np.array_equal(got, want), per_round
        assert rounds >= 1


def test_surface_samples_fill_the_plane_slab():
    c, w = 0.3, 0.02
    f = ofuncs.OracleFunc(plane_mlp([1.0, 0.0, 0.0], c))
    pts = osamp.sample_surface(f, LO, UP, 4000, w, 11, n_node_thresh=512)
    assert pts.shape == (4000, 3)
    assert np.all(np.abs(pts[:, 0] - c) < w + 1e-6)
    # uniform in the slab: x uniform on (c - w, c + w), y and z on (-1, 1)
    assert abs(pts[:, 0].mean() - c) < 0.05 * w
    assert abs(pts[:, 0].var() - (2 * w) ** 2 / 12) < 0.1 * (2 * w) ** 2 / 12
    assert np.all(np.abs(pts[:, 1:].mean(0)) < 0.05)
    assert np.all(np.abs(pts[:, 1:].var(0) - 1 / 3) < 0.03)
    # the whole-box sampler draws from the same slab
    pu = osamp.sample_surface_uniform(f, LO, UP, 1000, w, 12)
    assert np.all(np.abs(pu[:, 0] - c) < w + 1e-6)


def test_bulk_properties_of_an_axis_half_space():
    f = ofuncs.OracleFunc(plane_mlp([1.0, 0.0, 0.0], 0.3))
    mass, cen = osamp.bulk_properties(f, LO, UP, 5, n_expand=2000, n_sample=200000)
    assert abs(mass - 5.2) < 5e-3
    assert np.allclose(cen, [-0.35, 0.0, 0.0], atol=2e-3)


def irwin_hall_cut(s):
    """Volume and centroid of {x in [-1,1]^3 : x+y+z < s}: with x = 2u - 1 the
    region is U1+U2+U3 < t = (s+3)/2 for uniform U on [0,1]^3."""
    t = (s + 3.0) / 2.0

    def dens(x):   # Irwin-Hall density, n = 3
        if x < 0 or x > 3:
            return 0.0
        if x < 1:
            return x * x / 2
        if x < 2:
            return (-2 * x * x + 6 * x - 3) / 2
        return (3 - x) ** 2 / 2

    p = integrate.quad(dens, 0, t, points=[1, 2])[0]
    es = integrate.quad(lambda x: x * dens(x), 0, t, points=[1, 2])[0]
    eu = es / 3 / p                       # E[U1 | S < t] by symmetry
    return 8 * p, 2 * eu - 1


def test_bulk_properties_of_an_oblique_half_space():
    n = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    c = 0.2
    f = ofuncs.OracleFunc(plane_mlp(n, c))
    vol, cx = irwin_hall_cut(c * np.sqrt(3.0))
    mass, cen = osamp.bulk_properties(f, LO, UP, 9, n_expand=4000, n_sample=200000)
    assert abs(mass - vol) < 5e-3 * vol
    assert np.allclose(cen, [cx] * 3, atol=3e-3)


def test_bulk_sums_with_no_unknown_nodes():
    f = ofuncs.OracleFunc(plane_mlp([1.0, 0.0, 0.0], 0.3))
    il = np.array([[-1, -1, -1]], np.float32)
    iu = np.array([[0, 1, 1]], np.float32)
    m_int, c_int, n_in, mom = osamp.bulk_sums(f, np.zeros((0, 3)), np.zeros((0, 3)), 0, il, iu, 1, 1000, 1)
    assert m_int == 4.0 and np.allclose(c_int, [-2.0, 0, 0]) and n_in == 0
