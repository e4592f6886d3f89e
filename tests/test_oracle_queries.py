"""Pins for the oracle's queries (CPU): tree, rays, closest point, MC.

Tree: brute-force ancestor walk.  Rays: closed-form plane hits and no sign
change before the hit.  Closest point: plane distance within the reference's
node-centre tolerance.  MC: table invariants and the closed-form area of a
plane cut.
"""
import numpy as np
import pytest

from oracle import funcs as ofuncs, tree as otree, rays as orays, closest as ocp, mc as omc
from oracle import affine as oaff, mlp as omlp
from niq import mlp as nmlp, models

LO = np.array([-1., -1., -1.], np.float32)
UP = np.array([1., 1., 1.], np.float32)


def plane_mlp(n, c):
    """f(x) = n.x - c as an activation-free MLP (through one hidden layer)."""
    n = np.asarray(n, np.float32)
    A1 = np.concatenate([np.eye(3, dtype=np.float32), -np.eye(3, dtype=np.float32)], 1)
    A2 = np.concatenate([0.5 * n, -0.5 * n])[:, None].astype(np.float32)
    ops = [nmlp.dense(3, 6, A=A1, b=np.zeros(6)), nmlp.dense(6, 1, A=A2, b=[-c]), nmlp.squeeze_last()]
    return nmlp.build_spec(ops)


# ----------------------------------------------------------------- tree
def brute_force_tree(func, depth):
    """Walk every node of the full dyadic tree; a node is visited iff all of
    its ancestors classify UNKNOWN (src/kd_tree.py:61-96)."""
    unknown, interior, exterior = [], [], []
    level = [(LO.copy(), UP.copy())]
    for L in range(depth + 1):
        lo = np.array([a for a, _ in level], np.float32).reshape(-1, 3)
        up = np.array([b for _, b in level], np.float32).reshape(-1, 3)
        if lo.shape[0] == 0:
            break
        codes = func.classify_box(lo, up)
        interior += [tuple(np.r_[a, b]) for a, b, c in zip(lo, up, codes) if c == oaff.SIGN_NEGATIVE]
        exterior += [tuple(np.r_[a, b]) for a, b, c in zip(lo, up, codes) if c == oaff.SIGN_POSITIVE]
        keep = codes == oaff.SIGN_UNKNOWN
        if L == depth:
            unknown = [tuple(np.r_[a, b]) for a, b in zip(lo[keep], up[keep])]
            break
        nxt = []
        for a, b in zip(lo[keep], up[keep]):
            dim = int(np.argmax(b - a))
            mid = np.float32(0.5) * (a + b)
            a2, b1 = a.copy(), b.copy()
            b1[dim] = mid[dim]
            a2[dim] = mid[dim]
            nxt += [(a, b1), (a2, b)]
        level = nxt
    return set(unknown), set(interior), set(exterior)


def rows(lo, up, valid):
    return {tuple(np.r_[a, b]) for a, b in zip(lo[valid], up[valid])}


@pytest.mark.parametrize("mode", ["affine_fixed", "interval"])
def test_tree_matches_brute_force(mode):
    p = models.config_model("c1_relu_3x32")
    f = ofuncs.OracleFunc(p, mode)
    depth = 9
    out = otree.construct_uniform_unknown_levelset_tree(
        f, LO, UP, split_depth=depth, with_interior_nodes=True, with_exterior_nodes=True,
        batch_process_size=64)
    u, i, e = brute_force_tree(f, depth)
    assert rows(out["unknown_node_lower"], out["unknown_node_upper"], out["unknown_node_valid"]) == u
    assert rows(out["interior_node_lower"], out["interior_node_upper"], out["interior_node_valid"]) == i
    assert rows(out["exterior_node_lower"], out["exterior_node_upper"], out["exterior_node_valid"]) == e
    assert out["stats"]["n_unknown"] == len(u) > 0


def test_tree_order_and_padding():
    """A-children of a batch precede its B-children; arrays are bucketed."""
    p = models.config_model("c1_relu_3x32")
    f = ofuncs.OracleFunc(p, "affine_fixed")
    out = otree.construct_uniform_unknown_levelset_tree(f, LO, UP, split_depth=3)
    n = out["stats"]["n_unknown"]
    assert out["unknown_node_valid"].shape[0] == otree.get_next_bucket_size(n)
    assert out["unknown_node_valid"].sum() == n
    # depth 3 from a cube splits x, y, z once each: every leaf is an octant-sized cell
    w = out["unknown_node_upper"][:n] - out["unknown_node_lower"][:n]
    np.testing.assert_array_equal(w, np.ones_like(w))
    # the last split was along z: the first half of the list are A (lower z) children
    lz = out["unknown_node_lower"][:n, 2]
    assert np.all(lz[: n // 2] == -1.0) and np.all(lz[n // 2:] == 0.0)


def test_bucket_sizes():
    assert otree.get_next_bucket_size(1) == 128
    assert otree.get_next_bucket_size(129) == 256
    assert otree.get_next_bucket_size(2 ** 20) == 2 ** 20
    with pytest.raises(ValueError):
        otree.get_next_bucket_size(2 ** 30 + 1)
    # interior buffer doubling from B while free space < current level size
    assert otree.finished_capacity_sizes([1, 2, 4, 3000], [0, 0, 0, 1500], 2048) == 8192


# ----------------------------------------------------------------- rays
def test_rays_hit_plane_in_closed_form():
    nrm = np.array([0.2, 0.3, 1.0], np.float32)
    nrm /= np.linalg.norm(nrm)
    c = 0.25
    p = plane_mlp(nrm, c)
    f = ofuncs.OracleFunc(p, "affine_fixed")
    eye = np.array([0.1, -0.2, 3.0], np.float32)
    look, up, _ = orays.look_at(eye)
    roots, dirs = orays.generate_camera_rays(eye, look, up, res=24, fov_deg=60.)
    opts = orays.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, hid, cnt, ne = orays.cast_rays((f,), roots, dirs, opts)
    t_star = (c - roots @ nrm) / (dirs @ nrm)
    hit = (t_star > 0) & (t_star < 6.0)
    assert np.all(hid[hit] == 1)
    # a hit is reported at t with f(t) and f(t + eps) of different sign
    assert np.all(np.abs(t[hit] - t_star[hit]) <= 1.01e-3)
    assert np.all(hid[~hit] == 0) and np.all(t[~hit] > 6.0)
    assert ne >= roots.shape[0]


def test_rays_never_skip_a_sign_change():
    p = models.config_model("c1_relu_3x32")
    f = ofuncs.OracleFunc(p, "affine_fixed")
    eye = np.array([0., 0., 3.], np.float32)
    look, up, _ = orays.look_at(eye)
    roots, dirs = orays.generate_camera_rays(eye, look, up, res=16, fov_deg=60.)
    opts = orays.get_default_cast_opts()
    opts["max_dist"] = 6.0
    t, hid, cnt, _ = orays.cast_rays((f,), roots, dirs, opts)
    ts = np.linspace(0, 1, 400)[None, :] * np.minimum(t, 6.0)[:, None]
    pts = roots[:, None, :] + ts[:, :, None] * dirs[:, None, :]
    v = omlp.eval_points(p, pts.reshape(-1, 3)).reshape(pts.shape[:2])
    v0 = omlp.eval_points(p, roots)
    assert np.all(np.sign(v[:, :-1]) == np.sign(v0)[:, None]) or np.all(
        (np.sign(v[:, :-2]) == np.sign(v0)[:, None]).all(axis=1) | (hid > 0))
    assert np.all(cnt <= opts["n_max_step"])


# ----------------------------------------------------------- closest point
def test_closest_point_plane():
    nrm = np.array([0., 0., 1.], np.float32)
    p = plane_mlp(nrm, 0.1)
    f = ofuncs.OracleFunc(p, "affine_fixed")
    rng = np.random.default_rng(0)
    q = rng.uniform(-1, 1, size=(16, 3)).astype(np.float32)
    eps = 0.02
    dist, loc, info = ocp.closest_point(f, LO, UP, q, eps=eps, batch_process_size=64)
    true = np.abs(q[:, 2] - 0.1)
    # the result is the distance to a spanning node's centre plus its diagonal
    assert np.all(dist >= true - 1e-6)
    assert np.all(dist <= true + 4 * eps)
    assert np.all(np.abs(loc[:, 2] - 0.1) < 2 * eps)


# ---------------------------------------------------------------- MC
def test_mc_table_edges_match_sign_changes():
    tab = omc.tri_table()
    for case in range(256):
        neg = [(case >> v) & 1 for v in range(8)]
        cut = {e for e, (a, b) in enumerate(omc.EDGE_VERTS) if neg[a] != neg[b]}
        used = {int(x) for x in tab[case].reshape(-1) if x >= 0}
        assert used == cut, case
    n_tris = (tab[:, :, 0] >= 0).sum(1)
    assert n_tris[0] == 0 and n_tris[255] == 0 and n_tris.max() == 5
    # every cut edge is used by at least one triangle, each triangle has 3 distinct cut edges
    for case in range(256):
        for tri in tab[case]:
            if tri[0] >= 0:
                assert len(set(tri.tolist())) == 3


def test_mc_plane_area():
    p = plane_mlp(np.array([0., 0., 1.], np.float32), 0.3)
    f = ofuncs.OracleFunc(p, "affine_fixed")
    tris, _ = omc.hierarchical_marching_cubes(f, LO, UP, depth=4, n_subcell_depth=2)
    a = np.linalg.norm(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]), axis=-1) / 2
    assert abs(a.sum() - 4.0) < 1e-4
    np.testing.assert_allclose(tris[..., 2], 0.3, atol=1e-5)
