"""Pins for the oracle's queries (CPU): tree, rays, closest point, MC.

Tree: brute-force ancestor walk.  Rays: closed-form plane hits and no sign
change before the hit.  Closest point: plane distance within the reference's
node-centre tolerance.  MC: table invariants and the closed-form area of a
plane cut.
"""
import numpy as np
import pytest

from oracle import funcs as ofuncs, tree as otree, rays as orays, closest as ocp, mc as omc
from oracle import frustum as ofr, intersect as oisect
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


@pytest.mark.parametrize("mode", ["affine_fixed", "interval", "sdf"])
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


# ------------------------------------------------------------- frustum rays
def pixel_rays(eye, look, up, left, fov, res):
    """The sample ray of every pixel [x, y] of the frustum caster: a
    single-pixel frustum's corners coincide at 2 x / (res + 1) - 1."""
    xs, ys = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    f32 = np.float32
    tx = (f32(2) * xs.reshape(-1).astype(f32) / f32(res + 1) - f32(1)).astype(f32)
    ty = (f32(2) * ys.reshape(-1).astype(f32) / f32(res + 1) - f32(1)).astype(f32)
    tan = f32(np.tan(np.deg2rad(f32(fov)) / f32(2)))
    dirs = ofr.camera_ray(look, up, left, tan, tan, tx, ty)
    return np.tile(eye, (dirs.shape[0], 1)), dirs


def test_frustum_hits_plane_in_closed_form():
    nrm = np.array([0.2, 0.3, 1.0], np.float32)
    nrm /= np.linalg.norm(nrm)
    c = 0.25
    f = ofuncs.OracleFunc(plane_mlp(nrm, c), "affine_fixed")
    eye = np.array([0.1, -0.2, 3.0], np.float32)
    look, up, left = orays.look_at(eye)
    opts = orays.get_default_cast_opts()
    opts["max_dist"] = 6.0
    opts["n_side_init"] = 8
    res = 32
    t, hid, cnt, ne = ofr.cast_rays_frustum((f,), eye, look, up, left, 60.0, 60.0, res, res, opts)
    roots, dirs = pixel_rays(eye, look, up, left, 60.0, res)
    t_star = ((c - roots @ nrm) / (dirs @ nrm)).reshape(res, res)
    hit = (t_star > 0) & (t_star < 6.0)
    assert np.all(hid[hit] == 1) and np.all(hid[~hit] == 0)
    assert np.all(np.abs(t[hit] - t_star[hit]) <= 1.01e-3)
    assert ne >= 64


def test_frustum_matches_single_rays():
    """Every pixel's frustum result equals the certified march of its own
    sample ray (the ray oracle, pinned above): same hit function, t within
    hit_eps (both stop inside the first sign change's hit_eps window)."""
    p = models.config_model("c1_relu_3x32")
    f = ofuncs.OracleFunc(p, "affine_fixed")
    eye = np.array([0., 0., 3.], np.float32)
    look, up, left = orays.look_at(eye)
    opts = orays.get_default_cast_opts()
    opts["max_dist"] = 6.0
    opts["n_side_init"] = 8
    res = 32
    t, hid, cnt, ne = ofr.cast_rays_frustum((f,), eye, look, up, left, 60.0, 60.0, res, res, opts)
    roots, dirs = pixel_rays(eye, look, up, left, 60.0, res)
    rt, rh, _, _ = orays.cast_rays((f,), roots, dirs, opts)
    rt, rh = rt.reshape(res, res), rh.reshape(res, res)
    assert np.mean(rh == hid) >= 0.99
    both = (rh == hid) & (hid > 0)
    assert np.mean(np.abs(rt[both] - t[both]) <= 1.001e-3) >= 0.99


def test_frustum_subdivision_partitions_the_image():
    rng = np.array([[0, 0, 7, 4], [3, 5, 4, 9], [2, 2, 6, 6]], np.int32)
    A, B = ofr.subdivide(rng)
    for r, a, b in zip(rng, A, B):
        area = (r[2] - r[0]) * (r[3] - r[1])
        assert (a[2] - a[0]) * (a[3] - a[1]) + (b[2] - b[0]) * (b[3] - b[1]) == area
        assert a[0] == r[0] and a[1] == r[1] and b[2] == r[2] and b[3] == r[3]
    # the wider index axis is split (x on ties)
    assert A[0][2] == 3 and B[0][0] == 3 and A[1][3] == 7 and A[2][2] == 4


# --------------------------------------------------------- intersection
def test_intersection_of_half_spaces():
    """Two half-spaces f_A = n.x - a, f_B = -n.x - b (inside where negative)
    overlap iff a + b > 0; with exact (activation-free) bounds the search
    finds a point inside both exactly when they overlap."""
    nrm = np.array([0.3, 0.4, 0.86], np.float32)
    nrm /= np.linalg.norm(nrm)
    lo, up = -np.ones(3, np.float32), np.ones(3, np.float32)
    fa = ofuncs.OracleFunc(plane_mlp(nrm, 0.2), "affine_fixed")
    fb = ofuncs.OracleFunc(plane_mlp(-nrm, 0.1), "affine_fixed")
    found, ia, ib, loc, n = oisect.find_any_intersection(fa, fb, lo, up, 0.01)
    assert found and (ia, ib) == (1, 2)
    assert fa(loc[None])[0] < 0 and fb(loc[None])[0] < 0
    fb2 = ofuncs.OracleFunc(plane_mlp(-nrm, -0.5), "affine_fixed")   # n.x > 0.5: disjoint from n.x < 0.2
    found2, _, _, loc2, n2 = oisect.find_any_intersection(fa, fb2, lo, up, 0.01)
    assert not found2 and np.all(loc2 == -777.0)
    # touching half-spaces (n.x < 0.2 and n.x > 0.2): no sample is inside both,
    # the search converges on the shared plane through the near-surface test
    fb3 = ofuncs.OracleFunc(plane_mlp(-nrm, -0.2), "affine_fixed")
    found3, _, _, loc3, n3 = oisect.find_any_intersection(fa, fb3, lo, up, 0.05)
    assert found3 and abs(float(loc3 @ nrm) - 0.2) <= 0.05


def test_intersection_point_is_inside_or_near_both():
    a = models.random_mlp([32, 32], seed=11)
    b = models.random_mlp([32, 32], seed=12)
    fa, fb = ofuncs.OracleFunc(a), ofuncs.OracleFunc(b)
    lo, up = -np.ones(3, np.float32), np.ones(3, np.float32)
    eps = 0.02
    found, _, _, loc, n = oisect.find_any_intersection(fa, fb, lo, up, eps)
    if found:
        # a sample inside both is preferred; these nets overlap on a large set
        va, vb = fa(loc[None])[0], fb(loc[None])[0]
        assert va < 0 and vb < 0
    # brute force: a grid point inside both exists only if the search found one
    g = np.linspace(-1, 1, 33, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    both = (fa(pts) < 0) & (fb(pts) < 0)
    if both.any():
        assert found


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
