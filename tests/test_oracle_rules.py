"""Pins for the oracle's range-analysis rules (CPU).

No reference golden vectors exist for this path, so each rule is pinned by
what the mathematics fixes: closed forms (activation-free MLPs are affine, so
affine arithmetic is exact), Chebyshev identities, and sampled soundness in
FP64.  A dropped term, a wrong sign or a transposed operand fails at least one.
"""
import numpy as np
import pytest

from oracle import affine as oaff, slope as oslope, funcs as ofuncs, mlp as omlp
from niq import mlp as nmlp, models

RNG = np.random.default_rng(1234)


def linear_mlp(dims, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(len(dims) - 1):
        A = rng.normal(size=(dims[i], dims[i + 1]))
        b = rng.normal(size=(dims[i + 1],))
        ops.append(nmlp.dense(dims[i], dims[i + 1], A=A, b=b))
    ops.append(nmlp.squeeze_last())
    return nmlp.build_spec(ops)


def small_mlp(act, hidden=(16, 12), seed=0):
    return models.random_mlp(list(hidden), activation=act, seed=seed, median_shift=False)


def random_boxes(n, rng, scale=1.0):
    c = rng.uniform(-1, 1, size=(n, 3))
    h = rng.uniform(0.01, 0.5, size=(n, 3)) * scale
    return c - h, c + h


# ------------------------------------------------------------- closed forms
@pytest.mark.parametrize("mode", ["affine_fixed", "affine_all", "slope_interval"])
def test_linear_mlp_bounds_are_exact(mode):
    """f = c.x + d is affine; the affine and slope forms must give exactly
    f(center) +- sum_i |c_i| h_i (distinct widths catch a transposed A)."""
    p = linear_mlp([3, 7, 5, 1], seed=3)
    layers = nmlp.lower(p)
    c = np.eye(3)
    d = np.zeros(1)
    M = np.eye(3)
    off = np.zeros(3)
    for L in layers:
        off = off @ L.A.astype(np.float64) + (L.b if L.b is not None else 0)
        M = M @ L.A.astype(np.float64)
    coef = M[:, 0]
    lo, up = random_boxes(50, RNG)
    ctr = 0.5 * (lo + up)
    h = up - ctr
    f_ctr = ctr @ coef + off[0]
    rad = np.abs(h) @ np.abs(coef)
    f = ofuncs.OracleFunc(p, mode, dtype=np.float64)
    got_lo, got_hi = f.bounds_box(lo, up)
    np.testing.assert_allclose(got_lo, f_ctr - rad, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got_hi, f_ctr + rad, rtol=1e-12, atol=1e-12)


def test_interval_mode_on_linear_mlp_is_interval_arithmetic():
    """interval mode = |A|-propagated radius, layer by layer (looser than affine)."""
    p = linear_mlp([3, 6, 1], seed=5)
    layers = nmlp.lower(p)
    lo, up = random_boxes(20, RNG)
    ctr = 0.5 * (lo + up)
    rad = up - ctr
    for L in layers:
        ctr = ctr @ L.A.astype(np.float64) + L.b
        rad = rad @ np.abs(L.A.astype(np.float64))
    f = ofuncs.OracleFunc(p, "interval", dtype=np.float64)
    got_lo, got_hi = f.bounds_box(lo, up)
    np.testing.assert_allclose(got_lo, ctr[:, 0] - rad[:, 0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got_hi, ctr[:, 0] + rad[:, 0], rtol=1e-12, atol=1e-12)


def test_relu_chebyshev_special_cases():
    lo = np.array([0.5, -3.0, -2.0, -1.0])
    up = np.array([2.0, -1.0, 2.0, 3.0])
    a, b, d = oaff.relu_rule(lo, up)
    np.testing.assert_array_equal(a[:2], [1.0, 0.0])
    np.testing.assert_array_equal(b[:2], [0.0, 0.0])
    assert a[2] == 0.5 and b[2] == 0.5 and d[2] == 0.5          # l = -u: alpha 1/2, beta u/4
    assert a[3] == 0.75 and np.isclose(b[3], 0.375) and b[3] == d[3]


def test_relu_chebyshev_equioscillation():
    """For l < 0 < u the error relu(x) - (ax+b) is +d, -d, +d at l, 0, u."""
    lo = -RNG.uniform(0.1, 3, 100)
    up = RNG.uniform(0.1, 3, 100)
    a, b, d = oaff.relu_rule(lo, up)
    e = lambda x: np.maximum(x, 0) - (a * x + b)
    np.testing.assert_allclose(e(lo), d, rtol=1e-12)
    np.testing.assert_allclose(e(np.zeros_like(lo)), -d, rtol=1e-12)
    np.testing.assert_allclose(e(up), d, rtol=1e-12)


@pytest.mark.parametrize("name", ["relu", "elu", "sin", "tanh"])
def test_linearisation_is_sound(name):
    """|act(x) - (alpha x + beta)| <= delta on [l, u], sampled densely (FP64)."""
    rng = np.random.default_rng(7)
    n = 400
    c = rng.uniform(-4, 4, n)
    w = rng.uniform(1e-3, 6, n)
    lo, up = c - w, c + w
    a, b, d = oaff.RULES[name](lo, up)
    assert np.all(d >= 0)
    xs = lo[:, None] + (up - lo)[:, None] * np.linspace(0, 1, 2001)[None, :]
    act = {"relu": lambda x: np.maximum(x, 0), "elu": omlp.elu, "sin": np.sin, "tanh": np.tanh}[name]
    err = np.abs(act(xs) - (a[:, None] * xs + b[:, None]))
    assert np.all(err <= d[:, None] * (1 + 1e-9) + 1e-12)


def test_tanh_rule_is_tight_for_small_intervals():
    """For a tiny interval the secant is nearly the tangent: delta -> 0."""
    lo = np.array([0.3, -1.2]); up = lo + 1e-4
    a, b, d = oaff.tanh_rule(lo, up)
    assert np.all(d < 1e-8)
    np.testing.assert_allclose(a, 1 - np.tanh(lo) ** 2, rtol=1e-3)


@pytest.mark.parametrize("mode", ["interval", "affine_fixed", "affine_all", "affine_truncate",
                                  "affine_append", "slope_interval"])
@pytest.mark.parametrize("act", ["relu", "elu", "sin", "tanh"])
def test_bounds_are_sound(mode, act):
    """f(x) lies in [lower, upper] for dense samples of every box (FP64)."""
    p = small_mlp(act, seed=11)
    kw = {"affine_n_truncate": 8, "affine_n_append": 4}
    f = ofuncs.OracleFunc(p, mode, dtype=np.float64, **kw)
    rng = np.random.default_rng(5)
    lo, up = random_boxes(24, rng)
    blo, bhi = f.bounds_box(lo, up)
    t = rng.uniform(0, 1, size=(24, 300, 3))
    pts = lo[:, None, :] + t * (up - lo)[:, None, :]
    vals = omlp.eval_points(p, pts.reshape(-1, 3), np.float64).reshape(24, 300)
    assert np.all(vals >= blo[:, None] - 1e-9)
    assert np.all(vals <= bhi[:, None] + 1e-9)


def test_affine_tighter_than_interval_on_relu_net():
    p = small_mlp("relu", hidden=(32, 32), seed=2)
    lo, up = random_boxes(64, RNG, scale=0.3)
    w_int = np.subtract(*ofuncs.OracleFunc(p, "interval", np.float64).bounds_box(lo, up)[::-1])
    w_aff = np.subtract(*ofuncs.OracleFunc(p, "affine_fixed", np.float64).bounds_box(lo, up)[::-1])
    w_all = np.subtract(*ofuncs.OracleFunc(p, "affine_all", np.float64).bounds_box(lo, up)[::-1])
    assert np.median(w_aff / w_int) < 1.0
    assert np.all(w_all <= w_aff * (1 + 1e-9))   # keeping every term never loosens


def test_fp32_matches_fp64_closely():
    p = models.config_model("c1_relu_3x32")
    lo, up = random_boxes(500, RNG, scale=0.2)
    b32 = ofuncs.OracleFunc(p, "affine_fixed", np.float32).bounds_box(lo, up)
    b64 = ofuncs.OracleFunc(p, "affine_fixed", np.float64).bounds_box(lo, up)
    scale = np.abs(b64[0]) + np.abs(b64[1])
    for x32, x64 in zip(b32, b64):
        assert np.max(np.abs(x32 - x64) / scale) < 1e-5


def test_point_eval_closed_form_and_format(golden_dir):
    """The reference's own sample model loads with its key grammar; plain
    evaluation of a linear net equals the matrix product."""
    import os
    p = nmlp.load(os.path.join(golden_dir, "fox.npz"))
    assert omlp.n_ops(p) == 18
    name, args = omlp.get_op_data(p, 0)
    assert name == "dense" and args["A"].shape == (3, 32) and args["A"].dtype == np.float32
    assert omlp.get_op_data(p, 17)[0] == "squeeze_last"
    layers = nmlp.lower(p)
    assert [L.act for L in layers] == ["relu"] * 8 + ["none"]
    q = linear_mlp([3, 4, 1], seed=9)
    x = RNG.normal(size=(10, 3))
    L = nmlp.lower(q)
    want = (x @ L[0].A.astype(np.float64) + L[0].b) @ L[1].A.astype(np.float64) + L[1].b
    np.testing.assert_allclose(omlp.eval_points(q, x, np.float64), want[:, 0], rtol=1e-12)


def test_encode_lowering_is_exact():
    """pow2_frequency_encode lowered to a dense layer gives the same values."""
    ops = [nmlp.pow2_frequency_encode(3), nmlp.dense(18, 1, A=RNG.normal(size=(18, 1)), b=[0.1]),
           nmlp.squeeze_last()]
    p = nmlp.build_spec(ops)
    x = RNG.uniform(-1, 1, size=(50, 3)).astype(np.float32)
    ref = omlp.eval_points(p, x, np.float32)
    layers = nmlp.lower(p)
    h = x @ layers[0].A + layers[0].b
    h = h @ layers[1].A + layers[1].b
    np.testing.assert_array_equal(h[:, 0], ref)


def test_slope_bound_contains_affine_range_on_linear():
    """For a linear net the slope form is exact too (slopes are constant)."""
    p = linear_mlp([3, 5, 1], seed=2)
    lo, up = random_boxes(10, RNG)
    a = ofuncs.OracleFunc(p, "affine_fixed", np.float64).bounds_box(lo, up)
    s = ofuncs.OracleFunc(p, "slope_interval", np.float64).bounds_box(lo, up)
    np.testing.assert_allclose(a[0], s[0], rtol=1e-12)
    np.testing.assert_allclose(a[1], s[1], rtol=1e-12)
