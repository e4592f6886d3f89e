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


# ------------------------------------------------- sin slope (src/affine_layers.py:100-138)
def _brute_cos_midrange(lo, up, n=200001):
    """alpha by definition: the midpoint of [min cos, max cos] over [lo, up],
    found by dense sampling (independent of sin_bound's extremum tests)."""
    xs = np.linspace(lo, up, n)
    c = np.cos(xs)
    # the extrema of cos lie at multiples of pi: add those inside the interval
    k = np.arange(np.ceil(lo / np.pi), np.floor(up / np.pi) + 1)
    if k.size:
        c = np.concatenate([c, np.cos(k * np.pi)])
    return 0.5 * (c.min() + c.max())


def _brute_sin_residual(lo, up, alpha, n=200001):
    """min and max of sin(x) - alpha x over [lo, up] by dense sampling plus the
    critical points x = +-arccos(alpha) + 2 k pi inside the interval."""
    xs = np.linspace(lo, up, n)
    a0 = np.arccos(np.clip(alpha, -1, 1))
    crit = []
    for s in (a0, -a0):
        k = np.arange(np.ceil((lo - s) / (2 * np.pi)), np.floor((up - s) / (2 * np.pi)) + 1)
        crit.extend(list(s + 2 * np.pi * k))
    xs = np.concatenate([xs, np.array(crit, dtype=np.float64)])
    r = np.sin(xs) - alpha * xs
    return r.min(), r.max()


@pytest.mark.parametrize("lo,up,want_alpha", [
    (0.0, np.pi / 2, 0.5),                 # cos over [0,1]: the midrange 1/2, not the secant 2/pi
    (-np.pi / 2, np.pi / 2, 0.5),          # cos peaks inside: [0, 1]
    (0.0, np.pi, 0.0),                     # cos over [-1, 1]
    (np.pi / 2, 3 * np.pi / 2, -0.5),      # cos bottoms out at pi: [-1, 0]
    (6.0, 7.0, None),                      # contains 2 pi: [min(cos 6, cos 7), 1]
    (0.3, 0.5, None),                      # monotone piece
    (-2.5, 4.0, None),                     # wider than pi: [-1, 1] -> 0
    (-9.0, -8.2, None),
])
def test_sin_slope_is_the_midrange_of_cos(lo, up, want_alpha):
    """The reference's sin slope is 0.5 (min cos + max cos) over [l, u]
    (src/affine_layers.py:109-110), "not the Chebyshev value": pinned against
    the definition by dense sampling, and in closed form where one exists.
    A secant or tangent slope differs on [0, pi/2] (2/pi and 1 against 1/2)."""
    a, b, d = oaff.sin_rule(np.array([lo]), np.array([up]))
    ref = _brute_cos_midrange(lo, up)
    assert abs(a[0] - ref) < 1e-9
    if want_alpha is not None:
        assert abs(a[0] - want_alpha) < 1e-12
    # beta and delta centre and half-width the range of sin(x) - alpha x
    rmin, rmax = _brute_sin_residual(lo, up, a[0])
    assert abs(b[0] - 0.5 * (rmin + rmax)) < 1e-9
    assert abs(d[0] - 0.5 * (rmax - rmin)) < 1e-9


def test_sin_rule_closed_form_on_symmetric_interval():
    """[-pi/2, pi/2]: alpha = 1/2; sin(x) - x/2 peaks at x = pi/3 (cos x = 1/2)
    with value sqrt(3)/2 - pi/6 and is odd, so beta = 0, delta = sqrt(3)/2 - pi/6."""
    a, b, d = oaff.sin_rule(np.array([-np.pi / 2]), np.array([np.pi / 2]))
    assert abs(a[0] - 0.5) < 1e-15
    assert abs(b[0]) < 1e-15
    assert abs(d[0] - (np.sqrt(3) / 2 - np.pi / 6)) < 1e-14


def test_sin_rule_tiny_interval_tends_to_the_tangent():
    """On [x0 - h, x0 + h] away from cos's extrema the midrange of cos is
    cos(x0) + O(h^2) and the linearisation error is O(h^2)."""
    x0 = np.array([0.4, 1.1, 2.0, -0.7, 4.0])
    for h in (1e-2, 1e-3):
        a, b, d = oaff.sin_rule(x0 - h, x0 + h)
        assert np.all(np.abs(a - np.cos(x0)) <= 0.6 * h * h)
        assert np.all(d <= 0.6 * h * h)


# ------------------------------- truncate_affine (src/affine.py:127-162, top-n rows kept)
def _small_form(rng, N, v, w):
    base = rng.normal(size=(N, w))
    aff = rng.normal(size=(N, v, w))
    err = rng.uniform(0, 0.1, size=(N, w))
    return base, aff, err


@pytest.mark.parametrize("v,n_keep", [(6, 2), (6, 4), (5, 1), (7, 6)])
def test_truncate_keeps_the_top_rows_by_brute_force(v, n_keep):
    """The kept rows are the n_keep rows of largest sum |row|, in decreasing
    order; the dropped rows' |entries| are added to err, feature by feature.
    Checked against an exhaustive search over all row subsets."""
    import itertools
    rng = np.random.default_rng(v * 10 + n_keep)
    w = 3
    base, aff, err = _small_form(rng, 5, v, w)
    ctx = oaff.AffineContext("affine_truncate", truncate_count=n_keep)
    b2, a2, e2 = oaff.truncate_affine(ctx, base, aff, err)
    assert a2.shape == (5, n_keep, w)
    np.testing.assert_array_equal(b2, base)
    for i in range(5):
        mags = np.abs(aff[i]).sum(-1)
        best = max(itertools.combinations(range(v), n_keep), key=lambda s: sum(mags[list(s)]))
        keep = sorted(best, key=lambda r: -mags[r])
        np.testing.assert_array_equal(a2[i], aff[i][keep])
        drop = [r for r in range(v) if r not in best]
        np.testing.assert_allclose(e2[i], err[i] + np.abs(aff[i][drop]).sum(0), rtol=1e-14)


def test_truncate_preserves_the_radius_and_is_a_no_op_when_short():
    rng = np.random.default_rng(3)
    base, aff, err = _small_form(rng, 8, 9, 4)
    ctx = oaff.AffineContext("affine_truncate", truncate_count=3)
    _, a2, e2 = oaff.truncate_affine(ctx, base, aff, err)
    np.testing.assert_allclose(oaff.radius(a2, e2), oaff.radius(aff, err), rtol=1e-13)
    ctx_long = oaff.AffineContext("affine_truncate", truncate_count=9)
    b3, a3, e3 = oaff.truncate_affine(ctx_long, base, aff, err)
    assert a3 is aff and e3 is err and b3 is base


@pytest.mark.parametrize("act", ["relu", "tanh", "sin"])
def test_truncate_with_room_for_every_row_equals_affine_all(act):
    """With truncate_count at least the largest row count nothing is dropped:
    affine_truncate gives affine_all's bounds bit for bit."""
    p = small_mlp(act, hidden=(10, 8), seed=4)
    lo, up = random_boxes(40, np.random.default_rng(8), scale=0.5)
    for dt in (np.float32, np.float64):
        a = ofuncs.OracleFunc(p, "affine_all", dtype=dt).bounds_box(lo, up)
        t = ofuncs.OracleFunc(p, "affine_truncate", dtype=dt, affine_n_truncate=3 + 10 + 8).bounds_box(lo, up)
        np.testing.assert_array_equal(a[0], t[0])
        np.testing.assert_array_equal(a[1], t[1])


# --------------------------- affine_append (src/affine.py:183-191, top-k of delta appended)
@pytest.mark.parametrize("w,k", [(5, 2), (6, 3), (4, 1), (4, 4)])
def test_append_keeps_the_top_k_deltas_by_brute_force(w, k):
    """apply_linear_approx in affine_append: the k largest |delta| become new
    one-hot rows (largest first, value delta at its feature), the rest of the
    deltas are summed into every feature's err; alpha scales base, rows and
    err.  Checked against an exhaustive search over feature subsets."""
    import itertools
    rng = np.random.default_rng(w * 7 + k)
    N, v = 4, 3
    base, aff, err = _small_form(rng, N, v, w)
    alpha = rng.uniform(0, 1, size=(N, w))
    beta = rng.normal(size=(N, w))
    delta = rng.uniform(0, 1, size=(N, w))
    ctx = oaff.AffineContext("affine_append", n_append=k)
    b2, a2, e2 = oaff.apply_linear_approx(ctx, base, aff, err, alpha, beta, delta)
    np.testing.assert_allclose(b2, alpha * base + beta, rtol=1e-15)
    assert a2.shape == (N, v + k, w)
    np.testing.assert_allclose(a2[:, :v], alpha[:, None, :] * aff, rtol=1e-15)
    for i in range(N):
        best = max(itertools.combinations(range(w), k), key=lambda s: sum(delta[i][list(s)]))
        order = sorted(best, key=lambda j: -delta[i][j])
        want = np.zeros((k, w))
        for r, j in enumerate(order):
            want[r, j] = delta[i][j]
        np.testing.assert_array_equal(a2[i, v:], want)
        rest = sum(delta[i][j] for j in range(w) if j not in best)
        np.testing.assert_allclose(e2[i], alpha[i] * err[i] + rest, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("act", ["relu", "tanh", "sin"])
def test_append_of_every_delta_equals_affine_all(act):
    """With n_append equal to the layer width every delta gets its own row: the
    same rows as affine_all's diag(delta) in another order, so the bounds agree
    to rounding (FP64)."""
    p = small_mlp(act, hidden=(9, 9), seed=6)
    lo, up = random_boxes(40, np.random.default_rng(9), scale=0.5)
    a = ofuncs.OracleFunc(p, "affine_all", dtype=np.float64).bounds_box(lo, up)
    t = ofuncs.OracleFunc(p, "affine_append", dtype=np.float64, affine_n_append=9).bounds_box(lo, up)
    scale = np.abs(a[0]) + np.abs(a[1])
    assert np.max(np.abs(a[0] - t[0]) / scale) < 1e-12
    assert np.max(np.abs(a[1] - t[1]) / scale) < 1e-12
