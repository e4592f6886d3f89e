"""Seeded synthetic MLPs in the reference's ``.npz`` op-list format.

The benchmark models are ReLU / TanH MLPs drawn with NumPy's PCG64
(``default_rng(seed)``), per layer W then b:
  * ``he``:     W ~ N(0, 2 / fan_in)
  * ``xavier``: W ~ N(0, 2 / (fan_in + fan_out))
  * biases ~ N(0, 0.1^2)
and the output bias is shifted so that the median of f over a
``linspace(-1, 1, 64)^3`` grid ('ij' order) is 0, the median taken in FP64.
This puts the zero level set inside [-1, 1]^3 for every seed.
"""
import functools

import numpy as np

from . import mlp

CONFIGS = {
    # name: (hidden widths, activation, init, seed)
    "c1_relu_3x32": ([32, 32, 32], "relu", "he", 0),
    "c2_relu_256x6": ([256] * 6, "relu", "he", 0),
    "c3_tanh_128x8": ([128] * 8, "tanh", "xavier", 1),
    "c4_relu_512x8": ([512] * 8, "relu", "he", 2),
}


def _eval_f64(layers, x):
    """Plain FP64 forward of (A, b, act) layers at points x (N, 3)."""
    try:
        import torch
        if torch.cuda.is_available():
            h = torch.as_tensor(x, dtype=torch.float64, device="cuda")
            for A, b, act in layers:
                h = h @ torch.as_tensor(A, dtype=torch.float64, device="cuda") + \
                    torch.as_tensor(b, dtype=torch.float64, device="cuda")
                if act == "relu":
                    h = torch.clamp_min(h, 0.0)
                elif act == "tanh":
                    h = torch.tanh(h)
                elif act == "elu":
                    h = torch.nn.functional.elu(h)
                elif act == "sin":
                    h = torch.sin(h)
            return h[:, 0].cpu().numpy()
    except ImportError:
        pass
    h = np.asarray(x, np.float64)
    for A, b, act in layers:
        h = h @ A.astype(np.float64) + b.astype(np.float64)
        if act == "relu":
            h = np.maximum(h, 0.0)
        elif act == "tanh":
            h = np.tanh(h)
        elif act == "elu":
            h = np.where(h > 0, h, np.expm1(np.minimum(h, 0)))
        elif act == "sin":
            h = np.sin(h)
    return h[:, 0]


def median_grid(n=64):
    ax = np.linspace(-1.0, 1.0, n)
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    return g


def random_mlp(hidden, activation="relu", init="he", seed=0, bias_std=0.1, in_dim=3,
               median_shift=True):
    """Spec dict for in_dim -> hidden... -> 1 with the seeded init above."""
    rng = np.random.default_rng(seed)
    dims = [in_dim] + list(hidden) + [1]
    layers = []
    for i in range(len(dims) - 1):
        fi, fo = dims[i], dims[i + 1]
        std = np.sqrt(2.0 / fi) if init == "he" else np.sqrt(2.0 / (fi + fo))
        A = rng.normal(0.0, std, size=(fi, fo)).astype(np.float32)
        b = rng.normal(0.0, bias_std, size=(fo,)).astype(np.float32)
        act = activation if i + 2 < len(dims) else "none"
        layers.append([A, b, act])
    if median_shift:
        vals = _eval_f64([(A, b, act) for A, b, act in layers], median_grid())
        med = float(np.median(vals))
        layers[-1][1] = (layers[-1][1].astype(np.float64) - med).astype(np.float32)
    ops = []
    act_op = {"relu": mlp.relu, "tanh": mlp.tanh, "elu": mlp.elu, "sin": mlp.sin}
    for A, b, act in layers:
        ops.append(mlp.dense(A.shape[0], A.shape[1], A=A, b=b))
        if act != "none":
            ops.append(act_op[act]())
    ops.append(mlp.squeeze_last())
    return mlp.build_spec(ops)


@functools.lru_cache(maxsize=None)
def _config_cached(name):
    hidden, act, init, seed = CONFIGS[name]
    return random_mlp(hidden, act, init, seed)


def config_model(name):
    """A fresh copy of a named benchmark model's params dict."""
    return dict(_config_cached(name))


def flops_per_box(dims, rows):
    """Algorithmic FLOPs of one range-analysis forward: sum 2 * rows * in * out."""
    return sum(2 * rows * dims[i] * dims[i + 1] for i in range(len(dims) - 1))


# The serving mix: eight resident MLPs with seeds 10-17, each shape drawn
# (default_rng(0)) from the three single-GPU shapes above.
SERVE_SHAPES = (("c1 3->32x3->1 relu", [32] * 3, "relu", "he"),
                ("c2 3->256x6->1 relu", [256] * 6, "relu", "he"),
                ("c3 3->128x8->1 tanh", [128] * 8, "tanh", "xavier"))


def serving_models(seed=0):
    """[(shape name, params)] of the eight resident serving models, and the
    generator the shapes were drawn from (for drawing the request trace)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(10, 18):
        name, hidden, act, init = SERVE_SHAPES[int(rng.integers(0, 3))]
        out.append((name, random_mlp(hidden, activation=act, init=init, seed=s)))
    return out, rng


def concurrent_tree_models(n_models=16, seed=1):
    """The 16 TanH MLPs of the concurrent-tree workload (Xavier init, seeds
    20-35): widths from {32, 64, 128, 256} and 3-8 hidden layers drawn from
    default_rng(seed).  Returns [(width, hidden layers, params)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_models):
        w = int(rng.choice([32, 64, 128, 256]))
        nl = int(rng.integers(3, 9))
        out.append((w, nl, random_mlp([w] * nl, activation="tanh", init="xavier", seed=20 + i)))
    return out


def flops_per_box_affine_all(dims):
    """Algorithmic FLOPs of one affine_all forward: layer l contracts its
    1 + v_l rows (base and the error terms), with v_0 the input dimension and
    v growing by the layer's width at every hidden activation."""
    v = dims[0]
    total = 0
    for i in range(len(dims) - 1):
        total += 2 * (1 + v) * dims[i] * dims[i + 1]
        if i + 2 < len(dims):
            v += dims[i + 1]
    return total


def error_bytes_per_box_affine_all(dims):
    """HBM bytes of one affine_all forward whose error-term matrix lives in
    memory: each hidden layer reads its input rows and writes its output rows
    (v_l x width floats each)."""
    v = dims[0]
    total = 0
    for i in range(len(dims) - 2):
        v_out = v + dims[i + 1]
        total += 4 * (v * dims[i] + v_out * dims[i + 1])
        v = v_out
    return total
