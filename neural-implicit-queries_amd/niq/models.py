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
