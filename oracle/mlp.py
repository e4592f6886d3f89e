"""Oracle: MLP ``.npz`` op-list grammar and plain point evaluation.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``src/mlp.py`` of the reference:
  * key grammar ``"%04d.<op>.<arg>"``          -- src/mlp.py:14-24, 117-144
  * op-list interpreter ``func_from_spec``     -- src/mlp.py:96-113
  * default (point) ops                        -- src/mlp.py:253-258, 283-347
  * ``pow2_frequency_encode`` / spatial xform  -- src/mlp.py:304-322, 335-347
"""
import numpy as np

ACTIVATIONS = ("relu", "elu", "sin", "tanh")


def n_ops(params):
    """src/mlp.py:135-144."""
    n = 0
    for key in params:
        vals = key.split(".")
        try:
            i_op = int(vals[0])
        except ValueError:
            raise ValueError(f"Could not parse out key {key}.")
        n = max(n, i_op + 1)
    return n


def get_op_data(params, i_op):
    """src/mlp.py:117-133: op name and its argument arrays."""
    prefix = f"{i_op:04d}"
    name = ""
    args = {}
    for key in params:
        if key.startswith(prefix):
            tokens = key.split(".")
            name = tokens[1]
            if len(tokens) > 2:
                args[tokens[2]] = params[key]
    if name == "":
        raise ValueError(f"didn't find op {i_op}")
    args.pop("_", None)  # placeholder arrays are dropped, src/mlp.py:108-109
    return name, args


def op_list(params):
    return [get_op_data(params, i) for i in range(n_ops(params))]


# ---------------------------------------------------------------- activations
def elu(x):
    # jax.nn.elu: where(x > 0, x, expm1(x)) with alpha = 1 (src/mlp.py:291-292)
    return np.where(x > 0, x, np.expm1(np.minimum(x, 0)))


def act_point(name, x):
    """Default-mode activations, src/mlp.py:285-300 (tanh is an addition)."""
    if name == "relu":
        return np.maximum(x, x.dtype.type(0))
    if name == "elu":
        return elu(x)
    if name == "sin":
        return np.sin(x)
    if name == "tanh":
        return np.tanh(x)
    raise ValueError(name)


def spatial_inverse(R, t, dtype):
    """src/mlp.py:342-346: input points get the inverse of the shape's R,t."""
    R_inv = np.linalg.inv(np.asarray(R, dtype=dtype)).astype(dtype)
    t_inv = (R_inv @ (-np.asarray(t, dtype=dtype))).astype(dtype)
    return R_inv, t_inv


def eval_points(params, x, dtype=np.float32):
    """Plain MLP evaluation of a batch of points ``x`` (N, d) -> (N,).

    Mirrors ``func_from_spec('default')`` (src/mlp.py:96-113) vmapped over
    points; each op follows its default rule (src/mlp.py:253-347).
    """
    h = np.asarray(x, dtype=dtype)
    for name, args in op_list(params):
        if name == "dense":
            A = np.asarray(args["A"], dtype=dtype)
            h = h @ A
            if "b" in args:
                h = h + np.asarray(args["b"], dtype=dtype)
        elif name in ACTIVATIONS:
            h = act_point(name, h)
        elif name == "squeeze_last":
            h = h[..., 0]
        elif name == "pow2_frequency_encode":
            coefs = np.asarray(args["coefs"], dtype=dtype)
            h = h[:, :, None] * coefs[None, None, :]
            if "shift" in args:
                h = h + np.asarray(args["shift"], dtype=dtype)
            h = h.reshape(h.shape[0], -1)
        elif name == "spatial_transformation":
            R_inv, t_inv = spatial_inverse(args["R"], args["t"], dtype)
            h = h @ R_inv + t_inv
        else:
            raise ValueError(f"unsupported op {name}")
    return h


def act_deriv(name, x):
    """Derivatives as jax differentiates the default activations: jax.nn.relu
    has derivative 0 at 0; jax.nn.elu's is exp(x) for x <= 0; tanh's 1 - tanh^2."""
    t = x.dtype.type
    if name == "relu":
        return np.where(x > 0, t(1), t(0)).astype(x.dtype)
    if name == "elu":
        return np.where(x > 0, t(1), np.exp(np.minimum(x, t(0)))).astype(x.dtype)
    if name == "sin":
        return np.cos(x)
    if name == "tanh":
        return (t(1) - np.tanh(x) ** 2).astype(x.dtype)
    raise ValueError(name)


def eval_points_grad(params, x, dtype=np.float32):
    """Value and gradient of the MLP at points x (N, d): forward mode, as
    jax.jacfwd of func_from_spec('default') computes it (src/render.py:62-64)."""
    h = np.asarray(x, dtype=dtype)
    N, d = h.shape
    T = np.broadcast_to(np.eye(d, dtype=dtype), (N, d, d)).copy()   # (N, d tangents, features)
    for name, args in op_list(params):
        if name == "dense":
            A = np.asarray(args["A"], dtype=dtype)
            h = h @ A
            T = T @ A
            if "b" in args:
                h = h + np.asarray(args["b"], dtype=dtype)
        elif name in ACTIVATIONS:
            dv = act_deriv(name, h)
            T = (T * dv[:, None, :]).astype(dtype)
            h = act_point(name, h)
        elif name == "squeeze_last":
            h = h[..., 0]
            T = T[..., 0]
        elif name == "spatial_transformation":
            R_inv, t_inv = spatial_inverse(args["R"], args["t"], dtype)
            h = h @ R_inv + t_inv
            T = T @ R_inv
        else:
            raise ValueError(f"unsupported op {name}")
    return h, T
