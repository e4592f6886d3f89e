"""MLP specification in the reference's ``.npz`` dictionary format.

Keys are ``"%04d.<op>.<arg>"`` (src/mlp.py:14-24); op i is parsed by
``get_op_data`` (src/mlp.py:117-133).  Weights ``dense.A`` are shaped
(in, out) and applied as ``x @ A + b`` (src/mlp.py:218-258).  This module
keeps that format unchanged (load/save, spec builders) and lowers an op list
to the dense-layer list that the gfx950 kernels run (``lower``), uploading it
once per params dict (``device_model``).
"""
import collections
import ctypes
import threading

import numpy as np

from . import _lib

ACTIVATIONS = ("relu", "elu", "sin", "tanh")


# ---------------------------------------------------------------- spec format
def build_spec(mlp_op_list):
    """src/mlp.py:14-24."""
    out = {}
    for i_op, op in enumerate(mlp_op_list):
        for key, val in op.items():
            out[f"{i_op:04d}." + key] = val
    return out


def n_ops(params):
    """src/mlp.py:135-144."""
    n = 0
    for key in params:
        vals = key.split(".")
        try:
            i_op = int(vals[0])
        except ValueError:
            raise ValueError(f"Could not parse out key {key}. Is this a valid mlp spec? "
                             "Did you make a mistake passing params dictionaries around?")
        n = max(n, i_op + 1)
    return n


def get_op_data(params, i_op):
    """src/mlp.py:117-133."""
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
    return name, args


def dense(in_dim, out_dim, with_bias=True, A=None, b=None):
    """src/mlp.py:218-249 (explicit arrays required; no JAX RNG init)."""
    if not with_bias and b is not None:
        raise ValueError("cannot specifify 'b' and 'with_bias=False'")
    A = np.asarray(A, dtype=np.float32)
    if A.shape != (in_dim, out_dim):
        raise ValueError(f"A should have shape ({in_dim},{out_dim}). Has shape {A.shape}.")
    out = {"dense.A": A}
    if with_bias:
        b = np.asarray(b, dtype=np.float32)
        if b.shape != (out_dim,):
            raise ValueError(f"b should have shape ({out_dim}). Has shape {b.shape}.")
        out["dense.b"] = b
    return out


def relu():
    return {"relu._": np.zeros((0,), np.float32)}


def elu():
    return {"elu._": np.zeros((0,), np.float32)}


def sin():
    return {"sin._": np.zeros((0,), np.float32)}


def tanh():
    return {"tanh._": np.zeros((0,), np.float32)}


def squeeze_last():
    return {"squeeze_last._": np.zeros((0,), np.float32)}


def pow2_frequency_encode(count_pow2, start_pow=0, with_shift=True):
    """src/mlp.py:304-314."""
    pows = 2.0 ** np.arange(start_pow, start_pow + count_pow2, dtype=np.float64)
    coefs = (pows * np.pi).astype(np.float32)
    if with_shift:
        coefs = np.repeat(coefs, 2)
        shift = np.zeros_like(coefs)
        shift[1::2] = np.float32(np.pi)
        return {"pow2_frequency_encode.coefs": coefs, "pow2_frequency_encode.shift": shift}
    return {"pow2_frequency_encode.coefs": coefs}


def spatial_transformation():
    return {"spatial_transformation.R": np.eye(3, dtype=np.float32),
            "spatial_transformation.t": np.zeros(3, np.float32)}


def load(filename):
    """src/mlp.py:173-185: npz -> dict of float arrays (no pickle)."""
    out = {}
    with np.load(filename, allow_pickle=False) as data:
        for key, val in data.items():
            out[key] = np.asarray(val)
    return out


def save(filename, params):
    """src/mlp.py:187-196."""
    np.savez(filename, **{k: np.asarray(v) for k, v in params.items()})


# ------------------------------------------------------------------- lowering
class Layer:
    __slots__ = ("A", "b", "act")

    def __init__(self, A, b, act):
        self.A, self.b, self.act = A, b, act


def lower(params):
    """Lower the op list to dense layers (A (in,out), b, activation after).

    ``pow2_frequency_encode`` becomes a dense layer with A[i, i*c + j] =
    coefs[j] and bias shift[j] (exact: every other product is by zero);
    ``spatial_transformation`` becomes the dense layer (R^-1, R^-1 (-t))
    (src/affine_layers.py:175-180).  The last layer must have one output,
    followed by ``squeeze_last`` (src/mlp.py:92)."""
    layers = []
    pending_act = None
    saw_squeeze = False
    for i in range(n_ops(params)):
        name, args = get_op_data(params, i)
        if saw_squeeze:
            raise ValueError("ops after squeeze_last are not supported")
        if name == "dense":
            A = np.asarray(args["A"], dtype=np.float32)
            b = np.asarray(args["b"], dtype=np.float32) if "b" in args else None
            layers.append(Layer(A, b, "none"))
        elif name in ACTIVATIONS:
            if not layers or layers[-1].act != "none":
                d = layers[-1].A.shape[1] if layers else None
                if d is None:
                    raise ValueError("an activation needs a preceding dense layer")
                layers.append(Layer(np.eye(d, dtype=np.float32), None, "none"))
            layers[-1].act = name
        elif name == "pow2_frequency_encode":
            coefs = np.asarray(args["coefs"], dtype=np.float32)
            d = layers[-1].A.shape[1] if layers else 3
            c = coefs.shape[0]
            A = np.zeros((d, d * c), np.float32)
            for k in range(d):
                A[k, k * c:(k + 1) * c] = coefs
            b = np.tile(np.asarray(args["shift"], np.float32), d) if "shift" in args else None
            layers.append(Layer(A, b, "none"))
        elif name == "spatial_transformation":
            R = np.asarray(args["R"], dtype=np.float32)
            t = np.asarray(args["t"], dtype=np.float32)
            R_inv = np.linalg.inv(R).astype(np.float32)
            layers.append(Layer(R_inv, (R_inv @ (-t)).astype(np.float32), "none"))
        elif name == "squeeze_last":
            saw_squeeze = True
        else:
            raise ValueError(f"unsupported op {name}")
    if not layers or not saw_squeeze or layers[-1].A.shape[1] != 1:
        raise ValueError("the op list must end with a dense layer of width 1 and squeeze_last")
    if layers[-1].act != "none":
        raise ValueError("no activation may follow the last dense layer")
    return layers


def dims_of(layers):
    return [layers[0].A.shape[0]] + [L.A.shape[1] for L in layers]


# ------------------------------------------------------ device model handles
class DeviceModel:
    """Uploaded, kernel-ready copy of a params dict (immutable)."""

    def __init__(self, params):
        _lib.require_gpu()
        lib = _lib.load_library()
        layers = lower(params)
        self.layers = layers
        self.dims = dims_of(layers)
        n = len(layers)
        dims = (ctypes.c_int * (n + 1))(*self.dims)
        acts = (ctypes.c_int * n)(*[_lib.ACT[L.act] for L in layers])
        self._keep = [np.ascontiguousarray(L.A, dtype=np.float32) for L in layers]
        self._keepb = [np.ascontiguousarray(L.b, dtype=np.float32) if L.b is not None else None
                       for L in layers]
        W = (ctypes.c_void_p * n)(*[a.ctypes.data for a in self._keep])
        B = (ctypes.c_void_p * n)(*[b.ctypes.data if b is not None else None for b in self._keepb])
        torch_init_device()
        handle = ctypes.c_void_p()
        _lib.check(lib.niq_model_create(n, dims, acts, W, B, ctypes.byref(handle)))
        self.handle = handle
        self._lib = lib

    @property
    def device_bytes(self):
        return int(self._lib.niq_model_device_bytes(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.niq_model_destroy(h)
            except Exception:
                pass
            self.handle = None


def torch_init_device():
    import torch
    torch.cuda.init()
    torch.cuda.current_device()


_cache_lock = threading.Lock()
_cache = collections.OrderedDict()   # id(params) -> (params, fingerprint, DeviceModel), LRU order
CACHE_MAX_MODELS = 64                 # resident uploads kept for reuse
CACHE_MAX_BYTES = 8 << 30             # ... and their device bytes


def _fingerprint(params):
    """What identifies the uploaded weights of a params dict: its keys and the
    identity, shape and dtype of each array.  Replacing an entry changes it;
    writing into an array in place does not (params are treated as immutable,
    as JAX arrays are in the reference)."""
    return tuple((k, id(v), getattr(v, "shape", None), str(getattr(v, "dtype", ""))) for k, v in params.items())


def device_model(params):
    """The uploaded model for ``params``, cached by dict identity and content
    fingerprint.  The cache is bounded (``CACHE_MAX_MODELS`` uploads and
    ``CACHE_MAX_BYTES`` of device memory, least recently used evicted); an
    evicted model's device memory is freed once no caller holds it."""
    key = id(params)
    fp = _fingerprint(params)
    with _cache_lock:
        ent = _cache.get(key)
        if ent is not None and ent[0] is params and ent[1] == fp:
            _cache.move_to_end(key)
            return ent[2]
    m = DeviceModel(params)
    with _cache_lock:
        _cache[key] = (params, fp, m)
        _cache.move_to_end(key)
        total = sum(e[2].device_bytes for e in _cache.values())
        while len(_cache) > 1 and (len(_cache) > CACHE_MAX_MODELS or total > CACHE_MAX_BYTES):
            _, (_, _, old) = _cache.popitem(last=False)
            total -= old.device_bytes
    return m


def forget_device_model(params):
    """Drop the cached upload of ``params`` (its device memory is freed once no
    caller holds the model)."""
    with _cache_lock:
        ent = _cache.get(id(params))
        if ent is not None and ent[0] is params:
            del _cache[id(params)]
