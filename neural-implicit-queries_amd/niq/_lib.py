"""ctypes binding of ``libniq.so`` (the C ABI declared in include/niq.h).

The library is built in-tree by ``neural-implicit-queries_amd/build.py``.  There
is no fallback: if the library is missing or no GPU is visible, every query
raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libniq.so")

NIQ_OK = 0
NIQ_ERR_INVALID = -1
NIQ_ERR_UNSUPPORTED = -2
NIQ_ERR_CAPACITY = -3
NIQ_ERR_HIP = -4

ACT = {"none": 0, "relu": 1, "elu": 2, "sin": 3, "tanh": 4}
MODE = {"interval": 0, "affine_fixed": 1, "affine_truncate": 2, "affine_append": 3,
        "affine_all": 4, "slope_interval": 5, "sdf": 6}
PREC = {"fp32": 0, "bf16": 1}

EXPORTS = ("niq_model_create", "niq_model_destroy", "niq_model_device_bytes",
           "niq_classify_boxes", "niq_classify_general_boxes", "niq_tile_shape", "niq_eval_points",
           "niq_tree_workspace_bytes", "niq_tree_build", "niq_tree_enqueue",
           "niq_tree_counters_bytes", "niq_tree_decode", "niq_cast_workspace_bytes",
           "niq_cast_rays", "niq_cast_rays_frustum", "niq_find_any_intersection", "niq_closest_point",
           "niq_sample_workspace_bytes", "niq_sample_surface_nodes", "niq_bulk_sums_nodes",
           "niq_camera_rays", "niq_hit_positions", "niq_outward_normals", "niq_shade_image", "niq_mc_extract", "niq_mc_table",
           "niq_serve_create", "niq_serve_destroy", "niq_serve_submit_rays", "niq_serve_submit_tree",
           "niq_serve_submit_closest_point", "niq_serve_poll", "niq_serve_bytes_in_use",
           "niq_last_error", "niq_version")


class NiqError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"niq error {status}: {msg}")
        self.status = status


class CapacityError(NiqError):
    pass


class EvalCfg(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("precision", ctypes.c_int),
                ("truncate_count", ctypes.c_int), ("n_append", ctypes.c_int),
                ("recheck_tol", ctypes.c_float), ("variant", ctypes.c_int),
                ("sdf_lipschitz", ctypes.c_float), ("ray_variant", ctypes.c_int)]


class TreeParams(ctypes.Structure):
    _fields_ = [("lower", ctypes.c_float * 3), ("upper", ctypes.c_float * 3),
                ("split_depth", ctypes.c_int), ("node_terminate_thresh", ctypes.c_int64),
                ("offset", ctypes.c_float), ("batch_process_size", ctypes.c_int),
                ("with_interior_nodes", ctypes.c_int), ("with_exterior_nodes", ctypes.c_int)]


class TreeResult(ctypes.Structure):
    _fields_ = [("unknown_lower", ctypes.c_void_p), ("unknown_upper", ctypes.c_void_p),
                ("unknown_capacity", ctypes.c_int64),
                ("interior_lower", ctypes.c_void_p), ("interior_upper", ctypes.c_void_p),
                ("interior_capacity", ctypes.c_int64),
                ("exterior_lower", ctypes.c_void_p), ("exterior_upper", ctypes.c_void_p),
                ("exterior_capacity", ctypes.c_int64),
                ("n_unknown", ctypes.c_int64), ("n_interior", ctypes.c_int64),
                ("n_exterior", ctypes.c_int64), ("n_classified", ctypes.c_int64),
                ("n_recheck", ctypes.c_int64), ("n_levels", ctypes.c_int),
                ("level_n", ctypes.c_int32 * 64), ("level_interior", ctypes.c_int32 * 64),
                ("level_exterior", ctypes.c_int32 * 64)]


class CastOpts(ctypes.Structure):
    _fields_ = [("hit_eps", ctypes.c_float), ("max_dist", ctypes.c_float),
                ("n_max_step", ctypes.c_int), ("n_substeps", ctypes.c_int),
                ("safety_factor", ctypes.c_float), ("interval_grow_fac", ctypes.c_float),
                ("interval_shrink_fac", ctypes.c_float), ("init_step", ctypes.c_float)]


class FrustumCam(ctypes.Structure):
    _fields_ = [("root", ctypes.c_float * 3), ("look", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("left", ctypes.c_float * 3), ("fov_x", ctypes.c_float), ("fov_y", ctypes.c_float),
                ("res_x", ctypes.c_int), ("res_y", ctypes.c_int)]


class FrustumOpts(ctypes.Structure):
    _fields_ = [("cast", CastOpts), ("refine_width_fac", ctypes.c_float), ("n_side_init", ctypes.c_int)]


class ServeInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("n_evals", ctypes.c_int64), ("n_batched", ctypes.c_int32),
                ("graph_replay", ctypes.c_int32), ("n_processed", ctypes.c_int64)]


class ShadeOpts(ctypes.Structure):
    _fields_ = [("shading", ctypes.c_int), ("up_dir", ctypes.c_float * 3),
                ("matcaps", ctypes.c_void_p * 4), ("matcap_h", ctypes.c_int), ("matcap_w", ctypes.c_int),
                ("colors", (ctypes.c_float * 3) * 8), ("n_colors", ctypes.c_int),
                ("color_per_hit", ctypes.c_void_p), ("composite", ctypes.c_int), ("tonemap", ctypes.c_int),
                ("gamma", ctypes.c_float), ("white_level", ctypes.c_float), ("exposure", ctypes.c_float)]


class BulkSums(ctypes.Structure):
    _fields_ = [("interior_volume", ctypes.c_double), ("interior_moment", ctypes.c_double * 3),
                ("unknown_volume", ctypes.c_double), ("n_inside", ctypes.c_double),
                ("inside_moment", ctypes.c_double * 3)]


_lib = None


def load_library():
    """Load (once) and return the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with neural-implicit-queries_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, I64, F, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_size_t
    sig = {
        "niq_model_create": (I, [I, P, P, P, P, ctypes.POINTER(P)]),
        "niq_model_destroy": (None, [P]),
        "niq_model_device_bytes": (S, [P]),
        "niq_classify_boxes": (I, [P, P, P, P, I64, I, F, P, P, P]),
        "niq_classify_general_boxes": (I, [P, P, P, P, I64, I, I, F, P, P, P]),
        "niq_tile_shape": (I, [I, I, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]),
        "niq_eval_points": (I, [P, P, I64, I, P, P]),
        "niq_tree_workspace_bytes": (I, [P, P, P, ctypes.POINTER(S)]),
        "niq_tree_build": (I, [P, P, P, P, S, P, P]),
        "niq_tree_enqueue": (I, [P, P, P, P, S, P, ctypes.POINTER(S), P]),
        "niq_tree_counters_bytes": (S, []),
        "niq_tree_decode": (I, [P, P, P]),
        "niq_cast_workspace_bytes": (S, [P]),
        "niq_cast_rays": (I, [P, I, P, P, P, I64, P, P, P, P, P, S, ctypes.POINTER(I64), P]),
        "niq_closest_point": (I, [P, P, P, P, P, I64, F, I, P, P, ctypes.POINTER(I64), P]),
        "niq_cast_rays_frustum": (I, [P, I, P, P, P, P, P, P, ctypes.POINTER(I64), P]),
        "niq_find_any_intersection": (I, [P, P, P, P, P, P, F, ctypes.POINTER(ctypes.c_int32), P,
                                          ctypes.POINTER(I64), P]),
        "niq_mc_extract": (I, [P, P, P, I64, I, P, I64, ctypes.POINTER(I64), P]),
        "niq_mc_table": (I, [P]),
        "niq_sample_workspace_bytes": (S, [I64]),
        "niq_sample_surface_nodes": (I, [P, P, P, I64, F, ctypes.c_uint64, I64, I64, I64, P, I64,
                                         ctypes.POINTER(I64), ctypes.POINTER(I64), P, S, P]),
        "niq_bulk_sums_nodes": (I, [P, P, P, I64, P, P, I64, I64, ctypes.c_uint64, P, P]),
        "niq_camera_rays": (I, [P, P, P, P, I, F, P, P, P]),
        "niq_hit_positions": (I, [P, P, P, I64, P, P]),
        "niq_outward_normals": (I, [P, I, P, P, I64, F, I, P, P]),
        "niq_shade_image": (I, [P, P, P, P, I64, P, P]),
        "niq_serve_create": (I, [I, S, I64, ctypes.POINTER(P)]),
        "niq_serve_destroy": (I, [P]),
        "niq_serve_submit_rays": (I, [P, P, I, P, P, P, I64, P, P, P, P, ctypes.POINTER(ctypes.c_uint64)]),
        "niq_serve_submit_tree": (I, [P, P, P, P, P, ctypes.POINTER(ctypes.c_uint64)]),
        "niq_serve_submit_closest_point": (I, [P, P, P, P, P, P, I64, F, I, P, P,
                                               ctypes.POINTER(ctypes.c_uint64)]),
        "niq_serve_poll": (I, [P, ctypes.c_uint64, I, P, P]),
        "niq_serve_bytes_in_use": (S, [P]),
        "niq_last_error": (ctypes.c_char_p, []),
        "niq_version": (I, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != NIQ_OK:
        msg = load_library().niq_last_error().decode(errors="replace")
        if status == NIQ_ERR_CAPACITY:
            raise CapacityError(status, msg)
        raise NiqError(status, msg)


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("niq needs a ROCm GPU (gfx950); none is visible")


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def eval_cfg(mode="affine_fixed", precision="fp32", truncate_count=0, n_append=0, recheck_tol=0.0, variant=0,
             sdf_lipschitz=1.0):
    if mode not in MODE:
        raise ValueError("invalid mode")
    return EvalCfg(MODE[mode], PREC[precision], int(truncate_count), int(n_append),
                   float(recheck_tol), int(variant), float(sdf_lipschitz), 0)
