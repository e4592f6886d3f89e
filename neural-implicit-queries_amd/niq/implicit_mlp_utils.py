"""``generate_implicit_from_file`` (src/implicit_mlp_utils.py:12-64).

Same signature and modes; returns ``(func, params)`` where ``params`` is the
npz dictionary (float32 NumPy arrays) and ``func`` runs on the gfx950 kernels.
Extra keyword arguments of this build: ``precision`` ('fp32' | 'bf16') and
``recheck_tol`` (FP64 re-test band, relative to the bound's scale).
"""
from . import mlp, affine, sdf, slope_interval


def implicit_from_params(params, mode, **kwargs):
    precision = kwargs.get('precision', 'fp32')
    recheck_tol = kwargs.get('recheck_tol', 0.0)
    if mode == 'sdf':
        return sdf.WeakSDFImplicitFunction(None, lipschitz_bound=kwargs.get('sdf_lipschitz', 1.)), params
    elif mode == 'interval':
        ctx = affine.AffineContext('interval')
    elif mode == 'affine_fixed':
        ctx = affine.AffineContext('affine_fixed')
    elif mode == 'affine_truncate':
        ctx = affine.AffineContext('affine_truncate', truncate_count=kwargs['affine_n_truncate'],
                                   truncate_policy=kwargs['affine_truncate_policy'])
        if ctx.truncate_policy != 'absolute':
            # src/affine.py:146 divides a (v,) vector by a (w,) vector, which
            # only broadcasts when v == w; only 'absolute' is defined
            raise ValueError("only the 'absolute' truncate policy is supported")
    elif mode == 'affine_append':
        ctx = affine.AffineContext('affine_append', n_append=kwargs['affine_n_append'])
    elif mode == 'affine_all':
        ctx = affine.AffineContext('affine_all')
    elif mode == 'slope_interval':
        return slope_interval.SlopeIntervalImplicitFunction(None, recheck_tol), params
    else:
        raise RuntimeError("unrecognized mode")
    return affine.AffineImplicitFunction(None, ctx, precision, recheck_tol), params


def generate_implicit_from_file(input_path, mode, **kwargs):
    if input_path.endswith(".npz"):
        params = mlp.load(input_path)
    else:
        raise ValueError("unrecognized filetype")
    return implicit_from_params(params, mode, **kwargs)
