"""Oracle: uniform level-set kd-tree construction.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``construct_uniform_unknown_levelset_tree`` (src/kd_tree.py:103-218)
and its level step (src/kd_tree.py:19-100), with the bucketing helpers of
src/bucketing.py:7-36.  Output ordering is the reference's: within a level the
nodes are taken in batches of ``batch_process_size``; per batch, the A children
(lower half along the split axis) of the UNKNOWN nodes come first, then the B
children, both in node order (src/kd_tree.py:67-82, 93-95), and the batches are
concatenated and stably compacted (src/bucketing.py:16-32).  Interior/exterior
leaves are appended in node order (src/kd_tree.py:46-59).  Padded slots of the
returned arrays are zero here (the reference leaves arbitrary copies there;
only valid entries carry meaning, src/bucketing.py:19-29).
"""
import math

import numpy as np

from .affine import SIGN_UNKNOWN, SIGN_POSITIVE, SIGN_NEGATIVE

BUCKET_SIZES = [2 ** s for s in range(7, 31)]


def get_next_bucket_size(s):
    """src/bucketing.py:10-14."""
    for b in BUCKET_SIZES:
        if s <= b:
            return b
    raise ValueError("max bucket size exceeded")


def split_children(lower, upper):
    """src/kd_tree.py:38, 67-77: split at the midpoint of the widest axis
    (first axis on ties)."""
    dim = np.argmax(upper - lower, axis=-1)
    t = lower.dtype.type
    mid = t(0.5) * (lower + upper)
    mask = np.arange(lower.shape[-1])[None, :] == dim[:, None]
    a_lower = lower
    a_upper = np.where(mask, mid, upper)
    b_lower = np.where(mask, mid, lower)
    b_upper = upper
    return a_lower, a_upper, b_lower, b_upper


def finished_capacity_sizes(level_counts, level_finished, B):
    """Size of the interior/exterior buffers after the reference's doubling
    (src/kd_tree.py:126-131, 157-164): start at B rows, double while
    size - N_finished < N_curr_nodes before each level."""
    size = B
    for n_curr, n_fin in zip(level_counts, level_finished):
        while size - n_fin < n_curr:
            size *= 2
    return size


def construct_uniform_unknown_levelset_tree(func, lower, upper, node_terminate_thresh=None,
                                            split_depth=None, with_interior_nodes=False,
                                            with_exterior_nodes=False, offset=0.0,
                                            batch_process_size=2048, classify=None):
    """Returns the reference's dict plus a ``stats`` entry.

    ``func`` is an ``oracle.funcs.OracleFunc``; ``classify`` optionally
    overrides the per-batch classifier (lower, upper) -> codes."""
    for b in BUCKET_SIZES:
        if b > batch_process_size and (b // batch_process_size) * batch_process_size != b:
            raise ValueError("batch_process_size must be a factor of our bucket sizes")
    if node_terminate_thresh is None and split_depth is None:
        raise ValueError("must specify at least one of node_terminate_thresh or split_depth")
    if node_terminate_thresh is None:
        node_terminate_thresh = 9999999999
    dtype = func.dtype
    if classify is None:
        classify = lambda lo, up: func.classify_box(lo, up, offset=offset)
    B = batch_process_size

    node_lower = np.asarray(lower, dtype=dtype)[None, :]
    node_upper = np.asarray(upper, dtype=dtype)[None, :]
    N_curr = 1
    interior = ([], [])
    exterior = ([], [])
    n_int = 0
    n_ext = 0
    level_counts, level_int, level_ext = [], [], []
    n_classified = 0

    n_splits = 99999999 if split_depth is None else split_depth + 1
    for i_split in range(n_splits):
        quit_next = (N_curr >= node_terminate_thresh) or (i_split + 1 == n_splits)
        level_counts.append(N_curr)
        level_int.append(n_int)
        level_ext.append(n_ext)
        next_lower, next_upper = [], []
        n_occ = int(math.ceil(N_curr / min(B, max(N_curr, 1))))
        this_b = B if N_curr > B else N_curr
        for ib in range(n_occ):
            lo = node_lower[ib * this_b:(ib + 1) * this_b]
            up = node_upper[ib * this_b:(ib + 1) * this_b]
            codes = classify(lo, up)
            n_classified += lo.shape[0]
            neg = codes == SIGN_NEGATIVE
            pos = codes == SIGN_POSITIVE
            unk = codes == SIGN_UNKNOWN
            if with_interior_nodes:
                interior[0].append(lo[neg]); interior[1].append(up[neg])
            if with_exterior_nodes:
                exterior[0].append(lo[pos]); exterior[1].append(up[pos])
            n_int += int(neg.sum())
            n_ext += int(pos.sum())
            if not quit_next:
                al, au, bl, bu = split_children(lo[unk], up[unk])
                next_lower += [al, bl]
                next_upper += [au, bu]
            else:
                next_lower.append(lo[unk])
                next_upper.append(up[unk])
        node_lower = np.concatenate(next_lower) if next_lower else np.zeros((0, 3), dtype)
        node_upper = np.concatenate(next_upper) if next_upper else np.zeros((0, 3), dtype)
        N_curr = node_lower.shape[0]
        if quit_next:
            break

    def pack(lo, up, n, size):
        out_lo = np.zeros((size, 3), dtype)
        out_up = np.zeros((size, 3), dtype)
        out_lo[:n] = lo
        out_up[:n] = up
        return np.arange(size) < n, out_lo, out_up

    bucket = get_next_bucket_size(N_curr)
    v, l, u = pack(node_lower, node_upper, N_curr, bucket)
    out = {"unknown_node_valid": v, "unknown_node_lower": l, "unknown_node_upper": u}
    if with_interior_nodes:
        lo = np.concatenate(interior[0]) if interior[0] else np.zeros((0, 3), dtype)
        up = np.concatenate(interior[1]) if interior[1] else np.zeros((0, 3), dtype)
        size = finished_capacity_sizes(level_counts, level_int, B)
        v, l, u = pack(lo, up, n_int, size)
        out.update(interior_node_valid=v, interior_node_lower=l, interior_node_upper=u)
    if with_exterior_nodes:
        lo = np.concatenate(exterior[0]) if exterior[0] else np.zeros((0, 3), dtype)
        up = np.concatenate(exterior[1]) if exterior[1] else np.zeros((0, 3), dtype)
        size = finished_capacity_sizes(level_counts, level_ext, B)
        v, l, u = pack(lo, up, n_ext, size)
        out.update(exterior_node_valid=v, exterior_node_lower=l, exterior_node_upper=u)
    out["stats"] = {"n_classified": n_classified, "level_counts": level_counts,
                    "n_unknown": N_curr, "n_interior": n_int, "n_exterior": n_ext}
    return out
