"""Oracle: closest point on the level set by stack-based subdivision.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``closest_point`` / ``closest_point_iter`` (src/kd_tree.py:659-802)
including its quirks:
  * the per-node distance used for pruning and for the result is the distance
    to the node *centre* (src/kd_tree.py:700), plus the node diagonal for a
    spanning node (src/kd_tree.py:697-698, 715);
  * the 7 sign samples sit at centre +- (upper - lower) * e_i
    (src/kd_tree.py:703-704);
  * the returned location is the centre of the winning node (src/kd_tree.py:729);
  * pruning compares against the query's minimum *before* this batch's
    scatter-min (src/kd_tree.py:684, 716, 725);
  * work is popped B at a time from the stack top and children are pushed,
    interleaved A0 B0 A1 B1 ..., back at the pop position (src/kd_tree.py:679-760).
Duplicate-query writes of the location within one batch resolve to the last
batch entry (XLA's CPU scatter applies updates in order).
"""
import numpy as np

from .affine import SIGN_POSITIVE, SIGN_NEGATIVE

f32 = np.float32


def all_same_sign(vals):
    """src/utils.py:146-150 over the last axis."""
    return np.all(vals < 0, axis=-1) | np.all(vals > 0, axis=-1)


def closest_point(func, lower, upper, query_points, eps=0.001, batch_process_size=2048,
                  max_rounds=None, track_near=False):
    """Returns (dist (Q,), loc (Q, 3), info).  With ``track_near`` (func is a
    ``margins.TwinFunc``) info["near"] marks the queries one of whose node
    decisions (classification, sample signs) lay near its threshold,
    info["near_round"] is the first round in which any query's did, and
    info["last_round"] the last round that processed a node of each query.
    Queries share the stack: a flipped decision changes how many nodes are
    pushed, hence which nodes later rounds pop together and the minimum each
    pruning test sees, so a query processed after ``near_round`` may change
    without a near decision of its own.  With batch_process_size above the
    largest stack every round pops the whole stack and no query depends on
    another's decisions."""
    B = batch_process_size
    q = np.asarray(query_points, f32)
    Q = q.shape[0]
    d = q.shape[1]
    work_q = np.arange(Q, dtype=np.int64)
    work_lo = np.tile(np.asarray(lower, f32), (Q, 1))
    work_up = np.tile(np.asarray(upper, f32), (Q, 1))
    top = Q
    min_dist = np.full(Q, np.inf, f32)
    min_loc = np.full((Q, 3), f32(-777.0), f32)
    eps_cube_width = f32(eps) / np.sqrt(f32(d))
    offsets = np.concatenate((np.zeros((1, d)), np.eye(d), -np.eye(d)), 0).astype(f32)
    rounds = 0
    n_processed = 0
    near = np.zeros(Q, bool)
    near_round = None
    last_round = np.full(Q, -1, np.int64)
    while top > 0:
        pop = max(top - B, 0)
        n = top - pop
        bq = work_q[pop:top]
        blo = work_lo[pop:top]
        bup = work_up[pop:top]
        n_processed += n
        qloc = q[bq]
        qmin = min_dist[bq]
        ext = bup - blo
        width = np.max(ext, axis=-1)
        center = f32(0.5) * (blo + bup)
        diag = np.sqrt(np.sum(ext * ext, axis=-1))
        dc = np.sqrt(np.sum((qloc - center) ** 2, axis=-1))
        max_dist = dc + diag
        split_dim = np.argmax(ext, axis=-1)
        is_small = width < eps_cube_width
        samples = center[:, None, :] + ext[:, None, :] * offsets[None, :, :]
        code = func.classify_box(blo, bup)
        if track_near:
            nb = func.last_near.copy()
        outside = (code == SIGN_NEGATIVE) | (code == SIGN_POSITIVE)
        vals = func(samples.reshape(-1, d)).reshape(n, 7)
        if track_near:
            nb |= func.last_near.reshape(n, 7).any(axis=1)
            near[bq[nb]] = True
            if near_round is None and nb.any():
                near_round = rounds
            last_round[bq] = rounds
        spans = ~all_same_sign(vals)
        this_dist = np.where(spans, max_dist, f32(np.inf)).astype(f32)
        needs = (~outside) & (~is_small) & (dc < qmin)
        # scatter-min, then argmin-location (src/kd_tree.py:725-729)
        np.minimum.at(min_dist, bq, this_dist)
        new_min = min_dist[bq]
        has_new = this_dist == new_min
        for i in np.nonzero(has_new)[0]:
            min_loc[bq[i]] = center[i]
        # compact, split, interleave, push (src/kd_tree.py:732-760)
        sel = np.nonzero(needs)[0]
        slo, sup, sq = blo[sel], bup[sel], bq[sel]
        sdim = split_dim[sel]
        mid = f32(0.5) * (slo + sup)
        mask = np.arange(3)[None, :] == sdim[:, None]
        a_up = np.where(mask, mid, sup)
        b_lo = np.where(mask, mid, slo)
        m = sel.shape[0]
        new_lo = np.empty((2 * m, 3), f32)
        new_up = np.empty((2 * m, 3), f32)
        new_q = np.empty(2 * m, np.int64)
        new_lo[0::2], new_lo[1::2] = slo, b_lo
        new_up[0::2], new_up[1::2] = a_up, sup
        new_q[0::2], new_q[1::2] = sq, sq
        end = pop + 2 * m
        if end > work_q.shape[0]:
            grow = max(end, 2 * work_q.shape[0])
            work_q = np.resize(work_q, grow)
            work_lo = np.resize(work_lo, (grow, 3))
            work_up = np.resize(work_up, (grow, 3))
        work_q[pop:end] = new_q
        work_lo[pop:end] = new_lo
        work_up[pop:end] = new_up
        top = end
        rounds += 1
        if max_rounds is not None and rounds >= max_rounds:
            break
    info = {"rounds": rounds, "n_processed": n_processed}
    if track_near:
        info["near"] = near
        info["near_round"] = near_round
        info["last_round"] = last_round
    return min_dist, min_loc, info
