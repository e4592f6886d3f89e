"""Oracle: search for any point inside two implicit functions at once.

Test infrastructure only (see ``oracle/__init__.py``).

Restates find_any_intersection_iter / find_any_intersection
(src/kd_tree.py:402-655) for two functions, vectorised over the nodes of one
level.  Arithmetic is float32 as in the reference.
"""
import numpy as np

from .affine import SIGN_UNKNOWN, SIGN_NEGATIVE

f32 = np.float32


def sample_offsets(d=3):
    """[0; I; -I] (src/kd_tree.py:459)."""
    return np.concatenate((np.zeros((1, d)), np.eye(d), -np.eye(d)), axis=0).astype(f32)


def process_nodes(funcA, funcB, eps, lower, upper):
    """process_node (src/kd_tree.py:445-516) for every node of a level:
    (found (N,), loc (N, 3), needs_subdiv (N,), split_dim (N,))."""
    lower = np.asarray(lower, f32)
    upper = np.asarray(upper, f32)
    N, d = lower.shape
    eps_w = f32(eps) / np.sqrt(f32(d)).astype(f32)
    width = np.max(upper - lower, axis=-1)
    split_dim = np.argmax(upper - lower, axis=-1)
    small = width < eps_w
    center = (f32(0.5) * (lower + upper)).astype(f32)
    pts = (center[:, None, :] + eps_w * sample_offsets(d)[None, :, :]).astype(f32)   # (N, 7, d)
    codeA = funcA.classify_box(lower, upper)
    codeB = funcB.classify_box(lower, upper)
    vA = np.asarray(funcA(pts.reshape(-1, d))).reshape(N, -1)
    vB = np.asarray(funcB(pts.reshape(-1, d))).reshape(N, -1)
    sameA = np.all(vA < 0, axis=1) | np.all(vA > 0, axis=1)
    sameB = np.all(vB < 0, axis=1) | np.all(vB > 0, axis=1)
    nearA = small & ~sameA
    nearB = small & ~sameB
    found = np.zeros(N, bool)
    loc = np.full((N, d), f32(-777.0), f32)
    # both functions negative near a small node: the midpoint of their first
    # negative samples (src/kd_tree.py:480-489)
    negA = vA < 0
    negB = vB < 0
    iA = np.argmax(negA, axis=1)
    iB = np.argmax(negB, axis=1)
    anyA = negA[np.arange(N), iA]
    anyB = negB[np.arange(N), iB]
    near_neg = small & anyA & anyB
    found |= near_neg
    mid = (f32(0.5) * (pts[np.arange(N), iA] + pts[np.arange(N), iB])).astype(f32)
    loc = np.where(near_neg[:, None], mid, loc)
    # a sample inside both (preferred, so applied second; src/kd_tree.py:491-501)
    both = negA & negB
    ib = np.argmax(both, axis=1)
    have = both[np.arange(N), ib]
    found |= have
    loc = np.where(have[:, None], pts[np.arange(N), ib], loc)
    insideA = (codeA == SIGN_NEGATIVE) | ((codeA == SIGN_UNKNOWN) & ~nearA)
    insideB = (codeB == SIGN_NEGATIVE) | ((codeB == SIGN_UNKNOWN) & ~nearB)
    return found, loc, insideA & insideB, split_dim


def find_any_intersection(funcA, funcB, lower, upper, eps):
    """src/kd_tree.py:558-655 -> (found, 1, 2, loc (3,), nodes processed)."""
    node_lower = np.asarray(lower, f32)[None, :]
    node_upper = np.asarray(upper, f32)[None, :]
    processed = 0
    while True:
        processed += node_lower.shape[0]
        found, loc, subdiv, sd = process_nodes(funcA, funcB, eps, node_lower, node_upper)
        if found.any():                         # the first node in level order (src/kd_tree.py:521-524)
            i = int(np.argmax(found))
            return True, 1, 2, loc[i], processed
        if not subdiv.any():
            return False, 0, 0, np.full(3, f32(-777.0), f32), processed
        lo, up, sd = node_lower[subdiv], node_upper[subdiv], sd[subdiv]
        mid = (f32(0.5) * (lo + up)).astype(f32)
        mask = np.arange(lo.shape[1])[None, :] == sd[:, None]
        a_up = np.where(mask, mid, up)
        b_lo = np.where(mask, mid, lo)
        # children interleaved A0, B0, A1, B1, ... (src/kd_tree.py:549-551)
        node_lower = np.stack((lo, b_lo), axis=1).reshape(-1, lo.shape[1])
        node_upper = np.stack((a_up, up), axis=1).reshape(-1, lo.shape[1])
