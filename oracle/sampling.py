"""Oracle: rejection sampling of the level set and bulk mass properties.

Test infrastructure only (see ``oracle/__init__.py``).

Restates:
  * sample_surface_iter / sample_surface          -- src/kd_tree.py:220-272
  * sample_surface_uniform_iter / _uniform        -- src/kd_tree.py:276-336
  * bulk_properties_sample_mass / bulk_properties -- src/kd_tree.py:804-863

The reference draws its candidates from ``jax.random`` (threefry); those
streams cannot be reproduced without JAX, so the device code and this oracle
share a stream of their own instead: Philox4x32-10 (Salmon, Moraes, Dror,
Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011), pinned by the
Random123 known-answer vectors in the tests.  Candidate ``j`` of a query
(a global index over all rounds) is drawn from

    counter = (j mod 2^32, j div 2^32, tag, 0),  key = (seed mod 2^32, seed div 2^32)
    node    = floor(w0 * n_nodes / 2^32)
    u_k     = (w_{k+1} >> 8) * 2^-24                    (k = 0, 1, 2; in [0, 1))
    p_k     = lower_k + u_k * (upper_k - lower_k)       (float32, unfused)

with ``tag`` = 0x53555246 for surface samples and 0x4d415353 for mass
samples.  Rounds append their accepted candidates in candidate order and drop
what exceeds the output (src/kd_tree.py:242-246), so the result is the first
``n_samples`` accepted candidates of the stream whatever the round size is.
"""
import numpy as np

from .tree import construct_uniform_unknown_levelset_tree

f32 = np.float32

TAG_SURFACE = 0x53555246
TAG_MASS = 0x4D415353

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_LO32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Philox4x32 with 10 rounds on arrays of counters: ctr is 4 uint32 arrays,
    key 2 uint32 scalars/arrays; returns the 4 output words."""
    c = [np.asarray(x, np.uint32) for x in ctr]
    k0 = np.asarray(key[0], np.uint32)
    k1 = np.asarray(key[1], np.uint32)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = (k0 + _W0).astype(np.uint32)
                k1 = (k1 + _W1).astype(np.uint32)
            p0 = c[0].astype(np.uint64) * _M0
            p1 = c[2].astype(np.uint64) * _M1
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _LO32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _LO32).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return c


def seed_key(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.uint32(seed & 0xFFFFFFFF), np.uint32(seed >> 32)


def candidates(seed, tag, j0, n, node_lower, node_upper, n_nodes):
    """Candidates j0 .. j0+n-1 of the stream: (positions (n, 3) float32, node index (n,))."""
    j = np.arange(j0, j0 + n, dtype=np.uint64)
    ctr = ((j & _LO32).astype(np.uint32), (j >> np.uint64(32)).astype(np.uint32),
           np.full(n, tag, np.uint32), np.zeros(n, np.uint32))
    w = philox4x32_10(ctr, seed_key(seed))
    node = ((w[0].astype(np.uint64) * np.uint64(n_nodes)) >> np.uint64(32)).astype(np.int64)
    lo = np.asarray(node_lower, f32)[node]
    up = np.asarray(node_upper, f32)[node]
    u = np.stack([(w[k] >> np.uint32(8)).astype(f32) * f32(2.0 ** -24) for k in (1, 2, 3)], axis=1)
    pos = (lo + (u * (up - lo).astype(f32)).astype(f32)).astype(f32)
    return pos, node


def sample_nodes(func, node_lower, node_upper, n_nodes, n_samples, width, seed, n_per_round,
                 max_rounds=100000):
    """The round loop of sample_surface (src/kd_tree.py:256-270) over
    sample_surface_iter (src/kd_tree.py:220-246): returns (points (n_samples, 3),
    rounds run, candidate values of the consumed candidates)."""
    if n_nodes < 1:
        raise ValueError("no nodes to sample from")
    out = np.zeros((n_samples, 3), f32)
    found = 0
    rounds = 0
    vals_all = []
    while found < n_samples:
        if rounds >= max_rounds:
            raise RuntimeError("sampling did not finish")
        pos, _ = candidates(seed, TAG_SURFACE, rounds * n_per_round, n_per_round, node_lower, node_upper,
                            n_nodes)
        val = np.asarray(func(pos), f32).reshape(-1)
        vals_all.append(val)
        ok = np.abs(val) < f32(width)
        take = pos[ok][: n_samples - found]
        out[found:found + take.shape[0]] = take
        found += take.shape[0]
        rounds += 1
    return out, rounds, np.concatenate(vals_all)


def sample_surface(func, lower, upper, n_samples, width, seed, n_node_thresh=4096):
    """src/kd_tree.py:249-272 (tree over |f| <= width, then rejection sampling)."""
    tree = construct_uniform_unknown_levelset_tree(func, lower, upper, node_terminate_thresh=n_node_thresh,
                                                   offset=width)
    n = int(tree["unknown_node_valid"].sum())
    n_per_round = min(3 * n_samples, 100000)
    pts, _, _ = sample_nodes(func, tree["unknown_node_lower"][:n], tree["unknown_node_upper"][:n], n, n_samples,
                             width, seed, n_per_round)
    return pts


def sample_surface_uniform(func, lower, upper, n_samples, width, seed):
    """src/kd_tree.py:321-336: the whole box as one node."""
    lo = np.asarray(lower, f32)[None, :]
    up = np.asarray(upper, f32)[None, :]
    pts, _, _ = sample_nodes(func, lo, up, 1, n_samples, width, seed, min(10 * n_samples, 100000))
    return pts


def bulk_sums(func, unk_lower, unk_upper, n_unknown, int_lower, int_upper, n_interior, n_samples, seed,
              chunk=1 << 18):
    """bulk_properties_sample_mass (src/kd_tree.py:804-833) and the interior
    sums of bulk_properties (src/kd_tree.py:846-853), accumulated in float64:
    (mass_interior, moment_interior (3,), n_inside, moment of inside samples (3,))."""
    il = np.asarray(int_lower, np.float64)[:n_interior]
    iu = np.asarray(int_upper, np.float64)[:n_interior]
    vol = np.prod(iu - il, axis=1)
    m_int = float(vol.sum())
    c_int = (vol[:, None] * 0.5 * (il + iu)).sum(axis=0)
    n_in = 0
    mom = np.zeros(3)
    if n_unknown > 0:
        for j0 in range(0, n_samples, chunk):
            n = min(chunk, n_samples - j0)
            pos, _ = candidates(seed, TAG_MASS, j0, n, unk_lower, unk_upper, n_unknown)
            inside = np.asarray(func(pos), f32).reshape(-1) < 0
            n_in += int(inside.sum())
            mom += pos[inside].astype(np.float64).sum(axis=0)
    return m_int, c_int, n_in, mom


def bulk_properties(func, lower, upper, seed, n_expand=10000, n_sample=1000000):
    """src/kd_tree.py:835-863: (mass, centroid (3,))."""
    tree = construct_uniform_unknown_levelset_tree(func, lower, upper, with_interior_nodes=True,
                                                   node_terminate_thresh=n_expand)
    nu = int(tree["unknown_node_valid"].sum())
    ni = int(tree["interior_node_valid"].sum())
    ul, uu = tree["unknown_node_lower"][:nu], tree["unknown_node_upper"][:nu]
    m_int, c_int, n_in, mom = bulk_sums(func, ul, uu, nu, tree["interior_node_lower"],
                                        tree["interior_node_upper"], ni, n_sample, seed)
    area = float(np.prod(uu.astype(np.float64) - ul, axis=1).sum())
    vps = area / n_sample
    mass = m_int + vps * n_in
    centroid = (c_int + vps * mom) / mass
    return mass, centroid
