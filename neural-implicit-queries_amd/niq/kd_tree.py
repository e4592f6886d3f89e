"""Spatial queries over a neural implicit: level-set tree (device resident).

``construct_uniform_unknown_levelset_tree`` keeps the reference's signature and
return keys (src/kd_tree.py:103-218).  The whole breadth-first build runs on
the GPU (niq_tree_build): per level one fused classify kernel plus three small
compaction kernels, with node counts kept on the device, so the host waits
once per tree instead of twice per level.  Output order, padding sizes and
valid masks follow the reference (see ``oracle/tree.py`` for the restated
ordering rules).
"""
import ctypes

import torch

from . import _lib
from . import mlp as _mlp
from .bucketing import bucket_sizes, get_next_bucket_size

INVALID_IND = 2 ** 30


def _finished_size(level_counts, level_finished, B):
    """Size of the interior/exterior arrays after the reference's doubling
    (src/kd_tree.py:126-131, 157-164)."""
    size = B
    for n_curr, n_fin in zip(level_counts, level_finished):
        while size - n_fin < n_curr:
            size *= 2
    return size


def _pad_rows(src, n, size):
    out = torch.zeros((size, 3), dtype=torch.float32, device=src.device)
    m = min(n, size)
    if m > 0:
        out[:m] = src[:m]
    return out


def tree_request(func, params, lower, upper, node_terminate_thresh=None, split_depth=None,
                 with_interior_nodes=False, with_exterior_nodes=False, offset=0., batch_process_size=2048):
    """Validate a tree query (src/kd_tree.py:105-111) and lower it to the ABI:
    returns (device model, TreeParams, level capacity, finished-list capacity)."""
    for b in bucket_sizes:
        if b > batch_process_size and (b // batch_process_size) * batch_process_size != b:
            raise ValueError(f"batch_process_size must be a factor of our bucket sizes, is not a factor of {b} (try a power of 2)")
    if node_terminate_thresh is None and split_depth is None:
        raise ValueError("must specify at least one of node_terminate_thresh or split_depth as a terminating condition")
    _lib.require_gpu()
    m = _mlp.device_model(params)
    lower = [float(v) for v in torch.as_tensor(lower).reshape(-1).tolist()]
    upper = [float(v) for v in torch.as_tensor(upper).reshape(-1).tolist()]
    if len(lower) != 3 or len(upper) != 3:
        raise ValueError("the level-set tree is built over 3D boxes")
    p = _lib.TreeParams()
    for i in range(3):
        p.lower[i] = lower[i]
        p.upper[i] = upper[i]
    p.split_depth = -1 if split_depth is None else int(split_depth)
    p.node_terminate_thresh = 0 if node_terminate_thresh is None else int(node_terminate_thresh)
    p.offset = float(offset)
    p.batch_process_size = int(batch_process_size)
    p.with_interior_nodes = int(bool(with_interior_nodes))
    p.with_exterior_nodes = int(bool(with_exterior_nodes))
    # level capacity = max nodes of any level (the workspace holds two level lists)
    cap = 1 << p.split_depth if p.split_depth >= 0 else 2 * p.node_terminate_thresh
    if p.node_terminate_thresh > 0:
        cap = min(cap, 2 * p.node_terminate_thresh)
    cap_fin = min(2 * cap, 1 << 26)
    return m, p, cap, cap_fin


class TreeBuffers:
    """Caller-owned output buffers of one tree build and their ABI view."""

    def __init__(self, p, cap, cap_fin, dev):
        self.p = p
        self.unk = torch.empty((2, cap, 3), dtype=torch.float32, device=dev)
        self.fin_i = torch.empty((2, cap_fin, 3), dtype=torch.float32, device=dev) if p.with_interior_nodes else None
        self.fin_e = torch.empty((2, cap_fin, 3), dtype=torch.float32, device=dev) if p.with_exterior_nodes else None
        r = _lib.TreeResult()
        r.unknown_lower, r.unknown_upper, r.unknown_capacity = self.unk[0].data_ptr(), self.unk[1].data_ptr(), cap
        if self.fin_i is not None:
            r.interior_lower, r.interior_upper = self.fin_i[0].data_ptr(), self.fin_i[1].data_ptr()
            r.interior_capacity = cap_fin
        if self.fin_e is not None:
            r.exterior_lower, r.exterior_upper = self.fin_e[0].data_ptr(), self.fin_e[1].data_ptr()
            r.exterior_capacity = cap_fin
        self.r = r

    def outputs(self):
        """The reference's return dict (src/kd_tree.py:202-218) and build stats."""
        r, p = self.r, self.p
        dev = self.unk.device
        n_lv = r.n_levels
        level_n = list(r.level_n[:n_lv])
        B = int(p.batch_process_size)
        n_unk = int(r.n_unknown)
        bucket = get_next_bucket_size(n_unk)
        out = {
            'unknown_node_valid': torch.arange(bucket, device=dev) < n_unk,
            'unknown_node_lower': _pad_rows(self.unk[0], n_unk, bucket),
            'unknown_node_upper': _pad_rows(self.unk[1], n_unk, bucket),
        }
        if self.fin_i is not None:
            n = int(r.n_interior)
            size = _finished_size(level_n, list(r.level_interior[:n_lv]), B)
            out['interior_node_valid'] = torch.arange(size, device=dev) < n
            out['interior_node_lower'] = _pad_rows(self.fin_i[0], n, size)
            out['interior_node_upper'] = _pad_rows(self.fin_i[1], n, size)
        if self.fin_e is not None:
            n = int(r.n_exterior)
            size = _finished_size(level_n, list(r.level_exterior[:n_lv]), B)
            out['exterior_node_valid'] = torch.arange(size, device=dev) < n
            out['exterior_node_lower'] = _pad_rows(self.fin_e[0], n, size)
            out['exterior_node_upper'] = _pad_rows(self.fin_e[1], n, size)
        stats = {"n_classified": int(r.n_classified), "n_unknown": n_unk,
                 "n_interior": int(r.n_interior), "n_exterior": int(r.n_exterior),
                 "n_recheck": int(r.n_recheck), "level_counts": level_n}
        return out, stats


def build_tree(func, params, lower, upper, node_terminate_thresh=None, split_depth=None,
               with_interior_nodes=False, with_exterior_nodes=False, offset=0.,
               batch_process_size=2048, stream=None):
    """Device tree build; returns (out_dict, stats)."""
    m, p, cap, cap_fin = tree_request(func, params, lower, upper, node_terminate_thresh, split_depth,
                                      with_interior_nodes, with_exterior_nodes, offset, batch_process_size)
    lib = _lib.load_library()
    ws_bytes = ctypes.c_size_t()
    _lib.check(lib.niq_tree_workspace_bytes(m.handle, ctypes.byref(func.cfg), ctypes.byref(p),
                                           ctypes.byref(ws_bytes)))
    dev = torch.device("cuda")
    ws = torch.empty(ws_bytes.value, dtype=torch.uint8, device=dev)
    s = _lib.stream_handle(stream)
    while True:
        bufs = TreeBuffers(p, cap, cap_fin, dev)
        r = bufs.r
        st = lib.niq_tree_build(m.handle, ctypes.byref(func.cfg), ctypes.byref(p), _lib.ptr(ws),
                                ws_bytes.value, ctypes.byref(r), s)
        if st == _lib.NIQ_ERR_CAPACITY and max(r.n_interior, r.n_exterior) > cap_fin:
            cap_fin = 1 << int(max(r.n_interior, r.n_exterior) - 1).bit_length()
            continue
        _lib.check(st)
        break
    return bufs.outputs()


def construct_uniform_unknown_levelset_tree(func, params, lower, upper, node_terminate_thresh=None,
                                            split_depth=None, compress_after=False, with_childern=False,
                                            with_interior_nodes=False, with_exterior_nodes=False,
                                            offset=0., batch_process_size=2048):
    """src/kd_tree.py:103-218 (``compress_after`` and ``with_childern`` are
    unused there too)."""
    out, _ = build_tree(func, params, lower, upper, node_terminate_thresh, split_depth,
                        with_interior_nodes, with_exterior_nodes, offset, batch_process_size)
    return out


def closest_point(func, params, lower, upper, query_points, eps=0.001, batch_process_size=2048):
    """src/kd_tree.py:766-802: (query_min_dist (Q,), query_min_loc (Q, 3)).

    Runs the reference's batch-stack search on the device (niq_closest_point);
    the round order is the reference's, so results match it."""
    _lib.require_gpu()
    lib = _lib.load_library()
    m = _mlp.device_model(params)
    q = torch.as_tensor(query_points).to(device="cuda", dtype=torch.float32).contiguous()
    Q = q.shape[0]
    lo = (ctypes.c_float * 3)(*[float(v) for v in torch.as_tensor(lower).reshape(-1).tolist()])
    up = (ctypes.c_float * 3)(*[float(v) for v in torch.as_tensor(upper).reshape(-1).tolist()])
    dist = torch.empty(Q, dtype=torch.float32, device=q.device)
    loc = torch.empty((Q, 3), dtype=torch.float32, device=q.device)
    n_proc = ctypes.c_int64(0)
    _lib.check(lib.niq_closest_point(m.handle, ctypes.byref(func.cfg), lo, up, _lib.ptr(q), Q, float(eps),
                                     int(batch_process_size), _lib.ptr(dist), _lib.ptr(loc),
                                     ctypes.byref(n_proc), _lib.stream_handle()))
    closest_point.last_n_processed = int(n_proc.value)
    return dist, loc


def find_any_intersection(func_tuple, params_tuple, lower, upper, eps, viz_nodes=False):
    """src/kd_tree.py:558-655: (found, 1, 2, loc (3,)) for a pair of
    functions; loc is (-777, -777, -777) when none is found.  The node
    visualisation output (viz_nodes) is not provided."""
    if len(func_tuple) != 2 or len(params_tuple) != 2:
        raise ValueError("intersection supports pairwise only as written")
    if viz_nodes:
        raise NotImplementedError("viz_nodes output is not provided")
    _lib.require_gpu()
    lib = _lib.load_library()
    mA, mB = _mlp.device_model(params_tuple[0]), _mlp.device_model(params_tuple[1])
    lo = (ctypes.c_float * 3)(*[float(v) for v in torch.as_tensor(lower).reshape(-1).tolist()])
    up = (ctypes.c_float * 3)(*[float(v) for v in torch.as_tensor(upper).reshape(-1).tolist()])
    found = ctypes.c_int32(0)
    loc = (ctypes.c_float * 3)()
    n_proc = ctypes.c_int64(0)
    _lib.check(lib.niq_find_any_intersection(mA.handle, ctypes.byref(func_tuple[0].cfg), mB.handle,
                                             ctypes.byref(func_tuple[1].cfg), lo, up, float(eps),
                                             ctypes.byref(found), loc, ctypes.byref(n_proc),
                                             _lib.stream_handle()))
    find_any_intersection.last_n_processed = int(n_proc.value)
    if found.value:
        return True, 1, 2, torch.tensor(list(loc), dtype=torch.float32)
    return False, 0, 0, torch.tensor([-777., -777., -777.], dtype=torch.float32)


def _seed(rngkey):
    """A 64-bit seed from an int or a two-word key laid out like
    jax.random.PRNGKey's ([high word, low word])."""
    if isinstance(rngkey, int):
        return rngkey & 0xFFFFFFFFFFFFFFFF
    words = [int(v) & 0xFFFFFFFF for v in torch.as_tensor(rngkey).reshape(-1).tolist()]
    if len(words) != 2:
        raise ValueError("rngkey must be an int or a pair of 32-bit words")
    return (words[0] << 32) | words[1]


def _valid_nodes(node_valid, node_lower, node_upper):
    """The valid nodes come first (src/kd_tree.py:225-227)."""
    n = int(torch.as_tensor(node_valid).sum().item())
    lo = torch.as_tensor(node_lower).to(device="cuda", dtype=torch.float32)[:n].contiguous()
    up = torch.as_tensor(node_upper).to(device="cuda", dtype=torch.float32)[:n].contiguous()
    return n, lo, up


def sample_nodes(params, node_lower, node_upper, n_nodes, n_samples, width, rngkey, n_samples_per_round,
                 found_sample_points=None, found_start_ind=0, first_candidate=0, max_rounds=100000):
    """Rounds of sample_surface_iter (src/kd_tree.py:220-246) until n_samples
    points with |f| < width are found: returns (points (n_samples, 3), found,
    rounds).  Candidates are drawn from the Philox stream of
    niq_sample_surface_nodes; the points are the first accepted candidates of
    that stream whatever the round size."""
    _lib.require_gpu()
    lib = _lib.load_library()
    m = _mlp.device_model(params)
    if n_nodes < 1:
        raise ValueError("no nodes to sample from")
    out = found_sample_points
    if out is None:
        out = torch.zeros((n_samples, 3), dtype=torch.float32, device="cuda")
    if out.shape != (n_samples, 3) or out.dtype != torch.float32 or not out.is_cuda or not out.is_contiguous():
        raise ValueError("found_sample_points must be a contiguous (n_samples, 3) float32 device tensor")
    ws = torch.empty(lib.niq_sample_workspace_bytes(int(n_samples_per_round)), dtype=torch.uint8, device="cuda")
    found = ctypes.c_int64(int(found_start_ind))
    rounds = ctypes.c_int64(0)
    _lib.check(lib.niq_sample_surface_nodes(m.handle, _lib.ptr(node_lower), _lib.ptr(node_upper), int(n_nodes),
                                            float(width), _seed(rngkey), int(first_candidate),
                                            int(n_samples_per_round), int(max_rounds), _lib.ptr(out),
                                            int(n_samples), ctypes.byref(found), ctypes.byref(rounds),
                                            _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
    return out, int(found.value), int(rounds.value)


def _sample_all(params, lo, up, n, n_samples, width, rngkey, per_round, max_rounds=100000):
    """The reference's loop runs until the output is full (src/kd_tree.py:260-270);
    this one gives up after max_rounds rounds."""
    pts, found, rounds = sample_nodes(params, lo, up, n, n_samples, width, rngkey, per_round, max_rounds=max_rounds)
    if found < n_samples:
        raise _lib.CapacityError(_lib.NIQ_ERR_CAPACITY,
                                 f"found {found} of {n_samples} samples with |f| < {width} in {rounds} rounds")
    return pts, rounds


def sample_surface_iter(func, params, n_samples_per_round, width, rngkey, u_node_valid, u_node_lower, u_node_upper,
                        found_sample_points, found_start_ind):
    """One round (src/kd_tree.py:220-246): (found_sample_points, found_start_ind).
    Like the reference, the same key draws the same candidates; pass a new key
    per round."""
    n, lo, up = _valid_nodes(u_node_valid, u_node_lower, u_node_upper)
    pts, found, _ = sample_nodes(params, lo, up, n, found_sample_points.shape[0], width, rngkey,
                                 n_samples_per_round, found_sample_points, int(found_start_ind), max_rounds=1)
    return pts, found


def _round_size(n_samples, per_sample):
    # Any round size gives the same points; large rounds keep the chip busy.
    return max(1024, min(per_sample * n_samples, 1 << 22))


def sample_surface(func, params, lower, upper, n_samples, width, rngkey, n_node_thresh=4096):
    """src/kd_tree.py:249-272: a tree over |f| <= width (offset=width), then
    rejection sampling from its unknown nodes.  Returns (n_samples, 3)."""
    out = construct_uniform_unknown_levelset_tree(func, params, lower, upper, node_terminate_thresh=n_node_thresh,
                                                  offset=width)
    n, lo, up = _valid_nodes(out["unknown_node_valid"], out["unknown_node_lower"], out["unknown_node_upper"])
    pts, rounds = _sample_all(params, lo, up, n, n_samples, width, rngkey, _round_size(n_samples, 3))
    sample_surface.last_rounds = rounds
    return pts


def sample_surface_uniform_iter(func, params, n_samples_per_round, width, rngkey, lower, upper,
                                found_sample_points, found_start_ind):
    """One round over the whole box (src/kd_tree.py:276-301)."""
    lo = torch.as_tensor(lower).to(device="cuda", dtype=torch.float32).reshape(1, 3).contiguous()
    up = torch.as_tensor(upper).to(device="cuda", dtype=torch.float32).reshape(1, 3).contiguous()
    pts, found, _ = sample_nodes(params, lo, up, 1, found_sample_points.shape[0], width, rngkey,
                                 n_samples_per_round, found_sample_points, int(found_start_ind), max_rounds=1)
    return pts, found


def sample_surface_uniform(func, params, lower, upper, n_samples, width, rngkey):
    """src/kd_tree.py:321-336: rejection sampling over the whole box."""
    lo = torch.as_tensor(lower).to(device="cuda", dtype=torch.float32).reshape(1, 3).contiguous()
    up = torch.as_tensor(upper).to(device="cuda", dtype=torch.float32).reshape(1, 3).contiguous()
    pts, rounds = _sample_all(params, lo, up, 1, n_samples, width, rngkey, _round_size(n_samples, 10))
    sample_surface_uniform.last_rounds = rounds
    return pts


def bulk_sums(params, unknown_lower, unknown_upper, n_unknown, interior_lower, interior_upper, n_interior,
              n_samples, rngkey):
    """niq_bulk_sums_nodes: the float64 sums behind bulk_properties."""
    _lib.require_gpu()
    lib = _lib.load_library()
    m = _mlp.device_model(params)
    out = _lib.BulkSums()
    _lib.check(lib.niq_bulk_sums_nodes(m.handle, _lib.ptr(unknown_lower) if n_unknown else None,
                                       _lib.ptr(unknown_upper) if n_unknown else None, int(n_unknown),
                                       _lib.ptr(interior_lower) if n_interior else None,
                                       _lib.ptr(interior_upper) if n_interior else None, int(n_interior),
                                       int(n_samples), _seed(rngkey), ctypes.byref(out), _lib.stream_handle()))
    return out


def bulk_properties_sample_mass(func, params, node_valid, node_lower, node_upper, n_samples, rngkey):
    """src/kd_tree.py:804-833: (mass, centroid moment (3,)) of the samples
    inside (f < 0) drawn from the valid nodes; the centroid is not divided by
    the mass, as in the reference."""
    n, lo, up = _valid_nodes(node_valid, node_lower, node_upper)
    s = bulk_sums(params, lo, up, n, None, None, 0, n_samples, rngkey)
    vps = s.unknown_volume / n_samples if n_samples > 0 else 0.0
    mass = torch.tensor(vps * s.n_inside, dtype=torch.float32)
    centroid = torch.tensor([vps * v for v in s.inside_moment], dtype=torch.float32)
    return mass, centroid


def bulk_properties(func, params, lower, upper, rngkey, n_expand=int(1e4), n_sample=int(1e6)):
    """src/kd_tree.py:835-863: (mass, centroid (3,)) of the region f < 0:
    interior boxes exactly, unknown boxes by n_sample mass samples."""
    out = construct_uniform_unknown_levelset_tree(func, params, lower, upper, with_interior_nodes=True,
                                                  node_terminate_thresh=n_expand)
    nu, ulo, uup = _valid_nodes(out["unknown_node_valid"], out["unknown_node_lower"], out["unknown_node_upper"])
    ni, ilo, iup = _valid_nodes(out["interior_node_valid"], out["interior_node_lower"], out["interior_node_upper"])
    s = bulk_sums(params, ulo, uup, nu, ilo, iup, ni, n_sample, rngkey)
    vps = s.unknown_volume / n_sample if n_sample > 0 else 0.0
    mass = s.interior_volume + vps * s.n_inside
    # an empty region gives NaN like the reference's 0 / 0
    cen = [(s.interior_moment[k] + vps * s.inside_moment[k]) / mass if mass else float("nan") for k in range(3)]
    return torch.tensor(mass, dtype=torch.float32), torch.tensor(cen, dtype=torch.float32)


def extract_cells(params, node_lower, node_upper, n_leaves, n_subcell_depth):
    """Marching cubes over the first n_leaves boxes: (T, 3, 3) triangles."""
    lib = _lib.load_library()
    m = _mlp.device_model(params)
    lo = node_lower[:n_leaves].contiguous()
    up = node_upper[:n_leaves].contiguous()
    cap = max(1024, n_leaves * 8)
    while True:
        tri = torch.empty((cap, 3, 3), dtype=torch.float32, device=lo.device)
        n = ctypes.c_int64(0)
        st = lib.niq_mc_extract(m.handle, _lib.ptr(lo), _lib.ptr(up), n_leaves, int(n_subcell_depth),
                                _lib.ptr(tri), cap, ctypes.byref(n), _lib.stream_handle())
        if st == _lib.NIQ_ERR_CAPACITY:
            cap = int(n.value)
            continue
        _lib.check(st)
        return tri[:int(n.value)]


def hierarchical_marching_cubes(func, params, lower, upper, depth, n_subcell_depth=2,
                                extract_batch_max_tri_out=1000000):
    """src/kd_tree.py:357-399: tree to split depth 3 (depth - n_subcell_depth),
    then dense marching cubes on the 2^s subgrid of every unknown leaf.
    ``extract_batch_max_tri_out`` only sized the reference's batches; the
    device extraction chunks by memory instead."""
    out, st = build_tree(func, params, lower, upper, split_depth=3 * (depth - n_subcell_depth))
    n = st["n_unknown"]
    if n == 0:
        return torch.zeros((0, 3, 3), dtype=torch.float32, device="cuda")
    return extract_cells(params, out['unknown_node_lower'], out['unknown_node_upper'], n, n_subcell_depth)
