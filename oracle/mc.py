"""Oracle: hierarchical marching cubes over the unknown leaves of the tree.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``hierarchical_marching_cubes`` (src/kd_tree.py:338-399) and
``extract_triangles_from_subcells`` / ``extract_triangles_from_cell``
(src/extract_cell.py:314-421).  The triangle table is Paul Bourke's public
marching-cubes table (the convention src/extract_cell.py:11-270 uses), held
here as one string per case: each hex digit is an edge id, three per triangle.
The edge -> vertex pairs and the vertex logical coordinates follow
src/extract_cell.py:274-304.
"""
import numpy as np

from .tree import construct_uniform_unknown_levelset_tree, get_next_bucket_size

f32 = np.float32

TRI_TABLE_HEX = [
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98",
    "3b2", "0b28b0", "19023b", "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b",
    "478", "430734", "019847", "419471731", "12a847", "34730412a", "92a902847", "2a9297273794",
    "8473b2", "b47b24204", "90184723b", "47b94b9b2921", "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba",
    "954", "954083", "054150", "854835315", "12a954", "30812a495", "52a542402", "2a5325354348",
    "95423b", "0b208b495", "05401523b", "21525828b485", "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b",
    "978579", "930953573", "078017157", "153357", "978957a12", "a12950530573", "802825857a52", "2a5253357",
    "7957893b2", "95797292027b", "23b018178157", "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5",
    "a65", "0835a6", "9015a6", "1831985a6", "165261", "165126308", "965906026", "598582526328",
    "23ba65", "b08b20a65", "01923b5a6", "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98",
    "5a6478", "43047365a", "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269",
    "3b2784a65", "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847", "65969b4797b9",
    "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264", "024426", "832824426",
    "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63", "8b1810b61914641", "3b6360064", "648b68",
    "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173", "126168189867", "269291679093739", "780706602", "732672",
    "23ba68a89867", "20727b09767a9a7", "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6",
    "76b", "308b76", "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98",
    "723627", "708760620", "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9",
    "684b86", "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93", "a946a4",
    "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495", "76b54a42a402", "348354325a52b76",
    "723762549", "954086062687", "362376150540", "628687218485158", "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a",
    "6956b9b89", "36b063056095", "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a", "a56",
    "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7", "75272b592328982",
    "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175", "903935537", "987597",
    "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458", "04b0b345b2b151b", "0250592b5458b85", "9452b3",
    "25a352345384", "5a2524420", "3a235a385458019", "5a2524192942", "845853351", "045105", "845853905035", "945",
    "4b749b9ab", "0834979b79ab", "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743", "487",
    "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", "",
]

# edge e joins vertices EDGE_VERTS[e] (src/extract_cell.py:274-289)
EDGE_VERTS = np.array([(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                       (0, 4), (1, 5), (2, 6), (3, 7)], dtype=np.int64)
# vertex v sits at the logical corner VERT_LOGICAL[v] (src/extract_cell.py:293-304)
VERT_LOGICAL = np.array([(0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 0),
                         (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)], dtype=bool)


def tri_table():
    """(256, 5, 3) edge ids, -1 where a case has fewer triangles."""
    out = -np.ones((256, 5, 3), dtype=np.int64)
    for c, s in enumerate(TRI_TABLE_HEX):
        for i in range(len(s) // 3):
            out[c, i] = [int(ch, 16) for ch in s[3 * i:3 * i + 3]]
    return out


def extract_triangles_from_cells(vert_pos, vert_vals):
    """src/extract_cell.py:314-364, batched over cells.

    vert_pos (M, 8, 3), vert_vals (M, 8) -> tris (M, 5, 3, 3), valid (M, 5)."""
    tab = tri_table()
    a = EDGE_VERTS[:, 0]
    b = EDGE_VERTS[:, 1]
    va, vb = vert_vals[:, a], vert_vals[:, b]                 # (M, 12)
    pa, pb = vert_pos[:, a, :], vert_pos[:, b, :]             # (M, 12, 3)
    with np.errstate(divide="ignore", invalid="ignore"):
        tc = -va / (vb - va)
    tc = np.nan_to_num(tc, nan=0.0)
    tc = np.clip(tc, f32(0), f32(1)).astype(vert_vals.dtype)
    cross = (f32(1) - tc)[:, :, None] * pa + tc[:, :, None] * pb
    bits = (vert_vals < 0).astype(np.int64) * (1 << np.arange(8))[None, :]
    case = bits.sum(axis=1)
    tris = tab[case]                                          # (M, 5, 3)
    valid = tris[:, :, 0] != -1
    ids = np.maximum(tris, 0)
    M = vert_vals.shape[0]
    pos = cross[np.arange(M)[:, None, None], ids, :]           # (M, 5, 3, 3)
    return pos, valid


def extract_triangles_from_subcells(func, lower, upper, n_sub_depth):
    """src/extract_cell.py:366-421 batched over leaves (L, 3)."""
    L = lower.shape[0]
    n = 2 ** n_sub_depth
    npt = n + 1
    t = np.arange(npt, dtype=f32)
    step = (upper - lower) / f32(n)
    coords = lower[:, None, :] + t[None, :, None] * step[:, None, :]   # (L, npt, 3)
    coords[:, -1, :] = upper
    gx = coords[:, :, 0][:, :, None, None]
    gy = coords[:, :, 1][:, None, :, None]
    gz = coords[:, :, 2][:, None, None, :]
    shape = (L, npt, npt, npt)
    grid = np.stack(np.broadcast_arrays(gx, gy, gz), axis=-1).reshape(shape + (3,))
    vals = func(grid.reshape(-1, 3)).reshape(shape)
    ii = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"),
                  -1).reshape(-1, 3)                                     # (S, 3)
    delta = (upper - lower) / f32(n)
    sub_lo = lower[:, None, :] + ii[None, :, :].astype(f32) * delta[:, None, :]
    sub_up = sub_lo + delta[:, None, :]
    vi = ii[:, None, :] + VERT_LOGICAL[None, :, :].astype(np.int64)     # (S, 8, 3)
    vv = vals[:, vi[:, :, 0], vi[:, :, 1], vi[:, :, 2]]                 # (L, S, 8)
    vpos = np.where(VERT_LOGICAL[None, None, :, :], sub_up[:, :, None, :], sub_lo[:, :, None, :])
    S = ii.shape[0]
    pos, valid = extract_triangles_from_cells(vpos.reshape(L * S, 8, 3), vv.reshape(L * S, 8))
    return pos.reshape(L, S * 5, 3, 3), valid.reshape(L, S * 5)


def hierarchical_marching_cubes(func, lower, upper, depth, n_subcell_depth=2,
                                extract_batch_max_tri_out=1000000):
    """src/kd_tree.py:357-399: (T, 3, 3) triangle vertex positions."""
    tree = construct_uniform_unknown_levelset_tree(
        func, lower, upper, split_depth=3 * (depth - n_subcell_depth))
    n_valid = tree["stats"]["n_unknown"]
    lo = tree["unknown_node_lower"][:n_valid]
    up = tree["unknown_node_upper"][:n_valid]
    out = []
    chunk = 4096
    for i in range(0, n_valid, chunk):
        pos, valid = extract_triangles_from_subcells(func, lo[i:i + chunk], up[i:i + chunk],
                                                     n_subcell_depth)
        out.append(pos[valid])
    if not out:
        return np.zeros((0, 3, 3), f32), tree
    return np.concatenate(out).astype(f32), tree
