"""The oracle's BSP cluster lookup, PVS decompression and PVS-culled
MakeTransfers pinned by hand-derived cases and invariants (no GPU)."""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import scenes


def _bsp(planes, nodes, leaves):
    pl = np.zeros(len(planes), oracle.PLANE_DTYPE)
    for k, (n, d) in enumerate(planes):
        pl[k] = (n, d)
    nd = np.zeros(len(nodes), oracle.NODE_DTYPE)
    for k, (p, c0, c1) in enumerate(nodes):
        nd[k] = (p, (c0, c1))
    return pl, nd, np.array(leaves, np.int16)


# ---- ClusterFromPoint / PointInLeaf ---------------------------------------
def test_cluster_from_point_hand_cases():
    # node 0: x = 0, front -> leaf 0 (cluster 5), back -> node 1
    # node 1: y = 0, front -> leaf 1 (cluster -1), back -> leaf 2 (cluster 7)
    pl, nd, lc = _bsp([((1, 0, 0), 0.0), ((0, 1, 0), 0.0)],
                      [(0, -1, 1), (1, -2, -3)], [5, -1, 7])
    pts = [(1, 0, 0), (-1, 1, 0), (-1, -1, 0),
           (0.03125, -9, 0),        # exactly TEST_EPSILON: not in front -> both sides, leaf 0 first
           (0.0313, -9, 0),         # just past it: in front
           (-0.03125, -9, 0),       # within on the back side: leaf 0 (cluster 5) is tried first
           (-0.0313, -9, 0),        # behind: node 1, y = -9 -> 7
           (-0.0313, 0.02, 0)]      # behind, then within y: leaf 1 (-1), then leaf 2 -> 7
    got = oracle.cluster_from_points(pl, nd, lc, np.array(pts, np.float32))
    assert list(got) == [5, -1, 7, 5, 5, 5, 7, 7]


def test_cluster_from_point_minus_one_falls_through():
    # same tree with leaf 0 in solid (-1): a point on x = 0 continues behind
    pl, nd, lc = _bsp([((1, 0, 0), 0.0), ((0, 1, 0), 0.0)],
                      [(0, -1, 1), (1, -2, -3)], [-1, -1, 7])
    got = oracle.cluster_from_points(pl, nd, lc, np.array(
        [(0, -5, 0), (0, 0, 0), (0, 5, 0), (3, 0, 0)], np.float32))
    # (0,-5): 0 -> -1, then node 1 behind -> 7; (0,0): -1, -1, 7; (0,5): -1
    # then node 1 in front -> -1; (3,0): in front -> -1
    assert list(got) == [7, 7, -1, -1]
    # only -1 falls through: a leaf of cluster -2 is returned as found
    pl, nd, lc = _bsp([((1, 0, 0), 0.0)], [(0, -1, -2)], [-2, 4])
    assert list(oracle.cluster_from_points(pl, nd, lc, np.zeros((1, 3), np.float32))) == [-2]


def test_cluster_from_point_dot_order():
    # dist = ((x nx + y ny) + z nz) - d, one FP32 rounding per step: for
    # (2^24, 1, -2^24) and normal (1, 1, 1), 2^24 + 1 rounds to 2^24 (tie to
    # even), so the dot is 0 and dist = -0.5 against d = 0.5: behind (cluster
    # 9).  The exact dot is 1 (dist +0.5, in front, cluster 3), and so is the
    # order x + (y + z).
    pl, nd, lc = _bsp([((1, 1, 1), 0.5)], [(0, -1, -2)], [3, 9])
    got = oracle.cluster_from_points(pl, nd, lc, np.array([(2.0**24, 1.0, -2.0**24)], np.float32))
    assert list(got) == [9]


def test_cluster_from_point_brute_force_on_random_trees():
    """Against an independent evaluation: walk every within-epsilon branch
    in float64 of the FP32 dist and take the first non -1 leaf."""
    rng = np.random.default_rng(5)
    for _ in range(4):
        pl, nd, lc = H.random_bsp(rng)
        pts = H.bsp_points(rng, pl, 600)

        def walk(node, p):
            if node < 0:
                return int(lc[-1 - node])
            n = nd[node]
            pp = pl[n["planenum"]]
            v = np.float32(np.float32(p[0] * pp["normal"][0]) + np.float32(p[1] * pp["normal"][1]))
            v = np.float32(v + np.float32(p[2] * pp["normal"][2]))
            d = np.float32(v - pp["dist"])
            if d > np.float32(0.03125):
                return walk(int(n["children"][0]), p)
            if d < -np.float32(0.03125):
                return walk(int(n["children"][1]), p)
            r = walk(int(n["children"][0]), p)
            return r if r != -1 else walk(int(n["children"][1]), p)

        want = [walk(0, p) for p in pts]
        assert list(oracle.cluster_from_points(pl, nd, lc, pts)) == want


def test_patch_clusters_winding_fallback():
    # x = 0 splits cluster 2 (front) from solid (-1, back)
    pl, nd, lc = _bsp([((1, 0, 0), 0.0)], [(0, -1, -2)], [2, -1])
    p = np.zeros(3, oracle.PATCH_DTYPE)
    p["origin"] = [(-5, 0, 0), (-5, 0, 0), (7, 0, 0)]
    p["n_points"] = [4, 2, 0]
    p["points"][0, :4] = [(-9, 0, 0), (-1, 0, 0), (4, 0, 0), (-3, 0, 0)]  # third point is in 2
    p["points"][1, :2] = [(-9, 0, 0), (-1, 0, 0)]                        # none is
    assert list(oracle.patch_clusters(pl, nd, lc, p)) == [2, -1, 2]


def test_grid_bsp_puts_patches_in_their_rooms():
    s = scenes.level(seed=4, rooms_x=3, rooms_y=3, chop=64.0, tess=64.0)
    bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y)
    c = oracle.cluster_from_points(bsp.planes, bsp.nodes, bsp.leaf_cluster, s.p_origin)
    x, y = s.p_origin[:, 0].astype(np.float64), s.p_origin[:, 1].astype(np.float64)
    cx = np.searchsorted(s.grid_x, x, side="left") - 1
    cy = np.searchsorted(s.grid_y, y, side="left") - 1
    # points strictly inside a cell belong to it; those on a boundary line
    # (door tunnels) take the front side (the larger cell index)
    on_x = np.isin(x, s.grid_x)
    on_y = np.isin(y, s.grid_y)
    cx = np.where(on_x, np.searchsorted(s.grid_x, x, side="right") - 1, cx)
    cy = np.where(on_y, np.searchsorted(s.grid_y, y, side="right") - 1, cy)
    want = cy * 3 + cx
    assert np.array_equal(c.astype(np.int64), want)


# ---- DecompressVis ---------------------------------------------------------
def test_decompress_vis_known_answers():
    # 40 clusters: 5-byte rows
    data = np.array([0x01, 0x00, 0x03, 0x80,     # row 0: 01 00 00 00 80
                     0x00, 0x05,                 # row 1: all zero
                     0xFF, 0x00, 0x09,           # row 2: FF, then a run of 9 clipped to 4
                     0x12, 0x34, 0x56, 0x78, 0x9A], np.uint8)
    ofs = [0, 4, 6, 9, 0] + [9] * 35  # rows may share their data
    rows, over = oracle.decompress_vis(data, ofs, 40)
    assert rows[:5].tolist() == [[1, 0, 0, 0, 0x80], [0] * 5, [0xFF, 0, 0, 0, 0],
                                 [0x12, 0x34, 0x56, 0x78, 0x9A], [1, 0, 0, 0, 0x80]]
    assert (rows[5:] == rows[3]).all()
    assert over == 1
    with pytest.raises(ValueError, match="zero repeat"):
        oracle.decompress_vis(np.array([0, 0], np.uint8), [0], 8)
    with pytest.raises(ValueError, match="ends inside"):
        oracle.decompress_vis(np.array([1, 2], np.uint8), [0], 40)


def test_compress_decompress_round_trip():
    rng = np.random.default_rng(9)
    for nc, dens in ((1, 1.0), (8, 0.5), (9, 0.3), (300, 0.02), (700, 0.6), (2049, 0.001)):
        bits = rng.uniform(size=(nc, nc)) < dens
        rows = np.packbits(bits, axis=1, bitorder="little")
        data, ofs = scenes.compress_vis(rows)
        got, over = oracle.decompress_vis(data, ofs, nc)
        assert over == 0 and np.array_equal(got, rows)


# ---- MakeTransfers with a PVS ------------------------------------------------
@pytest.fixture(scope="module")
def small_level():
    s = scenes.level(seed=6, rooms_x=2, rooms_y=2, chop=128.0, tess=128.0, n_lights=1)
    bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y)
    cl = oracle.cluster_from_points(bsp.planes, bsp.nodes, bsp.leaf_cluster, s.p_origin)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    return s, vis, cl.astype(np.int32), tri, kd


def test_transfers_pvs_all_visible_equals_plain(small_level):
    s, vis, cl, tri, kd = small_level
    full = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area, threads=8)
    ones = np.full((vis.n_clusters, vis.n_clusters), True)
    rows = np.packbits(ones, axis=1, bitorder="little")
    got = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area, threads=8,
                                cluster=cl, pvs_rows=rows)
    for a, b in zip(full, got):
        assert np.array_equal(a, b)


def test_transfers_pvs_culls_exactly_the_invisible_cluster_pairs(small_level):
    s, vis, cl, tri, kd = small_level
    rp, col, coef, _ = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area,
                                             threads=8, normalize=False)
    # a PVS with only the diagonal and one cross pair (0 sees 3, one way)
    nc = vis.n_clusters
    bits = np.eye(nc, dtype=bool)
    bits[0, 3] = True
    rows = np.packbits(bits, axis=1, bitorder="little")
    g = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area, threads=8,
                              normalize=False, cluster=cl, pvs_rows=rows)
    row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    a, b = cl[row], cl[col]
    keep = (a == b) | ((a == 0) & (b == 3)) | ((a == 3) & (b == 0))
    assert np.array_equal(g[1], col[keep])
    assert np.array_equal(g[2].view(np.uint32), coef[keep].view(np.uint32))
    assert np.array_equal(np.diff(g[0]), np.bincount(row[keep], minlength=len(rp) - 1))
    # a negative cluster sees everything
    cl2 = cl.copy()
    cl2[::5] = -1
    g2 = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area, threads=8,
                               normalize=False, cluster=cl2, pvs_rows=rows)
    a, b = cl2[row], cl2[col]
    keep2 = (a == b) & (a >= 0) | ((a == 0) & (b == 3)) | ((a == 3) & (b == 0)) | (a < 0) | (b < 0)
    assert np.array_equal(g2[1], col[keep2])


def test_grid_bsp_structure():
    """The generated BSP is a tree within the walk's depth bound, every room
    is exactly one leaf, and each room's PVS row holds itself and its grid
    neighbours (one doorway)."""
    xs = np.array([0.0, 300, 700, 1000, 1600])
    ys = np.array([0.0, 500, 900])
    bsp, vis = scenes.grid_bsp(xs, ys, pvs_hops=1)
    nd = bsp.nodes
    seen, stack, depth = set(), [(0, 1)], 0
    while stack:
        n, d = stack.pop()
        assert n not in seen
        seen.add(n)
        depth = max(depth, d)
        for c in nd[n]["children"]:
            if c >= 0:
                stack.append((int(c), d + 1))
    assert depth <= 64
    assert sorted(c for c in bsp.leaf_cluster if c >= 0) == list(range(8))
    rows, over = oracle.decompress_vis(vis.data, vis.pvs_ofs, vis.n_clusters)
    bits = np.unpackbits(rows, axis=1, bitorder="little")[:, :8].astype(bool)
    for c in range(8):
        r, k = divmod(c, 4)
        want = {c} | {rr * 4 + kk for rr, kk in ((r, k - 1), (r, k + 1), (r - 1, k), (r + 1, k))
                      if 0 <= rr < 2 and 0 <= kk < 4}
        assert set(np.flatnonzero(bits[c])) == want
    # room centres land in their own cluster
    cx = (xs[:-1] + xs[1:]) / 2
    cy = (ys[:-1] + ys[1:]) / 2
    pts = np.array([(x, y, 100.0) for y in cy for x in cx], np.float32)
    got = oracle.cluster_from_points(bsp.planes, bsp.nodes, bsp.leaf_cluster, pts)
    assert list(got) == list(range(8))
