"""BSP clusters, PVS decompression and PVS-culled MakeTransfers on the GPU
against the CPU restatement: clusters and PVS rows byte for byte, transfers
bit for bit."""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, pvs, scenes
from vrad_amd.patches import MakePatches, MakeTransfers
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu


def _env(v9, ids):
    e = Environment()
    e.AddTriangles(v9, ids)
    e.SetupAccelerationStructure()
    return e


def test_gpu_cluster_from_point_random_trees():
    rng = np.random.default_rng(21)
    for k in range(6):
        pl, nd, lc = H.random_bsp(rng, n_clusters=int(rng.integers(1, 300)),
                                  max_depth=int(rng.integers(5, 16)))
        pts = H.bsp_points(rng, pl, 5000)
        print(f"[pvs] random BSP {k}: {len(nd)} nodes, {len(pts)} points", flush=True)
        got = pvs.ClusterFromPoint(pvs.Bsp(pl, nd, lc), pts)
        want = oracle.cluster_from_points(pl, nd, lc, pts)
        assert np.array_equal(got, want)


def test_gpu_cluster_from_point_deep_chain_and_errors():
    # a chain of 64 nodes, every plane through the origin: the origin takes
    # the within-epsilon branch at each, all children[0] leaves are solid,
    # so the walk pops all 64 pending children[1]; the last is cluster 11
    n = _lib.VR_BSP_MAX_DEPTH
    pl = np.zeros(1, _lib.PLANE_DTYPE)
    pl[0] = ((1, 0, 0), 0.0)
    nd = np.zeros(n, _lib.NODE_DTYPE)
    for k in range(n):
        nd[k] = (0, (-1, k + 1 if k + 1 < n else -2))
    lc = np.array([-1, 11], np.int16)
    pts = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0)], np.float32)
    got = pvs.ClusterFromPoint(pvs.Bsp(pl, nd, lc), pts)
    assert list(got) == [11, -1, 11]
    assert list(got) == list(oracle.cluster_from_points(pl, nd, lc, pts))
    # one more level is refused, as are a cycle and out-of-range indices
    nd2 = np.zeros(n + 1, _lib.NODE_DTYPE)
    for k in range(n + 1):
        nd2[k] = (0, (-1, k + 1 if k + 1 < n + 1 else -2))
    for bad in (pvs.Bsp(pl, nd2, lc),
                pvs.Bsp(pl, np.array([(0, (-1, 0))], _lib.NODE_DTYPE), lc),
                pvs.Bsp(pl, np.array([(1, (-1, -2))], _lib.NODE_DTYPE), lc),
                pvs.Bsp(pl, np.array([(0, (-1, -3))], _lib.NODE_DTYPE), lc),
                pvs.Bsp(pl, np.array([(0, (-1, 5))], _lib.NODE_DTYPE), lc)):
        with pytest.raises(_lib.VrError):
            pvs.ClusterFromPoint(bad, pts)


def test_gpu_patch_clusters():
    faces = scenes.faces_of("level", rooms_x=3, rooms_y=3, chop=64.0, tess=64.0, seed=4)
    patches = MakePatches(faces)
    s = scenes.level(seed=4, rooms_x=3, rooms_y=3, chop=64.0, tess=64.0)
    bsp, _ = scenes.grid_bsp(s.grid_x, s.grid_y)
    print(f"[pvs] patch clusters: {len(patches)} patches", flush=True)
    got = pvs.PatchClusters(bsp, patches)
    assert np.array_equal(got, oracle.patch_clusters(bsp.planes, bsp.nodes, bsp.leaf_cluster,
                                                     patches))
    assert (got >= 0).all()
    # random trees: many origins land in solid leaves and fall back to
    # winding points
    rng = np.random.default_rng(3)
    pl, nd, lc = H.random_bsp(rng, n_clusters=50)
    p = np.zeros(4000, _lib.PATCH_DTYPE)
    p["origin"] = H.bsp_points(rng, pl, 4000)
    p["n_points"] = rng.integers(0, 17, 4000)
    p["points"] = H.bsp_points(rng, pl, 4000 * 16).reshape(4000, 16, 3)
    got = pvs.PatchClusters(pvs.Bsp(pl, nd, lc), p)
    assert np.array_equal(got, oracle.patch_clusters(pl, nd, lc, p))


def test_gpu_decompress_vis():
    rng = np.random.default_rng(8)
    for nc, dens in ((1, 1.0), (9, 0.4), (513, 0.05), (4000, 0.002), (3000, 0.7)):
        bits = rng.uniform(size=(nc, nc)) < dens
        rows = np.packbits(bits, axis=1, bitorder="little")
        data, ofs = scenes.compress_vis(rows)
        print(f"[pvs] decompress {nc} clusters, {len(data)} bytes", flush=True)
        got, over = pvs.DecompressVis(pvs.Vis(data, ofs, nc))
        assert over == 0 and np.array_equal(got, rows)
    # clipped runs, shared rows, and the error cases
    data = np.array([0x01, 0x00, 0x03, 0x80, 0x00, 0x05, 0xFF, 0x00, 0x09,
                     0x12, 0x34, 0x56, 0x78, 0x9A], np.uint8)
    ofs = np.array([0, 4, 6, 9, 0] + [9] * 35, np.int32)
    got, over = pvs.DecompressVis(pvs.Vis(data, ofs, 40))
    want, wover = oracle.decompress_vis(data, ofs, 40)
    assert np.array_equal(got, want) and over == wover == 1
    for d, o, nc in ((np.array([0, 0], np.uint8), [0], 8),       # zero count
                     (np.array([1, 2], np.uint8), [0], 40),      # ends inside a row
                     (np.array([1], np.uint8), [-1], 8),         # visofs == -1
                     (np.array([1], np.uint8), [1], 8)):         # offset past the data
        with pytest.raises(_lib.VrError):
            pvs.DecompressVis(pvs.Vis(d, np.array(o, np.int32), nc))


@pytest.fixture(scope="module")
def pvs_level():
    s = scenes.level(seed=6, rooms_x=3, rooms_y=2, chop=96.0, tess=96.0, n_lights=1)
    bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y, pvs_hops=1)
    cl = pvs.ClusterFromPoint(bsp, s.p_origin).astype(np.int32)
    rows, _ = pvs.DecompressVis(vis)
    return s, cl, rows


def test_gpu_transfers_pvs_bitexact(pvs_level):
    s, cl, rows = pvs_level
    print(f"[pvs] transfers with PVS: {len(s.p_area)} patches", flush=True)
    env = _env(s.verts9, s.ids)
    cl2 = cl.copy()
    cl2[7::23] = -1  # some patches outside every cluster: they see all
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    for c, batch in ((cl, 0), (cl2, 5000)):
        xf = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=c, pvs_rows=rows,
                           batch_pairs=batch)
        rp, col, coef = xf.download()
        orp, ocol, ocoef, _ = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area,
                                                    threads=8, cluster=c, pvs_rows=rows)
        assert np.array_equal(rp, orp)
        assert np.array_equal(col, ocol)
        assert np.array_equal(coef.view(np.uint32), ocoef.view(np.uint32))
        # pairs considered: i < j with the clusters visible either way
        g = np.where(c < 0, -1, c)
        vis = np.unpackbits(rows, axis=1, bitorder="little")[:, : len(rows)].astype(bool)
        vis = vis | vis.T
        n = len(c)
        ii, jj = np.triu_indices(n, 1)
        ok = (g[ii] < 0) | (g[jj] < 0) | vis[np.maximum(g[ii], 0), np.maximum(g[jj], 0)]
        assert xf.info().pairs_considered == int(ok.sum())


def test_gpu_transfers_pvs_conservative_equals_unculled(pvs_level):
    """A PVS holding every cluster pair that exchanges a transfer changes
    nothing; one without a pair removes exactly that pair's transfers."""
    s, cl, _ = pvs_level
    assert (cl >= 0).all()  # every generated patch lies in a room
    env = _env(s.verts9, s.ids)
    full = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, normalize=False)
    rp, col, coef = full.download()
    row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    nc = int(cl.max()) + 1
    used = np.zeros((nc, nc), bool)
    used[cl[row], cl[col]] = True
    rows = np.packbits(used, axis=1, bitorder="little")
    x = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, normalize=False, cluster=cl,
                      pvs_rows=rows)
    g = x.download()
    assert np.array_equal(g[0], rp) and np.array_equal(g[1], col)
    assert np.array_equal(g[2].view(np.uint32), coef.view(np.uint32))
    assert x.info().segments_traced <= full.info().segments_traced
    # drop the cross pair exchanging the most transfers
    cross = [(int(((cl[row] == a) & (cl[col] == b)).sum()), a, b)
             for a in range(nc) for b in range(a + 1, nc) if used[a, b]]
    _, a, b = max(cross)
    used[a, b] = used[b, a] = False
    rows = np.packbits(used, axis=1, bitorder="little")
    g = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, normalize=False, cluster=cl,
                      pvs_rows=rows).download()
    keep = ~(((cl[row] == a) & (cl[col] == b)) | ((cl[row] == b) & (cl[col] == a)))
    assert np.array_equal(g[1], col[keep])
    assert np.array_equal(g[2].view(np.uint32), coef[keep].view(np.uint32))


def test_gpu_transfers_pvs_argument_errors(pvs_level):
    s, cl, rows = pvs_level
    env = _env(s.verts9, s.ids)
    bad = cl.copy()
    bad[3] = len(rows)  # a cluster without a PVS row
    with pytest.raises(_lib.VrError):
        MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=bad, pvs_rows=rows)
    with pytest.raises(ValueError):
        MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=cl[:-1], pvs_rows=rows)
    with pytest.raises(ValueError):
        MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=cl, pvs_rows=rows[:, :0])
    # zero clusters: every pair is considered, as without a PVS
    a = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=cl,
                      pvs_rows=np.zeros((0, 0), np.uint8))
    b = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area)
    assert all(np.array_equal(x, y) for x, y in zip(a.download(), b.download()))
    assert a.info().pairs_considered == b.info().pairs_considered


def test_gpu_transfers_pvs_full_level_conservative():
    """4x4-room level (99,730 patches): a PVS made of the cluster pairs that
    exchange a transfer reproduces the unculled transfers bit for bit, and
    considers fewer pairs."""
    s = scenes.level(seed=2, rooms_x=4, rooms_y=4, chop=32.0, tess=64.0)
    bsp, _ = scenes.grid_bsp(s.grid_x, s.grid_y)
    cl = pvs.ClusterFromPoint(bsp, s.p_origin).astype(np.int32)
    assert (cl >= 0).all()
    print(f"[pvs] full level: {len(cl)} patches", flush=True)
    env = _env(s.verts9, s.ids)
    full = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, normalize=False)
    rp, col, coef = full.download()
    row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    nc = int(cl.max()) + 1
    used = np.zeros((nc, nc), bool)
    used[cl[row], cl[col]] = True
    x = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, normalize=False, cluster=cl,
                      pvs_rows=np.packbits(used, axis=1, bitorder="little"))
    g = x.download()
    print(f"[pvs] full level: {len(col)} transfers, pairs {full.info().pairs_considered} -> "
          f"{x.info().pairs_considered}", flush=True)
    assert np.array_equal(g[0], rp) and np.array_equal(g[1], col)
    assert np.array_equal(g[2].view(np.uint32), coef.view(np.uint32))
    assert x.info().pairs_considered < full.info().pairs_considered
