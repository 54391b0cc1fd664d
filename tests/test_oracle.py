"""The CPU oracle pinned against hand-derived fixtures, closed forms and
invariants (no GPU)."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


# ---- ChangeIntoIntersectionFormat -------------------------------------------
def test_precompute_golden(golden_dir):
    g = json.load(open(os.path.join(golden_dir, "tri_precompute.json")))
    for t in g["triangles"]:
        rec = oracle.precompute(np.array(t["v"], np.float32), np.array([t["id"]], np.int32))[0]
        assert list(_bits([rec["nx"], rec["ny"], rec["nz"]])) == t["N_bits"]
        assert int(_bits(rec["d"])) == t["D_bits"]
        assert list(_bits(rec["e"])) == t["E_bits"], t
        assert (rec["c0"], rec["c1"]) == (t["c0"], t["c1"])
        assert rec["id"] == t["id"]


def test_precompute_golden_general_position(golden_dir):
    """Triangles where every operation rounds: the expected bits come from an
    exact-rational float32 emulation of mgl32's Cross/Len/Normalize/Dot, and
    each triangle is one that a plausible mistake (N / Len, a float64 Len,
    another association, a fused multiply-add) gets wrong."""
    g = json.load(open(os.path.join(golden_dir, "tri_precompute_general.json")))
    assert len(g["triangles"]) == 12
    caught = set()
    for t in g["triangles"]:
        rec = oracle.precompute(np.array(t["v"], np.float32), np.array([t["id"]], np.int32))[0]
        assert list(_bits([rec["nx"], rec["ny"], rec["nz"]])) == t["N_bits"], t["v"]
        assert int(_bits(rec["d"])) == t["D_bits"], t["v"]
        assert list(_bits(rec["e"])) == t["E_bits"], t["v"]
        assert (rec["c0"], rec["c1"]) == (t["c0"], t["c1"])
        caught |= set(t["catches"])
    assert len(caught) == 5


def test_precompute_edge_invariants(room_scene):
    """E1.(p,1): 0 at p1 and p2, 1 at p3; E2.(p,1): 0 at p2, p3, 1 at p1."""
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    v = s.verts9.reshape(-1, 3, 3).astype(np.float64)
    idx = np.arange(len(tri))
    c0, c1 = tri["c0"].astype(int), tri["c1"].astype(int)
    e = tri["e"].astype(np.float64)

    def ev(E, p):
        return E[:, 0] * p[idx, c0] + E[:, 1] * p[idx, c1] + E[:, 2]
    p1, p2, p3 = v[:, 0], v[:, 1], v[:, 2]
    E1, E2 = e[:, :3], e[:, 3:]
    scale = 1e-4
    assert np.all(np.abs(ev(E1, p1)) < scale) and np.all(np.abs(ev(E1, p2)) < scale)
    assert np.all(np.abs(ev(E1, p3) - 1) < scale)
    assert np.all(np.abs(ev(E2, p2)) < scale) and np.all(np.abs(ev(E2, p3)) < scale)
    assert np.all(np.abs(ev(E2, p1) - 1) < scale)


def test_precompute_degenerate_never_hits():
    v9 = np.array([[0, 0, 0, 1, 1, 1, 2, 2, 2]], np.float32)  # collinear
    tri = oracle.precompute(v9, np.array([5], np.int32))
    assert np.isnan(tri["nx"][0])
    o = np.array([[1, 1, -5], [0.5, 0.5, 0.5]], np.float32)
    d = np.array([[0, 0, 1], [1, 0, 0]], np.float32)
    hit, _ = oracle.trace_brute(tri, o, d, 0.0, 1e30)
    assert list(hit) == [-1, -1]


# ---- hit test ------------------------------------------------------------------
def _box_scene(golden_dir):
    g = json.load(open(os.path.join(golden_dir, "analytic_rays.json")))
    v9, ids = H.to_arrays(H.box(g["box_min"], g["box_max"], g["id"]))
    return g, v9, ids


def test_analytic_rays_brute_and_kd(golden_dir):
    g, v9, ids = _box_scene(golden_dir)
    tri = oracle.precompute(v9, ids)
    kd = oracle.KdTree(v9)
    for r in g["rays"]:
        o = np.array([r["o"]], np.float32)
        d = np.array([r["d"]], np.float32)
        tmin, tmax = r.get("tmin", 0.0), r.get("tmax", 1e30)
        for hit, t in (oracle.trace_brute(tri, o, d, tmin, tmax),
                       oracle.trace_kd(kd, tri, o, d, tmin, tmax)):
            assert hit[0] == r["hit"], r
            assert t[0] == np.float32(r["t"]), r


def test_skip_id(golden_dir):
    g, v9, ids = _box_scene(golden_dir)
    ids = ids.copy()
    ids[1] = 77
    tri = oracle.precompute(v9, ids)
    o = np.array([[100, 200, 10]], np.float32)
    d = np.array([[0, 0, 1]], np.float32)
    assert oracle.trace_brute(tri, o, d, 0.0, 1e30, skip=77)[0][0] == -1
    assert oracle.trace_brute(tri, o, d, 0.0, 1e30, skip=-1)[0][0] == 1


def test_f64_agreement(room_scene):
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    rng = np.random.default_rng(3)
    o, d = H.random_rays(rng, 3000, [1, 1, 1], [1023, 1023, 255])
    h32, t32 = oracle.trace_brute(tri, o, d, 0.0, 1e30)
    h64, t64 = oracle.trace_brute_f64(s.verts9, s.ids, o, d, 0.0, 1e30)
    same = h32 == h64
    assert same.mean() > 0.995          # differences only on shared edges
    m = same & (h32 >= 0)
    rel = np.abs(t32[m] - t64[m]) / np.abs(t64[m])
    assert rel.max() < 1e-5
    # a hit-index difference must be a tie of two triangles at the same t
    diff = ~same
    assert np.all(np.abs(t32[diff] - t64[diff]) / t64[diff] < 1e-5)


# ---- kd-tree ----------------------------------------------------------------
def _check_tree(children, split, idx, ntri, v9, bmin, bmax):
    """Encoding and containment: each leaf's triangles touch the leaf cell,
    every triangle is referenced, children indices are in range."""
    n = len(children)
    seen = np.zeros(ntri, bool)
    stack = [(0, np.array(bmin, np.float64), np.array(bmax, np.float64), 0)]
    lo = v9.reshape(-1, 3, 3).min(axis=1)
    hi = v9.reshape(-1, 3, 3).max(axis=1)
    depth = 0
    while stack:
        node, mn, mx, dep = stack.pop()
        depth = max(depth, dep)
        ch = int(children[node])
        typ, v = ch & 3, ch >> 2
        if typ == 3:
            cnt = int(split[node])
            assert float(cnt) == split[node]
            ts = idx[v:v + cnt]
            assert v + cnt <= len(idx)
            seen[ts] = True
            assert np.all(lo[ts] <= mx + 1e-6) and np.all(hi[ts] >= mn - 1e-6)
        else:
            assert v + 1 < n
            s = float(split[node])
            lmx, rmn = mx.copy(), mn.copy()
            lmx[typ] = s
            rmn[typ] = s
            stack.append((v, mn, lmx, dep + 1))
            stack.append((v + 1, rmn, mx, dep + 1))
    assert seen.all()
    return depth


def test_kd_tree_structure(room_scene):
    s = room_scene
    kd = oracle.KdTree(s.verts9)
    depth = _check_tree(kd.children, kd.split, kd.idx, len(s.ids), s.verts9, kd.bmin, kd.bmax)
    assert depth == kd.max_depth <= 22
    assert len(kd.children) > 100      # a real tree, not the reference's single leaf


def test_kd_small_lists_are_leaves():
    v9 = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [5, 0, 0, 6, 0, 0, 5, 1, 0]], np.float32)
    kd = oracle.KdTree(v9)
    assert list(kd.children) == [3] and list(kd.split) == [2.0]   # n < 3 -> one leaf


def test_kd_sah_splits_in_gap():
    # two clusters of 12 boxes' triangles far apart on x: the root split lies in the gap
    tris = []
    for k in range(6):
        tris += H.box((k * 3.0, 0, 0), (k * 3.0 + 2, 2, 2))
        tris += H.box((1000 + k * 3.0, 0, 0), (1000 + k * 3.0 + 2, 2, 2))
    v9, ids = H.to_arrays(tris)
    kd = oracle.KdTree(v9)
    assert kd.children[0] & 3 == 0
    assert 17.0 <= kd.split[0] <= 1000.0


def test_kd_traversal_equals_brute(room_scene):
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    rng = np.random.default_rng(7)
    o, d = H.random_rays(rng, 4000, [-50, -50, -50], [1100, 1100, 300])
    # include axis-aligned directions and zero components
    d[:500] = np.eye(3, dtype=np.float32)[rng.integers(0, 3, 500)] * rng.choice([-1, 1], (500, 1))
    hb, tb = oracle.trace_brute(tri, o, d, 0.0, 1e30)
    hk, tk = oracle.trace_kd(kd, tri, o, d, 0.0, 1e30, threads=4)
    assert np.array_equal(hb, hk)
    assert np.array_equal(_bits(tb), _bits(tk))


# ---- TestLine ----------------------------------------------------------------
def test_testline_occluder_and_sky():
    tris = H.box((40, -10, -10), (60, 10, 10))                       # occluder
    tris += H.quad(0x01000000, (140, -50, -50), (140, 50, -50), (140, 50, 50), (140, -50, 50))
    v9, ids = H.to_arrays(tris)
    tri = oracle.precompute(v9, ids)
    kd = oracle.KdTree(v9)
    s = np.array([[0, 0, 0], [0, 30, 0], [100, 0, 0], [100, 0, 0], [0, 0, 0]], np.float32)
    e = np.array([[100, 0, 0], [100, 30, 0], [200, 0, 0], [130, 0, 0], [30, 0, 0]], np.float32)
    for k in (None, kd):
        vis, sky = oracle.testline(k, tri, s, e)
        assert list(vis) == [0, 1, 1, 1, 1]   # box blocks; above it clear; sky not an occluder
        assert list(sky) == [0, 0, 1, 0, 0]


def test_testline_work_counts():
    # one leaf of two triangles (x in [0,6]): a crossing segment visits the leaf
    # once and tests both; one that misses the scene box visits nothing; a
    # skipped id is not tested
    v9 = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [5, 0, 0, 6, 0, 0, 5, 1, 0]], np.float32)
    ids = np.array([7, 9], np.int32)
    tri = oracle.precompute(v9, ids)
    kd = oracle.KdTree(v9)
    cross = (np.array([[0.2, 0.2, -1]], np.float32), np.array([[0.2, 0.2, 1]], np.float32))
    miss = (np.array([[0, 50, -1]], np.float32), np.array([[0, 50, 1]], np.float32))
    vis, nv, nt = oracle.testline_stats(kd, tri, *cross)
    assert (list(vis), nv, nt) == ([0], 1, 2)
    vis, nv, nt = oracle.testline_stats(kd, tri, *miss)
    assert (list(vis), nv, nt) == ([1], 0, 0)
    vis, nv, nt = oracle.testline_stats(kd, tri, *cross, skip=7)
    assert (list(vis), nv, nt) == ([1], 1, 1)


def test_testline_work_counts_real_tree(room_scene):
    # on a deep tree the counts agree with testline's answers and stay far below
    # the brute-force work (every triangle for every segment)
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    rng = np.random.default_rng(3)
    a = rng.uniform([0, 0, 0], [1024, 1024, 256], (2000, 3)).astype(np.float32)
    b = rng.uniform([0, 0, 0], [1024, 1024, 256], (2000, 3)).astype(np.float32)
    vis, nv, nt = oracle.testline_stats(kd, tri, a, b, threads=4)
    vis2, _ = oracle.testline(kd, tri, a, b, threads=4)
    assert np.array_equal(vis, vis2)
    assert 2000 <= nv <= 2000 * len(kd.children)
    assert 0 < nt < 2000 * len(s.ids) // 10


# ---- transfers -------------------------------------------------------------------
def test_form_factor_parallel_squares():
    n = 16
    pa = H.square_patches((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), n)
    pb = H.square_patches((0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 0, -1), n)
    o, nrm, a = [np.concatenate(x) for x in zip(pa, pb)]
    rp, col, coef, rs = oracle.make_transfers(None, np.zeros(0, oracle.TRI_DTYPE), o, nrm, a,
                                              normalize=False)
    m = n * n
    F_ab = np.sum(rs[:m] * a[:m]) / np.sum(a[:m])
    assert abs(F_ab - H.howell_parallel(1, 1, 1)) < 0.01 * H.howell_parallel(1, 1, 1)


def test_form_factor_perpendicular_squares():
    n = 24
    pa = H.square_patches((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), n)     # floor
    pb = H.square_patches((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0), n)     # wall x=0
    o, nrm, a = [np.concatenate(x) for x in zip(pa, pb)]
    rp, col, coef, rs = oracle.make_transfers(None, np.zeros(0, oracle.TRI_DTYPE), o, nrm, a,
                                              normalize=False)
    m = n * n
    F = np.sum(rs[:m] * a[:m]) / np.sum(a[:m])
    ref = H.howell_perpendicular(1, 1, 1)
    assert abs(F - ref) < 0.03 * ref


def test_transfers_reciprocity_and_enclosure():
    # closed cube of 6 faces x 6x6 patches, no occluders: A_i F_ij = A_j F_ji
    n = 6
    faces = [((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)), ((0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, -1)),
             ((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0)), ((1, 0, 0), (0, 0, 1), (0, 1, 0), (-1, 0, 0)),
             ((0, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 1, 0), (1, 0, 0), (0, 0, 1), (0, -1, 0))]
    parts = [H.square_patches(*f, n) for f in faces]
    o, nrm, a = [np.concatenate(x) for x in zip(*parts)]
    rp, col, coef, rs = oracle.make_transfers(None, np.zeros(0, oracle.TRI_DTYPE), o, nrm, a,
                                              normalize=False)
    N = len(a)
    F = np.zeros((N, N))
    for i in range(N):
        F[i, col[rp[i]:rp[i + 1]]] = coef[rp[i]:rp[i + 1]]
    AF = a[:, None] * F
    assert np.allclose(AF, AF.T, rtol=1e-5, atol=0)
    assert np.all(np.diff(col[rp[0]:rp[1]]) > 0)                  # ascending columns
    # enclosure: row sums -> 1 as patches shrink (edge error ~ 1/n)
    err6 = abs(rs.mean() - 1.0)
    parts12 = [H.square_patches(*f, 12) for f in faces]
    o12, n12, a12 = [np.concatenate(x) for x in zip(*parts12)]
    rs12 = oracle.make_transfers(None, np.zeros(0, oracle.TRI_DTYPE), o12, n12, a12,
                                 normalize=False)[3]
    err12 = abs(rs12.mean() - 1.0)
    assert err12 < 0.6 * err6 and err12 < 0.05 and rs12.min() > 0.99
    # normalisation caps the row sums at 1 (in FP64)
    rp2, col2, coef2, rs2 = oracle.make_transfers(None, np.zeros(0, oracle.TRI_DTYPE), o, nrm, a)
    for i in range(N):
        assert coef2[rp2[i]:rp2[i + 1]].astype(np.float64).sum() <= 1.0 + 1e-6


def test_row_sum_wave_order():
    rng = np.random.default_rng(1)
    for L in (0, 1, 7, 256, 1000, 5003):
        c = rng.random(L).astype(np.float32)
        assert abs(oracle.row_sum_f64(c) - math.fsum(c.astype(np.float64))) <= 1e-12 * max(1, L)


# ---- direct light ----------------------------------------------------------------
def _pt_light(pos, inten=10000.0, q=1.0):
    L = np.zeros(1, oracle.LIGHT_DTYPE)
    L["type"] = 1
    L["origin"] = pos
    L["intensity"] = inten
    L["quadratic_attn"] = q
    L["end_fade"] = -1
    L["cap_distance"] = 1e22
    return L


def test_point_light_closed_form():
    L = _pt_light((0, 0, 100))
    pos = np.array([[0, 0, 0], [100, 0, 0], [0, 0, 200]], np.float32)
    nrm = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32)
    rgb = oracle.gather_light(None, np.zeros(0, oracle.TRI_DTYPE), pos, nrm, L)
    assert rgb[0, 0] == 1.0                                  # 10000 * 1 / 100^2
    assert abs(rgb[1, 0] - 10000 * (100 / math.sqrt(2e4)) / 2e4) < 1e-6
    assert rgb[2, 0] == 0.0                                  # light behind the surface


def test_point_light_occluded():
    v9, ids = H.to_arrays(H.box((-10, -10, 40), (10, 10, 60)))
    tri = oracle.precompute(v9, ids)
    kd = oracle.KdTree(v9)
    L = _pt_light((0, 0, 100))
    pos = np.array([[0, 0, 0], [200, 0, 0]], np.float32)
    nrm = np.array([[0, 0, 1], [0, 0, 1]], np.float32)
    rgb = oracle.gather_light(kd, tri, pos, nrm, L)
    assert rgb[0, 0] == 0.0 and rgb[1, 0] > 0.0


def test_spot_cone():
    L = _pt_light((0, 0, 100))
    L["type"] = 2
    L["normal"] = (0, 0, -1)
    L["stopdot"] = math.cos(math.radians(20))
    L["stopdot2"] = math.cos(math.radians(40))
    pos = np.array([[0, 0, 0], [100 * math.tan(math.radians(30)), 0, 0],
                    [100 * math.tan(math.radians(50)), 0, 0]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (3, 1))
    rgb = oracle.gather_light(None, np.zeros(0, oracle.TRI_DTYPE), pos, nrm, L)
    assert rgb[0, 0] == 1.0
    d2 = 100 ** 2 + pos[1, 0] ** 2
    cosr = 100 / math.sqrt(d2)
    ramp = (math.cos(math.radians(30)) - math.cos(math.radians(40))) / \
           (math.cos(math.radians(20)) - math.cos(math.radians(40)))
    assert abs(rgb[1, 0] - 10000 * cosr / d2 * ramp) < 1e-5
    assert rgb[2, 0] == 0.0


# ---- bounce + energy ----------------------------------------------------------------
def _furnace(n=64, m=8, rho=0.5, seed=0):
    rng = np.random.default_rng(seed)
    rp = np.arange(n + 1, dtype=np.int64) * m
    col = np.concatenate([np.sort(rng.choice(n, m, replace=False)) for _ in range(n)]).astype(np.int32)
    coef = np.full(n * m, 1.0 / m, np.float32)
    e0 = np.ones((n, 3), np.float32)
    refl = np.full((n, 3), rho, np.float32)
    area = np.ones(n, np.float32)
    return rp, col, coef, e0, refl, area


def test_furnace_closed_form():
    rp, col, coef, e0, refl, area = _furnace()
    tot, en, k = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=100, stop_frac=1e-3)
    # add_k = rho^(k-1) exactly; stop when 2^(1-k) <= 1e-3 (2 - 2^(1-k)) -> k = 10
    assert k == 10
    assert np.all(tot == np.float32(2.0 - 2.0 ** -9))
    assert np.array_equal(en, 3.0 * 64 * 0.5 ** np.arange(10))
    tot64, en64, k64 = oracle.bounce_f64(rp, col, coef, e0, refl, area, max_bounces=100,
                                         stop_frac=1e-3)
    assert k64 == 10 and np.allclose(tot64, 2.0 - 2.0 ** -9, rtol=0, atol=1e-15)


def test_bounce_matches_linear_solve():
    scipy = pytest.importorskip("scipy.sparse")
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(5)
    n = 40
    rows = []
    for i in range(n):
        c = np.sort(rng.choice(np.delete(np.arange(n), i), 10, replace=False))
        w = rng.random(10)
        rows.append((c, (w / w.sum() * 0.9).astype(np.float32)))
    rp = np.concatenate([[0], np.cumsum([len(c) for c, _ in rows])]).astype(np.int64)
    col = np.concatenate([c for c, _ in rows]).astype(np.int32)
    coef = np.concatenate([w for _, w in rows])
    e0 = rng.random((n, 3)).astype(np.float32)
    refl = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    area = np.ones(n, np.float32)
    tot, en, k = oracle.bounce_f64(rp, col, coef, e0, refl, area, max_bounces=400, stop_frac=0.0)
    F = scipy.csr_matrix((coef.astype(np.float64), col, rp), shape=(n, n))
    for ch in range(3):
        A = scipy.identity(n) - F @ scipy.diags(refl[:, ch].astype(np.float64))
        # total = sum_k add_k, add_1 = F e0, add_{k+1} = F diag(rho) add_k
        ref = spla.spsolve(A.tocsc(), F @ e0[:, ch].astype(np.float64))
        assert np.allclose(tot[:, ch], ref, rtol=1e-9)
    tot32, _, _ = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=400, stop_frac=0.0)
    assert np.max(np.abs(tot32 - tot) / np.abs(tot)) < 1e-5


def test_bounce_orders_agree():
    """The kernel summation orders differ only by FP32 rounding."""
    rng = np.random.default_rng(8)
    n = 3000
    L = rng.integers(1, 400, n)
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in L]).astype(np.int32)
    coef = (rng.random(rp[-1]) / 400).astype(np.float32)
    e0 = rng.random((n, 3)).astype(np.float32)
    refl = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    area = np.ones(n, np.float32)
    ref, _, _ = oracle.bounce_f64(rp, col, coef, e0, refl, area, max_bounces=5, stop_frac=0.0)
    outs = []
    for kw in (dict(order="wave"), dict(order="sequential"), dict(order="tiled"),
               dict(order="tiled", tile_cols=512, group=16), dict(order="stream"),
               dict(order="stream", tile_cols=512), dict(order="stream-runs"),
               dict(order="stream-by-index")):
        t, _, k = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=5, stop_frac=0.0,
                                    **kw)
        assert k == 5 and np.max(np.abs(t - ref) / ref) < 2e-6
        outs.append(t)
    assert not np.array_equal(outs[0], outs[1])      # the orders really differ
    assert not np.array_equal(outs[2], outs[4])
    assert not np.array_equal(outs[4], outs[6])      # the row placement changes the order
    assert not np.array_equal(outs[4], outs[7])


def _dyadic_rows(rng, n, lens):
    """Rows of power-of-two lengths with coefficients 2^-m: every FP32 partial
    sum of emit values k / 64 is exact, so any summation order that adds each
    entry exactly once gives the same bits as the FP64 sum."""
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    coef = np.concatenate([np.full(l, 1.0 / l) for l in lens]).astype(np.float32)
    return rp, col, coef


@pytest.mark.parametrize("order", ["stream", "stream-runs", "stream-by-index"])
@pytest.mark.parametrize("tile", [512, 4096])
def test_stream_order_adds_every_entry_once(tile, order):
    # uneven rows (1 .. 2048 entries), several row chunks and column tiles, and
    # segments that cross passes: the stream order must equal the exact sum
    rng = np.random.default_rng(11)
    n = 6000
    lens = 2 ** rng.integers(0, 12, n)
    rp, col, coef = _dyadic_rows(rng, n, lens)
    e0 = (rng.integers(0, 64, (n, 3)) / 64.0).astype(np.float32)
    refl = np.full((n, 3), 0.5, np.float32)
    area = np.ones(n, np.float32)
    t, en, k = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=1, stop_frac=0.0,
                                 order=order, tile_cols=tile, threads=4)
    t64, _, _ = oracle.bounce_f64(rp, col, coef, e0, refl, area, max_bounces=1, stop_frac=0.0)
    assert np.array_equal(t.astype(np.float64), t64)
    # FP16 storage with row scale 1: coefficients 2^-m are exact in FP16 too
    raw = np.array([oracle.f32_to_f16(v) for v in np.unique(coef)], np.uint16)[
        np.searchsorted(np.unique(coef), coef)]
    t16, _, _ = oracle.bounce_f32(rp, col, raw, e0, refl, area, max_bounces=1, stop_frac=0.0,
                                  order=order, tile_cols=tile, kind=1,
                                  row_scale=np.ones(n, np.float32), threads=4)
    assert np.array_equal(t16.astype(np.float64), t64)


@pytest.mark.parametrize("order", ["stream", "stream-runs", "stream-by-index"])
def test_stream_order_beyond_256_full_chunks(order):
    # more rows than 256 chunks of 1024 hold: 512 chunks, still every entry once
    rng = np.random.default_rng(12)
    n = 270_000
    lens = 2 ** rng.integers(0, 4, n)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    coef = np.concatenate([np.full(l, 1.0 / l) for l in lens]).astype(np.float32)
    e0 = (rng.integers(0, 64, (n, 3)) / 64.0).astype(np.float32)
    refl = np.full((n, 3), 0.5, np.float32)
    area = np.ones(n, np.float32)
    t, _, _ = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=1, stop_frac=0.0,
                                order=order, threads=8)
    t64, _, _ = oracle.bounce_f64(rp, col, coef, e0, refl, area, max_bounces=1, stop_frac=0.0)
    assert np.array_equal(t.astype(np.float64), t64)


def test_stream_order_independent_of_threads():
    rng = np.random.default_rng(5)
    n = 30000
    L = rng.integers(1, 60, n)
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in L]).astype(np.int32)
    coef = (rng.random(rp[-1]) / 60).astype(np.float32)
    e0 = rng.random((n, 3)).astype(np.float32)
    refl = np.full((n, 3), 0.5, np.float32)
    area = np.ones(n, np.float32)
    a, _, _ = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=2, stop_frac=0.0,
                                order="stream", threads=1)
    b, _, _ = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=2, stop_frac=0.0,
                                order="stream", threads=8)
    assert np.array_equal(a, b)


def test_energy_reduce_exact():
    rng = np.random.default_rng(2)
    for n in (0, 1, 5, 2048, 2049, 100000):
        e = rng.integers(0, 1000, n).astype(np.float64)
        assert oracle.energy_reduce(e) == float(e.sum())
        x = rng.random(n)
        assert abs(oracle.energy_reduce(x) - math.fsum(x)) <= 1e-12 * max(1.0, math.fsum(x))


def test_half_conversions():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.random(20000).astype(np.float32),
                        (rng.random(2000) * 1e-6).astype(np.float32),
                        np.array([0.0, 1.0, 65504.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8], np.float32)])
    want = x.astype(np.float16).view(np.uint16)
    got = np.array([oracle.f32_to_f16(v) for v in x], np.uint16)
    assert np.array_equal(got, want)
    torch = pytest.importorskip("torch")
    wb = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    gb = np.array([oracle.f32_to_bf16(v) for v in x], np.uint16)
    assert np.array_equal(gb, wb)


# ---- patch generation (MakePatchForFace + SubdividePatches) ----------------
def _quad_face(x0, y0, x1, y1, z=0.0, luxel=16.0, flags=0):
    f = np.zeros(1, oracle.FACE_DTYPE)
    f["n_points"] = 4
    f["points"][0, :4] = [[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]]
    f["normal"] = [0, 0, 1]
    f["lux_s"] = [1.0 / luxel, 0, 0]
    f["lux_t"] = [0, 1.0 / luxel, 0]
    f["flags"] = flags
    return f


def _check_patch_tree(p):
    """Structure of the reference's patch array: children in pairs created in
    depth-first order after every base patch; leaves tile their parent."""
    for i, q in enumerate(p):
        if q["child1"] >= 0:
            c1, c2 = q["child1"], q["child2"]
            assert c2 == c1 + 1 and c1 > i
            assert p[c1]["parent"] == i and p[c2]["parent"] == i
            assert np.isclose(p[c1]["area"] + p[c2]["area"], q["area"], rtol=1e-5)
        else:
            assert q["child2"] == -1


def test_patches_closed_form_square():
    # 256 x 256 at 16 units per luxel: 16 luxels wide; a patch is split while
    # it spans >= 4 luxels (maxChop = minChop = 4), so leaves are 32 x 32
    p = oracle.make_patches(_quad_face(0, 0, 256, 256))
    leaves = p[p["child1"] < 0]
    assert len(p) == 127 and len(leaves) == 64
    assert np.all(leaves["area"] == 1024.0)
    assert np.all(p["lux_scale"] == np.float32(1.0 / 16.0)) and np.all(p["chop"] == 4.0)
    c = np.sort(np.round(leaves["origin"][:, :2], 3), axis=0)
    assert np.allclose(np.unique(c[:, 0]), 16 + 32 * np.arange(8), atol=1e-3)
    assert np.all(leaves["normal"] == [0, 0, 1])
    _check_patch_tree(p)
    # the first split is along x at 128 (the widest axis, x before y on a
    # tie); the front side (x > 128) is child 1 and is subdivided first
    assert p[0]["child1"] == 1 and p[1]["mins"][0] == 128 and p[2]["maxs"][0] == 128
    assert p[3]["parent"] == 1


def test_patches_special_faces():
    f = np.concatenate([_quad_face(0, 0, 256, 256, flags=1),      # sky: never subdivided
                        _quad_face(0, 0, 0, 256),                  # zero area: no patch
                        _quad_face(0, 0, 256, 256, flags=2),       # PreventSubdivision
                        _quad_face(0, 0, 60, 60)])                 # under 4 luxels: a leaf
    p = oracle.make_patches(f)
    assert len(p) == 3 and list(p["face"]) == [0, 2, 3]
    assert np.all(p["child1"] == -1) and np.all(p["parent"] == -1)
    # WindingCenter: the points' average
    assert np.array_equal(p[2]["origin"], np.float32([30, 30, 0]))


def test_patches_triangle_and_ramp():
    f = np.zeros(2, oracle.FACE_DTYPE)
    f["n_points"] = [3, 4]
    f["points"][0, :3] = [[0, 0, 0], [300, 0, 0], [0, 200, 0]]
    f["normal"][0] = [0, 0, 1]
    s = np.float32(1 / np.sqrt(2))
    f["points"][1, :4] = [[0, 0, 0], [200, 0, 200], [200, 100, 200], [0, 100, 0]]  # 45 degree ramp
    f["normal"][1] = [-s, 0, s]
    f["lux_s"] = [1 / 16, 0, 0]
    f["lux_t"] = [0, 1 / 16, 0]
    p = oracle.make_patches(f)
    _check_patch_tree(p)
    for face, area in ((0, 30000.0), (1, 200 * np.sqrt(2) * 100)):
        leaves = p[(p["face"] == face) & (p["child1"] < 0)]
        assert len(leaves) > 4
        assert np.isclose(leaves["area"].astype(np.float64).sum(), area, rtol=1e-5)
        # children lie inside the face's bounds
        base = p[(p["face"] == face) & (p["parent"] < 0)][0]
        assert np.all(leaves["mins"] >= base["mins"] - 1e-3)
        assert np.all(leaves["maxs"] <= base["maxs"] + 1e-3)


def test_patches_generated_level_invariants():
    from vrad_amd import scenes
    f = scenes.faces_of("level", rooms_x=3, rooms_y=3, chop=64.0, tess=64.0, n_lights=2)
    p = oracle.make_patches(f)
    _check_patch_tree(p)
    base = p[p["parent"] < 0]
    leaves = p[p["child1"] < 0]
    assert len(base) == len(f)
    per_face = np.bincount(leaves["face"], weights=leaves["area"].astype(np.float64),
                           minlength=len(f))
    assert np.allclose(per_face, base["area"], rtol=1e-5)
    # leaves are at most 4 luxels (64 units at 16 units per luxel) on a side
    ext = (leaves["maxs"] - leaves["mins"]).max(1) * leaves["lux_scale"]
    assert ext.max() < 4.0 + 1e-4
