"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

TRI_DTYPE = np.dtype([("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4"), ("d", "<f4"), ("id", "<i4"),
                      ("e", "<f4", 6), ("c0", "u1"), ("c1", "u1"), ("flags", "u1"),
                      ("unused", "u1")])
assert TRI_DTYPE.itemsize == 48
LIGHT_DTYPE = np.dtype([("type", "<i4"), ("origin", "<f4", 3), ("intensity", "<f4", 3),
                        ("normal", "<f4", 3), ("constant_attn", "<f4"), ("linear_attn", "<f4"),
                        ("quadratic_attn", "<f4"), ("stopdot", "<f4"), ("stopdot2", "<f4"),
                        ("exponent", "<f4"), ("start_fade", "<f4"), ("end_fade", "<f4"),
                        ("cap_distance", "<f4")])

# patch generation records (layout of or_face_t / or_patch_t)
FACE_DTYPE = np.dtype([("n_points", "<i4"), ("points", "<f4", (16, 3)), ("normal", "<f4", 3),
                       ("lux_s", "<f4", 3), ("lux_t", "<f4", 3), ("base_light", "<f4", 3),
                       ("flags", "<u4")])
PATCH_DTYPE = np.dtype([("face", "<i4"), ("parent", "<i4"), ("child1", "<i4"), ("child2", "<i4"),
                        ("origin", "<f4", 3), ("normal", "<f4", 3), ("area", "<f4"),
                        ("chop", "<f4"), ("lux_scale", "<f4"), ("mins", "<f4", 3),
                        ("maxs", "<f4", 3), ("n_points", "<i4"), ("points", "<f4", (16, 3))])
assert FACE_DTYPE.itemsize == 248 and PATCH_DTYPE.itemsize == 272

FP = C.POINTER(C.c_float)
P = C.c_void_p


class _Rays(C.Structure):
    _fields_ = [(n, FP) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class _Segs(C.Structure):
    _fields_ = [(n, FP) for n in ("sx", "sy", "sz", "ex", "ey", "ez")]


class _Patches(C.Structure):
    _fields_ = [(n, FP) for n in ("ox", "oy", "oz", "nx", "ny", "nz", "area")]


class _Samples(C.Structure):
    _fields_ = [(n, FP) for n in ("px", "py", "pz", "nx", "ny", "nz")]


class _Kd(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("children", C.POINTER(C.c_uint32)), ("split", FP),
                ("n_idx", C.c_int64), ("idx", C.POINTER(C.c_int32)), ("bmin", C.c_float * 3),
                ("bmax", C.c_float * 3), ("max_depth", C.c_int)]


class _Csr(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("nnz", C.c_int64), ("row_ptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int32)), ("coef", FP), ("row_sum", C.POINTER(C.c_double))]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "or_precompute": (None, [P, P, P, C.c_int64, P]),
        "or_trace_brute": (None, [P, C.c_int64, C.POINTER(_Rays), C.c_int64, C.c_int32, P, P]),
        "or_trace_brute_f64": (None, [P, P, C.c_int64, C.POINTER(_Rays), C.c_int64, C.c_int32,
                                      P, P]),
        "or_kd_build": (C.c_int, [P, C.c_int64, C.POINTER(_Kd)]),
        "or_kd_free": (None, [C.POINTER(_Kd)]),
        "or_trace_kd": (None, [C.POINTER(_Kd), P, C.POINTER(_Rays), C.c_int64, C.c_int32, P, P,
                               C.c_int]),
        "or_testline": (None, [C.POINTER(_Kd), P, C.c_int64, C.POINTER(_Segs), C.c_int64,
                               C.c_int32, P, P, C.c_int]),
        "or_testline_stats": (None, [C.POINTER(_Kd), P, C.POINTER(_Segs), C.c_int64, C.c_int32,
                                     P, C.c_int, P, P]),
        "or_make_transfers": (C.c_int, [C.POINTER(_Kd), P, C.c_int64, C.POINTER(_Patches),
                                        C.c_int64, C.c_float, C.c_int32, C.c_int, C.c_int,
                                        C.POINTER(_Csr)]),
        "or_csr_free": (None, [C.POINTER(_Csr)]),
        "or_row_sum_f64": (C.c_double, [P, C.c_int64]),
        "or_gather_light": (None, [C.POINTER(_Kd), P, C.c_int64, C.POINTER(_Samples), C.c_int64,
                                   P, C.c_int32, C.c_float, C.c_int32, P, C.c_int]),
        "or_sky_dirs": (None, [P]),
        "or_can_leaf_trace_to_sky": (None, [C.POINTER(_Kd), P, C.c_int64, P, P, C.c_int64, P,
                                            C.c_int]),
        "or_bounce_f32": (C.c_int, [C.c_int64, P, P, P, C.c_int, P, P, P, P, P, C.c_int,
                                    C.c_double, P, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_bounce_f64": (C.c_int, [C.c_int64, P, P, P, P, P, P, P, C.c_int, C.c_double, P]),
        "or_energy_reduce": (C.c_double, [P, C.c_int64]),
        "or_f32_to_f16": (C.c_uint16, [C.c_float]),
        "or_f16_to_f32": (C.c_float, [C.c_uint16]),
        "or_f32_to_bf16": (C.c_uint16, [C.c_float]),
        "or_bf16_to_f32": (C.c_float, [C.c_uint16]),
        "or_make_patches": (C.c_int, [P, C.c_int64, C.c_float, C.c_float, P, C.c_int64,
                                      C.POINTER(C.c_int64)]),
        "or_make_patches_phong": (C.c_int, [P, C.c_int64, C.c_float, C.c_float, P, P,
                                            C.c_int64, C.POINTER(C.c_int64)]),
        "or_pair_edges": (None, [P, C.c_int64, C.c_float, P]),
        "or_phong_normal": (None, [P, P, P, P, P]),
        "or_calc_face_extents": (C.c_int64, [P, C.c_int64, P, P]),
        "or_face_luxels": (None, [P, C.c_int64, P, P, P, C.c_int64, P, P]),
        "or_rgbexp32": (C.c_uint32, [C.c_float, C.c_float, C.c_float]),
        "or_lighting_lump": (C.c_int64, [P, P, C.c_int64, P, P]),
        "or_cluster_from_point": (C.c_int16, [P, P, P, P]),
        "or_patch_clusters": (None, [P, P, P, P, C.c_int64, P]),
        "or_decompress_vis": (C.c_int, [P, C.c_int64, P, C.c_int32, P, C.POINTER(C.c_int64)]),
        "or_make_transfers_pvs": (C.c_int, [C.POINTER(_Kd), P, C.c_int64, C.POINTER(_Patches),
                                            C.c_int64, C.c_float, C.c_int32, C.c_int, C.c_int,
                                            P, P, C.c_int32, C.POINTER(_Csr)]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f(a):
    return a.ctypes.data_as(FP)


def _cols(arr, k):
    a = np.ascontiguousarray(arr, np.float32).reshape(-1, k)
    return [np.ascontiguousarray(a[:, i]) for i in range(k)]


def precompute(v9, ids, flags=None):
    v9 = np.ascontiguousarray(v9, np.float32).reshape(-1, 9)
    ids = np.ascontiguousarray(ids, np.int32)
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    out = np.zeros(len(ids), TRI_DTYPE)
    lib.or_precompute(_p(v9), _p(ids), _p(fl), len(ids), _p(out))
    return out


class KdTree:
    def __init__(self, v9):
        self.v9 = np.ascontiguousarray(v9, np.float32).reshape(-1, 9)
        self.kd = _Kd()
        rc = lib.or_kd_build(_p(self.v9), len(self.v9), C.byref(self.kd))
        if rc:
            raise MemoryError("oracle kd build failed")
        n, m = self.kd.n_nodes, self.kd.n_idx
        self.children = np.ctypeslib.as_array(self.kd.children, (n,)).copy()
        self.split = np.ctypeslib.as_array(self.kd.split, (n,)).copy()
        self.idx = np.ctypeslib.as_array(self.kd.idx, (max(m, 1),))[:m].copy()
        self.max_depth = self.kd.max_depth
        self.bmin = np.array(self.kd.bmin[:], np.float32)
        self.bmax = np.array(self.kd.bmax[:], np.float32)

    def __del__(self):
        try:
            lib.or_kd_free(C.byref(self.kd))
        except Exception:
            pass


def _rays(o, d, tmin, tmax):
    n = len(np.asarray(o).reshape(-1, 3))
    cols = _cols(o, 3) + _cols(d, 3) + [np.ascontiguousarray(np.broadcast_to(tmin, (n,)), np.float32),
                                        np.ascontiguousarray(np.broadcast_to(tmax, (n,)), np.float32)]
    return _Rays(*[_f(c) for c in cols]), cols, n


def trace_brute(tris, o, d, tmin, tmax, skip=-1):
    r, keep, n = _rays(o, d, tmin, tmax)
    hit = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    lib.or_trace_brute(_p(tris), len(tris), C.byref(r), n, skip, _p(hit), _p(t))
    return hit, t


def trace_brute_f64(v9, ids, o, d, tmin, tmax, skip=-1):
    v9 = np.ascontiguousarray(v9, np.float32)
    ids = np.ascontiguousarray(ids, np.int32)
    r, keep, n = _rays(o, d, tmin, tmax)
    hit = np.zeros(n, np.int32)
    t = np.zeros(n, np.float64)
    lib.or_trace_brute_f64(_p(v9), _p(ids), len(ids), C.byref(r), n, skip, _p(hit), _p(t))
    return hit, t


def trace_kd(kd, tris, o, d, tmin, tmax, skip=-1, threads=1):
    r, keep, n = _rays(o, d, tmin, tmax)
    hit = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    lib.or_trace_kd(C.byref(kd.kd), _p(tris), C.byref(r), n, skip, _p(hit), _p(t), threads)
    return hit, t


def testline(kd, tris, s, e, skip=-1, threads=1):
    cols = _cols(s, 3) + _cols(e, 3)
    g = _Segs(*[_f(c) for c in cols])
    n = len(cols[0])
    vis = np.zeros(n, np.uint8)
    sky = np.zeros(n, np.uint8)
    lib.or_testline(C.byref(kd.kd) if kd is not None else None, _p(tris), len(tris), C.byref(g),
                    n, skip, _p(vis), _p(sky), threads)
    return vis, sky


def testline_stats(kd, tris, s, e, skip=-1, threads=1):
    """testline on a kd-tree plus the work it took: (visible, node visits,
    triangle tests), the counts summed over all segments."""
    cols = _cols(s, 3) + _cols(e, 3)
    g = _Segs(*[_f(c) for c in cols])
    n = len(cols[0])
    vis = np.zeros(n, np.uint8)
    cnt = np.zeros(2, np.int64)
    lib.or_testline_stats(C.byref(kd.kd), _p(tris), C.byref(g), n, skip, _p(vis), threads,
                          _p(cnt[:1]), _p(cnt[1:]))
    return vis, int(cnt[0]), int(cnt[1])


def make_transfers(kd, tris, origin, normal, area, eps=0.1, skip=-1, normalize=True, threads=1,
                   cluster=None, pvs_rows=None):
    cols = _cols(origin, 3) + _cols(normal, 3) + [np.ascontiguousarray(area, np.float32)]
    pt = _Patches(*[_f(c) for c in cols])
    n = len(cols[0])
    csr = _Csr()
    cl = None if cluster is None else np.ascontiguousarray(cluster, np.int32)
    rows = None if pvs_rows is None else np.ascontiguousarray(pvs_rows, np.uint8)
    rc = lib.or_make_transfers_pvs(C.byref(kd.kd) if kd is not None else None, _p(tris),
                                   len(tris), C.byref(pt), n, eps, skip, 1 if normalize else 0,
                                   threads, _p(cl), _p(rows), 0 if rows is None else len(rows),
                                   C.byref(csr))
    if rc:
        raise MemoryError("oracle transfers failed")
    nnz = csr.nnz
    rp = np.ctypeslib.as_array(csr.row_ptr, (n + 1,)).copy()
    col = np.ctypeslib.as_array(csr.col, (max(nnz, 1),))[:nnz].copy()
    coef = np.ctypeslib.as_array(csr.coef, (max(nnz, 1),))[:nnz].copy()
    rs = np.ctypeslib.as_array(csr.row_sum, (max(n, 1),))[:n].copy()
    lib.or_csr_free(C.byref(csr))
    return rp, col, coef, rs


def gather_light(kd, tris, pos, nrm, lights, eps=0.1, skip=-1, threads=1):
    cols = _cols(pos, 3) + _cols(nrm, 3)
    s = _Samples(*[_f(c) for c in cols])
    n = len(cols[0])
    L = np.ascontiguousarray(lights).view(LIGHT_DTYPE) if lights.dtype != LIGHT_DTYPE else \
        np.ascontiguousarray(lights)
    rgb = np.zeros((n, 3), np.float32)
    lib.or_gather_light(C.byref(kd.kd) if kd is not None else None, _p(tris), len(tris),
                        C.byref(s), n, _p(L), len(L), eps, skip, _p(rgb), threads)
    return rgb


def sky_dirs():
    d = np.zeros((162, 3), np.float32)
    lib.or_sky_dirs(_p(d))
    return d


def can_leaf_trace_to_sky(kd, tris, mins, maxs, threads=1):
    """CanLeafTraceToSky per leaf (int16 bounds, n x 3 each): uint8 0/1."""
    mn = np.ascontiguousarray(mins, np.int16).reshape(-1, 3)
    mx = np.ascontiguousarray(maxs, np.int16).reshape(-1, 3)
    out = np.zeros(len(mn), np.uint8)
    lib.or_can_leaf_trace_to_sky(C.byref(kd.kd) if kd is not None else None, _p(tris),
                                 len(tris), _p(mn), _p(mx), len(mn), _p(out), threads)
    return out


ORDERS = {"wave": 0, "sequential": 1, "tiled": 2, "stream": 3, "stream-runs": 4,
          "stream-by-index": 5}


def bounce_f32(rp, col, coef, emit0, refl, area, total=None, max_bounces=100, stop_frac=1e-3,
               kind=0, row_scale=None, threads=1, order="stream", tile_cols=4096, group=None):
    """FP32 bounce in a kernel's summation order: "stream" (the default
    gather, 4096-patch tiles), "tiled" (16-lane groups for FP32 coefficients,
    8-lane groups for FP16/BF16 unless `group` says otherwise), "wave" or
    "sequential"."""
    if group is None:
        group = 16 if kind == 0 else 8
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    coef = np.ascontiguousarray(coef, np.float32 if kind == 0 else np.uint16)
    sc = None if row_scale is None else np.ascontiguousarray(row_scale, np.float32)
    e0 = np.ascontiguousarray(emit0, np.float32).reshape(n, 3)
    rf = np.ascontiguousarray(refl, np.float32).reshape(n, 3)
    ar = np.ascontiguousarray(area, np.float32).reshape(n)
    tot = np.zeros((n, 3), np.float32) if total is None else \
        np.ascontiguousarray(total, np.float32).reshape(n, 3).copy()
    en = np.zeros(max(max_bounces, 1), np.float64)
    k = lib.or_bounce_f32(n, _p(rp), _p(col), _p(coef), kind, _p(sc), _p(e0), _p(rf), _p(ar),
                          _p(tot), max_bounces, stop_frac, _p(en), ORDERS[order], tile_cols, group,
                          threads)
    return tot, en[:k], k


def bounce_f64(rp, col, coef, emit0, refl, area, total=None, max_bounces=100, stop_frac=1e-3):
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    coef = np.ascontiguousarray(coef, np.float32)
    e0 = np.ascontiguousarray(emit0, np.float32).reshape(n, 3)
    rf = np.ascontiguousarray(refl, np.float32).reshape(n, 3)
    ar = np.ascontiguousarray(area, np.float32).reshape(n)
    tot = np.zeros((n, 3), np.float64) if total is None else \
        np.ascontiguousarray(total, np.float64).reshape(n, 3).copy()
    en = np.zeros(max(max_bounces, 1), np.float64)
    k = lib.or_bounce_f64(n, _p(rp), _p(col), _p(coef), _p(e0), _p(rf), _p(ar), _p(tot),
                          max_bounces, stop_frac, _p(en))
    return tot, en[:k], k


def energy_reduce(e):
    e = np.ascontiguousarray(e, np.float64)
    return lib.or_energy_reduce(_p(e), len(e))


def row_sum_f64(coef):
    c = np.ascontiguousarray(coef, np.float32)
    return lib.or_row_sum_f64(_p(c), len(c))


def f32_to_f16(x):
    return lib.or_f32_to_f16(float(x))


def f32_to_bf16(x):
    return lib.or_f32_to_bf16(float(x))


def make_patches(faces, max_chop=4.0, min_chop=4.0, vertex_normals=None):
    """MakePatchForFace + SubdividePatches over faces (FACE_DTYPE): patches
    (PATCH_DTYPE) in the reference's array order.  vertex_normals (n_faces x
    16 x 3 float32, e.g. pair_edges()) gives children GetPhongNormal normals."""
    f = np.ascontiguousarray(faces, FACE_DTYPE)
    vn = None
    if vertex_normals is not None:
        vn = np.ascontiguousarray(vertex_normals, np.float32)
        assert vn.shape == (len(f), 16, 3)
    n = C.c_int64()
    if lib.or_make_patches_phong(_p(f), len(f), max_chop, min_chop, _p(vn), None, 0,
                                 C.byref(n)) != 0:
        raise ValueError("a winding exceeds 16 points")
    out = np.zeros(n.value, PATCH_DTYPE)
    if lib.or_make_patches_phong(_p(f), len(f), max_chop, min_chop, _p(vn), _p(out), len(out),
                                 C.byref(n)) != 0:
        raise ValueError("a winding exceeds 16 points")
    return out


SMOOTHING_THRESHOLD = np.float32(0.7071067)  # rad/lightmap/lightmap.go:29


def phong_normal(face, centroid, vertex_normals, spot):
    """GetPhongNormal for one face (FACE_DTYPE record), its centroid, its
    16 x 3 vertex normals and a point: the bent unit normal (float32[3])."""
    f = np.ascontiguousarray(np.atleast_1d(face)[:1], FACE_DTYPE)
    c = np.ascontiguousarray(centroid, np.float32)
    vn = np.ascontiguousarray(vertex_normals, np.float32).reshape(16, 3)
    sp = np.ascontiguousarray(spot, np.float32)
    out = np.zeros(3, np.float32)
    lib.or_phong_normal(_p(f), _p(c), _p(vn), _p(sp), _p(out))
    return out


def pair_edges(faces, threshold=SMOOTHING_THRESHOLD):
    """PairEdges' per-vertex normals: n_faces x 16 x 3 float32."""
    f = np.ascontiguousarray(faces, FACE_DTYPE)
    out = np.zeros((len(f), 16, 3), np.float32)
    lib.or_pair_edges(_p(f), len(f), float(threshold), _p(out))
    return out


PLANE_DTYPE = np.dtype([("normal", "<f4", 3), ("dist", "<f4")])
NODE_DTYPE = np.dtype([("planenum", "<i4"), ("children", "<i4", 2)])


def cluster_from_points(planes, nodes, leaf_cluster, points):
    """ClusterFromPoint per point (recursive PointInLeaf, as written)."""
    pl = np.ascontiguousarray(planes, PLANE_DTYPE)
    nd = np.ascontiguousarray(nodes, NODE_DTYPE)
    lc = np.ascontiguousarray(leaf_cluster, np.int16)
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros(len(pts), np.int16)
    for k in range(len(pts)):
        out[k] = lib.or_cluster_from_point(_p(pl), _p(nd), _p(lc), _p(pts[k]))
    return out


def patch_clusters(planes, nodes, leaf_cluster, patches):
    pl = np.ascontiguousarray(planes, PLANE_DTYPE)
    nd = np.ascontiguousarray(nodes, NODE_DTYPE)
    lc = np.ascontiguousarray(leaf_cluster, np.int16)
    p = np.ascontiguousarray(patches, PATCH_DTYPE)
    out = np.zeros(len(p), np.int32)
    lib.or_patch_clusters(_p(pl), _p(nd), _p(lc), _p(p), len(p), _p(out))
    return out


def decompress_vis(data, ofs, n_clusters):
    """(rows, overruns); raises ValueError on a zero count or a short buffer."""
    d = np.ascontiguousarray(data, np.uint8)
    o = np.ascontiguousarray(ofs, np.int32)
    rows = np.zeros((n_clusters, (n_clusters + 7) >> 3), np.uint8)
    over = C.c_int64()
    rc = lib.or_decompress_vis(_p(d), len(d), _p(o), n_clusters, _p(rows), C.byref(over))
    if rc:
        raise ValueError({-1: "zero repeat", -2: "data ends inside a row"}[rc])
    return rows, over.value


FACE_LIGHTMAP_DTYPE = np.dtype([("mins", "<i4", 2), ("size", "<i4", 2), ("luxel_ofs", "<i8"),
                                ("n_luxels", "<i4"), ("flags", "<u4")])


def calc_face_extents(faces, lux_ofs=None):
    """CalcFaceExtents per face: (FACE_LIGHTMAP_DTYPE records, luxel total)."""
    f = np.ascontiguousarray(faces, FACE_DTYPE)
    o = None if lux_ofs is None else np.ascontiguousarray(lux_ofs, np.float32).reshape(-1)
    out = np.zeros(len(f), FACE_LIGHTMAP_DTYPE)
    n = lib.or_calc_face_extents(_p(f), len(f), _p(o), _p(out))
    return out, int(n)


def face_luxels(faces, lm, patches, lux_ofs=None):
    """(positions (n, 3), chosen patch per luxel)."""
    f = np.ascontiguousarray(faces, FACE_DTYPE)
    o = None if lux_ofs is None else np.ascontiguousarray(lux_ofs, np.float32).reshape(-1)
    L = np.ascontiguousarray(lm, FACE_LIGHTMAP_DTYPE)
    p = np.ascontiguousarray(patches, PATCH_DTYPE)
    n = int(L["n_luxels"].sum())
    pos = np.zeros((n, 3), np.float32)
    pat = np.zeros(n, np.int32)
    lib.or_face_luxels(_p(f), len(f), _p(o), _p(L), _p(p), len(p), _p(pos), _p(pat))
    return pos, pat


def rgbexp32(r, g, b):
    """ColorRGBExp32 of one colour as its 4 bytes (r, g, b, exponent)."""
    v = lib.or_rgbexp32(float(r), float(g), float(b))
    return [(v >> s) & 0xFF for s in (0, 8, 16, 24)]


def lighting_lump(luxel_rgb, lm):
    """(lump bytes (uint8), lightofs (int32 per face))."""
    v = np.ascontiguousarray(luxel_rgb, np.float32).reshape(-1, 3)
    L = np.ascontiguousarray(lm, FACE_LIGHTMAP_DTYPE)
    nb = lib.or_lighting_lump(_p(v), _p(L), len(L), None, None)
    lump = np.zeros(nb // 4, np.uint32)
    ofs = np.zeros(len(L), np.int32)
    lib.or_lighting_lump(_p(v), _p(L), len(L), _p(lump), _p(ofs))
    return lump.view(np.uint8), ofs
