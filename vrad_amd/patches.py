"""rad/patches on the MI355X path: MakeTransfers.

The reference declares the result records (Transfer, common/types/transfer.go
:3-6; Patch.NumTransfers / Transfers, common/types/patch.go:60-61) but never
builds them.  Here the transfers stay resident on the GPU as a padded CSR
(Transfers), with download() for inspection.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, fptr, f32

ON_EPSILON = 0.1  # vmath/constants.go:8


class Transfers:
    def __init__(self, handle):
        self._h = handle

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = _lib.XfersInfo()
        check(lib.vr_xfers_get_info(self._h, C.byref(inf)))
        return inf

    def set_precision(self, precision):
        check(lib.vr_xfers_set_precision(self._h, int(precision)))

    def download(self, raw16=False):
        """Logical CSR: (row_ptr int64, col int32, coef float32[, raw uint16,
        row_scale float32]).  16-bit coefficients come back decoded and
        without the row scale."""
        inf = self.info()
        rp = np.zeros(inf.n_rows + 1, np.int64)
        col = np.zeros(inf.nnz, np.int32)
        coef = np.zeros(inf.nnz, np.float32)
        if inf.precision != _lib.VR_COEF_F32:
            raw = np.zeros(inf.nnz, np.uint16)
            sc = np.zeros(inf.n_rows, np.float32)
            check(lib.vr_xfers_download(self._h, ptr(rp), ptr(col), ptr(coef), ptr(raw), ptr(sc)))
            return (rp, col, coef, raw, sc) if raw16 else (rp, col, coef)
        check(lib.vr_xfers_download(self._h, ptr(rp), ptr(col), ptr(coef), None, None))
        return rp, col, coef

    def close(self):
        if self._h is not None:
            lib.vr_xfers_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def from_csr(row_ptr, col, coef, precision=_lib.VR_COEF_F32):
        rp = np.ascontiguousarray(row_ptr, np.int64)
        c = np.ascontiguousarray(col, np.int32)
        w = f32(coef)
        h = C.c_void_p()
        check(lib.vr_xfers_from_csr(C.byref(h), len(rp) - 1, ptr(rp), ptr(c), ptr(w),
                                    int(precision)))
        return Transfers(h)

    @staticmethod
    def synthetic(seed, n_rows, mean_len=2000.0, sigma=1.0, max_len=40000,
                  precision=_lib.VR_COEF_F32, columns="uniform"):
        """The precision workload's synthetic CSR; columns "uniform" (drawn
        without replacement) or "stratified" (one per equal stratum)."""
        mode = {"uniform": 0, "stratified": 1}[columns]
        h = C.c_void_p()
        check(lib.vr_xfers_synthetic(C.byref(h), int(seed), int(n_rows), float(mean_len),
                                     float(sigma), int(max_len), int(precision), mode))
        return Transfers(h)


def patch_struct(origin, normal, area):
    o = f32(origin).reshape(-1, 3)
    nrm = f32(normal).reshape(-1, 3)
    a = f32(area).reshape(-1)
    cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + \
           [np.ascontiguousarray(nrm[:, k]) for k in range(3)] + [a]
    return _lib.Patches(*[fptr(c) for c in cols]), cols


def MakeTransfers(env, origin, normal, area, normal_offset=ON_EPSILON, skip_id=-1,
                  normalize=True, batch_pairs=0, cluster=None, pvs_rows=None):
    """Leaf patches (origin (n,3), normal (n,3), area (n,)) -> Transfers.

    With cluster (int32 per patch) and pvs_rows (the (n_clusters, row) PVS
    rows of pvs.DecompressVis), only pairs whose clusters pass PVSCheck in
    either direction are considered (a negative cluster passes)."""
    p, keep = patch_struct(origin, normal, area)
    prm = _lib.XferParams(float(normal_offset), int(skip_id), 1 if normalize else 0,
                          int(batch_pairs))
    h = C.c_void_p()
    n = len(keep[-1])
    if pvs_rows is None:
        check(lib.vr_make_transfers(env.handle, C.byref(p), n, C.byref(prm), C.byref(h)))
    else:
        cl = np.ascontiguousarray(cluster, np.int32)
        rows = np.ascontiguousarray(pvs_rows, np.uint8)
        if cl.shape != (n,):
            raise ValueError("cluster must hold one int per patch")
        nc = rows.shape[0]
        if rows.ndim != 2 or rows.shape[1] != (nc + 7) >> 3:
            raise ValueError("pvs_rows must be (n_clusters, (n_clusters + 7) >> 3) bytes")
        check(lib.vr_make_transfers_pvs(env.handle, C.byref(p), n, C.byref(prm), ptr(cl),
                                        ptr(rows), nc, C.byref(h)))
    return Transfers(h)


def MakeTransfersBatch(items, normal_offset=ON_EPSILON, skip_id=-1, normalize=True,
                       batch_pairs=0):
    """MakeTransfers for many scenes in shared visibility launches: items is a
    list of (env, origin, normal, area); returns one Transfers per item, each
    equal to MakeTransfers on that item alone."""
    k = len(items)
    if k == 0:
        return []
    structs, keep = (_lib.Patches * k)(), []
    hs = (C.c_void_p * k)()
    n = np.zeros(k, np.int64)
    for q, (env, origin, normal, area) in enumerate(items):
        p, cols = patch_struct(origin, normal, area)
        structs[q] = p
        keep.append(cols)
        n[q] = len(cols[-1])
        hs[q] = env.handle.value if isinstance(env.handle, C.c_void_p) else env.handle
    prm = _lib.XferParams(float(normal_offset), int(skip_id), 1 if normalize else 0,
                          int(batch_pairs))
    out = (C.c_void_p * k)()
    check(lib.vr_make_transfers_batch(hs, structs, ptr(n), k, C.byref(prm), out))
    return [Transfers(C.c_void_p(out[q])) for q in range(k)]


MAX_CHOP = 4.0  # maxChop, rad/patches/face.go:24 (luxel widths)
MIN_CHOP = 4.0  # minChop, rad/patches/face.go:25


SMOOTHING_THRESHOLD = np.float32(0.7071067)  # smoothingThreshold, rad/lightmap/lightmap.go:29


def PairEdges(faces, smoothing_threshold=SMOOTHING_THRESHOLD):
    """PairEdges (rad/lightmap/lightmap.go:37-228): per-vertex normals of the
    faces (_lib.FACE_DTYPE), n_faces x VR_FACE_MAX_POINTS x 3 float32 -- the
    normalised sum of the face's own normal and those of faces sharing the
    vertex within the smoothing threshold."""
    f = np.ascontiguousarray(faces, _lib.FACE_DTYPE)
    out = np.zeros((len(f), _lib.VR_FACE_MAX_POINTS, 3), np.float32)
    check(lib.vr_pair_edges(ptr(f), len(f), float(smoothing_threshold), ptr(out)))
    return out


def MakePatches(faces, max_chop=MAX_CHOP, min_chop=MIN_CHOP, vertex_normals=None):
    """MakePatches / MakePatchForFace (rad/patches/build.go:21-68,
    rad/patches/face.go:30-197) and SubdividePatches (rad/patches/subdivide.go
    :22-438) on the GPU: faces is an array of _lib.FACE_DTYPE; returns the
    patches (_lib.PATCH_DTYPE) in the reference's array order: a base patch
    per face of positive area, then each base patch's children in creation
    order.  Leaves (child1 == -1) are the patches the bake uses.

    vertex_normals (PairEdges' output) gives each child the GetPhongNormal
    normal at its centre (rad/lightmap/normallist.go:51-144, called at
    subdivide.go:387); without it children keep the face normal."""
    f = np.ascontiguousarray(faces, _lib.FACE_DTYPE)
    vn = None
    if vertex_normals is not None:
        vn = np.ascontiguousarray(vertex_normals, np.float32)
        if vn.shape != (len(f), _lib.VR_FACE_MAX_POINTS, 3):
            raise ValueError(f"vertex_normals: shape {vn.shape}, want "
                             f"({len(f)}, {_lib.VR_FACE_MAX_POINTS}, 3)")
    n = C.c_int64()
    check(lib.vr_make_patches_phong(ptr(f), len(f), float(max_chop), float(min_chop), ptr(vn),
                                    None, 0, C.byref(n)))
    out = np.zeros(n.value, _lib.PATCH_DTYPE)
    check(lib.vr_make_patches_phong(ptr(f), len(f), float(max_chop), float(min_chop), ptr(vn),
                                    ptr(out), len(out), C.byref(n)))
    return out
