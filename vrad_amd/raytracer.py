"""raytracer.Environment on the MI355X path.

Mirrors the reference's Go API (raytracer/environment.go): triangles are
appended with AddTriangle / AddQuad / AddAxisAlignedRectangularSolid,
SetupAccelerationStructure builds the SAH kd-tree and uploads the scene to
HBM, Trace4Rays traces a FourRays packet.  TraceRays is the batched form the
GPU is meant for (one call per batch, not per packet of four).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, fptr, f32

# raytracer/constants.go:5-18
RTE_FLAGS_FAST_TREE_GENERATION = 1
RTE_FLAGS_DONT_STORE_TRIANGLE_COLORS = 2
RTE_FLAGS_DONT_STORE_TRIANGLE_MATERIALS = 4
TRACE_ID_SKY = 0x01000000
TRACE_ID_OPAQUE = 0x02000000
TRACE_ID_STATICPROP = 0x04000000
KDNODE_STATE_XSPLIT, KDNODE_STATE_YSPLIT, KDNODE_STATE_ZSPLIT, KDNODE_STATE_LEAF = 0, 1, 2, 3

# TriIntersectData, raytracer/cache/triangle/triintersectdata.go:3-22
TRI_INTERSECT_DTYPE = np.dtype([("FlNx", "<f4"), ("FlNy", "<f4"), ("FlNz", "<f4"), ("FlD", "<f4"),
                                ("NTriangleID", "<i4"), ("ProjectedEdgeEquations", "<f4", 6),
                                ("NCoordSelect0", "u1"), ("NCoordSelect1", "u1"),
                                ("NFlags", "u1"), ("Unused0", "u1")])
assert TRI_INTERSECT_DTYPE.itemsize == 48


class FourRays:
    """raytracer/types/fourrays.go:8-11: Origin/Direction as FourVectors
    (arrays of shape (3, 4): X, Y, Z rows, one column per ray)."""

    def __init__(self, origin=None, direction=None):
        self.Origin = f32(np.zeros((3, 4)) if origin is None else origin, (3, 4))
        self.Direction = f32(np.zeros((3, 4)) if direction is None else direction, (3, 4))


class RayTracingResult:
    """raytracer/types/result.go:8-12."""

    def __init__(self):
        self.SurfaceNormal = np.zeros((3, 4), np.float32)
        self.HitIds = np.full(4, -1, np.int32)
        self.HitDistance = np.zeros(4, np.float32)


class Environment:
    """raytracer/environment.go:28-39."""

    def __init__(self, flags=0):
        self.Flags = flags
        self.MinBound = np.zeros(3, np.float32)
        self.MaxBound = np.zeros(3, np.float32)
        # ingested triangles as chunks (one per call), joined at build time
        self._v9, self._ids, self._flags = [], [], []
        self.TriangleColors, self.TriangleMaterials = [], []
        self._scene = None
        # the built scene's tree and triangle records, copied out on first use
        self._kd = None    # (children uint32, split float32, triangle index list int32)
        self._tri = None   # TriIntersectData records

    # ---- ingest: raytracer/environment.go:41-117 ----
    def AddTriangle(self, id, v1, v2, v3, colour=(1.0, 1.0, 1.0)):
        self.AddTriangleWithMaterial(id, v1, v2, v3, colour, 0, 0)

    def AddTriangleWithMaterial(self, id, v1, v2, v3, colour, flags, materialIndex):
        self._v9.append(np.concatenate([f32(v1, 3), f32(v2, 3), f32(v3, 3)]).reshape(1, 9))
        self._ids.append(np.array([id], np.int32))
        self._flags.append(np.array([flags & 0xFF], np.uint8))
        if not (flags & RTE_FLAGS_DONT_STORE_TRIANGLE_COLORS):
            self.TriangleColors.append(f32(colour, 3))
        if not (flags & RTE_FLAGS_DONT_STORE_TRIANGLE_MATERIALS):
            self.TriangleMaterials.append(int(materialIndex))

    def AddQuad(self, id, v1, v2, v3, v4, colour=(1.0, 1.0, 1.0)):
        self.AddTriangle(id, v1, v2, v3, colour)
        self.AddTriangle(id + 1, v1, v3, v4, colour)

    def AddAxisAlignedRectangularSolid(self, id, minCoord, maxCoord, colour=(1.0, 1.0, 1.0)):
        a, b = f32(minCoord, 3), f32(maxCoord, 3)
        x0, y0, z0 = a
        x1, y1, z1 = b
        self.AddQuad(id, (x0, y1, z1), (x1, y1, z1), (x1, y0, z1), (x0, y0, z1), colour)  # far
        self.AddQuad(id, (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (x0, y0, z0), colour)  # near
        self.AddQuad(id, (x0, y1, z1), (x0, y1, z0), (x0, y0, z0), (x0, y0, z1), colour)  # left
        self.AddQuad(id, (x1, y1, z1), (x1, y1, z0), (x1, y0, z0), (x1, y0, z1), colour)  # right
        self.AddQuad(id, (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0), colour)  # top
        self.AddQuad(id, (x0, y0, z1), (x1, y0, z1), (x1, y0, z0), (x0, y0, z0), colour)  # bot

    def AddTriangles(self, verts9, ids, flags=None):
        """Bulk append (n x 9 float32 VertexCoordData, n ids)."""
        v = f32(verts9).reshape(-1, 9)
        i = np.asarray(ids, np.int32).reshape(-1)
        if len(i) != len(v):
            raise ValueError(f"{len(v)} triangles but {len(i)} ids")
        fl = np.zeros(len(v), np.uint8) if flags is None else np.asarray(flags, np.uint8).reshape(-1)
        # copies: later changes to the caller's arrays do not reach the scene
        self._v9.append(v.copy())
        self._ids.append(i.copy())
        self._flags.append(np.array(fl, np.uint8))

    def triangle_arrays(self):
        if not self._ids:
            return np.zeros((0, 9), np.float32), np.zeros(0, np.int32), np.zeros(0, np.uint8)
        def join(parts):
            return parts[0] if len(parts) == 1 else np.concatenate(parts)
        v9 = f32(join(self._v9).reshape(-1, 9))
        ids = np.ascontiguousarray(join(self._ids), np.int32)
        fl = np.ascontiguousarray(join(self._flags), np.uint8)
        return v9, ids, fl

    # ---- raytracer/environment.go:119-138 ----
    def SetupAccelerationStructure(self, n_threads=0):
        v9, ids, fl = self.triangle_arrays()
        self.close()
        h = C.c_void_p()
        check(lib.vr_scene_create(C.byref(h), ptr(v9), ptr(ids), ptr(fl), len(ids)))
        self._scene = h
        check(lib.vr_scene_build(h, n_threads))
        info = self.info()
        self.MinBound = np.array(info.bmin[:], np.float32)
        self.MaxBound = np.array(info.bmax[:], np.float32)

    def _tree(self):
        if self._kd is None and self._scene is not None:
            info = self.info()
            ch = np.zeros(info.n_nodes, np.uint32)
            sp = np.zeros(info.n_nodes, np.float32)
            idx = np.zeros(info.n_leaf_refs, np.int32)
            check(lib.vr_scene_get_kdtree(self._scene, ptr(ch), ptr(sp), ptr(idx)))
            self._kd = (ch, sp, idx)
        return self._kd

    @property
    def OptimizedKDTree(self):
        """(children, split) of the built tree (optimisedkdnode.go:15-54)."""
        t = self._tree()
        return None if t is None else (t[0], t[1])

    @property
    def TriangleIndexList(self):
        t = self._tree()
        return None if t is None else t[2]

    @property
    def OptimizedTriangleList(self):
        """TriIntersectData records of the built scene, in ingest order."""
        if self._tri is None and self._scene is not None:
            n = self.info().n_tri
            tri = np.zeros(n, TRI_INTERSECT_DTYPE)
            if n:
                check(lib.vr_scene_get_triangles(self._scene, ptr(tri)))
            self._tri = tri
        return self._tri

    def info(self):
        inf = _lib.SceneInfo()
        check(lib.vr_scene_get_info(self._scene, C.byref(inf)))
        return inf

    @property
    def handle(self):
        if self._scene is None:
            raise RuntimeError("SetupAccelerationStructure has not been called")
        return self._scene

    # ---- raytracer/environment.go:140-145 (a stub in the reference) ----
    def Trace4Rays(self, rays, TMin, TMax, resultOut, skipId, callback=None):
        if callback is not None:
            raise _lib.VrError(_lib.VR_E_UNSUPPORTED,
                               "transparent-triangle callbacks are not supported on the GPU path")
        o, d = rays.Origin, rays.Direction
        hit, t, nrm = self.TraceRays(o.T, d.T, np.broadcast_to(f32(TMin), (4,)),
                                     np.broadcast_to(f32(TMax), (4,)), skipId, normals=True)
        resultOut.HitIds[:] = hit
        resultOut.HitDistance[:] = t
        resultOut.SurfaceNormal[:] = nrm.T

    def TraceRays(self, origins, directions, tmin, tmax, skip_id=-1, normals=False):
        """Batched closest hit: origins/directions (n,3), tmin/tmax (n,).
        Returns (hit_idx int32, hit_t float32[, normals (n,3)])."""
        o = f32(origins).reshape(-1, 3)
        d = f32(directions).reshape(-1, 3)
        n = o.shape[0]
        cols = [np.ascontiguousarray(o[:, 0]), np.ascontiguousarray(o[:, 1]),
                np.ascontiguousarray(o[:, 2]), np.ascontiguousarray(d[:, 0]),
                np.ascontiguousarray(d[:, 1]), np.ascontiguousarray(d[:, 2]),
                f32(np.broadcast_to(tmin, (n,))), f32(np.broadcast_to(tmax, (n,)))]
        r = _lib.Rays(*[fptr(c) for c in cols])
        hit = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        nrm = np.empty((n, 3), np.float32) if normals else None
        check(lib.vr_trace_closest(self.handle, C.byref(r), n, int(skip_id), ptr(hit), ptr(t),
                                   ptr(nrm)))
        return (hit, t, nrm) if normals else (hit, t)

    def TestLines(self, starts, ends, skip_id=-1):
        """Batched TestLine: returns (visible uint8, hit_sky uint8)."""
        s = f32(starts).reshape(-1, 3)
        e = f32(ends).reshape(-1, 3)
        n = s.shape[0]
        cols = [np.ascontiguousarray(a[:, k]) for a in (s, e) for k in range(3)]
        g = _lib.Segs(*[fptr(c) for c in cols])
        vis = np.empty(n, np.uint8)
        sky = np.empty(n, np.uint8)
        check(lib.vr_test_line(self.handle, C.byref(g), n, int(skip_id), ptr(vis), ptr(sky)))
        return vis, sky

    # ---- raytracer/environment.go:422-432 ----
    def GetTriangle(self, triangleId):
        return self.OptimizedTriangleList[triangleId]

    def GetTriangleMaterial(self, triangleId):
        return self.TriangleMaterials[triangleId]

    def GetTriangleColor(self, triangleId):
        return self.TriangleColors[triangleId]

    def close(self):
        self._kd = self._tri = None
        if self._scene is not None:
            lib.vr_scene_destroy(self._scene)
            self._scene = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_environment = None


def GetEnvironment():
    """raytracer/environment.go:17-25 (process-wide environment)."""
    global _environment
    if _environment is None:
        _environment = Environment()
    return _environment
