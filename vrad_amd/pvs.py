"""BSP clusters and the PVS on the GPU, mirroring rad/clustertable and
rad/lightmap/vis.go.

A BSP is given as its three lumps (planes, nodes, leaf clusters); the PVS as
the visibility lump's run-length data with one PVS offset per cluster."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


@dataclass
class Bsp:
    planes: np.ndarray        # PLANE_DTYPE: normal, dist (dplane_t)
    nodes: np.ndarray         # NODE_DTYPE: planenum, children (< 0: leaf -1-child)
    leaf_cluster: np.ndarray  # int16 per leaf (dleaf_t.cluster)

    def struct(self):
        self.planes = np.ascontiguousarray(self.planes, _lib.PLANE_DTYPE)
        self.nodes = np.ascontiguousarray(self.nodes, _lib.NODE_DTYPE)
        self.leaf_cluster = np.ascontiguousarray(self.leaf_cluster, np.int16)
        return _lib.Bsp(ptr(self.planes), len(self.planes), ptr(self.nodes), len(self.nodes),
                        ptr(self.leaf_cluster), len(self.leaf_cluster))


@dataclass
class Vis:
    data: np.ndarray      # uint8 run-length data (the visibility lump's bytes)
    pvs_ofs: np.ndarray   # int32 per cluster: ByteOffset[cluster][DVIS_PVS]
    n_clusters: int

    @property
    def row_bytes(self):
        return (self.n_clusters + 7) >> 3


def ClusterFromPoint(bsp, points):
    """ClusterFromPoint (rad/clustertable/point.go:10-36) for (n, 3) points:
    int16 clusters."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros(len(pts), np.int16)
    b = bsp.struct()
    check(lib.vr_cluster_from_points(C.byref(b), ptr(pts), len(pts), ptr(out)))
    return out


def PatchClusters(bsp, patches):
    """Patch.ClusterNumber as SubdividePatches sets it (rad/patches/
    subdivide.go:100-115) for patches of _lib.PATCH_DTYPE: int32 clusters."""
    p = np.ascontiguousarray(patches, _lib.PATCH_DTYPE)
    out = np.zeros(len(p), np.int32)
    b = bsp.struct()
    check(lib.vr_patch_clusters(C.byref(b), ptr(p), len(p), ptr(out)))
    return out


def DecompressVis(vis):
    """GetVisCache / DecompressVis (rad/lightmap/vis.go:9-94) for every
    cluster: (n_clusters, row_bytes) uint8 PVS rows and the number of runs
    clipped at a row's end."""
    data = np.ascontiguousarray(vis.data, np.uint8)
    ofs = np.ascontiguousarray(vis.pvs_ofs, np.int32)
    rows = np.zeros((vis.n_clusters, vis.row_bytes), np.uint8)
    over = C.c_int64()
    check(lib.vr_decompress_vis(ptr(data), len(data), ptr(ofs), int(vis.n_clusters), ptr(rows),
                                C.byref(over)))
    return rows, over.value


def PVSCheck(rows, pvs_cluster, cluster):
    """PVSCheck (rad/lightmap/lightmap.go:413-422) against the row of
    pvs_cluster (a negative pvs_cluster has the all-ones row GetVisCache
    gives it); a negative cluster passes."""
    if cluster < 0 or pvs_cluster < 0:
        return 1
    return int(rows[pvs_cluster, cluster >> 3] & (1 << (cluster & 7)))
