"""Procedural scenes (seeded) used as inputs for tests and benchmarks."""
import ctypes as C
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, LIGHT_DTYPE


@dataclass
class Scene:
    verts9: np.ndarray    # (n_tri, 9) float32
    ids: np.ndarray       # (n_tri,) int32 NTriangleID
    flags: np.ndarray     # (n_tri,) uint8
    p_origin: np.ndarray  # (n_patch, 3)
    p_normal: np.ndarray  # (n_patch, 3)
    p_area: np.ndarray    # (n_patch,)
    p_refl: np.ndarray    # (n_patch, 3)
    lights: np.ndarray    # LIGHT_DTYPE
    grid_x: np.ndarray = field(default=None)  # room cell boundaries along x (nx + 1)
    grid_y: np.ndarray = field(default=None)  # along y (ny + 1); room r * nx + c


def _take(h):
    nt, npch, nl = C.c_int64(), C.c_int64(), C.c_int32()
    check(lib.vr_gen_counts(h, C.byref(nt), C.byref(npch), C.byref(nl)))
    nt, npch, nl = nt.value, npch.value, nl.value
    s = Scene(np.zeros((nt, 9), np.float32), np.zeros(nt, np.int32), np.zeros(nt, np.uint8),
              np.zeros((npch, 3), np.float32), np.zeros((npch, 3), np.float32),
              np.zeros(npch, np.float32), np.zeros((npch, 3), np.float32),
              np.zeros(nl, LIGHT_DTYPE))
    check(lib.vr_gen_copy(h, ptr(s.verts9), ptr(s.ids), ptr(s.flags), ptr(s.p_origin),
                          ptr(s.p_normal), ptr(s.p_area), ptr(s.p_refl), ptr(s.lights)))
    gx, gy = C.c_int32(), C.c_int32()
    check(lib.vr_gen_grid(h, None, C.byref(gx), None, C.byref(gy)))
    s.grid_x, s.grid_y = np.zeros(gx.value + 1), np.zeros(gy.value + 1)
    check(lib.vr_gen_grid(h, ptr(s.grid_x), C.byref(gx), ptr(s.grid_y), C.byref(gy)))
    lib.vr_gen_destroy(h)
    return s


def room(seed=1, n_pillars=40, chop=64.0, tess=32.0):
    """Closed room 1024x1024x256, seeded pillars, one point light."""
    h = C.c_void_p()
    check(lib.vr_gen_room(C.byref(h), seed, n_pillars, chop, tess))
    return _take(h)


def level(seed=2, rooms_x=8, rooms_y=8, chop=32.0, tess=64.0, n_lights=16, sun=True):
    """Grid of connected rooms with pillars, ramps, skylights, lights, sun."""
    h = C.c_void_p()
    check(lib.vr_gen_level(C.byref(h), seed, rooms_x, rooms_y, chop, tess, n_lights,
                           1 if sun else 0))
    return _take(h)


def faces_of(kind="level", luxel_size=16.0, **kw):
    """The rectangles a generated scene's patches are chopped from, as
    vr_make_patches faces (lightmap vectors of 1/luxel_size along the edges)."""
    h = C.c_void_p()
    if kind == "room":
        a = dict(seed=1, n_pillars=40, chop=64.0, tess=32.0)
        a.update(kw)
        check(lib.vr_gen_room(C.byref(h), a["seed"], a["n_pillars"], a["chop"], a["tess"]))
    else:
        a = dict(seed=2, rooms_x=8, rooms_y=8, chop=32.0, tess=64.0, n_lights=16, sun=True)
        a.update(kw)
        check(lib.vr_gen_level(C.byref(h), a["seed"], a["rooms_x"], a["rooms_y"], a["chop"],
                               a["tess"], a["n_lights"], 1 if a["sun"] else 0))
    try:
        n = C.c_int64()
        check(lib.vr_gen_faces(h, float(luxel_size), None, C.byref(n)))
        out = np.zeros(n.value, _lib.FACE_DTYPE)
        check(lib.vr_gen_faces(h, float(luxel_size), ptr(out), C.byref(n)))
    finally:
        lib.vr_gen_destroy(h)
    return out


def compress_vis(rows):
    """Run-length encode PVS rows in the format DecompressVis reads (rad/
    lightmap/vis.go:53-94): nonzero bytes verbatim, a run of up to 255 zero
    bytes as 0 followed by its length.  Returns (data, offsets)."""
    data = bytearray()
    ofs = np.zeros(len(rows), np.int32)
    for c, row in enumerate(np.asarray(rows, np.uint8)):
        ofs[c] = len(data)
        j, n = 0, len(row)
        while j < n:
            if row[j]:
                data.append(int(row[j]))
                j += 1
                continue
            run = 1
            while j + run < n and row[j + run] == 0 and run < 255:
                run += 1
            data += bytes((0, run))
            j += run
    return np.frombuffer(bytes(data), np.uint8).copy(), ofs


def grid_bsp(grid_x, grid_y, pvs_hops=1, z_top=4096.0):
    """A BSP and PVS for a generated scene's room grid: planes z = 0 (the
    floor, so floor patches take the within-epsilon branch), x and y at the
    outer walls and z = z_top bound the rooms; inside, axis planes at the
    room boundaries split the grid down to one leaf per room (cluster r * nx
    + c); leaf 0 (cluster -1) is outside.  A room's PVS holds the rooms at
    most pvs_hops doorways away (rooms share a doorway with each grid
    neighbour).  Returns (pvs.Bsp, pvs.Vis)."""
    from .pvs import Bsp, Vis
    xs, ys = np.asarray(grid_x, np.float64), np.asarray(grid_y, np.float64)
    nx, ny = len(xs) - 1, len(ys) - 1
    planes, nodes, leaves = [], [], [-1]

    def plane(n, d):
        planes.append((n, d))
        return len(planes) - 1

    def node(pl):
        nodes.append([pl, 0, 0])
        return len(nodes) - 1

    def grid(c0, c1, r0, r1):
        if c1 - c0 == 1 and r1 - r0 == 1:
            leaves.append(r0 * nx + c0)
            return -1 - (len(leaves) - 1)
        if c1 - c0 >= r1 - r0:
            m = (c0 + c1) // 2
            k = node(plane((1, 0, 0), xs[m]))
            nodes[k][1], nodes[k][2] = grid(m, c1, r0, r1), grid(c0, m, r0, r1)
        else:
            m = (r0 + r1) // 2
            k = node(plane((0, 1, 0), ys[m]))
            nodes[k][1], nodes[k][2] = grid(c0, c1, m, r1), grid(c0, c1, r0, m)
        return k

    # bounding planes: front (children[0]) is the inside, back is leaf 0
    bounds = [((0, 0, 1), 0.0), ((0, 0, -1), -z_top), ((1, 0, 0), xs[0]),
              ((-1, 0, 0), -xs[-1]), ((0, 1, 0), ys[0]), ((0, -1, 0), -ys[-1])]
    chain = [node(plane(n, d)) for n, d in bounds]
    for a, b in zip(chain, chain[1:]):
        nodes[a][1], nodes[a][2] = b, -1
    nodes[chain[-1]][1], nodes[chain[-1]][2] = grid(0, nx, 0, ny), -1
    pl = np.zeros(len(planes), _lib.PLANE_DTYPE)
    for k, (n, d) in enumerate(planes):
        pl[k] = (n, d)
    nd = np.zeros(len(nodes), _lib.NODE_DTYPE)
    for k, (p, c0, c1) in enumerate(nodes):
        nd[k] = (p, (c0, c1))
    bsp = Bsp(pl, nd, np.array(leaves, np.int16))
    nc = nx * ny
    rows = np.zeros((nc, (nc + 7) >> 3), np.uint8)
    for src in range(nc):
        dist = {src: 0}
        q = deque([src])
        while q:
            u = q.popleft()
            if dist[u] == pvs_hops:
                continue
            r, c = divmod(u, nx)
            for rr, cc in ((r, c - 1), (r, c + 1), (r - 1, c), (r + 1, c)):
                if 0 <= rr < ny and 0 <= cc < nx and rr * nx + cc not in dist:
                    dist[rr * nx + cc] = dist[u] + 1
                    q.append(rr * nx + cc)
        for v in dist:
            rows[src, v >> 3] |= 1 << (v & 7)
    data, ofs = compress_vis(rows)
    return bsp, Vis(data, ofs, nc)
