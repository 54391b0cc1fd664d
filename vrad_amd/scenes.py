"""Procedural scenes (seeded) used as inputs for tests and benchmarks."""
import ctypes as C
from dataclasses import dataclass

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
