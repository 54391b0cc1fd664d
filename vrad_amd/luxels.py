"""The luxel stage on the GPU: lightmap extents per face (CalcFaceExtents,
rad/world/face.go:14-90) and each face's luxels -- positions on the face
plane, the leaf patch each takes its bounced light from -- plus the final
luxel light (direct light gathered at the luxel, bounced light from its
patch).  The reference has no patch-to-luxel step (FinalLightFace is
absent); the rule here is the build's."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


def _ofs(lux_ofs, n):
    if lux_ofs is None:
        return None
    o = np.ascontiguousarray(lux_ofs, np.float32).reshape(-1)
    if o.shape != (2 * n,):
        raise ValueError("lux_ofs must hold 2 floats per face")
    return o


def CalcFaceExtents(faces, lux_ofs=None):
    """Lightmap extents of every face (faces as _lib.FACE_DTYPE; lux_ofs
    the lightmapVecs[j][3] offsets, 2 per face).  Returns the
    _lib.FACE_LIGHTMAP_DTYPE records (mins, size, luxel_ofs, n_luxels,
    flags) and the luxel total."""
    f = np.ascontiguousarray(faces, _lib.FACE_DTYPE)
    o = _ofs(lux_ofs, len(f))
    out = np.zeros(len(f), _lib.FACE_LIGHTMAP_DTYPE)
    n = C.c_int64()
    check(lib.vr_calc_face_extents(ptr(f), len(f), ptr(o), ptr(out), C.byref(n)))
    return out, n.value


def FaceLuxels(faces, lm, patches, patch_value=None, lux_ofs=None):
    """Luxels of every face: dict of pos (n, 3), normal (n, 3), patch (n,)
    (the leaf patch the luxel takes its value from, -1 for none) and, with
    patch_value (3 per patch), value (n, 3)."""
    f = np.ascontiguousarray(faces, _lib.FACE_DTYPE)
    o = _ofs(lux_ofs, len(f))
    L = np.ascontiguousarray(lm, _lib.FACE_LIGHTMAP_DTYPE)
    p = np.ascontiguousarray(patches, _lib.PATCH_DTYPE)
    n = int(L["n_luxels"].sum())
    out = {"pos": np.zeros((n, 3), np.float32), "normal": np.zeros((n, 3), np.float32),
           "patch": np.zeros(n, np.int32)}
    pv = None
    if patch_value is not None:
        pv = np.ascontiguousarray(patch_value, np.float32).reshape(len(p), 3)
        out["value"] = np.zeros((n, 3), np.float32)
    check(lib.vr_face_luxels(ptr(f), len(f), ptr(o), ptr(L), ptr(p), len(p), ptr(pv),
                             ptr(out["pos"]), ptr(out["normal"]), ptr(out["patch"]),
                             ptr(out.get("value"))))
    return out


def FinalLightFace(env, faces, patches, patch_total, patch_direct, lights, lux_ofs=None):
    """Luxel light of every face: GatherSampleLight at the luxel (its
    position and the face normal) plus the bounced light of its patch
    (patch_total - patch_direct, 3 per patch).  Returns (lm records, luxel
    light (n, 3), luxels dict)."""
    from .lightmap import GatherSampleLight
    lm, _ = CalcFaceExtents(faces, lux_ofs)
    bounced = (np.asarray(patch_total, np.float32) - np.asarray(patch_direct, np.float32))
    lx = FaceLuxels(faces, lm, patches, bounced, lux_ofs)
    direct = GatherSampleLight(env, lx["pos"], lx["normal"], lights)
    return lm, (direct + lx["value"]).astype(np.float32), lx


def LightingLump(luxel_rgb, lm):
    """The BSP lighting lump of the luxel light (luxel_rgb (n, 3) in the
    luxel order of lm, the CalcFaceExtents records): (lump bytes (uint8),
    lightofs (int32 per face, -1 for a face without luxels)).  Per face: its
    average colour, then its luxels, as ColorRGBExp32."""
    v = np.ascontiguousarray(luxel_rgb, np.float32).reshape(-1, 3)
    L = np.ascontiguousarray(lm, _lib.FACE_LIGHTMAP_DTYPE)
    nb = C.c_int64()
    ofs = np.zeros(len(L), np.int32)
    check(lib.vr_lighting_lump(ptr(v), len(v), ptr(L), len(L), None, 0, ptr(ofs), C.byref(nb)))
    lump = np.zeros(nb.value, np.uint8)
    check(lib.vr_lighting_lump(ptr(v), len(v), ptr(L), len(L), ptr(lump), len(lump), ptr(ofs),
                               C.byref(nb)))
    return lump, ofs


def decode_rgbexp32(lump):
    """Colours of ColorRGBExp32 entries (4 bytes each): bytes x 2^exponent."""
    b = np.asarray(lump, np.uint8).reshape(-1, 4)
    e = b[:, 3].view(np.int8).astype(np.int32)
    return (b[:, :3].astype(np.float64) * np.ldexp(1.0, e)[:, None]).astype(np.float32)
