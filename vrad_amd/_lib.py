"""ctypes binding of libvrad_hip.so (the C-ABI in include/vrad_hip.h).

The product path has no fallback: if the shared library is missing the
import of this module raises, and every entry point that returns a non-zero
status raises VrError with the library's message.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvrad_hip.so")

VR_OK = 0
VR_E_INVALID, VR_E_NOMEM, VR_E_HIP, VR_E_UNSUPPORTED = -1, -2, -3, -4
VR_E_STATE, VR_E_NODEVICE, VR_E_CAPACITY = -5, -6, -7
VR_PATH_OPT, VR_PATH_PLAIN = 0, 1
VR_COEF_F32, VR_COEF_F16, VR_COEF_BF16 = 0, 1, 2
VR_EMIT_SURFACE, VR_EMIT_POINT, VR_EMIT_SPOTLIGHT = 0, 1, 2
VR_EMIT_SKYLIGHT, VR_EMIT_SKYAMBIENT = 3, 5
VR_SERVE_BATCHED, VR_SERVE_PLAIN = 0, 1


class VrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vrad_hip error {code}: {msg}")
        self.code = code


FP = C.POINTER(C.c_float)


class Rays(C.Structure):
    _fields_ = [(n, FP) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class Segs(C.Structure):
    _fields_ = [(n, FP) for n in ("sx", "sy", "sz", "ex", "ey", "ez")]


class Patches(C.Structure):
    _fields_ = [(n, FP) for n in ("ox", "oy", "oz", "nx", "ny", "nz", "area")]


class Samples(C.Structure):
    _fields_ = [(n, FP) for n in ("px", "py", "pz", "nx", "ny", "nz")]


class XferParams(C.Structure):
    _fields_ = [("normal_offset", C.c_float), ("skip_id", C.c_int32),
                ("normalize", C.c_int32), ("batch_pairs", C.c_int64)]


class XfersInfo(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("nnz", C.c_int64), ("stored_nnz", C.c_int64),
                ("stored_bytes", C.c_int64), ("precision", C.c_int32),
                ("pairs_considered", C.c_int64), ("segments_traced", C.c_int64)]


class SceneInfo(C.Structure):
    _fields_ = [("n_tri", C.c_int64), ("n_nodes", C.c_int64), ("n_leaf_refs", C.c_int64),
                ("max_depth", C.c_int32), ("has_sky", C.c_int32),
                ("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("device_bytes", C.c_int64)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("origin", C.c_float * 3), ("intensity", C.c_float * 3),
                ("normal", C.c_float * 3), ("constant_attn", C.c_float),
                ("linear_attn", C.c_float), ("quadratic_attn", C.c_float),
                ("stopdot", C.c_float), ("stopdot2", C.c_float), ("exponent", C.c_float),
                ("start_fade", C.c_float), ("end_fade", C.c_float),
                ("cap_distance", C.c_float)]


# vr_face / vr_patch records
VR_FACE_MAX_POINTS = 16  # points per vr_face / vr_patch winding
FACE_DTYPE = np.dtype([("n_points", "<i4"), ("points", "<f4", (16, 3)), ("normal", "<f4", 3),
                       ("lux_s", "<f4", 3), ("lux_t", "<f4", 3), ("base_light", "<f4", 3),
                       ("flags", "<u4")])
PATCH_DTYPE = np.dtype([("face", "<i4"), ("parent", "<i4"), ("child1", "<i4"), ("child2", "<i4"),
                        ("origin", "<f4", 3), ("normal", "<f4", 3), ("area", "<f4"),
                        ("chop", "<f4"), ("lux_scale", "<f4"), ("mins", "<f4", 3),
                        ("maxs", "<f4", 3), ("n_points", "<i4"), ("points", "<f4", (16, 3))])
assert FACE_DTYPE.itemsize == 248 and PATCH_DTYPE.itemsize == 272
VR_FACE_SKY, VR_FACE_NOCHOP = 1, 2

VR_FACE_NOLIGHT = 4
VR_LIGHTMAP_MAX_DIM = 125
VR_LM_UNLIT, VR_LM_TOO_BIG, VR_LM_NO_POINTS, VR_LM_DEGENERATE = 1, 2, 4, 8
FACE_LIGHTMAP_DTYPE = np.dtype([("mins", "<i4", 2), ("size", "<i4", 2), ("luxel_ofs", "<i8"),
                                ("n_luxels", "<i4"), ("flags", "<u4")])
assert FACE_LIGHTMAP_DTYPE.itemsize == 32

# BSP lumps for ClusterFromPoint: vr_bsp_plane, vr_bsp_node
PLANE_DTYPE = np.dtype([("normal", "<f4", 3), ("dist", "<f4")])
NODE_DTYPE = np.dtype([("planenum", "<i4"), ("children", "<i4", 2)])
VR_BSP_MAX_DEPTH = 64


class Bsp(C.Structure):
    _fields_ = [("planes", C.c_void_p), ("n_planes", C.c_int64), ("nodes", C.c_void_p),
                ("n_nodes", C.c_int64), ("leaf_cluster", C.c_void_p), ("n_leafs", C.c_int64)]

LIGHT_DTYPE = np.dtype([("type", "<i4"), ("origin", "<f4", 3), ("intensity", "<f4", 3),
                        ("normal", "<f4", 3), ("constant_attn", "<f4"), ("linear_attn", "<f4"),
                        ("quadratic_attn", "<f4"), ("stopdot", "<f4"), ("stopdot2", "<f4"),
                        ("exponent", "<f4"), ("start_fade", "<f4"), ("end_fade", "<f4"),
                        ("cap_distance", "<f4")])
assert LIGHT_DTYPE.itemsize == C.sizeof(Light)


class BounceParams(C.Structure):
    _fields_ = [("max_bounces", C.c_int32), ("stop_frac", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("max_active", C.c_int32), ("hbm_budget", C.c_int64),
                ("bounces", C.c_int32), ("normal_offset", C.c_float)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "completed", "iterations",
                                         "max_active_seen", "pending", "active")]


class BounceJob(C.Structure):
    _fields_ = [("xfers", C.c_void_p), ("emit0", FP), ("refl", FP), ("area", FP),
                ("total", FP), ("energy", C.POINTER(C.c_double)), ("n_bounces", C.c_int32)]


# name -> (restype, argtypes); every function returns int status unless noted
_P = C.c_void_p
_SIGS = {
    "vr_last_error": (C.c_char_p, []),
    "vr_init": (C.c_int, [C.c_int]),
    "vr_set_path": (C.c_int, [C.c_int]),
    "vr_set_tuning": (C.c_int, [C.c_char_p, C.c_int]),
    "vr_get_tuning": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
    "vr_profile_enable": (C.c_int, [C.c_int]),
    "vr_profile_reset": (C.c_int, []),
    "vr_profile_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "vr_measure_read_bandwidth": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "vr_scene_create": (C.c_int, [C.POINTER(_P), _P, _P, _P, C.c_int64]),
    "vr_scene_build": (C.c_int, [_P, C.c_int32]),
    "vr_scene_build_host": (C.c_int, [_P, C.c_int32]),
    "vr_scene_get_info": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "vr_scene_get_kdtree": (C.c_int, [_P, _P, _P, _P]),
    "vr_scene_get_triangles": (C.c_int, [_P, _P]),
    "vr_scene_destroy": (C.c_int, [_P]),
    "vr_trace_closest": (C.c_int, [_P, C.POINTER(Rays), C.c_int64, C.c_int32, _P, _P, _P]),
    "vr_test_line": (C.c_int, [_P, C.POINTER(Segs), C.c_int64, C.c_int32, _P, _P]),
    "vr_trace_closest_dev": (C.c_int, [_P, C.POINTER(Rays), C.c_int64, C.c_int32, _P, _P, _P, _P]),
    "vr_test_line_dev": (C.c_int, [_P, C.POINTER(Segs), C.c_int64, C.c_int32, _P, _P, _P]),
    "vr_make_transfers": (C.c_int, [_P, C.POINTER(Patches), C.c_int64, C.POINTER(XferParams),
                                    C.POINTER(_P)]),
    "vr_xfers_from_csr": (C.c_int, [C.POINTER(_P), C.c_int64, _P, _P, _P, C.c_int32]),
    "vr_xfers_synthetic": (C.c_int, [C.POINTER(_P), C.c_uint64, C.c_int64, C.c_double,
                                     C.c_double, C.c_int64, C.c_int32, C.c_int32]),
    "vr_xfers_set_precision": (C.c_int, [_P, C.c_int32]),
    "vr_xfers_get_info": (C.c_int, [_P, C.POINTER(XfersInfo)]),
    "vr_xfers_download": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "vr_xfers_destroy": (C.c_int, [_P]),
    "vr_gather_sample_light": (C.c_int, [_P, C.POINTER(Samples), C.c_int64, _P, C.c_int32,
                                         C.c_float, C.c_int32, _P]),
    "vr_can_leaf_trace_to_sky": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "vr_lighting_lump": (C.c_int, [_P, C.c_int64, _P, C.c_int64, _P, C.c_int64, _P,
                                   C.POINTER(C.c_int64)]),
    "vr_make_transfers_batch": (C.c_int, [_P, _P, _P, C.c_int32, C.POINTER(XferParams), _P]),
    "vr_gather_sample_light_batch": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32, C.c_float,
                                               C.c_int32, _P]),
    "vr_bounce_light": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(BounceParams), _P,
                                  C.POINTER(C.c_int32)]),
    "vr_bounce_light_noalloc": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(BounceParams), _P,
                                          C.POINTER(C.c_int32)]),
    "vr_bounce_light_batch": (C.c_int, [C.POINTER(BounceJob), C.c_int32, C.POINTER(BounceParams),
                                        C.c_int32]),
    "vr_gen_room": (C.c_int, [C.POINTER(_P), C.c_uint64, C.c_int32, C.c_float, C.c_float]),
    "vr_gen_level": (C.c_int, [C.POINTER(_P), C.c_uint64, C.c_int32, C.c_int32, C.c_float,
                               C.c_float, C.c_int32, C.c_int32]),
    "vr_gen_counts": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32)]),
    "vr_gen_copy": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "vr_gen_faces": (C.c_int, [_P, C.c_float, _P, C.POINTER(C.c_int64)]),
    "vr_make_patches": (C.c_int, [_P, C.c_int64, C.c_float, C.c_float, _P, C.c_int64,
                                  C.POINTER(C.c_int64)]),
    "vr_make_patches_phong": (C.c_int, [_P, C.c_int64, C.c_float, C.c_float, _P, _P,
                                        C.c_int64, C.POINTER(C.c_int64)]),
    "vr_pair_edges": (C.c_int, [_P, C.c_int64, C.c_float, _P]),
    "vr_gen_grid": (C.c_int, [_P, _P, C.POINTER(C.c_int32), _P, C.POINTER(C.c_int32)]),
    "vr_gen_destroy": (C.c_int, [_P]),
    "vr_calc_face_extents": (C.c_int, [_P, C.c_int64, _P, _P, C.POINTER(C.c_int64)]),
    "vr_face_luxels": (C.c_int, [_P, C.c_int64, _P, _P, _P, C.c_int64, _P, _P, _P, _P, _P]),
    "vr_cluster_from_points": (C.c_int, [C.POINTER(Bsp), _P, C.c_int64, _P]),
    "vr_patch_clusters": (C.c_int, [C.POINTER(Bsp), _P, C.c_int64, _P]),
    "vr_decompress_vis": (C.c_int, [_P, C.c_int64, _P, C.c_int32, _P, C.POINTER(C.c_int64)]),
    "vr_make_transfers_pvs": (C.c_int, [_P, C.POINTER(Patches), C.c_int64,
                                        C.POINTER(XferParams), _P, _P, C.c_int32,
                                        C.POINTER(_P)]),
    "vr_server_create": (C.c_int, [C.POINTER(_P), C.POINTER(ServerParams)]),
    "vr_server_add_scene": (C.c_int, [_P, _P, _P, C.POINTER(Samples), _P, _P, _P, C.c_int32,
                                      C.POINTER(C.c_int32)]),
    "vr_server_submit": (C.c_int, [_P, C.c_int32, _P, _P, C.c_int32, C.POINTER(C.c_int64)]),
    "vr_server_wait": (C.c_int, [_P, C.c_int64, C.c_int32, _P, _P, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double)]),
    "vr_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "vr_server_destroy": (C.c_int, [_P]),
}
EXPORTED = tuple(_SIGS)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C vrad_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    if rc != VR_OK:
        raise VrError(rc, lib.vr_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return C.c_void_p(a.ctypes.data)


def fptr(a):
    return None if a is None else a.ctypes.data_as(FP)


def f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a
