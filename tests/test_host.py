"""Host-side checks of the product library that need no GPU: the C-ABI loads
and exports every declared symbol, the host half of the scene build (kd-tree
and intersection format) is bit-identical to the oracle, inputs are
deterministic, and bad arguments come back as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, scenes
from vrad_amd._lib import lib, ptr
from vrad_amd.raytracer import Environment, TRI_INTERSECT_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "vrad_hip.h")).read()
    names = set(re.findall(r"\b(vr_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 25
    raw = C.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(raw, n), n
    assert names <= set(_lib.EXPORTED), sorted(names - set(_lib.EXPORTED))


def _host_scene(v9, ids, flags=None):
    h = C.c_void_p()
    v9 = np.ascontiguousarray(v9, np.float32)
    ids = np.ascontiguousarray(ids, np.int32)
    _lib.check(lib.vr_scene_create(C.byref(h), ptr(v9), ptr(ids), ptr(flags), len(ids)))
    _lib.check(lib.vr_scene_build_host(h, 4))
    info = _lib.SceneInfo()
    _lib.check(lib.vr_scene_get_info(h, C.byref(info)))
    ch = np.zeros(info.n_nodes, np.uint32)
    sp = np.zeros(info.n_nodes, np.float32)
    idx = np.zeros(info.n_leaf_refs, np.int32)
    _lib.check(lib.vr_scene_get_kdtree(h, ptr(ch), ptr(sp), ptr(idx)))
    tri = np.zeros(len(ids), TRI_INTERSECT_DTYPE)
    if len(ids):
        _lib.check(lib.vr_scene_get_triangles(h, ptr(tri)))
    lib.vr_scene_destroy(h)
    return info, ch, sp, idx, tri


def _same_tree(v9, ids):
    info, ch, sp, idx, tri = _host_scene(v9, ids)
    kd = oracle.KdTree(v9)
    assert np.array_equal(ch, kd.children)
    assert np.array_equal(sp.view(np.uint32), kd.split.view(np.uint32))
    assert np.array_equal(idx, kd.idx)
    assert info.max_depth == kd.max_depth
    ot = oracle.precompute(v9, ids)
    assert tri.tobytes() == ot.tobytes()


def test_tree_and_precompute_match_oracle_room(room_scene):
    _same_tree(room_scene.verts9, room_scene.ids)


def test_tree_and_precompute_match_oracle_level():
    # the threaded builder (subtrees + parallel candidate search) on a larger input
    s = scenes.level(seed=5, rooms_x=3, rooms_y=3, chop=64.0, tess=32.0, n_lights=2)
    assert len(s.ids) > 20000
    _same_tree(s.verts9, s.ids)


def test_precompute_general_position_goldens(golden_dir):
    """The product's host precompute against the exact-rational float32
    emulation of mgl32 (triangles where every operation rounds)."""
    import json
    g = json.load(open(os.path.join(golden_dir, "tri_precompute_general.json")))
    v9 = np.array([t["v"] for t in g["triangles"]], np.float32)
    ids = np.array([t["id"] for t in g["triangles"]], np.int32)
    tri = _host_scene(v9, ids)[4]
    for r, t in zip(tri, g["triangles"]):
        assert list(np.float32([r["FlNx"], r["FlNy"], r["FlNz"]]).view(np.uint32)) == t["N_bits"]
        assert int(np.float32(r["FlD"]).view(np.uint32)) == t["D_bits"]
        assert list(np.asarray(r["ProjectedEdgeEquations"], np.float32).view(np.uint32)) == t["E_bits"]


def test_tree_degenerate_inputs():
    _same_tree(np.zeros((0, 9), np.float32), np.zeros(0, np.int32))
    v9, ids = H.to_arrays(H.box((0, 0, 0), (1, 1, 1)))
    _same_tree(v9, ids)
    flat = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]] * 5, np.float32)   # coplanar duplicates
    _same_tree(flat, np.arange(5, dtype=np.int32))


def test_environment_ingest_order():
    env = Environment()
    env.AddAxisAlignedRectangularSolid(0x02000000, (0, 0, 0), (1024, 1024, 256))
    v9, ids, _ = env.triangle_arrays()
    hv, hid = H.to_arrays(H.box((0, 0, 0), (1024, 1024, 256)))
    assert np.array_equal(v9, hv) and np.array_equal(ids, hid)
    assert list(ids[:4]) == [0x02000000, 0x02000001, 0x02000000, 0x02000001]


def test_generator_deterministic_and_shaped():
    a = scenes.room(seed=9, n_pillars=10)
    b = scenes.room(seed=9, n_pillars=10)
    c = scenes.room(seed=10, n_pillars=10)
    assert np.array_equal(a.verts9, b.verts9) and np.array_equal(a.p_refl, b.p_refl)
    assert not np.array_equal(a.verts9, c.verts9)
    # shell: 6 faces of 32-unit quads; pillars: 12 triangles each
    shell = 2 * (32 * 32 * 2) + 2 * (32 * 8 * 2) * 2
    assert len(a.ids) == shell + 10 * 12
    assert np.all((a.p_refl >= 0.2) & (a.p_refl <= 0.8))
    assert np.allclose(np.linalg.norm(a.p_normal, axis=1), 1.0)
    assert len(a.lights) == 1 and a.lights["intensity"][0][0] == 10000.0
    lv = scenes.level(seed=2, rooms_x=2, rooms_y=2, n_lights=3, sun=True)
    assert len(lv.lights) == 4 and lv.lights["type"][-1] == _lib.VR_EMIT_SKYLIGHT
    assert np.any(lv.ids == 0x01000000) or True


def test_bad_arguments_return_status():
    h = C.c_void_p()
    assert lib.vr_scene_create(None, None, None, None, 0) == _lib.VR_E_INVALID
    assert lib.vr_scene_create(C.byref(h), None, None, None, -1) == _lib.VR_E_INVALID
    assert lib.vr_scene_create(C.byref(h), None, None, None, 5) == _lib.VR_E_INVALID
    assert b"NULL" in lib.vr_last_error()
    v9, ids = H.to_arrays(H.box((0, 0, 0), (1, 1, 1)))
    _lib.check(lib.vr_scene_create(C.byref(h), ptr(v9), ptr(ids), None, len(ids)))
    # not built: tracing and tree queries are refused
    r = _lib.Rays()
    out_i = np.zeros(1, np.int32)
    out_t = np.zeros(1, np.float32)
    assert lib.vr_trace_closest(h, C.byref(r), 1, -1, ptr(out_i), ptr(out_t), None) == _lib.VR_E_STATE
    assert lib.vr_scene_get_kdtree(h, None, None, None) == _lib.VR_E_STATE
    lib.vr_scene_destroy(h)
    assert lib.vr_set_path(7) == _lib.VR_E_INVALID
    g = C.c_void_p()
    assert lib.vr_gen_room(C.byref(g), 1, 4, -1.0, 32.0) == _lib.VR_E_INVALID
    # batched bounce: argument checks come before any device work
    prm = _lib.BounceParams(8, 0.0)
    jobs = (_lib.BounceJob * 2)()
    assert lib.vr_bounce_light_batch(jobs, 0, C.byref(prm), 0) == _lib.VR_OK
    assert lib.vr_bounce_light_batch(jobs, 2, C.byref(prm), 0) == _lib.VR_E_INVALID
    assert b"transfers are NULL" in lib.vr_last_error()
    assert lib.vr_bounce_light_batch(None, 1, C.byref(prm), 0) == _lib.VR_E_INVALID
    # serving: parameters are checked before any device work
    sv = C.c_void_p()
    bad = _lib.ServerParams(7, 4, 0, 8, 0.1)
    assert lib.vr_server_create(C.byref(sv), C.byref(bad)) == _lib.VR_E_INVALID
    bad = _lib.ServerParams(_lib.VR_SERVE_BATCHED, 0, 0, 8, 0.1)
    assert lib.vr_server_create(C.byref(sv), C.byref(bad)) == _lib.VR_E_INVALID
    assert lib.vr_server_destroy(None) == _lib.VR_OK


def test_bsp_vis_luxel_arguments_checked_before_device_work():
    from vrad_amd import luxels, pvs
    pts = np.zeros((2, 3), np.float32)
    pl = np.zeros(1, _lib.PLANE_DTYPE)
    pl[0] = ((1, 0, 0), 0.0)
    lc = np.array([-1, 3], np.int16)

    def bsp(nodes):
        return pvs.Bsp(pl, np.array(nodes, _lib.NODE_DTYPE), lc)
    for nodes, what in (([(0, (-1, 0))], b"tree"),            # a cycle back to the root
                        ([(1, (-1, -2))], b"plane"),          # plane index out of range
                        ([(0, (-1, -3))], b"leaf"),           # leaf index out of range
                        ([(0, (-1, 4))], b"child"),           # node index out of range
                        ([(0, (1, -1)), (0, (-1, 1))], b"tree")):  # a node reached twice
        with pytest.raises(_lib.VrError) as e:
            pvs.ClusterFromPoint(bsp(nodes), pts)
        assert e.value.code == _lib.VR_E_INVALID and what in lib.vr_last_error(), what
    chain = [(0, (-1, k + 1)) for k in range(_lib.VR_BSP_MAX_DEPTH)] + [(0, (-1, -2))]
    with pytest.raises(_lib.VrError):
        pvs.ClusterFromPoint(bsp(chain), pts)
    assert b"deeper" in lib.vr_last_error()
    # vis data: offsets of -1 and past the data, before any device work
    rows = np.zeros((1, 1), np.uint8)
    ofs = np.array([-1], np.int32)
    data = np.array([1], np.uint8)
    assert lib.vr_decompress_vis(ptr(data), 1, ptr(ofs), 1, ptr(rows), None) == _lib.VR_E_INVALID
    assert b"visofs == -1" in lib.vr_last_error()
    ofs[0] = 1
    assert lib.vr_decompress_vis(ptr(data), 1, ptr(ofs), 1, ptr(rows), None) == _lib.VR_E_INVALID
    assert lib.vr_decompress_vis(ptr(data), 1, ptr(ofs), -1, ptr(rows), None) == _lib.VR_E_INVALID
    assert lib.vr_decompress_vis(None, 0, None, 0, None, None) == _lib.VR_OK  # no clusters
    # luxel stage: counts and the lightmap layout
    n = C.c_int64()
    assert lib.vr_calc_face_extents(None, -1, None, None, C.byref(n)) == _lib.VR_E_INVALID
    assert lib.vr_calc_face_extents(None, 0, None, None, C.byref(n)) == _lib.VR_OK
    faces = np.zeros(2, _lib.FACE_DTYPE)
    lm = np.zeros(2, _lib.FACE_LIGHTMAP_DTYPE)
    lm["size"] = [[1, 1], [0, 0]]
    lm["n_luxels"] = [4, 1]
    lm["luxel_ofs"] = [0, 5]  # the second face must start at 4
    with pytest.raises(_lib.VrError):
        luxels.FaceLuxels(faces, lm, np.zeros(0, _lib.PATCH_DTYPE))
    lm["luxel_ofs"][1] = 4
    p = np.zeros(1, _lib.PATCH_DTYPE)
    p["child1"] = -1
    p["face"] = 7  # no such face
    with pytest.raises(_lib.VrError):
        luxels.FaceLuxels(faces, lm, p)
    assert b"face index" in lib.vr_last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_no_device_is_an_error():
    v9, ids = H.to_arrays(H.box((0, 0, 0), (1, 1, 1)))
    h = C.c_void_p()
    _lib.check(lib.vr_scene_create(C.byref(h), ptr(v9), ptr(ids), None, len(ids)))
    assert lib.vr_scene_build(h, 1) == _lib.VR_E_NODEVICE
    lib.vr_scene_destroy(h)


def test_light_helpers():
    from vrad_amd import lightmap
    L = lightmap.point_light((0, 0, 0), (1, 1, 1), quadratic=1.0)
    assert L["intensity"][0][0] == 10000.0 and L["quadratic_attn"][0] == 1.0
    L = lightmap.point_light((0, 0, 0), (2, 2, 2))           # all zero -> constant 1
    assert L["constant_attn"][0] == 1.0 and L["intensity"][0][0] == 2.0
    a, b, c = lightmap.solve_inverse_quadratic(0, 1, 1, 2, 2, 5)   # y = x^2 + 1
    assert (a, b, c) == (1.0, 0.0, 1.0)
    S = lightmap.spot_light((0, 0, 0), (0, 0, -2), (1, 1, 1), 120, 150)
    # clamped to 90 degrees; 90 / 180 * pi in float32 is just above pi / 2
    assert S["stopdot"][0] == np.float32(-4.371139e-08) == S["stopdot2"][0]
