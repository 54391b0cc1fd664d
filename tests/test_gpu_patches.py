"""Patch generation on the GPU (vr_make_patches) against the CPU restatement
of MakePatchForFace + SubdividePatches: every field of every patch, bit for
bit, in the reference's array order."""
import time

import numpy as np
import pytest

import oracle
from vrad_amd import _lib, scenes
from vrad_amd.patches import MakePatches

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 0], ids=["level-parallel", "thread-per-face"])
def build(request):
    """Both patch builders (knob patch_build) must give the same patches."""
    print(f"[patches] patch_build={request.param}", flush=True)
    _lib.check(_lib.lib.vr_set_tuning(b"patch_build", request.param))
    yield request.param
    _lib.check(_lib.lib.vr_set_tuning(b"patch_build", 1))


def _same(faces, label, **kw):
    print(f"[patches] {label}: {len(faces)} faces", flush=True)
    t0 = time.perf_counter()
    g = MakePatches(faces, **kw)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    o = oracle.make_patches(faces, **kw)
    dto = time.perf_counter() - t0
    print(f"[patches] {label}: {len(g)} patches ({int((g['child1'] < 0).sum())} leaves), "
          f"GPU {dt * 1e3:.1f} ms, CPU restatement {dto * 1e3:.1f} ms", flush=True)
    assert len(g) == len(o)
    assert g.tobytes() == o.tobytes()
    return g


def _random_convex(rng, n):
    """Planar convex polygons of 3..8 points in random orientations, sizes
    and lightmap scales."""
    f = np.zeros(n, _lib.FACE_DTYPE)
    for k in range(n):
        m = int(rng.integers(3, 9))
        ang = np.sort(rng.uniform(0, 2 * np.pi, m))
        r = rng.uniform(20, 400)
        pts2 = np.stack([np.cos(ang) * r * rng.uniform(0.5, 1.5), np.sin(ang) * r], 1)
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        b = np.cross(a, rng.normal(size=3))
        b /= np.linalg.norm(b)
        c = rng.uniform(-2000, 2000, 3)
        if k % 3 == 0:  # axis-aligned, on integer coordinates
            a, b = np.eye(3)[k % 3], np.eye(3)[(k + 1) % 3]
            pts2 = np.round(pts2)
            c = np.round(c)
        P = c + pts2[:, :1] * a + pts2[:, 1:] * b
        f["n_points"][k] = m
        f["points"][k, :m] = P
        f["normal"][k] = np.cross(a, b)
        lux = rng.choice([8.0, 16.0, 32.0])
        f["lux_s"][k] = a / lux
        f["lux_t"][k] = b / lux
        if k % 7 == 0:
            f["base_light"][k] = [1, 1, 1]  # light surface: no edge refinement
    return f


def test_gpu_patches_generated_scenes(build):
    _same(scenes.faces_of("room"), "room")
    _same(scenes.faces_of("level"), "level 8x8")
    _same(scenes.faces_of("level", luxel_size=8.0), "level 8x8, 8 units per luxel")


def test_gpu_patches_random_polygons_and_chops(build):
    rng = np.random.default_rng(17)
    f = _random_convex(rng, 3000)
    f["flags"][::11] = _lib.VR_FACE_SKY
    f["flags"][5::13] = _lib.VR_FACE_NOCHOP
    f["n_points"][3::97] = 2  # degenerate: no patch
    _same(f, "random convex polygons")
    # maxChop above minChop exercises the chop halving rules
    _same(f, "random convex polygons, maxChop 16 / minChop 4", max_chop=16.0, min_chop=4.0)


def test_gpu_patches_count_only_and_capacity(build):
    f = scenes.faces_of("room")
    import ctypes as C
    n = C.c_int64()
    fa = np.ascontiguousarray(f, _lib.FACE_DTYPE)
    _lib.check(_lib.lib.vr_make_patches(_lib.ptr(fa), len(fa), 4.0, 4.0, None, 0, C.byref(n)))
    small = np.zeros(max(n.value - 1, 1), _lib.PATCH_DTYPE)
    rc = _lib.lib.vr_make_patches(_lib.ptr(fa), len(fa), 4.0, 4.0, _lib.ptr(small), len(small),
                                  C.byref(n))
    assert rc == _lib.VR_E_CAPACITY


def test_gpu_patches_phong_normals(build):
    """Children take GetPhongNormal normals from PairEdges' vertex normals
    (rad/lightmap/normallist.go:51-144): bit for bit against the oracle, on
    a generated level with some normals jittered into shallow creases and on
    random polygons with random unit vertex normals."""
    from vrad_amd.patches import PairEdges
    rng = np.random.default_rng(23)
    f = scenes.faces_of("level", rooms_x=3, rooms_y=2, chop=64.0, tess=64.0)
    pick = rng.random(len(f)) < 0.3
    nn = f["normal"][pick] + rng.normal(scale=0.2, size=(pick.sum(), 3)).astype(np.float32)
    f["normal"][pick] = (nn / np.linalg.norm(nn, axis=1, keepdims=True)).astype(np.float32)
    vn = PairEdges(f)
    assert np.array_equal(vn.view(np.uint32), oracle.pair_edges(f).view(np.uint32))
    g = _same(f, "level 3x2, phong normals", vertex_normals=vn)
    assert (g["normal"][g["parent"] >= 0] != f["normal"][g["face"][g["parent"] >= 0]]).any()
    r = _random_convex(rng, 1500)
    rv = rng.normal(size=(len(r), 16, 3))
    rv = (rv / np.linalg.norm(rv, axis=2, keepdims=True)).astype(np.float32)
    _same(r, "random convex polygons, random vertex normals", vertex_normals=rv)
