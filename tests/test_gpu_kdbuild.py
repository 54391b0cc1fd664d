"""kd-tree built on the GPU (vr_scene_build with kd_build = 1) against the host
builder (kd_build = 0), whose tree the oracle pins: the same nodes, split
values, triangle index list, bounds and depth, bit for bit.  Visibility on a
device-built tree is checked by every GPU visibility test, since scenes of
2048 triangles and more build on the GPU by default."""
import ctypes as C
import time

import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, autotune, scenes
from vrad_amd._lib import lib, ptr

pytestmark = pytest.mark.gpu


def _tree(v9, ids, mode):
    v9 = np.ascontiguousarray(v9, np.float32)
    ids = np.ascontiguousarray(ids, np.int32)
    old = autotune.get_knob("kd_build")
    autotune.set_knob("kd_build", mode)
    h = C.c_void_p()
    try:
        _lib.check(lib.vr_scene_create(C.byref(h), ptr(v9), ptr(ids), None, len(ids)))
        t0 = time.perf_counter()
        _lib.check(lib.vr_scene_build(h, 16))
        dt = time.perf_counter() - t0
        info = _lib.SceneInfo()
        _lib.check(lib.vr_scene_get_info(h, C.byref(info)))
        ch = np.zeros(info.n_nodes, np.uint32)
        sp = np.zeros(info.n_nodes, np.float32)
        idx = np.zeros(info.n_leaf_refs, np.int32)
        _lib.check(lib.vr_scene_get_kdtree(h, ptr(ch), ptr(sp), ptr(idx)))
    finally:
        autotune.set_knob("kd_build", old)
        if h.value:
            lib.vr_scene_destroy(h)
    return info, ch, sp, idx, dt


def _same(v9, ids, label, against_oracle=False):
    print(f"[kdbuild] {label}: {len(ids)} triangles", flush=True)
    hi, hch, hsp, hidx, ht = _tree(v9, ids, 0)
    di, dch, dsp, didx, dt = _tree(v9, ids, 1)
    print(f"[kdbuild] {label}: host {ht * 1e3:.1f} ms, device {dt * 1e3:.1f} ms, "
          f"{len(dch)} nodes, {len(didx)} refs, depth {di.max_depth}", flush=True)
    assert len(dch) == len(hch) and len(didx) == len(hidx)
    assert np.array_equal(dch, hch)
    assert np.array_equal(dsp.view(np.uint32), hsp.view(np.uint32))
    assert np.array_equal(didx, hidx)
    assert di.max_depth == hi.max_depth
    assert np.array_equal(np.array(di.bmin[:], np.float32).view(np.uint32),
                          np.array(hi.bmin[:], np.float32).view(np.uint32))
    assert np.array_equal(np.array(di.bmax[:], np.float32).view(np.uint32),
                          np.array(hi.bmax[:], np.float32).view(np.uint32))
    if against_oracle:
        kd = oracle.KdTree(np.ascontiguousarray(v9, np.float32))
        assert np.array_equal(dch, kd.children)
        assert np.array_equal(dsp.view(np.uint32), kd.split.view(np.uint32))
        assert np.array_equal(didx, kd.idx)


def test_device_tree_degenerate_inputs():
    _same(np.zeros((0, 9), np.float32), np.zeros(0, np.int32), "empty", True)
    one = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    _same(one, np.zeros(1, np.int32), "one triangle", True)
    _same(np.repeat(one, 2, 0), np.arange(2, dtype=np.int32), "two triangles", True)
    _same(np.repeat(one, 3, 0), np.arange(3, dtype=np.int32), "three coplanar", True)
    _same(np.repeat(one, 200, 0), np.arange(200, dtype=np.int32), "200 identical", True)
    v9, ids = H.to_arrays(H.box((0, 0, 0), (1, 1, 1)))
    _same(v9, ids, "box", True)


def test_device_tree_signed_zeros_and_flat():
    rng = np.random.default_rng(7)
    n = 3000
    v9 = rng.integers(-4, 5, (n, 9)).astype(np.float32) * 0.5
    # -0 and +0 coordinates mixed (equal under comparison, different bits)
    z = v9 == 0
    v9[z] = np.where(rng.random(z.sum()) < 0.5, np.float32(-0.0), np.float32(0.0))
    # a third of the triangles flat on an axis plane (lo == hi on that axis)
    ax = rng.integers(0, 3, n)
    flat = rng.random(n) < 0.33
    for k in np.nonzero(flat)[0]:
        v9[k, ax[k]::3] = v9[k, ax[k]]
    _same(v9, np.arange(n, dtype=np.int32), "integer lattice with signed zeros", True)


def test_device_tree_random_soup():
    rng = np.random.default_rng(11)
    n = 20000
    c = rng.random((n, 1, 3)).astype(np.float32) * 1000.0
    v9 = (c + rng.normal(0, 5, (n, 3, 3)).astype(np.float32)).reshape(n, 9)
    _same(v9.astype(np.float32), np.arange(n, dtype=np.int32), "random soup")


def test_device_tree_room(room_scene):
    _same(room_scene.verts9, room_scene.ids, "room", True)


def test_device_tree_levels():
    s = scenes.level(seed=5, rooms_x=3, rooms_y=3, chop=64.0, tess=32.0, n_lights=2)
    _same(s.verts9, s.ids, "level 3x3")
    s = scenes.level(seed=2, rooms_x=8, rooms_y=8, chop=32.0, tess=64.0, n_lights=16, sun=True)
    _same(s.verts9, s.ids, "level 8x8")
