"""Batched traversal across scenes on the GPU: MakeTransfersBatch and
GatherSampleLightBatch equal the per-scene calls bit for bit (one shared
visibility launch per round through a table of scene descriptors), and
sampled scenes equal the CPU oracle."""
import numpy as np
import pytest

import oracle
from vrad_amd import _lib, lightmap, scenes
from vrad_amd.patches import MakeTransfers, MakeTransfersBatch
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu


def _env(s):
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure()
    return env


@pytest.fixture(scope="module")
def mixed():
    """Scenes of uneven size, with and without sun and sky triangles."""
    sc = [scenes.room(seed=11, n_pillars=6, chop=128.0, tess=64.0),
          scenes.level(seed=12, rooms_x=2, rooms_y=1, chop=96.0, tess=96.0, n_lights=3, sun=True),
          scenes.room(seed=13, n_pillars=20, chop=64.0, tess=64.0),
          scenes.level(seed=14, rooms_x=2, rooms_y=2, chop=128.0, tess=128.0, n_lights=5,
                       sun=False),
          scenes.room(seed=15, n_pillars=2, chop=256.0, tess=128.0)]
    return [(s, _env(s)) for s in sc]


def _same_csr(a, b):
    ra, ca, wa = a.download()
    rb, cb, wb = b.download()
    return (np.array_equal(ra, rb) and np.array_equal(ca, cb)
            and np.array_equal(wa.view(np.uint32), wb.view(np.uint32)))


@pytest.mark.parametrize("batch_pairs", [0, 500000, 40000], ids=["default", "split", "many-rounds"])
def test_make_transfers_batch_equals_single(mixed, batch_pairs):
    """Scenes with batch_pairs candidate pairs or more run on their own, the
    rest share launches: 0 (32M) shares all, 40000 runs most on their own."""
    print(f"[gpu] MakeTransfersBatch {len(mixed)} scenes batch_pairs={batch_pairs}", flush=True)
    items = [(env, s.p_origin, s.p_normal, s.p_area) for s, env in mixed]
    # an empty patch set and a single patch ride along
    s0, env0 = mixed[0]
    items.insert(2, (env0, s0.p_origin[:0], s0.p_normal[:0], s0.p_area[:0]))
    items.insert(4, (env0, s0.p_origin[:1], s0.p_normal[:1], s0.p_area[:1]))
    got = MakeTransfersBatch(items, batch_pairs=batch_pairs)
    n = np.array([len(a) for _, _, _, a in items], np.int64)
    solo = n * (n - 1) // 2 >= (batch_pairs or 32 << 20)
    print(f"[gpu] {int(solo.sum())} of {len(items)} scenes on their own", flush=True)
    if batch_pairs == 500000:
        assert 0 < solo.sum() < len(items) - 2
    for k, (env, o, nrm, a) in enumerate(items):
        one = MakeTransfers(env, o, nrm, a, batch_pairs=batch_pairs)
        assert _same_csr(got[k], one), k
        gi, oi = got[k].info(), one.info()
        assert (gi.pairs_considered, gi.segments_traced) == (oi.pairs_considered,
                                                             oi.segments_traced)


def test_make_transfers_batch_vs_oracle(mixed):
    pick = [0, 1, 4]
    items = [(mixed[k][1], mixed[k][0].p_origin, mixed[k][0].p_normal, mixed[k][0].p_area)
             for k in pick]
    print(f"[gpu] MakeTransfersBatch vs oracle on scenes {pick}", flush=True)
    got = MakeTransfersBatch(items, batch_pairs=60000)
    for k, xf in zip(pick, got):
        s = mixed[k][0]
        rp, col, coef = xf.download()
        orp, ocol, ocoef, _ = oracle.make_transfers(oracle.KdTree(s.verts9),
                                                    oracle.precompute(s.verts9, s.ids),
                                                    s.p_origin, s.p_normal, s.p_area, threads=8)
        assert np.array_equal(np.diff(rp), np.diff(orp))
        assert np.array_equal(col, ocol) and np.array_equal(coef.view(np.uint32),
                                                            ocoef.view(np.uint32))


def test_gather_batch_equals_single_and_oracle(mixed):
    print(f"[gpu] GatherSampleLightBatch {len(mixed)} scenes", flush=True)
    items = []
    for s, env in mixed:
        amb = lightmap.make_lights(1)
        amb["type"] = _lib.VR_EMIT_SKYAMBIENT
        amb["intensity"] = (0.1, 0.2, 0.3)
        items.append((env, s.p_origin, s.p_normal, np.concatenate([s.lights, amb])))
    s0, env0 = mixed[0]
    items.insert(1, (env0, s0.p_origin[:0], s0.p_normal[:0], s0.lights))      # no samples
    items.insert(3, (env0, s0.p_origin[:70], s0.p_normal[:70], s0.lights[:0]))  # no lights
    got = lightmap.GatherSampleLightBatch(items)
    for k, (env, pos, nrm, L) in enumerate(items):
        one = lightmap.GatherSampleLight(env, pos, nrm, L)
        assert np.array_equal(got[k].view(np.uint32), one.view(np.uint32)), k
    assert not got[3].any()
    for k in (0, 2):
        env, pos, nrm, L = items[k]
        s = mixed[0 if k == 0 else 1][0]
        want = oracle.gather_light(oracle.KdTree(s.verts9), oracle.precompute(s.verts9, s.ids),
                                   pos, nrm, L, threads=8)
        assert np.array_equal(got[k].view(np.uint32), want.view(np.uint32))


def test_batch_calls_refuse_bad_scenes(mixed):
    import ctypes as C

    class Unbuilt:                      # a scene handle created but never built
        def __init__(self, v9, ids):
            self.handle = C.c_void_p()
            _lib.check(_lib.lib.vr_scene_create(C.byref(self.handle), _lib.ptr(v9),
                                                _lib.ptr(ids), None, len(ids)))

    s, env = mixed[0]
    bad = Unbuilt(np.ascontiguousarray(s.verts9[:4]), np.ascontiguousarray(s.ids[:4]))
    try:
        items = [(env, s.p_origin, s.p_normal, s.p_area)]
        with pytest.raises(_lib.VrError) as ei:
            MakeTransfersBatch(items + [(bad, s.p_origin, s.p_normal, s.p_area)])
        assert ei.value.code == _lib.VR_E_STATE and "scene 1" in str(ei.value)
        with pytest.raises(_lib.VrError) as ei:
            lightmap.GatherSampleLightBatch([(bad, s.p_origin, s.p_normal, s.lights)])
        assert ei.value.code == _lib.VR_E_STATE and "scene 0" in str(ei.value)
    finally:
        _lib.lib.vr_scene_destroy(bad.handle)
