"""Edge cases on the GPU: empty and tiny inputs, zero bounces, lights that
are absent, a request that moves no light, and server errors."""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, lightmap, scenes
from vrad_amd.patches import MakeTransfers, Transfers
from vrad_amd.raytracer import Environment
from vrad_amd.serve import BATCHED, Server

pytestmark = pytest.mark.gpu


def _box_env():
    v9, ids = H.to_arrays(H.box((0, 0, 0), (64, 64, 64)))
    env = Environment()
    env.AddTriangles(v9, ids)
    env.SetupAccelerationStructure()
    return env


def test_make_transfers_empty_and_tiny():
    print("[gpu] edges: transfers", flush=True)
    env = _box_env()
    z = np.zeros((0, 3), np.float32)
    xf = MakeTransfers(env, z, z, np.zeros(0, np.float32))
    assert xf.info().n_rows == 0 and xf.info().nnz == 0
    xf.close()
    one = MakeTransfers(env, np.array([[32, 32, 1]], np.float32), np.array([[0, 0, 1]], np.float32),
                        np.ones(1, np.float32))
    rp, col, coef = one.download()
    assert list(rp) == [0, 0] and len(col) == 0
    one.close()
    # two patches facing each other inside the box: one transfer each way
    o = np.array([[32, 32, 1], [32, 32, 63]], np.float32)
    n = np.array([[0, 0, 1], [0, 0, -1]], np.float32)
    two = MakeTransfers(env, o, n, np.full(2, 4.0, np.float32), normalize=False)
    rp, col, coef = two.download()
    assert list(rp) == [0, 1, 2] and list(col) == [1, 0]
    assert coef[0] == coef[1] > 0  # equal areas: F_01 = F_10
    two.close()


def test_gather_without_lights_or_samples():
    print("[gpu] edges: gather", flush=True)
    env = _box_env()
    pos = np.array([[32, 32, 1]], np.float32)
    nrm = np.array([[0, 0, 1]], np.float32)
    none = np.zeros(0, dtype=_lib.LIGHT_DTYPE)
    rgb = lightmap.GatherSampleLight(env, pos, nrm, none)
    assert rgb.shape == (1, 3) and np.all(rgb == 0)
    lights = lightmap.point_light((32, 32, 32), (1, 1, 1), quadratic=1.0)
    empty = lightmap.GatherSampleLight(env, np.zeros((0, 3), np.float32),
                                       np.zeros((0, 3), np.float32), lights)
    assert empty.shape == (0, 3)


def test_bounce_zero_bounces_and_empty_rows():
    print("[gpu] edges: bounce", flush=True)
    xf = Transfers.from_csr(np.array([0, 0, 1, 1], np.int64), np.array([0], np.int32),
                            np.array([0.5], np.float32))
    e0 = np.ones((3, 3), np.float32)
    refl = np.full((3, 3), 0.5, np.float32)
    area = np.ones(3, np.float32)
    tot0 = np.full((3, 3), 2.0, np.float32)
    tot, en, k = lightmap.BounceLight(xf, e0, refl, area, total=tot0, max_bounces=0)
    assert k == 0 and len(en) == 0 and np.array_equal(tot, tot0)
    tot, en, k = lightmap.BounceLight(xf, e0, refl, area, total=tot0, max_bounces=3, stop_frac=0.0)
    ot, oe, ok = oracle.bounce_f32(*xf.download(), e0, refl, area, total=tot0, max_bounces=3,
                                   stop_frac=0.0)
    assert k == ok and np.array_equal(tot.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(en, oe)
    b = lightmap.BounceLightBatch([(xf, e0, refl, area, tot0)], max_bounces=0)
    assert b[0][2] == 0 and np.array_equal(b[0][0], tot0)
    xf.close()


def test_server_request_without_moves_and_errors():
    print("[gpu] edges: server", flush=True)
    s = scenes.room(seed=8, n_pillars=4, chop=128.0, tess=128.0)
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure()
    xf = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area)
    xf.set_precision(_lib.VR_COEF_F16)
    sv = Server(mode=BATCHED, max_active=2, bounces=4)
    try:
        sid = sv.add_scene(env, xf, s.p_origin, s.p_normal, s.p_refl, s.p_area, s.lights)
        tot, en, lat = sv.wait(sv.submit(sid, []))
        # unmoved lights: the same as a direct gather plus 4 bounces
        direct = lightmap.GatherSampleLight(env, s.p_origin, s.p_normal, s.lights)
        emit0 = (direct * s.p_refl).astype(np.float32)
        t2, e2, k2 = lightmap.BounceLight(xf, emit0, s.p_refl, s.p_area, total=direct,
                                          max_bounces=4, stop_frac=0.0)
        assert k2 == 4 and np.array_equal(tot.view(np.uint32), t2.view(np.uint32))
        with pytest.raises(_lib.VrError):
            sv.submit(sid + 5, [])                   # unknown scene
        with pytest.raises(_lib.VrError):
            sv.submit(sid, [(99, (0.0, 0.0, 1.0))])  # light index out of range
        with pytest.raises(_lib.VrError):
            sv.wait((123456, sid))                   # unknown ticket
        with pytest.raises(_lib.VrError):            # scenes are added before requests
            sv.add_scene(env, xf, s.p_origin, s.p_normal, s.p_refl, s.p_area, s.lights)
    finally:
        sv.close()
        xf.close()
