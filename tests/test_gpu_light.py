"""GatherSampleLight on the GPU vs the CPU oracle and closed forms."""
import math

import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, lightmap, scenes
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu


def _env(v9, ids):
    env = Environment()
    env.AddTriangles(v9, ids)
    env.SetupAccelerationStructure()
    return env


@pytest.mark.parametrize("path", [_lib.VR_PATH_OPT, _lib.VR_PATH_PLAIN], ids=["opt", "plain"])
def test_room_direct_light_bitexact(room_scene, path):
    s = room_scene
    print(f"[gpu] room direct light path={path}", flush=True)
    _lib.check(_lib.lib.vr_set_path(path))
    try:
        env = _env(s.verts9, s.ids)
        rgb = lightmap.GatherSampleLight(env, s.p_origin, s.p_normal, s.lights)
    finally:
        _lib.check(_lib.lib.vr_set_path(_lib.VR_PATH_OPT))
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    want = oracle.gather_light(kd, tri, s.p_origin, s.p_normal, s.lights, threads=8)
    assert np.array_equal(rgb.view(np.uint32), want.view(np.uint32))
    lit = (rgb[:, 0] > 0).mean()          # one low light among 40 full-height pillars
    assert 0.05 < lit < 0.95


def test_closed_forms_and_light_kinds():
    v9, ids = H.to_arrays(H.box((-10, -10, 40), (10, 10, 60)))
    env = _env(v9, ids)
    L = np.concatenate([lightmap.point_light((0, 0, 100), (1, 1, 1), quadratic=1.0),
                        lightmap.spot_light((300, 0, 100), (0, 0, -1), (1, 1, 1), 20, 40,
                                            exponent=2.0, quadratic=1.0)])
    pos = np.array([[0, 0, 0], [100, 0, 0], [300, 0, 0], [300 + 100 * math.tan(math.radians(30)), 0, 0]],
                   np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    rgb = lightmap.GatherSampleLight(env, pos, nrm, L)
    assert rgb[0, 0] == 0.0                                 # behind the box
    d2 = 100.0 ** 2 + 100.0 ** 2
    assert abs(rgb[1, 0] - 10000 * (100 / math.sqrt(d2)) / d2
               - 0.0) < 2e-6 + 1e-12                        # point light only (spot out of reach)
    tri = oracle.precompute(v9, ids)
    kd = oracle.KdTree(v9)
    want = oracle.gather_light(kd, tri, pos, nrm, L)
    # exponent != 0/1 uses powf, which may differ by ulps between libraries
    assert np.allclose(rgb, want, rtol=1e-6, atol=0)


def test_level_sun_and_sky():
    s = scenes.level(seed=2, rooms_x=4, rooms_y=4, chop=64.0, tess=64.0, n_lights=4, sun=True)
    print(f"[gpu] level sun/sky {len(s.ids)} tris {len(s.p_area)} patches", flush=True)
    env = _env(s.verts9, s.ids)
    amb = lightmap.make_lights(1)
    amb["type"] = _lib.VR_EMIT_SKYAMBIENT
    amb["intensity"] = (0.2, 0.25, 0.3)
    L = np.concatenate([s.lights, amb])
    idx = np.arange(0, len(s.p_area), 7)
    rgb = lightmap.GatherSampleLight(env, s.p_origin[idx], s.p_normal[idx], L)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    want = oracle.gather_light(kd, tri, s.p_origin[idx], s.p_normal[idx], L, threads=8)
    assert np.array_equal(rgb.view(np.uint32), want.view(np.uint32))
