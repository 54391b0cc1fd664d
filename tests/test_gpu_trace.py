"""Visibility on the GPU vs the CPU oracle: hit index and hit t bit-exact."""
import json
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, autotune, scenes
from vrad_amd.raytracer import Environment, FourRays, RayTracingResult
from vrad_amd.trace import TestLineDoesHitSky

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _env(v9, ids):
    env = Environment()
    env.AddTriangles(v9, ids)
    env.SetupAccelerationStructure()
    return env


@pytest.fixture(params=[(_lib.VR_PATH_OPT, 10, 1), (_lib.VR_PATH_OPT, 8, 1), (_lib.VR_PATH_OPT, 4, 1),
                        (_lib.VR_PATH_OPT, 24, 1), (_lib.VR_PATH_OPT, 2, 1), (_lib.VR_PATH_OPT, 2, 8),
                        (_lib.VR_PATH_OPT, 6, 2), (_lib.VR_PATH_OPT, 8, 2), (_lib.VR_PATH_OPT, 4, 4),
                        (_lib.VR_PATH_OPT, 10, 4), (_lib.VR_PATH_OPT, 8, 8), (_lib.VR_PATH_PLAIN, 8, 1)],
                ids=["opt-lds10", "opt-lds8", "opt-lds4", "opt-lds24", "opt-lds2", "opt-lds2-r8", "opt-lds6-r2",
                     "opt-lds8-r2", "opt-lds4-r4", "opt-lds10-r4", "opt-lds8-r8",
                     "plain"])
def path(request):
    p, lds, rays = request.param
    print(f"[gpu] path={p} trace_lds_entries={lds} rays_per_lane={rays}", flush=True)
    old = autotune.get_knob("trace_lds_entries")
    _lib.check(_lib.lib.vr_set_path(p))
    _lib.check(_lib.lib.vr_set_tuning(b"trace_lds_entries", lds))
    _lib.check(_lib.lib.vr_set_tuning(b"trace_rays_per_lane", rays))
    yield p
    _lib.check(_lib.lib.vr_set_path(_lib.VR_PATH_OPT))
    _lib.check(_lib.lib.vr_set_tuning(b"trace_lds_entries", old))
    _lib.check(_lib.lib.vr_set_tuning(b"trace_rays_per_lane", 1))


@pytest.fixture(scope="module")
def room_env(room_scene):
    return _env(room_scene.verts9, room_scene.ids)


def test_analytic_rays(golden_dir, path):
    g = json.load(open(os.path.join(golden_dir, "analytic_rays.json")))
    v9, ids = H.to_arrays(H.box(g["box_min"], g["box_max"], g["id"]))
    env = _env(v9, ids)
    for r in g["rays"]:
        hit, t = env.TraceRays(np.array([r["o"]]), np.array([r["d"]]), r.get("tmin", 0.0),
                               r.get("tmax", 1e30))
        assert hit[0] == r["hit"] and t[0] == np.float32(r["t"]), r


def test_random_rays_bitexact_vs_brute(room_scene, room_env, path):
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    rng = np.random.default_rng(11)
    o, d = H.random_rays(rng, 20000, [-20, -20, -20], [1044, 1044, 276])
    d[:2000] = np.eye(3, dtype=np.float32)[rng.integers(0, 3, 2000)] * \
        rng.choice([-1, 1], (2000, 1)).astype(np.float32)
    tmin = rng.uniform(0, 50, 20000).astype(np.float32)
    tmax = rng.uniform(100, 2000, 20000).astype(np.float32)
    print(f"[gpu] 20000 rays, {len(s.ids)} triangles", flush=True)
    hg, tg, ng = room_env.TraceRays(o, d, tmin, tmax, normals=True)
    hb, tb = oracle.trace_brute(tri, o, d, tmin, tmax)
    assert np.array_equal(hg, hb)
    assert np.array_equal(_bits(tg), _bits(tb))
    m = hb >= 0
    want_n = np.stack([tri["nx"], tri["ny"], tri["nz"]], 1)[hb[m]]
    assert np.array_equal(ng[m], want_n) and np.all(ng[~m] == 0)


def test_special_rays(room_scene, room_env, path):
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    o = np.array([[512, 512, 128]] * 8 + [[np.nan, 0, 0], [1e30, 5, 5], [512, 512, 128]],
                 np.float32)
    d = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                  [0, 0, 0], [1e-30, 1, 0], [1, 0, 0], [-1, 0, 0], [np.inf, 0, 0]], np.float32)
    tmin = np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], np.float32)
    tmax = np.array([1e30] * 6 + [1e30, 1e30, 1e30, 1e30, 1e30], np.float32)
    hg, tg = room_env.TraceRays(o, d, tmin, tmax)
    hb, tb = oracle.trace_brute(tri, o, d, tmin, tmax)
    assert np.array_equal(hg, hb) and np.array_equal(_bits(tg), _bits(tb))
    # tmin > tmax and skip ids
    hg, _ = room_env.TraceRays(o[:1], d[:1], 10.0, 5.0)
    assert hg[0] == -1
    hg, tg = room_env.TraceRays(o[:6], d[:6], 0.0, 1e30, skip_id=0x02000000)
    hb, tb = oracle.trace_brute(tri, o[:6], d[:6], 0.0, 1e30, skip=0x02000000)
    assert np.array_equal(hg, hb)


def test_level_rays_vs_oracle_kd(path):
    s = scenes.level(seed=2, rooms_x=8, rooms_y=8, chop=32.0, tess=64.0, n_lights=16)
    print(f"[gpu] level {len(s.ids)} triangles", flush=True)
    env = _env(s.verts9, s.ids)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    mn, mx = s.verts9.reshape(-1, 3).min(0), s.verts9.reshape(-1, 3).max(0)
    rng = np.random.default_rng(12)
    o, d = H.random_rays(rng, 100000, mn, mx)
    hg, tg = env.TraceRays(o, d, 0.0, 3000.0)
    hk, tk = oracle.trace_kd(kd, tri, o, d, 0.0, 3000.0, threads=8)
    assert np.array_equal(hg, hk) and np.array_equal(_bits(tg), _bits(tk))
    hb, tb = oracle.trace_brute(tri, o[:300], d[:300], 0.0, 3000.0)
    assert np.array_equal(hg[:300], hb) and np.array_equal(_bits(tg[:300]), _bits(tb))


def test_trace4rays_packet(room_scene, room_env):
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    rays = FourRays(np.array([[512, 100, 300, 700], [512, 900, 300, 40], [128, 20, 200, 250]]),
                    np.array([[1, 0, 0.6, 0], [0, -1, 0.8, 0], [0, 0, 0, -1]]))
    res = RayTracingResult()
    room_env.Trace4Rays(rays, 0.0, 5000.0, res, -1)
    hb, tb = oracle.trace_brute(tri, rays.Origin.T, rays.Direction.T, 0.0, 5000.0)
    assert np.array_equal(res.HitIds, hb) and np.array_equal(_bits(res.HitDistance), _bits(tb))
    with pytest.raises(_lib.VrError):
        room_env.Trace4Rays(rays, 0.0, 5000.0, res, -1, callback=object())


def test_test_line_closed_form(path):
    tris = H.box((40, -10, -10), (60, 10, 10))
    tris += H.quad(0x01000000, (140, -50, -50), (140, 50, -50), (140, 50, 50), (140, -50, 50))
    v9, ids = H.to_arrays(tris)
    env = _env(v9, ids)
    s = np.array([[0, 0, 0], [0, 30, 0], [100, 0, 0], [100, 0, 0], [0, 0, 0]], np.float32)
    e = np.array([[100, 0, 0], [100, 30, 0], [200, 0, 0], [130, 0, 0], [30, 0, 0]], np.float32)
    vis, sky = env.TestLines(s, e)
    assert list(vis) == [0, 1, 1, 1, 1] and list(sky) == [0, 0, 1, 0, 0]
    fv = TestLineDoesHitSky(s[:4].T, e[:4].T, env=env)
    assert list(fv) == [0.0, 1.0, 1.0, 1.0]


def test_test_line_random_vs_oracle(room_scene, room_env, path):
    s = room_scene
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    rng = np.random.default_rng(13)
    a = rng.uniform([1, 1, 1], [1023, 1023, 255], (30000, 3)).astype(np.float32)
    b = rng.uniform([1, 1, 1], [1023, 1023, 255], (30000, 3)).astype(np.float32)
    b[:10] = a[:10]                                  # zero-length segments
    vg, sg = room_env.TestLines(a, b)
    vo, so = oracle.testline(kd, tri, a, b, threads=8)
    assert np.array_equal(vg, vo) and np.array_equal(sg, so)
    vb, _ = oracle.testline(None, tri, a[:500], b[:500])
    assert np.array_equal(vg[:500], vb)


def test_empty_scene(path):
    env = _env(np.zeros((0, 9), np.float32), np.zeros(0, np.int32))
    hit, t = env.TraceRays(np.zeros((3, 3)), np.eye(3), 0.0, 1e30)
    assert np.all(hit == -1)
    vis, sky = env.TestLines(np.zeros((2, 3)), np.ones((2, 3)))
    assert np.all(vis == 1) and np.all(sky == 0)


def test_test_line_rows_with_sky_vs_oracle(path):
    # MakeTransfers-order segments (one origin patch per run of rows) in a
    # level with skylights: consecutive segments are blocked by the same
    # triangles, so with several rays per lane the occluder of a lane's
    # previous segment is tested first; results must not change
    from vrad_amd import scenes
    s = scenes.level(seed=5, rooms_x=3, rooms_y=3, chop=64.0, tess=64.0, n_lights=4, sun=True)
    assert np.any(s.ids & 0x01000000)
    env = _env(s.verts9, s.ids)
    o, nrm = s.p_origin, s.p_normal
    i = np.repeat(np.arange(0, len(o), 7), 200)[:120000]
    j = (i + 1 + np.arange(len(i)) % 200 * 3) % len(o)
    a = (o[i] + np.float32(0.1) * nrm[i]).astype(np.float32)
    b = (o[j] + np.float32(0.1) * nrm[j]).astype(np.float32)
    # every 4th run of 64 goes straight up out of the level (skylights: sky hits)
    up = (np.arange(len(a)) // 64) % 4 == 3
    b[up] = a[up] + np.array([0.0, 0.0, 5000.0], np.float32)
    vg, sg = env.TestLines(a, b)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    vo, so = oracle.testline(kd, tri, a, b, threads=8)
    print(f"[gpu] rows-with-sky: {len(a)} segments, visible {vo.mean():.3f}, sky {so.mean():.3f}",
          flush=True)
    assert 0.0 < vo.mean() < 1.0 and so.any()
    assert np.array_equal(vg, vo) and np.array_equal(sg, so)


@pytest.mark.parametrize("xg", [0, 16, 256, 1000])
def test_xcd_group_dealing_keeps_results(xg):
    """trace_xcd_group deals the visibility kernels' workgroups to the XCDs in
    groups of xg (0: one contiguous range per XCD).  300k segments and rays
    give whole rounds of 8 groups and a tail; every mapping equals the oracle."""
    s = scenes.level(seed=2, rooms_x=4, rooms_y=4, chop=32.0, tess=64.0, n_lights=4)
    env = _env(s.verts9, s.ids)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    mn, mx = s.verts9.reshape(-1, 3).min(0), s.verts9.reshape(-1, 3).max(0)
    rng = np.random.default_rng(77 + xg)
    o, d = H.random_rays(rng, 300000, mn, mx)
    a = o + d * 400.0
    old = autotune.get_knob("trace_xcd_group")
    print(f"[gpu] trace_xcd_group={xg}", flush=True)
    try:
        _lib.check(_lib.lib.vr_set_tuning(b"trace_xcd_group", xg))
        vis, sky = env.TestLines(o, a)
        hg, tg = env.TraceRays(o, d, 0.0, 3000.0)
    finally:
        _lib.check(_lib.lib.vr_set_tuning(b"trace_xcd_group", old))
    vo, so = oracle.testline(kd, tri, o, a, threads=8)
    assert np.array_equal(vis, vo) and np.array_equal(sky, so)
    hk, tk = oracle.trace_kd(kd, tri, o, d, 0.0, 3000.0, threads=8)
    assert np.array_equal(hg, hk) and np.array_equal(_bits(tg), _bits(tk))
    assert 0.05 < vis.mean() < 0.95
