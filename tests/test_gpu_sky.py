"""Sky visibility and the light branches on the GPU: CanLeafTraceToSky, sun,
sky ambient, surface and spot lights against the oracle (bit for bit) and
closed forms; the triangle precompute against the general-position
goldens."""
import json
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, autotune, lightmap, scenes
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu
OPAQUE, _light, open_box = H.OPAQUE, H.light, H.open_box


def _env(tris):
    v9, ids = H.to_arrays(tris)
    env = Environment()
    env.AddTriangles(v9, ids)
    env.SetupAccelerationStructure()
    return env, v9, ids


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_precompute_general_goldens(golden_dir):
    g = json.load(open(os.path.join(golden_dir, "tri_precompute_general.json")))
    v9 = np.array([t["v"] for t in g["triangles"]], np.float32)
    ids = np.array([t["id"] for t in g["triangles"]], np.int32)
    print(f"[gpu] precompute of {len(ids)} general-position triangles", flush=True)
    env = Environment()
    env.AddTriangles(v9, ids)
    env.SetupAccelerationStructure()
    rec = env.OptimizedTriangleList
    for k, t in enumerate(g["triangles"]):
        r = rec[k]
        assert list(_bits([r["FlNx"], r["FlNy"], r["FlNz"]])) == t["N_bits"]
        assert int(_bits(r["FlD"])) == t["D_bits"]
        assert list(_bits(r["ProjectedEdgeEquations"])) == t["E_bits"]
        assert (r["NCoordSelect0"], r["NCoordSelect1"]) == (t["c0"], t["c1"])


@pytest.mark.parametrize("lds", [4, 10, 24])
def test_leaf_sky_closed_forms(lds):
    print(f"[gpu] CanLeafTraceToSky closed forms lds={lds}", flush=True)
    old = autotune.get_knob("trace_lds_entries")
    _lib.check(_lib.lib.vr_set_tuning(b"trace_lds_entries", lds))
    try:
        ctr = np.array([[48, 48, 18]], np.int16)
        cases = [(H.box((0, 0, 0), (100, 100, 100)), False), (open_box(), True),
                 (open_box(sky_quad=False), True), (open_box(roof=True), False),
                 (open_box(roof=True, open_wall=True), True)]
        for tris, want in cases:
            env, _, _ = _env(tris)
            got = lightmap.CanLeafTraceToSky(env, ctr, ctr + 4)
            assert list(got) == [want]
    finally:
        _lib.check(_lib.lib.vr_set_tuning(b"trace_lds_entries", old))


def test_leaf_sky_level_vs_oracle():
    s = scenes.level(seed=2, rooms_x=4, rooms_y=4, chop=64.0, tess=64.0, n_lights=4, sun=True)
    print(f"[gpu] CanLeafTraceToSky on the 4x4 level: {len(s.ids)} tris", flush=True)
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure()
    lo = s.verts9.reshape(-1, 3).min(0)
    hi = s.verts9.reshape(-1, 3).max(0)
    rng = np.random.default_rng(11)
    mins = rng.uniform(lo - 50, hi + 50, (3000, 3)).astype(np.int16)
    maxs = (mins + rng.integers(0, 96, (3000, 3))).astype(np.int16)
    got = lightmap.CanLeafTraceToSky(env, mins, maxs)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    want = oracle.can_leaf_trace_to_sky(kd, tri, mins, maxs, threads=8).astype(bool)
    assert np.array_equal(got, want)
    assert 0.05 < got.mean() < 0.95            # both outcomes occur
    assert list(lightmap.CanLeafTraceToSky(env, mins[:0], maxs[:0])) == []


def test_sun_sky_surface_closed_forms():
    print("[gpu] sun / sky ambient / surface closed forms", flush=True)
    tris = open_box()
    env, v9, ids = _env(tris)
    sun = _light(3, normal=(0, 0, -1))
    amb = _light(5, inten=(1, 1, 1))
    surf = _light(0, origin=(50, 50, 90), normal=(0, 0, -1))
    L = np.concatenate([sun, amb, surf])
    # the second sample hangs inside the box facing the floor: no sun
    # (dot < 0), no sky (every downward direction meets the floor), and the
    # surface light is behind it
    pos = np.array([[50, 50, 0], [50, 50, 50], [5, 5, 0]], np.float32)
    nrm = np.array([[0, 0, 1], [0, 0, -1], [0, 0, 1]], np.float32)
    got = lightmap.GatherSampleLight(env, pos, nrm, L.astype(_lib.LIGHT_DTYPE))
    want = oracle.gather_light(oracle.KdTree(v9), oracle.precompute(v9, ids), pos, nrm, L)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the first sample sees the sun (I * 1), part of the sky and the surface
    # light; the second gets nothing
    assert got[0, 0] > 1.0 and got[1].tolist() == [0, 0, 0]
    roofed, rv9, rids = _env(open_box(roof=True))
    got2 = lightmap.GatherSampleLight(roofed, pos[:1], nrm[:1],
                                      np.concatenate([sun, amb]).astype(_lib.LIGHT_DTYPE))
    assert got2.tolist() == [[0, 0, 0]]        # roofed: no sun, no sky


def test_sky_ambient_half_space_gpu():
    big = 1e5
    wall = H.quad(OPAQUE, (10, -big, -big), (10, big, -big), (10, big, big), (10, -big, big))
    env, v9, ids = _env(wall)
    L = _light(5, inten=(1, 1, 1)).astype(_lib.LIGHT_DTYPE)
    got = lightmap.GatherSampleLight(env, np.zeros((1, 3), np.float32),
                                     np.array([[0, 0, 1]], np.float32), L)
    d = oracle.sky_dirs().astype(np.float64)
    up = d[:, 2] > 0
    want = d[up & (d[:, 0] <= 0), 2].sum() / d[up, 2].sum()
    assert abs(float(got[0, 0]) - want) < 2e-6


@pytest.mark.parametrize("exponent", [0.0, 1.0, 2.0, 1.7])
def test_spot_exponent_bitexact(exponent):
    print(f"[gpu] spot exponent {exponent}", flush=True)
    env, v9, ids = _env(H.box((400, 400, 0), (420, 420, 20)))
    L = lightmap.spot_light((0, 0, 100), (0.1, 0.05, -1), (1, 1, 1), 25, 50, exponent=exponent,
                            quadratic=1.0)
    rng = np.random.default_rng(3)
    pos = np.column_stack([rng.uniform(-150, 150, 2000), rng.uniform(-150, 150, 2000),
                           np.zeros(2000)]).astype(np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (2000, 1))
    got = lightmap.GatherSampleLight(env, pos, nrm, L)
    want = oracle.gather_light(oracle.KdTree(v9), oracle.precompute(v9, ids), pos, nrm, L)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (got[:, 0] > 0).sum() > 100 and (got[:, 0] == 0).sum() > 100
