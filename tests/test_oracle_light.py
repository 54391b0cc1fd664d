"""The oracle's direct-light branches, sky directions and CanLeafTraceToSky
pinned by closed forms, a brute-force count and the icosphere the direction
table samples (no GPU); plus the light-parameter helpers of the product's
Python mirror (host code, no GPU)."""
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H

SKY, OPAQUE, MTL, _light, open_box = H.SKY, H.OPAQUE, H.MTL, H.light, H.open_box


def _scene(tris):
    if not tris:
        return None, np.zeros(0, oracle.TRI_DTYPE)
    v9, ids = H.to_arrays(tris)
    return oracle.KdTree(v9), oracle.precompute(v9, ids)


def _gather(tris, pos, nrm, L):
    kd, tri = _scene(tris)
    return oracle.gather_light(kd, tri, np.asarray(pos, np.float32).reshape(-1, 3),
                               np.asarray(nrm, np.float32).reshape(-1, 3), L)


# ---- the 162 sky directions ---------------------------------------------------------
def icosphere4():
    """Vertices of the frequency-4 geodesic icosphere on the icosahedron with
    vertices (+-1, 0, +-phi) and their cyclic permutations: each face's
    barycentric grid (i, j, k) / 4 projected to the unit sphere."""
    phi = (1 + 5 ** 0.5) / 2
    ico = []
    for a in (-1, 1):
        for b in (-1, 1):
            ico += [(a, 0, b * phi), (0, b * phi, a), (b * phi, a, 0)]
    ico = np.array(ico, np.float64)
    ico /= np.linalg.norm(ico, axis=1, keepdims=True)
    d = np.linalg.norm(ico[:, None] - ico[None], axis=2)
    edge = d[d > 1e-9].min()
    faces = [f for f in itertools.combinations(range(12), 3)
             if all(abs(d[i, j] - edge) < 1e-9 for i, j in itertools.combinations(f, 2))]
    assert len(faces) == 20
    pts = []
    for f in faces:
        A, B, C = ico[list(f)]
        for i in range(5):
            for j in range(5 - i):
                p = (i * A + j * B + (4 - i - j) * C) / 4
                pts.append(p / np.linalg.norm(p))
    pts = np.array(pts)
    keep = []
    for p in pts:
        if not any(np.linalg.norm(p - q) < 1e-9 for q in keep):
            keep.append(p)
    return np.array(keep)


def test_sky_dirs_are_the_frequency4_icosphere():
    """Anorms (vmath/constants.go:15-184) is this icosphere to 6 decimals:
    each direction lies within 1.5e-6 of a distinct vertex."""
    d = oracle.sky_dirs().astype(np.float64)
    G = icosphere4()
    assert len(G) == 162
    dist = np.linalg.norm(d[:, None, :] - G[None], axis=2)
    j = dist.argmin(axis=1)
    assert len(set(j.tolist())) == 162
    assert dist.min(axis=1).max() < 1.5e-6
    # table order is fixed: the first entry and the +z pole at index 5
    assert np.array_equal(d[0], np.float32([-0.525731, 0.0, 0.850651]).astype(np.float64))
    assert np.array_equal(d[5], [0.0, 0.0, 1.0])
    # closed under negation, bit for bit
    neg = {tuple(x) for x in (-oracle.sky_dirs()).tolist()}
    assert neg == {tuple(x) for x in oracle.sky_dirs().tolist()}


# ---- sun (EMIT_SKYLIGHT) -------------------------------------------------------------
def test_sun_open_box_closed_form():
    L = _light(3, normal=(0, 0, -1))                      # light travels straight down
    pos, nrm = [50, 50, 0], [0, 0, 1]
    assert np.array_equal(_gather(open_box(), pos, nrm, L)[0], [1.0, 2.0, 3.0])   # I * 1
    assert np.array_equal(_gather(open_box(sky_quad=False), pos, nrm, L)[0], [1, 2, 3])
    assert np.array_equal(_gather(open_box(roof=True), pos, nrm, L)[0], [0, 0, 0])
    # a tilted receiver: I * dot with dot = -(n . normal) = cos 60
    tilt = [0.0, float(np.float32(math.sqrt(3) / 2)), 0.5]
    got = _gather(open_box(), [50, 50, 0], tilt, L)[0]
    assert np.array_equal(got, np.float32([1, 2, 3]) * np.float32(0.5))
    # low sun from +x is stopped by the x = hi wall, passes the open wall
    Llow = _light(3, normal=(-0.98, 0.0, -0.2))
    assert np.array_equal(_gather(open_box(), pos, nrm, Llow)[0], [0, 0, 0])
    assert _gather(open_box(open_wall=True), pos, nrm, Llow)[0, 0] > 0
    # surface facing away from the sun
    assert np.array_equal(_gather(open_box(), pos, [0, 0, -1], L)[0], [0, 0, 0])


# ---- sky ambient (EMIT_SKYAMBIENT) ---------------------------------------------------
def test_sky_ambient_extremes():
    L = _light(5)
    up = [0, 0, 1]
    assert np.array_equal(_gather([], [0, 0, 0], up, L)[0], [1.0, 2.0, 3.0])      # a == w
    closed = H.box((-50, -50, -50), (50, 50, 50))
    assert np.array_equal(_gather(closed, [0, 0, 0], up, L)[0], [0, 0, 0])
    # open box, no roof: directions through the opening only
    assert 0 < _gather(open_box(), [50, 50, 0], up, L)[0, 0] < 1


def _ambient_expected(blocked, n):
    d = oracle.sky_dirs().astype(np.float64)
    c = d @ np.asarray(n, np.float64)
    up = c > 0
    return c[up & ~blocked(d)].sum() / c[up].sum(), (up & ~blocked(d)).sum() / up.sum()


def test_sky_ambient_blockers_vs_brute_count():
    """A sample at the origin facing +z: the cosine-weighted visible fraction
    of the 162 directions, against a float64 count of which directions the
    blocker stops."""
    L = _light(5, inten=(1.0, 1.0, 1.0))
    big = 1e5
    # an opaque wall at x = 10: every direction with d.x > 0 hits it
    wall = H.quad(OPAQUE, (10, -big, -big), (10, big, -big), (10, big, big), (10, -big, big))
    want, _ = _ambient_expected(lambda d: d[:, 0] > 0, (0, 0, 1))
    got = float(_gather(wall, [0, 0, 0], [0, 0, 1], L)[0, 0])
    assert abs(got - want) < 2e-6
    # a half-ceiling z = 10 over x > 1, seen from (0, 0, 0.1) (the sample
    # moved 0.1 along its normal): a direction d with d.z > 0 meets z = 10 at
    # x = 9.9 d.x / d.z.  Weighting matters here (the plain count differs).
    ceil = H.quad(OPAQUE, (1, -big, 10), (big, -big, 10), (big, big, 10), (1, big, 10))

    def blocked(d):
        with np.errstate(divide="ignore", invalid="ignore"):
            x = np.where(d[:, 2] > 0, 9.9 * d[:, 0] / d[:, 2], -np.inf)
        assert np.abs(x[np.isfinite(x)] - 1.0).min() > 1e-3      # no direction grazes the edge
        return x > 1.0
    want, count = _ambient_expected(blocked, (0, 0, 1))
    got = float(_gather(ceil, [0, 0, 0], [0, 0, 1], L)[0, 0])
    assert abs(got - want) < 2e-6 and abs(count - want) > 3e-3


# ---- surface light (EMIT_SURFACE) ----------------------------------------------------
def test_surface_light_closed_form():
    nrm = [0, 0, 1]
    L = _light(0, origin=(0, 0, 100), normal=(0, 0, -1))          # facing the sample
    assert np.array_equal(_gather([], [0, 0, 0], nrm, L)[0],
                          np.float32([1, 2, 3]) * np.float32(1e-4))  # dot = dot2 = 1, d^2 = 1e4
    L = _light(0, origin=(100, 0, 100), normal=(0, 0, -1))
    got = _gather([], [0, 0, 0], nrm, L)[0]
    want = 0.5 / 2e4                                              # cos 45 * cos 45 / d^2
    assert np.allclose(got, np.array([1, 2, 3]) * want, rtol=2e-7, atol=0)
    # emitter facing away, receiver facing away, and an occluder: all 0
    assert not _gather([], [0, 0, 0], nrm, _light(0, (0, 0, 100), (0, 0, 1))).any()
    assert not _gather([], [0, 0, 0], [0, 0, -1], _light(0, (0, 0, 100), (0, 0, -1))).any()
    occ = H.box((-5, -5, 40), (5, 5, 60))
    assert not _gather(occ, [0, 0, 0], nrm, _light(0, (0, 0, 100), (0, 0, -1))).any()


# ---- CanLeafTraceToSky --------------------------------------------------------------------
def _leaf_sky(tris, centres):
    kd, tri = _scene(tris)
    c = np.asarray(centres, np.int16).reshape(-1, 3)
    return oracle.can_leaf_trace_to_sky(kd, tri, c - 2, c + 2).astype(bool)


def test_can_leaf_trace_to_sky_closed_forms():
    ctr = [50, 50, 20]
    assert list(_leaf_sky([], [ctr])) == [True]                              # every ray misses
    assert list(_leaf_sky(H.box((0, 0, 0), (100, 100, 100)), [ctr])) == [False]
    assert list(_leaf_sky(open_box(), [ctr])) == [True]                      # sky through the top
    assert list(_leaf_sky(open_box(sky_quad=False), [ctr])) == [True]        # a miss counts
    assert list(_leaf_sky(open_box(roof=True), [ctr])) == [False]
    assert list(_leaf_sky(open_box(roof=True, open_wall=True), [ctr])) == [True]
    # the centre is (mins + maxs) / 2 in float: odd sums land on .5
    kd, tri = _scene(H.box((0, 0, 0), (100, 100, 100)))
    mins = np.array([[10, 10, 10], [-300, 10, 10]], np.int16)
    maxs = np.array([[11, 11, 11], [-201, 11, 11]], np.int16)
    assert list(oracle.can_leaf_trace_to_sky(kd, tri, mins, maxs)) == [0, 1]


def test_can_leaf_trace_to_sky_equals_any_direction():
    """Per leaf: true iff some direction's TestLine segment is visible."""
    t = open_box(roof=True) + H.box((200, 0, 0), (300, 100, 100))
    t += H.quad(OPAQUE, (0, 0, 100), (100, 0, 100), (100, 100, 100), (0, 100, 100))
    kd, tri = _scene(open_box(roof=False))
    rng = np.random.default_rng(7)
    c = rng.integers(-20, 140, (60, 3)).astype(np.int16)
    got = oracle.can_leaf_trace_to_sky(kd, tri, c, c)
    d = oracle.sky_dirs()
    for i in range(len(c)):
        ctr = c[i].astype(np.float32)
        ends = (d * np.float32(-MTL) + ctr).astype(np.float32)
        vis, _ = oracle.testline(kd, tri, np.tile(ctr, (162, 1)), ends)
        assert got[i] == int(vis.any())


# ---- light parameters (the Python mirror's host helpers) ----------------------------
def test_spot_cosine_rounds_in_float32():
    """lights.go:249-250 forms deg / 180 * PI in float32 before the float64
    cosine; for 25 degrees that argument differs from the float64 one, and the
    stored cosine differs by one ulp."""
    from vrad_amd import lightmap
    deg = 25.0
    arg32 = np.float32(np.float32(deg) / np.float32(180.0)) * np.float32(math.pi)
    assert float(arg32) != deg / 180.0 * math.pi
    assert lightmap.spot_cosine(deg) == np.float32(math.cos(float(arg32)))
    assert lightmap.spot_cosine(deg) != np.float32(math.cos(deg / 180.0 * math.pi))
    L = lightmap.spot_light((0, 0, 0), (0, 0, -1), (1, 1, 1), deg, 30.0)
    assert L["stopdot"][0] == lightmap.spot_cosine(deg)
    assert L["stopdot2"][0] == lightmap.spot_cosine(30.0)


def test_fifty_percent_falloff_closed_form():
    """d50 = 1, d0 = 2: the quadratic through (0, 1), (1, 2), (2, 256) is
    126.5 x^2 - 125.5 x + 1, monotonic at the first try (2a + b = 127.5), and
    its value at d50 is exactly 2, so no rescale."""
    from vrad_amd import lightmap
    c, l, q, start, end, cap = lightmap._falloff_fifty_percent(1.0, 2.0, False)
    assert (q, l, c) == (126.5, -125.5, 1.0)
    fl_max = np.float32(-125.5) / (np.float32(-2.0) * np.float32(126.5))
    assert (start, end, cap) == (fl_max, np.float32(10) * fl_max, fl_max)
    c, l, q, start, end, cap = lightmap._falloff_fifty_percent(1.0, 2.0, True)
    assert (start, end, cap) == (np.float32(1.75), np.float32(2.0), np.float32(1e22))
    L = lightmap.point_light((0, 0, 0), (5, 5, 5), fifty_percent_distance=1.0,
                             zero_percent_distance=2.0, hardfalloff=True)
    assert L["intensity"][0].tolist() == [5, 5, 5]                 # no 100-unit rescale
    assert (L["start_fade"][0], L["end_fade"][0]) == (1.75, 2.0)


def test_fifty_percent_falloff_blends_to_monotonic():
    """d50 = 100, d0 = 200: the plain fit (a = 0.01265, b = -1.255) turns
    down at the start, so the middle point is blended toward the chord until
    2a + b >= 0; the result is then rescaled so that 1/(c + b d50 + a d50^2)
    is 1/2 (to float32 rounding), with the light's intensity unscaled."""
    from vrad_amd import lightmap
    raw = lightmap.solve_inverse_quadratic(0, 1, 100, 2, 200, 256)
    assert abs(raw[0] - 0.01265) < 1e-6 and abs(raw[1] + 1.255) < 1e-5
    sol = lightmap.solve_inverse_quadratic_monotonic(0, 1, 100, 2, 200, 256)
    a, b, c = sol
    assert np.float32(2) * a + b >= 0
    # the blend factor is the first multiple of 0.05 (float64 counter) that works
    f, k = 0.0, 0
    while True:
        y2 = (1 - f) * 2.0 + f * (1 + (256 - 1) * (100 - 0) / (200 - 0))
        A = (y2 - 1 - (256 - 1) / 2) / (100 * 100 - 100 * 200)       # float64 fit
        B = (256 - 1 - 200 * 200 * A) / 200
        if 2 * A + B >= 0:
            break
        f += 0.05
        k += 1
    assert 0 < k < 20
    assert abs(a - A) < 1e-6 * abs(A) + 1e-9 and abs(b - B) < 1e-5 * abs(B) + 1e-9
    c2, l2, q2, start, end, cap = lightmap._falloff_fifty_percent(100.0, 200.0, False)
    v50 = float(c2) + 100.0 * (float(l2) + 100.0 * float(q2))
    assert abs(v50 - 2.0) < 1e-5
    # the unsorted call sorts its points first (the intent of vSwap)
    assert lightmap.solve_inverse_quadratic_monotonic(100, 2, 0, 1, 200, 256) == sol
