"""Re-bake serving on the GPU: each request's result equals the CPU
restatement of the same re-bake, is the same alone and in the mix, and the
plain sequential mode agrees with the plain-order restatement."""
import numpy as np
import pytest

import oracle
from vrad_amd import _lib, lightmap, scenes
from vrad_amd.patches import MakeTransfers
from vrad_amd.raytracer import Environment
from vrad_amd.serve import BATCHED, PLAIN, Server

pytestmark = pytest.mark.gpu

BOUNCES = 8


class _Baked:
    def __init__(self, s, precision):
        self.s = s
        self.env = Environment()
        self.env.AddTriangles(s.verts9, s.ids)
        self.env.SetupAccelerationStructure()
        self.xf = MakeTransfers(self.env, s.p_origin, s.p_normal, s.p_area)
        self.rp, self.col, self.coef32 = self.xf.download()
        if precision != _lib.VR_COEF_F32:
            self.xf.set_precision(precision)
            _, _, _, self.raw, self.scale = self.xf.download(raw16=True)
        self.tri = oracle.precompute(s.verts9, s.ids)
        self.kd = oracle.KdTree(s.verts9)


@pytest.fixture(scope="module")
def scenes16():
    a = _Baked(scenes.room(seed=5, n_pillars=12, chop=96.0, tess=64.0), _lib.VR_COEF_F16)
    b = _Baked(scenes.level(seed=3, rooms_x=2, rooms_y=2, chop=96.0, tess=128.0, n_lights=4,
                            sun=True), _lib.VR_COEF_F16)
    return [a, b]


def _moved(lights, moves):
    L = lights.copy()
    for i, d in moves:
        L["origin"][i] += np.asarray(d, np.float32)
    return L


def _expected(b, moves, precision, order="stream"):
    s = b.s
    L = _moved(s.lights, moves)
    direct = oracle.gather_light(b.kd, b.tri, s.p_origin, s.p_normal, L, threads=8)
    emit0 = (direct * s.p_refl).astype(np.float32)
    if precision == _lib.VR_COEF_F32:
        return oracle.bounce_f32(b.rp, b.col, b.coef32, emit0, s.p_refl, s.p_area, total=direct,
                                 max_bounces=BOUNCES, stop_frac=0.0, threads=8, order=order)
    return oracle.bounce_f32(b.rp, b.col, b.raw, emit0, s.p_refl, s.p_area, total=direct,
                             max_bounces=BOUNCES, stop_frac=0.0, kind=precision,
                             row_scale=b.scale, threads=8, order=order)


def _server(mode, scs, **kw):
    sv = Server(mode=mode, bounces=BOUNCES, **kw)
    for b in scs:
        s = b.s
        sv.add_scene(b.env, b.xf, s.p_origin, s.p_normal, s.p_refl, s.p_area, s.lights)
    return sv


def _requests(scs, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        sid = int(rng.integers(len(scs)))
        nl = len(scs[sid].s.lights)
        k = int(rng.integers(1, min(4, nl) + 1))
        idx = rng.choice(nl, size=k, replace=False)
        out.append((sid, [(int(i), rng.uniform(-64, 64, 3).astype(np.float32)) for i in idx]))
    return out


def test_serve_batched_matches_restatement(scenes16):
    print("[gpu] serve batched vs oracle", flush=True)
    sv = _server(BATCHED, scenes16, max_active=4)
    try:
        for sid, moves in _requests(scenes16, 4, seed=1):
            tot, en, lat = sv.wait(sv.submit(sid, moves))
            ot, oe, ok = _expected(scenes16[sid], moves, _lib.VR_COEF_F16)
            assert ok == BOUNCES == len(en) and lat > 0
            assert np.array_equal(tot.view(np.uint32), ot.view(np.uint32))
            assert np.array_equal(en, oe)
    finally:
        sv.close()


def test_serve_alone_equals_mix(scenes16):
    reqs = _requests(scenes16, 16, seed=2)
    sv = _server(BATCHED, scenes16, max_active=8)
    try:
        print("[gpu] serve mix", flush=True)
        tickets = [sv.submit(sid, mv) for sid, mv in reqs]       # a burst: batched together
        mix = [sv.wait(t) for t in tickets]
        st = sv.stats()
        assert st["max_active_seen"] > 1 and st["completed"] == len(reqs)
        print("[gpu] serve alone", flush=True)
        alone = [sv.wait(sv.submit(sid, mv)) for sid, mv in reqs]  # one at a time
    finally:
        sv.close()
    for (t1, e1, _), (t2, e2, _) in zip(mix, alone):
        assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
        assert np.array_equal(e1, e2)


def test_serve_admission_by_footprint(scenes16):
    # a budget that holds one request of the larger scene at a time
    n = max(len(b.s.p_area) for b in scenes16)
    budget = 3 * ((12 * n + 16 + 255) // 256 * 256) + (8 * n + 255) // 256 * 256 + 4 * 4096
    sv = _server(BATCHED, scenes16, max_active=8, hbm_budget=budget)
    try:
        reqs = [(1, mv) for _, mv in _requests(scenes16, 4, seed=3)]
        got = [sv.wait(t) for t in [sv.submit(sid, mv) for sid, mv in reqs]]
        assert sv.stats()["max_active_seen"] == 1
    finally:
        sv.close()
    for (sid, mv), (tot, en, _) in zip(reqs, got):
        ot, _, _ = _expected(scenes16[sid], mv, _lib.VR_COEF_F16)
        assert np.array_equal(tot.view(np.uint32), ot.view(np.uint32))


def test_serve_plain_matches_sequential_restatement(scenes16):
    print("[gpu] serve plain", flush=True)
    sv = _server(PLAIN, scenes16)
    try:
        for sid, moves in _requests(scenes16, 3, seed=4):
            tot, en, lat = sv.wait(sv.submit(sid, moves))
            ot, oe, ok = _expected(scenes16[sid], moves, _lib.VR_COEF_F32, order="sequential")
            assert ok == len(en)
            assert np.array_equal(tot.view(np.uint32), ot.view(np.uint32))
            assert np.array_equal(en, oe)
    finally:
        sv.close()


def test_served_transfers_refuse_relayout(scenes16):
    """While a server holds a transfer set, calls that would reallocate or
    free its buffers fail with VR_E_STATE before any work, a bounce on the
    layout the server uses still runs, and the server keeps serving the same
    results; after the server is gone the calls work again."""
    print("[gpu] served transfers refuse re-layout", flush=True)
    b = scenes16[0]
    s = b.s
    sv = _server(BATCHED, [b], max_active=2)
    lib = _lib.lib
    try:
        h = b.xf.handle
        assert lib.vr_xfers_set_precision(h, _lib.VR_COEF_BF16) == _lib.VR_E_STATE
        assert lib.vr_xfers_destroy(h) == _lib.VR_E_STATE
        e0 = (s.p_refl * 0.5).astype(np.float32)
        tot, _, k = lightmap.BounceLight(b.xf, e0, s.p_refl, s.p_area, max_bounces=2,
                                         stop_frac=0.0)                  # same layout: runs
        assert k == 2
        _lib.check(lib.vr_set_tuning(b"bounce_tile_cols", 2048))
        try:
            with pytest.raises(_lib.VrError) as ei:
                lightmap.BounceLight(b.xf, e0, s.p_refl, s.p_area, max_bounces=2, stop_frac=0.0)
            assert ei.value.code == _lib.VR_E_STATE
        finally:
            _lib.check(lib.vr_set_tuning(b"bounce_tile_cols", 4096))
        _lib.check(lib.vr_set_tuning(b"bounce_placement", 1))   # deals the rows again
        try:
            with pytest.raises(_lib.VrError) as ei:
                lightmap.BounceLight(b.xf, e0, s.p_refl, s.p_area, max_bounces=2, stop_frac=0.0)
            assert ei.value.code == _lib.VR_E_STATE
        finally:
            _lib.check(lib.vr_set_tuning(b"bounce_placement", 0))
        moves = _requests([b], 1, seed=9)[0][1]
        tot, en, _ = sv.wait(sv.submit(0, moves))
        ot, oe, _ = _expected(b, moves, _lib.VR_COEF_F16)
        assert np.array_equal(tot.view(np.uint32), ot.view(np.uint32))
    finally:
        sv.close()
    _lib.check(lib.vr_xfers_set_precision(b.xf.handle, _lib.VR_COEF_F16))
