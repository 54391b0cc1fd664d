"""Batched BounceLight (one launch per bounce over many transfer sets, with
and without a replayed hipGraph) against the single-set call and the CPU
restatement: bit-exact per job."""
import numpy as np
import pytest

import oracle
from vrad_amd import _lib, lightmap
from vrad_amd.patches import Transfers

pytestmark = pytest.mark.gpu


def _inputs(n, seed, refl_hi=0.8):
    rng = np.random.default_rng(seed)
    e0 = rng.random((n, 3)).astype(np.float32)
    refl = rng.uniform(0.2, refl_hi, (n, 3)).astype(np.float32)
    area = rng.uniform(1, 4, n).astype(np.float32)
    return e0, refl, area


def _sets(prec):
    sizes = [(3000, 80.0), (20000, 200.0), (9000, 40.0), (45000, 150.0)]
    out = []
    for i, (n, mean) in enumerate(sizes):
        xf = Transfers.synthetic(seed=20 + i, n_rows=n, mean_len=mean, sigma=1.0, max_len=2000)
        if prec != _lib.VR_COEF_F32:
            xf.set_precision(prec)
        out.append((xf,) + _inputs(n, 40 + i, refl_hi=(0.3, 0.5, 0.8, 0.99)[i]))
    # a one-row set without transfers
    empty = Transfers.from_csr(np.array([0, 0], np.int64), np.zeros(0, np.int32),
                               np.zeros(0, np.float32))
    if prec != _lib.VR_COEF_F32:
        empty.set_precision(prec)
    out.append((empty,) + _inputs(1, 99))
    return out


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("prec", [_lib.VR_COEF_F32, _lib.VR_COEF_F16], ids=["f32", "f16"])
def test_batch_equals_single(prec, use_graph):
    sets = _sets(prec)
    print(f"[gpu] batch prec={prec} graph={use_graph} jobs={len(sets)}", flush=True)
    # stop at different bounces per job (energy rule), K not a multiple of 8
    single = [lightmap.BounceLight(xf, e0, rf, ar, max_bounces=21, stop_frac=2e-3)
              for xf, e0, rf, ar in sets]
    batch = lightmap.BounceLightBatch([(xf, e0, rf, ar, None) for xf, e0, rf, ar in sets],
                                      max_bounces=21, stop_frac=2e-3, use_graph=use_graph)
    ks = []
    for (t1, en1, k1), (t2, en2, k2) in zip(single, batch):
        assert k1 == k2
        assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
        assert np.array_equal(en1.view(np.uint64), en2.view(np.uint64))
        ks.append(k1)
    assert len(set(ks[:-1])) > 1 or max(ks) == 21  # jobs stop at different bounces
    # tie one job to the CPU restatement of the default (stream) order
    xf, e0, rf, ar = sets[1]
    if prec == _lib.VR_COEF_F32:
        rp, col, coef = xf.download()
        ot, oe, ok = oracle.bounce_f32(rp, col, coef, e0, rf, ar, max_bounces=21, stop_frac=2e-3,
                                       threads=8)
    else:
        rp, col, _, raw, sc = xf.download(raw16=True)
        ot, oe, ok = oracle.bounce_f32(rp, col, raw, e0, rf, ar, max_bounces=21, stop_frac=2e-3,
                                       kind=prec, row_scale=sc, threads=8)
    t2, en2, k2 = batch[1]
    assert k2 == ok and np.array_equal(t2.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(en2, oe)
    for s in sets:
        s[0].close()


def test_batch_graph_replays_repeatably():
    sets = _sets(_lib.VR_COEF_F32)[:3]
    jobs = [(xf, e0, rf, ar, None) for xf, e0, rf, ar in sets]
    ref = lightmap.BounceLightBatch(jobs, max_bounces=30, stop_frac=0.0, use_graph=False)
    for rep in range(3):  # each call replays its graph 4 times, host syncs between
        got = lightmap.BounceLightBatch(jobs, max_bounces=30, stop_frac=0.0, use_graph=True)
        for (t1, e1, k1), (t2, e2, k2) in zip(ref, got):
            assert k1 == k2 == 30
            assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32)), f"replay {rep}"
            assert np.array_equal(e1, e2)
    for s in sets:
        s[0].close()
