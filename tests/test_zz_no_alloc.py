"""The allocation-free bounce (vr_bounce_light_noalloc) against
vr_bounce_light through the C-ABI: the same totals, energies and bounce
counts bit for bit, on two calls in a row (the second smaller), for FP32 and
FP16 transfers; refused calls do no work."""
import ctypes as C

import numpy as np
import pytest

from vrad_amd import _lib
from vrad_amd._lib import lib, ptr
from vrad_amd.patches import Transfers

pytestmark = pytest.mark.gpu


def _call(fn, xf, e0, refl, area, total, K, stop):
    tot = total.copy()
    en = np.zeros(max(K, 1), np.float64)
    nb = C.c_int32(-1)
    prm = _lib.BounceParams(int(K), float(stop))
    rc = fn(xf.handle, ptr(e0), ptr(refl), ptr(area), ptr(tot), C.byref(prm), ptr(en),
            C.byref(nb))
    return rc, tot, en[:max(nb.value, 0)], nb.value


def _case(n, seed, prec):
    xf = Transfers.synthetic(seed=seed, n_rows=n, mean_len=300.0, sigma=1.0, max_len=3000)
    if prec != _lib.VR_COEF_F32:
        xf.set_precision(prec)
    rng = np.random.default_rng(seed)
    e0 = rng.random((n, 3)).astype(np.float32)
    refl = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    area = rng.uniform(1, 4, n).astype(np.float32)
    total = rng.random((n, 3)).astype(np.float32)
    return xf, e0, refl, area, total


@pytest.mark.parametrize("prec", [_lib.VR_COEF_F32, _lib.VR_COEF_F16], ids=["f32", "f16"])
def test_noalloc_equals_bounce_light(prec):
    cases = [(_case(60000, 31, prec), 6, 0.0), (_case(7000, 32, prec), 100, 1e-3)]
    refs = []
    for (xf, e0, refl, area, total), K, stop in cases:
        print(f"[gpu] noalloc prec={prec} n={len(area)} K={K}", flush=True)
        # the gather layout is not built yet: refused before any work
        rc, tot, _, nb = _call(lib.vr_bounce_light_noalloc, xf, e0, refl, area, total, K, stop)
        assert rc == _lib.VR_E_STATE and nb == -1 and np.array_equal(tot, total)
        rc, tot, en, nb = _call(lib.vr_bounce_light, xf, e0, refl, area, total, K, stop)
        assert rc == _lib.VR_OK and nb > 0
        refs.append((tot, en, nb))
    # two calls in a row, the second on a smaller transfer set
    for ((xf, e0, refl, area, total), K, stop), (rtot, ren, rnb) in zip(cases, refs):
        rc, tot, en, nb = _call(lib.vr_bounce_light_noalloc, xf, e0, refl, area, total, K, stop)
        assert rc == _lib.VR_OK, lib.vr_last_error()
        assert nb == rnb and np.array_equal(en, ren)
        assert np.array_equal(tot.view(np.uint32), rtot.view(np.uint32))
    assert refs[1][2] < 100                  # the second case stopped on its energy


def test_noalloc_refuses_beyond_its_limits():
    xf, e0, refl, area, total = _case(3000, 33, _lib.VR_COEF_F32)
    rc, _, _, _ = _call(lib.vr_bounce_light, xf, e0, refl, area, total, 2, 0.0)
    assert rc == _lib.VR_OK
    rc, tot, _, nb = _call(lib.vr_bounce_light_noalloc, xf, e0, refl, area, total, 1025, 0.0)
    assert rc == _lib.VR_E_CAPACITY and nb == -1 and np.array_equal(tot, total)
    big = Transfers.synthetic(seed=34, n_rows=(1 << 19) + 1, mean_len=1.0, sigma=0.0, max_len=1)
    n = (1 << 19) + 1
    z = np.zeros((n, 3), np.float32)
    rc, _, _, _ = _call(lib.vr_bounce_light_noalloc, big, z, z, np.ones(n, np.float32), z, 2, 0.0)
    assert rc == _lib.VR_E_CAPACITY
