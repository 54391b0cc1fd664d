"""The precision workload at its stated shape on the GPU: 250,000 rows,
lognormal lengths (mean 2,000, sigma 1, at most 40,000; about 5e8
transfers), columns uniform without replacement.  FP32, FP16 and BF16
bounces bit for bit against the oracle's stream order, and the error of each
against the FP64 recurrence; the synthetic generator's columns against a
restatement of its draw process."""
import numpy as np
import pytest

import oracle
from vrad_amd import _lib, lightmap
from vrad_amd.patches import Transfers

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _u01(seed, a, b):
    """The generator's hash to (0, 1]: splitmix64 of seed + golden * (a p + b + 1)."""
    z = _mix64((seed + 0x9E3779B97F4A7C15 * ((a * 0x100000001B3 + b + 1) & M64)) & M64)
    return ((z >> 11) + 1.0) * 2.0 ** -53


def _uniform_row(seed, i, L, n):
    """Columns of row i: the first L distinct values of floor(u_d (n - 1)),
    d = 0, 1, ..., mapped past i, sorted."""
    m = n - 1
    got, d = set(), 0
    while len(got) < L:
        c = int(_u01(seed ^ 0x5bd1e995, i, d) * m)
        got.add(min(c, m - 1))
        d += 1
    return np.array(sorted(c + 1 if c >= i else c for c in got), np.int32)


def test_synthetic_uniform_columns_restated():
    n, seed = 4000, 11
    print(f"[gpu] synthetic uniform columns n={n}", flush=True)
    xf = Transfers.synthetic(seed=seed, n_rows=n, mean_len=150.0, sigma=1.0, max_len=3000)
    rp, col, coef = xf.download()
    L = np.diff(rp)
    assert L.min() >= 1 and L.max() <= 3000 and abs(L.mean() - 150) < 15
    for i in list(range(0, n, 331)) + [int(np.argmax(L))]:
        assert np.array_equal(col[rp[i]:rp[i + 1]], _uniform_row(seed, i, int(L[i]), n)), i
    # every row sorted, unique, off the diagonal
    rows = np.repeat(np.arange(n), L)
    assert np.all(col != rows) and col.min() >= 0 and col.max() < n
    same_row = rows[1:] == rows[:-1]
    assert np.all(np.diff(col)[same_row] > 0)
    # gaps: uniform rows have gaps well above the strata width, stratified never
    xs = Transfers.synthetic(seed=seed, n_rows=n, mean_len=150.0, sigma=1.0, max_len=3000,
                             columns="stratified")
    rps, cols, _ = xs.download()
    assert np.array_equal(rps, rp)
    wide_u = wide_s = 0
    for i in range(0, n, 7):
        if L[i] < 20:
            continue
        lim = 2.0 * (n - 1) / L[i] + 1
        wide_u += np.diff(col[rp[i]:rp[i + 1]]).max() > lim
        wide_s += np.diff(cols[rp[i]:rp[i + 1]]).max() > lim
    assert wide_s == 0 and wide_u > 50


def _inputs(n, seed=5):
    rng = np.random.default_rng(seed)
    e0 = rng.random((n, 3)).astype(np.float32)
    refl = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    area = rng.uniform(1, 4, n).astype(np.float32)
    return e0, refl, area


def max_rel(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    z = ref == 0
    assert np.all(a[z] == 0)
    return float(np.max(np.abs(a[~z] - ref[~z]) / np.abs(ref[~z])))


def test_precision_workload_full_shape():
    n = 250_000
    print(f"[gpu] precision workload n={n} mean 2000 max 40000", flush=True)
    xf = Transfers.synthetic(seed=3, n_rows=n, mean_len=2000.0, sigma=1.0, max_len=40000)
    rp, col, coef = xf.download()
    L = np.diff(rp)
    T = int(rp[-1])
    assert abs(T / n - 2000) < 20 and L.max() <= 40000 and L.max() > 20000 and L.min() >= 1
    e0, refl, area = _inputs(n)
    tot, en, k = lightmap.BounceLight(xf, e0, refl, area, max_bounces=3, stop_frac=0.0)
    print(f"[gpu] T={T} FP32 done; oracle", flush=True)
    otot, oen, ok = oracle.bounce_f32(rp, col, coef, e0, refl, area, max_bounces=3,
                                      stop_frac=0.0, threads=16)
    assert k == ok == 3 and np.array_equal(tot.view(np.uint32), otot.view(np.uint32))
    assert np.array_equal(en, oen)
    t64, e64, _ = oracle.bounce_f64(rp, col, coef, e0, refl, area, max_bounces=3, stop_frac=0.0)
    err = {"f32": max_rel(tot, t64)}
    for name, prec, bound in (("f16", _lib.VR_COEF_F16, 1e-3), ("bf16", _lib.VR_COEF_BF16, 5e-3)):
        xf.set_precision(prec)
        t16, en16, k16 = lightmap.BounceLight(xf, e0, refl, area, max_bounces=3, stop_frac=0.0)
        _, _, _, raw, scale = xf.download(raw16=True)
        print(f"[gpu] {name} done; oracle", flush=True)
        o16, oen16, _ = oracle.bounce_f32(rp, col, raw, e0, refl, area, max_bounces=3,
                                          stop_frac=0.0, kind=prec, row_scale=scale, threads=16)
        assert k16 == 3 and np.array_equal(t16.view(np.uint32), o16.view(np.uint32))
        assert np.array_equal(en16, oen16)
        err[name] = max_rel(t16, t64)
        assert err[name] < bound
        del raw
    print(f"[gpu] max relative error vs FP64 after 3 bounces: {err}", flush=True)
    assert err["f32"] < 1e-4
