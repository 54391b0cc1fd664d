"""MakeTransfers on the GPU vs the CPU oracle (bit-exact CSR) and the
closed-form form factors."""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from vrad_amd import _lib, scenes
from vrad_amd.patches import MakeTransfers
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu


def _env(v9, ids):
    env = Environment()
    env.AddTriangles(v9, ids)
    env.SetupAccelerationStructure()
    return env


@pytest.mark.parametrize("path", [_lib.VR_PATH_OPT, _lib.VR_PATH_PLAIN], ids=["opt", "plain"])
@pytest.mark.parametrize("batch", [0, 50000], ids=["batch-default", "batch50k"])
def test_room_transfers_bitexact(room_scene, path, batch):
    s = room_scene
    print(f"[gpu] room transfers path={path} batch={batch} patches={len(s.p_area)}", flush=True)
    _lib.check(_lib.lib.vr_set_path(path))
    try:
        env = _env(s.verts9, s.ids)
        xf = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, batch_pairs=batch)
        rp, col, coef = xf.download()
        info = xf.info()
    finally:
        _lib.check(_lib.lib.vr_set_path(_lib.VR_PATH_OPT))
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    orp, ocol, ocoef, ors = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area,
                                                  threads=8)
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(coef.view(np.uint32), ocoef.view(np.uint32))
    n = len(s.p_area)
    assert info.pairs_considered == n * (n - 1) // 2
    # segments traced = pairs i < j passing the facing cull, in the cull
    # kernel's float32 arithmetic (the rounds' counts add up exactly, on one
    # stream or two)
    o = s.p_origin.astype(np.float32)
    nr = s.p_normal.astype(np.float32)
    want = 0
    for i in range(n - 1):
        d = o[i + 1:] - o[i]
        di = (nr[i, 0] * d[:, 0] + nr[i, 1] * d[:, 1]) + nr[i, 2] * d[:, 2]
        dj = -((nr[i + 1:, 0] * d[:, 0] + nr[i + 1:, 1] * d[:, 1]) + nr[i + 1:, 2] * d[:, 2])
        want += int(np.count_nonzero((di > 0) & (dj > 0)))
    assert info.segments_traced == want
    assert info.nnz == len(col) and info.stored_nnz % 8 == 0


def test_howell_parallel_on_gpu():
    n = 16
    pa = H.square_patches((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), n)
    pb = H.square_patches((0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 0, -1), n)
    o, nrm, a = [np.concatenate(x) for x in zip(pa, pb)]
    env = _env(np.zeros((0, 9), np.float32), np.zeros(0, np.int32))
    rp, col, coef = MakeTransfers(env, o, nrm, a, normalize=False).download()
    rs = np.add.reduceat(coef.astype(np.float64), rp[:-1])[: n * n]
    F = np.sum(rs * a[: n * n]) / np.sum(a[: n * n])
    ref = H.howell_parallel(1, 1, 1)
    assert abs(F - ref) < 0.01 * ref


def test_transfers_occluder_blocks():
    # two facing squares with a slab between them: no transfers
    n = 4
    pa = H.square_patches((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), n)
    pb = H.square_patches((0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 0, -1), n)
    o, nrm, a = [np.concatenate(x) for x in zip(pa, pb)]
    v9, ids = H.to_arrays(H.box((-5, -5, 0.4), (5, 5, 0.6)))
    env = _env(v9, ids)
    rp, col, coef = MakeTransfers(env, o, nrm, a, normal_offset=0.01).download()
    assert len(col) == 0


@pytest.mark.parametrize("layers,visible", [
    ("sky", True),             # the closest hit is sky: visible
    ("sky-then-slab", True),   # a sky quad before an opaque slab: still visible
    ("slab-then-sky", False),  # an opaque slab before the sky quad: occluded
], ids=["sky", "sky-then-slab", "slab-then-sky"])
def test_transfers_sky_layers_vs_oracle(layers, visible):
    # two facing squares; between them a sky quad and/or an opaque slab.  A
    # segment is occluded iff its closest hit is not sky, so a slab behind
    # the sky quad does not occlude (MakeTransfers walks the full tree for
    # any opaque hit first, then the sky tree, then the full tree again)
    n = 4
    pa = H.square_patches((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), n)
    pb = H.square_patches((0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 0, -1), n)
    o, nrm, a = [np.concatenate(x) for x in zip(pa, pb)]
    sky_z, slab = {"sky": (0.5, None), "sky-then-slab": (0.4, (0.6, 0.7)),
                   "slab-then-sky": (0.6, (0.3, 0.4))}[layers]
    tris = H.quad(H.SKY, (-5, -5, sky_z), (5, -5, sky_z), (5, 5, sky_z), (-5, 5, sky_z))
    if slab is not None:
        tris += H.box((-5, -5, slab[0]), (5, 5, slab[1]))
    v9, ids = H.to_arrays(tris)
    print(f"[gpu] transfers sky layers={layers}", flush=True)
    rp, col, coef = MakeTransfers(_env(v9, ids), o, nrm, a).download()
    tri = oracle.precompute(v9, ids)
    kd = oracle.KdTree(v9)
    orp, ocol, ocoef, _ = oracle.make_transfers(kd, tri, o, nrm, a, threads=4)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(coef.view(np.uint32), ocoef.view(np.uint32))
    # every bottom patch sees every top patch, or none does
    real = col[coef != 0]
    assert (len(real) == 2 * (n * n) ** 2) if visible else (len(real) == 0)
