"""The largest generated level: the level generator at
1,551,028 triangles and 264,363 patches (16 x 16 rooms, chop 72, tess 40).
The GPU-built kd-tree equals the host builder's and the oracle's bit for
bit, and TestLine on MakeTransfers-order segments and closest-hit rays from
patch centres equal the oracle's kd traversal on samples of each batch."""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from tests.test_gpu_kdbuild import _same as same_tree
from vrad_amd import scenes
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu

LEVEL_16X16 = dict(seed=3, rooms_x=16, rooms_y=16, chop=72.0, tess=40.0, n_lights=16, sun=True)


@pytest.fixture(scope="module")
def big_level():
    s = scenes.level(**LEVEL_16X16)
    print(f"[level16x16] {len(s.ids)} triangles, {len(s.p_area)} patches", flush=True)
    assert 1_400_000 <= len(s.ids) <= 1_700_000 and 230_000 <= len(s.p_area) <= 300_000
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure(n_threads=16)
    kd = oracle.KdTree(s.verts9)
    tri = oracle.precompute(s.verts9, s.ids)
    yield s, env, kd, tri
    env.close()


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def test_level16x16_device_tree_equals_host_and_oracle(big_level):
    s, env, kd, _ = big_level
    same_tree(s.verts9, s.ids, "level 16x16", against_oracle=False)
    ch, sp = env.OptimizedKDTree
    assert np.array_equal(ch, kd.children)
    assert np.array_equal(sp.view(np.uint32), kd.split.view(np.uint32))
    assert np.array_equal(env.TriangleIndexList, kd.idx)


def test_level16x16_test_line_vs_oracle(big_level):
    s, env, kd, tri = big_level
    o, nrm = s.p_origin, s.p_normal
    n = len(o)
    rng = np.random.default_rng(31)
    # rows of neighbouring patches (mostly one room, partly visible) and random
    # far pairs (mostly occluded), facing cull as MakeTransfers applies it
    i = np.concatenate([np.repeat(rng.integers(0, n - 400, 600), 400),
                        rng.integers(0, n, 200_000)])
    j = np.concatenate([i[:240_000] + np.tile(np.arange(1, 401), 600),
                        rng.integers(0, n, 200_000)])
    dv = o[j] - o[i]
    keep = (i != j) & (np.einsum("ij,ij->i", nrm[i], dv) > 0) & \
        (np.einsum("ij,ij->i", nrm[j], -dv) > 0)
    i, j = i[keep], j[keep]
    a = (o[i] + np.float32(0.1) * nrm[i]).astype(np.float32)
    b = (o[j] + np.float32(0.1) * nrm[j]).astype(np.float32)
    print(f"[level16x16] TestLine on {len(a)} segments", flush=True)
    vis, sky = env.TestLines(a, b)
    vo, so = oracle.testline(kd, tri, a, b, threads=16)
    assert np.array_equal(vis, vo) and np.array_equal(sky, so)
    assert 0.02 < vis.mean() < 0.98


def test_level16x16_closest_hit_vs_oracle(big_level):
    s, env, kd, tri = big_level
    mn, mx = s.verts9.reshape(-1, 3).min(0), s.verts9.reshape(-1, 3).max(0)
    rng = np.random.default_rng(32)
    o, d = H.random_rays(rng, 150_000, mn, mx)
    print(f"[level16x16] closest hit on {len(o)} rays", flush=True)
    hg, tg = env.TraceRays(o, d, 0.0, 1.732050807569 * 32768.0)
    hk, tk = oracle.trace_kd(kd, tri, o, d, 0.0, 1.732050807569 * 32768.0, threads=16)
    assert np.array_equal(hg, hk) and np.array_equal(_bits(tg), _bits(tk))
    assert (hg >= 0).mean() > 0.5
