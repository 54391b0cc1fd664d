"""The luxel stage on the GPU against the CPU restatement: lightmap extents
byte for byte, luxel positions bit for bit, the same patch per luxel; and the
final luxel light = direct light gathered at the luxel + its patch's bounced
light."""
import numpy as np
import pytest

import oracle
from vrad_amd import _lib, luxels, scenes
from vrad_amd.lightmap import GatherSampleLight
from vrad_amd.patches import MakePatches
from vrad_amd.raytracer import Environment

pytestmark = pytest.mark.gpu


def _random_faces(rng, n):
    """Planar convex polygons in random orientations with random lightmap
    vectors (scaled, slightly skewed from the polygon's plane axes) and
    offsets; some sky / nolight / degenerate."""
    f = np.zeros(n, _lib.FACE_DTYPE)
    ofs = np.zeros((n, 2), np.float32)
    for k in range(n):
        m = int(rng.integers(3, 9))
        ang = np.sort(rng.uniform(0, 2 * np.pi, m))
        r = rng.uniform(10, 600)
        pts2 = np.stack([np.cos(ang) * r, np.sin(ang) * r * rng.uniform(0.3, 1.0)], 1)
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        b = np.cross(a, rng.normal(size=3))
        b /= np.linalg.norm(b)
        c = rng.uniform(-3000, 3000, 3)
        if k % 3 == 0:
            a, b = np.eye(3)[k % 3], np.eye(3)[(k + 1) % 3]
            pts2, c = np.round(pts2), np.round(c)
        f["n_points"][k] = m
        f["points"][k, :m] = c + pts2[:, :1] * a + pts2[:, 1:] * b
        f["normal"][k] = np.cross(a, b)
        lux = rng.choice([4.0, 8.0, 16.0, 32.0])
        f["lux_s"][k] = (a + 0.1 * rng.normal(size=3) * (k % 2)) / lux
        f["lux_t"][k] = (b + 0.1 * rng.normal(size=3) * (k % 2)) / lux
        ofs[k] = rng.uniform(-50, 50, 2) * (k % 4 == 1)
    f["flags"][::17] = _lib.VR_FACE_SKY
    f["flags"][3::19] = _lib.VR_FACE_NOLIGHT
    f["n_points"][5::41] = 0
    return f, ofs


def _same(faces, ofs, label):
    print(f"[luxels] {label}: {len(faces)} faces", flush=True)
    lm, n = luxels.CalcFaceExtents(faces, ofs)
    olm, on = oracle.calc_face_extents(faces, ofs)
    assert n == on and lm.tobytes() == olm.tobytes()
    patches = MakePatches(faces)
    val = np.arange(3 * len(patches), dtype=np.float32).reshape(-1, 3)
    lx = luxels.FaceLuxels(faces, lm, patches, val, ofs)
    pos, pat = oracle.face_luxels(faces, olm, patches, ofs)
    print(f"[luxels] {label}: {n} luxels, {len(patches)} patches", flush=True)
    assert np.array_equal(lx["pos"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(lx["patch"], pat)
    has = pat >= 0
    assert np.array_equal(lx["value"][has], val[pat[has]])
    assert not lx["value"][~has].any()
    fi = np.repeat(np.arange(len(faces)), lm["n_luxels"])
    assert np.array_equal(lx["normal"], faces["normal"][fi])
    # every chosen patch is a leaf of the luxel's face
    assert (patches["child1"][pat[has]] == -1).all()
    assert np.array_equal(patches["face"][pat[has]], fi[has])
    return lm, lx, patches


def test_gpu_luxels_generated_scenes():
    _same(scenes.faces_of("room"), None, "room")
    f = scenes.faces_of("level", rooms_x=3, rooms_y=3, seed=4)
    ofs = np.tile(np.float32([0.5, -0.25]), (len(f), 1))
    _same(f, ofs, "3x3 level, offset lightmaps")


def test_gpu_luxels_random_faces():
    rng = np.random.default_rng(31)
    f, ofs = _random_faces(rng, 600)
    lm, lx, patches = _same(f, ofs, "random polygons")
    assert (lm["flags"] & _lib.VR_LM_UNLIT).any() and (lm["flags"] & _lib.VR_LM_NO_POINTS).any()


def test_gpu_luxels_layout_refused():
    f = scenes.faces_of("room")
    lm, _ = luxels.CalcFaceExtents(f)
    patches = MakePatches(f)
    bad = lm.copy()
    bad["luxel_ofs"][3] += 1
    with pytest.raises(_lib.VrError):
        luxels.FaceLuxels(f, bad, patches)
    bad = lm.copy()
    bad["n_luxels"][0] += 1
    with pytest.raises(_lib.VrError):
        luxels.FaceLuxels(f, bad, patches)


def test_gpu_final_light_face():
    """Room faces: luxel light = GatherSampleLight at the luxel + (total -
    direct) of its patch; a constant bounced light reaches every luxel."""
    s = scenes.room(seed=1, n_pillars=40, chop=64.0, tess=32.0)
    faces = scenes.faces_of("room")
    patches = MakePatches(faces)
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure()
    rng = np.random.default_rng(2)
    direct = rng.uniform(0, 1, (len(patches), 3)).astype(np.float32)
    total = direct + np.float32(0.25)
    lm, light, lx = luxels.FinalLightFace(env, faces, patches, total, direct, s.lights)
    d = GatherSampleLight(env, lx["pos"], lx["normal"], s.lights)
    assert np.array_equal(lx["value"], (total - direct)[lx["patch"]])
    assert np.array_equal(light, (d + lx["value"]).astype(np.float32))
    assert (lx["patch"] >= 0).all()


def test_gpu_lighting_lump():
    """The room's final luxel light encoded as the lighting lump: equal to the
    restatement byte for byte, lightofs -1 exactly on the faces without
    luxels, every luxel within one unit of its light, and hand-chosen luxels
    encoded to their hand-derived bytes."""
    s = scenes.room(seed=1, n_pillars=40, chop=64.0, tess=32.0)
    faces = scenes.faces_of("room")
    patches = MakePatches(faces)
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure()
    rng = np.random.default_rng(3)
    direct = rng.uniform(0, 4, (len(patches), 3)).astype(np.float32)
    lm, light, _ = luxels.FinalLightFace(env, faces, patches, direct * np.float32(1.5), direct,
                                         s.lights)
    print(f"[gpu] lighting lump: {len(faces)} faces, {len(light)} luxels", flush=True)
    lump, ofs = luxels.LightingLump(light, lm)
    o_lump, o_ofs = oracle.lighting_lump(light, lm)
    assert np.array_equal(lump, o_lump) and np.array_equal(ofs, o_ofs)
    assert np.array_equal(ofs < 0, lm["n_luxels"] == 0)
    for f in np.flatnonzero(lm["n_luxels"] > 0)[:50]:
        a, n = int(ofs[f]), int(lm["n_luxels"][f])
        dec = luxels.decode_rgbexp32(lump[a:a + 4 * n]).astype(np.float64)
        src = light[lm["luxel_ofs"][f]:lm["luxel_ofs"][f] + n].astype(np.float64)
        unit = np.ldexp(1.0, lump[a + 3:a + 4 * n:4].view(np.int8).astype(int))
        assert np.all(dec <= src) and np.all(src - dec < unit[:, None])
    hand = np.array([[1.0, 0.5, 0.25], [255.0, 0.0, 0.0], [256.0, 1.0, 0.0], [1e-3, 2e-3, 0.0]],
                    np.float32)
    lmh = np.zeros(1, _lib.FACE_LIGHTMAP_DTYPE)
    lmh["n_luxels"] = 4
    lh, oh = luxels.LightingLump(hand, lmh)
    assert list(oh) == [4]
    assert lh[4:].tolist() == [128, 64, 32, 0xF9, 255, 0, 0, 0, 128, 0, 0, 1, 65, 131, 0, 0xF0]


def test_gpu_luxels_full_level():
    """The 8x8-room level's 2,000 faces at 16 units per luxel (1.46 M
    luxels): GPU == restatement, every luxel on its face's plane."""
    f = scenes.faces_of("level", seed=2, rooms_x=8, rooms_y=8, chop=32.0, tess=64.0)
    lm, lx, patches = _same(f, None, "8x8 level")
    fi = np.repeat(np.arange(len(f)), lm["n_luxels"])
    d = np.einsum("ij,ij->i", lx["pos"].astype(np.float64), f["normal"][fi].astype(np.float64))
    d0 = np.einsum("ij,ij->i", f["points"][fi, 0].astype(np.float64),
                   f["normal"][fi].astype(np.float64))
    assert np.max(np.abs(d - d0)) < 1e-2  # FP32 rounding of coordinates up to ~1e4
