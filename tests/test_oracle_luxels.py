"""The oracle's luxel stage (CalcFaceExtents, luxel positions, the patch each
luxel takes) pinned by hand-derived cases and independent float64 checks (no
GPU)."""
import numpy as np

import oracle


def _face(points, normal, lux_s, lux_t, flags=0):
    f = np.zeros(1, oracle.FACE_DTYPE)
    f["n_points"] = len(points)
    if len(points):
        f["points"][0, :len(points)] = points
    f["normal"] = normal
    f["lux_s"] = lux_s
    f["lux_t"] = lux_t
    f["flags"] = flags
    return f


def _square(x0, y0, side, lux=1 / 16, flags=0):
    return _face([(x0, y0, 0), (x0 + side, y0, 0), (x0 + side, y0 + side, 0), (x0, y0 + side, 0)],
                 (0, 0, 1), (lux, 0, 0), (0, lux, 0), flags)


def test_extents_closed_forms():
    lm, n = oracle.calc_face_extents(_square(0, 0, 256))
    assert list(lm[0]["mins"]) == [0, 0] and list(lm[0]["size"]) == [16, 16]
    assert lm[0]["n_luxels"] == 17 * 17 and n == 289 and lm[0]["flags"] == 0
    # offsets: s in [0.5, 16.5] -> floor 0, ceil 17; t in [-0.25, 15.75] -> -1, 16
    lm, _ = oracle.calc_face_extents(_square(0, 0, 256), lux_ofs=[0.5, -0.25])
    assert list(lm[0]["mins"]) == [0, -1] and list(lm[0]["size"]) == [17, 17]
    # negative coordinates: -100/16 = -6.25 .. -36/16 = -2.25 -> -7, size 5
    lm, _ = oracle.calc_face_extents(_square(-100, -100, 64))
    assert list(lm[0]["mins"]) == [-7, -7] and list(lm[0]["size"]) == [5, 5]


def test_extents_limits_and_skips():
    faces = np.concatenate([_square(0, 0, 2016), _square(0, 0, 2032), _square(0, 0, 64, flags=1),
                            _square(0, 0, 64, flags=4), _square(0, 0, 64, flags=2),
                            _face([], (0, 0, 1), (1, 0, 0), (0, 1, 0))])
    lm, n = oracle.calc_face_extents(faces)
    # 2016 / 16 = 126 luxels = MAX_LIGHTMAP_DIM_WITHOUT_BORDER + 1: allowed;
    # 127 is "Bad surface extents"
    assert list(lm["size"][0]) == [126, 126] and lm["flags"][0] == 0
    assert lm["flags"][1] == 2 and lm["n_luxels"][1] == 0
    assert list(lm["flags"][2:]) == [1, 1, 0, 4]          # sky, nolight, nochop (lit), no points
    assert list(lm["n_luxels"][2:]) == [0, 0, 25, 0]
    assert list(lm["luxel_ofs"]) == [0, 127 * 127] + [127 * 127] * 3 + [127 * 127 + 25]
    assert n == 127 * 127 + 25


def test_extents_fp32_order():
    # ((x lx + y ly) + z lz) + o in FP32: 2^24 + 1 rounds to 2^24, so the value
    # is 0 (exactly 1 in real arithmetic, or in the order x lx + (y ly + o))
    f = _face([(2.0**24, 1.0, 0.0)], (0, 0, 1), (1, 1, 0), (0, 0, 1))
    lm, _ = oracle.calc_face_extents(f, lux_ofs=[-2.0**24, 0.0])
    assert list(lm[0]["mins"]) == [0, 0] and list(lm[0]["size"]) == [0, 0]


def test_luxel_positions_closed_form():
    lm, _ = oracle.calc_face_extents(_square(0, 0, 256), lux_ofs=[0.5, -0.25])
    pos, _ = oracle.face_luxels(_square(0, 0, 256), lm, np.zeros(0, oracle.PATCH_DTYPE),
                                lux_ofs=[0.5, -0.25])
    s, t = np.meshgrid(np.arange(18), np.arange(18))
    want = np.stack([16 * (s.ravel() + 0 - 0.5), 16 * (t.ravel() - 1 + 0.25),
                     np.zeros(s.size)], 1)
    assert np.array_equal(pos, want.astype(np.float32))
    # a wall on x = 5: s along y at 1/16, t along z at 1/8
    f = _face([(5, 0, 0), (5, 64, 0), (5, 64, 32), (5, 0, 32)], (1, 0, 0), (0, 1 / 16, 0),
              (0, 0, 1 / 8))
    lm, _ = oracle.calc_face_extents(f)
    assert list(lm[0]["size"]) == [4, 4]
    pos, _ = oracle.face_luxels(f, lm, np.zeros(0, oracle.PATCH_DTYPE))
    s, t = np.meshgrid(np.arange(5), np.arange(5))
    want = np.stack([np.full(25, 5.0), 16.0 * s.ravel(), 8.0 * t.ravel()], 1)
    assert np.array_equal(pos, want.astype(np.float32))


def test_luxel_patch_choice_independent_check():
    """On a chopped square the chosen leaf is the first (in patch order)
    whose square contains the luxel, inclusive of edges; outside every
    square it is the nearest centre (ties to the lower index)."""
    faces = np.concatenate([_square(0, 0, 256), _square(300, 40, 100, lux=1 / 8)])
    patches = oracle.make_patches(faces)
    ofs = np.array([0.5, 0.5, 0.0, 0.25], np.float32)  # some luxels of face 0 lie outside
    lm, _ = oracle.calc_face_extents(faces, lux_ofs=ofs)
    pos, pat = oracle.face_luxels(faces, lm, patches, lux_ofs=ofs)
    leaves = np.flatnonzero(patches["child1"] == -1)
    outside = 0
    for f in range(len(faces)):
        mine = leaves[patches["face"][leaves] == f]
        lo, hi = patches["mins"][mine, :2].astype(np.float64), patches["maxs"][mine, :2]
        ctr = patches["origin"][mine, :2].astype(np.float64)
        for k in range(lm["luxel_ofs"][f], lm["luxel_ofs"][f] + lm["n_luxels"][f]):
            p = pos[k, :2].astype(np.float64)
            inside = np.flatnonzero(np.all((p >= lo) & (p <= hi), axis=1))
            if len(inside):
                assert pat[k] == mine[inside[0]]
            else:
                outside += 1
                d = np.sum((p - ctr) ** 2, axis=1)
                assert pat[k] == mine[np.flatnonzero(d == d.min())[0]]
    assert outside > 0
