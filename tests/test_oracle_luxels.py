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


# ---- the lighting lump (ColorRGBExp32) ------------------------------------------------
def test_rgbexp32_hand_encoded():
    """Hand-derived: (1, .5, .25): max 1 = 2^0, e = -7, x 128 -> 128, 64, 32;
    255 fits with e = 0; 256 needs e = 1 (1 -> 0.5 truncates to 0); 2e-3:
    floor(log2) = -9, e = -16, x 65536 -> 131 and 65 (1e-3)."""
    assert oracle.rgbexp32(1.0, 0.5, 0.25) == [128, 64, 32, 0xF9]
    assert oracle.rgbexp32(255.0, 0.0, 0.0) == [255, 0, 0, 0]
    assert oracle.rgbexp32(256.0, 1.0, 0.0) == [128, 0, 0, 1]
    assert oracle.rgbexp32(1e-3, 2e-3, 0.0) == [65, 131, 0, 0xF0]
    assert oracle.rgbexp32(0.0, 0.0, 0.0) == [0, 0, 0, 0]
    assert oracle.rgbexp32(-1.0, -2.0, -3.0) == [0, 0, 0, 0]
    assert oracle.rgbexp32(1e-39, 0, 0) == [0, 0, 0, 0]           # e below -127
    assert oracle.rgbexp32(float("inf"), 1, 1) == [0, 0, 0, 0]


def test_rgbexp32_round_trip_bound():
    """Decoded = bytes x 2^e: the largest channel lands in [128, 256) (so its
    relative error is below 1/128), every channel is at most the input and
    within one unit 2^e of it."""
    from vrad_amd.luxels import decode_rgbexp32
    rng = np.random.default_rng(5)
    c = (rng.random((5000, 3)) * np.exp(rng.uniform(-20, 20, (5000, 1)))).astype(np.float32)
    enc = np.array([oracle.rgbexp32(*v) for v in c], np.uint8)
    dec = decode_rgbexp32(enc).astype(np.float64)
    e = enc[:, 3].view(np.int8).astype(np.float64)
    m = c.max(axis=1).astype(np.float64)
    assert np.all(enc[:, :3].max(axis=1) >= 128)
    assert np.all(dec <= c + 0.0) and np.all(c - dec < np.ldexp(1.0, e.astype(int))[:, None])
    assert np.all((m - dec.max(axis=1)) / m < 1.0 / 128)


def test_lighting_lump_layout():
    """Two lit faces around an unlit one: each lit face's average then its
    luxels; lightofs points past the average (-1 when unlit).  Constant luxels
    average to themselves."""
    lm = np.zeros(3, oracle.FACE_LIGHTMAP_DTYPE)
    lm["n_luxels"] = [4, 0, 70]
    lm["luxel_ofs"] = [0, 4, 4]
    lux = np.zeros((74, 3), np.float32)
    lux[:4] = (2.0, 1.0, 0.5)
    lux[4:] = np.linspace(0, 8, 70 * 3, dtype=np.float32).reshape(70, 3)
    lump, ofs = oracle.lighting_lump(lux, lm)
    assert list(ofs) == [4, -1, 4 + 16 + 4] and len(lump) == 4 * (5 + 71)
    words = lump.reshape(-1, 4)
    assert list(words[0]) == oracle.rgbexp32(2.0, 1.0, 0.5)        # the average
    assert all(list(words[1 + k]) == oracle.rgbexp32(2.0, 1.0, 0.5) for k in range(4))
    avg = lux[4:].astype(np.float64).mean(axis=0).astype(np.float32)
    assert list(words[5]) == oracle.rgbexp32(*avg)
    assert all(list(words[6 + k]) == oracle.rgbexp32(*lux[4 + k]) for k in range(70))


def test_lighting_lump_sizing_on_host():
    """The product sizes the lump and fills lightofs on the host (no GPU)."""
    import ctypes as C
    from vrad_amd import _lib
    lm = np.zeros(4, _lib.FACE_LIGHTMAP_DTYPE)
    lm["n_luxels"] = [9, 0, 1, 25]
    lm["luxel_ofs"] = [0, 9, 9, 10]
    nb = C.c_int64()
    ofs = np.zeros(4, np.int32)
    _lib.check(_lib.lib.vr_lighting_lump(None, 35, _lib.ptr(lm), 4, None, 0, _lib.ptr(ofs),
                                         C.byref(nb)))
    o_lump, o_ofs = oracle.lighting_lump(np.zeros((35, 3), np.float32), lm)
    assert nb.value == len(o_lump) == 4 * (10 + 2 + 26) and np.array_equal(ofs, o_ofs)
    lm["luxel_ofs"][3] = 11                                    # runs past the luxels
    assert _lib.lib.vr_lighting_lump(None, 35, _lib.ptr(lm), 4, None, 0, _lib.ptr(ofs),
                                     C.byref(nb)) == _lib.VR_E_INVALID
