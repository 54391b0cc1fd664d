"""PairEdges and GetPhongNormal (rad/lightmap/lightmap.go:37-228,
rad/lightmap/normallist.go:51-144) in the CPU oracle, pinned by closed forms:
coplanar neighbours leave the face normal, a shallow crease bends the shared
vertices' normals to the bisector and a steep one does not, the threshold is
inclusive, vertices are shared only on identical positions, and the phong
normal is the face normal at the centroid, the vertex normal at a vertex and
the mean of the two at an edge midpoint.  The product's host PairEdges must
equal the oracle bit for bit."""
import numpy as np
import pytest

import oracle
from vrad_amd import _lib
from vrad_amd.patches import PairEdges

T = oracle.SMOOTHING_THRESHOLD


def _faces(polys, normals):
    f = np.zeros(len(polys), oracle.FACE_DTYPE)
    for k, (p, n) in enumerate(zip(polys, normals)):
        f["n_points"][k] = len(p)
        f["points"][k, :len(p)] = p
        f["normal"][k] = n
        f["lux_s"][k] = [1 / 16, 0, 0]
        f["lux_t"][k] = [0, 1 / 16, 0]
    return f


def _crease(deg):
    """Face A: the square [0,100]^2 in z = 0 (normal +z).  Face B: a 100 x 100
    flap hinged on A's edge x = 100, lifted by deg degrees."""
    th = np.radians(deg)
    A = [[0, 0, 0], [100, 0, 0], [100, 100, 0], [0, 100, 0]]
    far = 100 + 100 * np.cos(th)
    up = 100 * np.sin(th)
    B = [[100, 0, 0], [far, 0, up], [far, 100, up], [100, 100, 0]]
    nb = [-np.sin(th), 0, np.cos(th)]
    return _faces([A, B], [[0, 0, 1], nb])


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_pair_edges_coplanar_neighbours_keep_face_normal():
    f = _faces([[[0, 0, 0], [100, 0, 0], [100, 100, 0], [0, 100, 0]],
                [[100, 0, 0], [200, 0, 0], [200, 100, 0], [100, 100, 0]],
                [[0, 100, 0], [100, 100, 0], [100, 200, 0], [0, 200, 0]]], [[0, 0, 1]] * 3)
    vn = oracle.pair_edges(f)
    for k in range(3):
        assert np.array_equal(vn[k, :4], np.tile(np.float32([0, 0, 1]), (4, 1)))
        assert not vn[k, 4:].any()


def test_pair_edges_shallow_crease_gives_bisector():
    f = _crease(30.0)
    vn = oracle.pair_edges(f)
    bis = _unit(f["normal"][0].astype(np.float64) + f["normal"][1])
    # the bisector of +z and the flap's normal leans 15 degrees toward -x
    assert np.allclose(bis, [-np.sin(np.radians(15)), 0, np.cos(np.radians(15))], atol=1e-7)
    shared_a, shared_b = [1, 2], [0, 3]
    for k, sh in ((0, shared_a), (1, shared_b)):
        for j in range(4):
            want = bis if j in sh else _unit(f["normal"][k])
            assert np.allclose(vn[k, j], want, atol=2e-7), (k, j)


def test_pair_edges_steep_crease_does_not_smooth():
    f = _crease(90.0)                     # dot 0 < 0.7071067
    vn = oracle.pair_edges(f)
    for k in range(2):
        for j in range(4):
            assert np.allclose(vn[k, j], _unit(f["normal"][k]), atol=1e-7)
    assert np.array_equal(vn[0, :4], np.tile(np.float32([0, 0, 1]), (4, 1)))


def test_pair_edges_threshold_is_inclusive():
    # face normals set directly: the dot with +z is the neighbour's z
    base = [[0, 0, 0], [100, 0, 0], [100, 100, 0], [0, 100, 0]]
    flap = [[100, 0, 0], [200, 0, 0], [200, 100, 0], [100, 100, 0]]
    for z, smoothed in ((T, True), (np.nextafter(T, np.float32(0)), False)):
        x = np.float32(np.sqrt(1.0 - float(z) ** 2))
        f = _faces([base, flap], [[0, 0, 1], [-x, 0, z]])
        vn = oracle.pair_edges(f)
        bent = not np.array_equal(vn[0, 1], np.float32([0, 0, 1]))
        assert bent == smoothed, z


def test_pair_edges_shares_vertices_by_exact_position():
    f = _crease(30.0)
    f["points"][1, 0, 1] = np.nextafter(np.float32(0), np.float32(1))  # one ulp off A's (100,0,0)
    vn = oracle.pair_edges(f)
    assert np.array_equal(vn[0, 1], np.float32([0, 0, 1]))    # no longer shared
    assert not np.array_equal(vn[0, 2], np.float32([0, 0, 1]))  # (100,100,0) still is


def test_phong_closed_forms():
    f = _crease(30.0)
    vn = oracle.pair_edges(f)
    c = np.float32([50, 50, 0])           # WindingCenter of A
    a = f[0]
    # at the centroid: the face normal
    assert np.array_equal(oracle.phong_normal(a, c, vn[0], c), np.float32([0, 0, 1]))
    # at each vertex: that vertex's normal
    for j in range(4):
        got = oracle.phong_normal(a, c, vn[0], a["points"][j])
        assert np.allclose(got, _unit(vn[0, j]), atol=2e-7), j
    # at the midpoint of the shared edge (100, 50, 0): the bisector, 15 degrees
    got = oracle.phong_normal(a, c, vn[0], [100, 50, 0])
    assert np.allclose(got, [-np.sin(np.radians(15)), 0, np.cos(np.radians(15))], atol=3e-7)
    # at the midpoint of an unshared edge: the face normal
    assert np.allclose(oracle.phong_normal(a, c, vn[0], [0, 50, 0]), [0, 0, 1], atol=1e-7)
    # halfway from the centroid to the shared edge's midpoint: normalised
    # 1/2 face normal + 1/4 + 1/4 of the bisector
    want = _unit(0.5 * np.float64([0, 0, 1]) + 0.5 * _unit(vn[0, 1].astype(np.float64)))
    assert np.allclose(oracle.phong_normal(a, c, vn[0], [75, 50, 0]), want, atol=3e-7)


def test_phong_patches_bend_toward_the_crease():
    f = _crease(30.0)
    vn = oracle.pair_edges(f)
    plain = oracle.make_patches(f)
    bent = oracle.make_patches(f, vertex_normals=vn)
    assert len(plain) == len(bent)
    same = [n for n in plain.dtype.names if n != "normal"]
    for n in same:
        assert np.array_equal(plain[n], bent[n]), n
    base = bent["parent"] < 0
    assert np.array_equal(bent["normal"][base], plain["normal"][base])  # base patches: face normal
    leaves = bent[(bent["child1"] < 0) & (bent["face"] == 0)]
    n = leaves["normal"].astype(np.float64)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    tilt = -n[:, 0]                        # toward -x, up to sin 15 degrees
    assert tilt.min() >= -1e-7 and tilt.max() <= np.sin(np.radians(15)) + 1e-6
    # the tilt grows toward the crease at x = 100
    x = leaves["origin"][:, 0]
    assert tilt[np.argmax(x)] > tilt[np.argmin(x)] + 0.05
    assert abs(n[:, 1]).max() < 1e-6        # nothing bends along y


def test_phong_flat_room_keeps_face_normals():
    from vrad_amd import scenes
    f = scenes.faces_of("room")
    vn = oracle.pair_edges(f)
    a, b = oracle.make_patches(f), oracle.make_patches(f, vertex_normals=vn)
    # the room's walls meet at 90 degrees: nothing is smoothed
    assert np.allclose(a["normal"], b["normal"], atol=1e-7)


def test_product_pair_edges_equals_oracle():
    """Host code in the product library: no GPU needed."""
    from vrad_amd import scenes
    rng = np.random.default_rng(5)
    for f in (_crease(30.0), _crease(44.0), scenes.faces_of("room"),
              scenes.faces_of("level", rooms_x=2, rooms_y=2, chop=64.0, tess=64.0)):
        f = np.ascontiguousarray(f, oracle.FACE_DTYPE)
        # jitter some normals so that shallow creases appear
        g = f.copy()
        pick = rng.random(len(g)) < 0.3
        nn = g["normal"][pick] + rng.normal(scale=0.2, size=(pick.sum(), 3)).astype(np.float32)
        g["normal"][pick] = (nn / np.linalg.norm(nn, axis=1, keepdims=True)).astype(np.float32)
        for faces in (f, g):
            want = oracle.pair_edges(faces)
            got = PairEdges(faces.view(_lib.FACE_DTYPE))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_product_pair_edges_rejects_bad_points():
    f = _crease(30.0).view(_lib.FACE_DTYPE).copy()
    f["n_points"][1] = 17
    with pytest.raises(_lib.VrError) as ei:
        PairEdges(f)
    assert ei.value.code == _lib.VR_E_INVALID
