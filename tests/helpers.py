"""Test-side geometry helpers written from the reference's definitions."""
import math

import numpy as np


def quad(id_, v1, v2, v3, v4):
    """Environment.AddQuad (raytracer/environment.go:71-75)."""
    return [(id_, v1, v2, v3), (id_ + 1, v1, v3, v4)]


def box(mn, mx, id_=0x02000000):
    """AddAxisAlignedRectangularSolid (raytracer/environment.go:77-117)."""
    x0, y0, z0 = mn
    x1, y1, z1 = mx
    t = []
    t += quad(id_, (x0, y1, z1), (x1, y1, z1), (x1, y0, z1), (x0, y0, z1))
    t += quad(id_, (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (x0, y0, z0))
    t += quad(id_, (x0, y1, z1), (x0, y1, z0), (x0, y0, z0), (x0, y0, z1))
    t += quad(id_, (x1, y1, z1), (x1, y1, z0), (x1, y0, z0), (x1, y0, z1))
    t += quad(id_, (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0))
    t += quad(id_, (x0, y0, z1), (x1, y0, z1), (x1, y0, z0), (x0, y0, z0))
    return t


SKY = 0x01000000     # TRACE_ID_SKY (raytracer/constants.go:9)
OPAQUE = 0x02000000  # TRACE_ID_OPAQUE
MTL = float(np.float32(1.732050807569 * 32768.0))  # MAX_TRACE_LENGTH as float32


def light(kind, origin=(0, 0, 0), normal=(0, 0, 0), inten=(1.0, 2.0, 3.0)):
    """One light record of the oracle's layout (no fade, no cap)."""
    import oracle
    L = np.zeros(1, oracle.LIGHT_DTYPE)
    L["type"] = kind
    L["origin"] = origin
    L["normal"] = normal
    L["intensity"] = inten
    L["end_fade"] = -1
    L["cap_distance"] = 1e22
    return L


def open_box(lo=0.0, hi=100.0, roof=False, sky_quad=True, open_wall=False):
    """Floor and four walls of the cube [lo, hi]^3 (no top), optionally an
    opaque roof at z = hi, a sky quad far above, and the x = hi wall left out."""
    a, b = lo, hi
    t = []
    t += quad(OPAQUE, (a, a, a), (b, a, a), (b, b, a), (a, b, a))            # floor
    t += quad(OPAQUE, (a, a, a), (a, b, a), (a, b, b), (a, a, b))            # x = lo
    if not open_wall:
        t += quad(OPAQUE, (b, a, a), (b, b, a), (b, b, b), (b, a, b))        # x = hi
    t += quad(OPAQUE, (a, a, a), (b, a, a), (b, a, b), (a, a, b))            # y = lo
    t += quad(OPAQUE, (a, b, a), (b, b, a), (b, b, b), (a, b, b))            # y = hi
    if roof:
        t += quad(OPAQUE, (a, a, b), (b, a, b), (b, b, b), (a, b, b))
    if sky_quad:
        t += quad(SKY, (-1e4, -1e4, 4 * hi), (1e4, -1e4, 4 * hi), (1e4, 1e4, 4 * hi),
                  (-1e4, 1e4, 4 * hi))
    return t


def to_arrays(tris):
    v9 = np.array([[*a, *b, *c] for _, a, b, c in tris], np.float32).reshape(-1, 9)
    ids = np.array([i for i, *_ in tris], np.int32)
    return v9, ids


def square_patches(origin, u, v, normal, n):
    """n x n patches of the parallelogram origin + a u + b v."""
    o = np.asarray(origin, np.float64)
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    pts, nrm, area = [], [], []
    a = np.linalg.norm(np.cross(u, v)) / (n * n)
    for j in range(n):
        for i in range(n):
            pts.append(o + u * (i + 0.5) / n + v * (j + 0.5) / n)
            nrm.append(normal)
            area.append(a)
    return (np.array(pts, np.float32), np.array(nrm, np.float32), np.array(area, np.float32))


def howell_parallel(a, b, c):
    """Parallel, directly opposed rectangles a x b at distance c (Howell C-11)."""
    X, Y = a / c, b / c
    s1, s2 = math.sqrt(1 + X * X), math.sqrt(1 + Y * Y)
    return 2.0 / (math.pi * X * Y) * (
        math.log(math.sqrt((1 + X * X) * (1 + Y * Y) / (1 + X * X + Y * Y)))
        + X * s2 * math.atan(X / s2) + Y * s1 * math.atan(Y / s1)
        - X * math.atan(X) - Y * math.atan(Y))


def howell_perpendicular(w, h, l):
    """Rectangles w x l and h x l at 90 degrees sharing the edge l (Howell C-14),
    from the w-rectangle to the h-rectangle."""
    W, H = w / l, h / l
    A = (1 + W * W) * (1 + H * H) / (1 + W * W + H * H)
    B = (W * W * (1 + W * W + H * H) / ((1 + W * W) * (W * W + H * H))) ** (W * W)
    Cc = (H * H * (1 + H * H + W * W) / ((1 + H * H) * (H * H + W * W))) ** (H * H)
    r = math.sqrt(H * H + W * W)
    return 1.0 / (math.pi * W) * (W * math.atan(1 / W) + H * math.atan(1 / H)
                                  - r * math.atan(1 / r) + 0.25 * math.log(A * B * Cc))


def random_rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def random_bsp(rng, n_clusters=20, max_depth=10, p_leaf=0.15):
    """A random BSP as (planes, nodes, leaf_cluster) structured arrays: half
    the planes axis-aligned at integer distances, half arbitrary; leaves take
    clusters in [-2, n_clusters) (-1 the most common negative)."""
    import oracle
    planes, nodes, leaves = [], [], []

    def leaf():
        u = rng.uniform()
        leaves.append(-1 if u < 0.25 else (-2 if u < 0.3 else int(rng.integers(0, n_clusters))))
        return -1 - (len(leaves) - 1)

    def build(depth):
        if depth > 1 and (depth >= max_depth or rng.uniform() < p_leaf):
            return leaf()
        if rng.uniform() < 0.5:
            n = np.zeros(3, np.float32)
            n[int(rng.integers(0, 3))] = rng.choice([-1.0, 1.0])
            d = float(rng.integers(-40, 41))
        else:
            n = rng.normal(size=3)
            n = (n / np.linalg.norm(n)).astype(np.float32)
            d = float(np.float32(rng.uniform(-40, 40)))
        planes.append((n, d))
        k = len(nodes)
        nodes.append([len(planes) - 1, 0, 0])
        nodes[k][1] = build(depth + 1)
        nodes[k][2] = build(depth + 1)
        return k

    build(1)
    pl = np.zeros(len(planes), oracle.PLANE_DTYPE)
    for k, (n, d) in enumerate(planes):
        pl[k] = (n, d)
    nd = np.zeros(len(nodes), oracle.NODE_DTYPE)
    for k, (p, c0, c1) in enumerate(nodes):
        nd[k] = (p, (c0, c1))
    return pl, nd, np.array(leaves, np.int16)


def bsp_points(rng, planes, n):
    """Points in [-50, 50]^3, a third of them moved onto a random plane or to
    within / just past TEST_EPSILON of it."""
    p = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    for k in range(0, n, 3):
        pl = planes[int(rng.integers(0, len(planes)))]
        nrm = pl["normal"].astype(np.float64)
        off = rng.choice([0.0, 0.03125, -0.03125, 0.02, -0.03, 0.0313, -0.0313])
        q = p[k].astype(np.float64)
        q = q - (q @ nrm - pl["dist"] - off) * nrm
        if np.count_nonzero(nrm) == 1:  # axis plane: exact coordinate
            ax = int(np.flatnonzero(nrm)[0])
            q[ax] = (pl["dist"] + off) * nrm[ax]
        p[k] = q.astype(np.float32)
    return p
