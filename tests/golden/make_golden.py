"""Writes the hand-derived golden fixtures used by the oracle and GPU tests.

Nothing here calls the oracle or the product: every expected value is worked
out below from the geometry (exactly representable inputs), so the fixtures
pin the implementations instead of restating them.

  tri_precompute.json  ChangeIntoIntersectionFormat outputs
                       (raytracer/cache/optimisedtriangle.go:30-78,
                       vmath/polygon/edge.go:5-25) as float bit patterns, so
                       signed zeros are pinned too.
  analytic_rays.json   rays inside the 12 triangles of
                       AddAxisAlignedRectangularSolid((0,0,0),(1024,1024,256))
                       (raytracer/environment.go:77-117) with the exact hit
                       triangle and distance.

  tri_precompute_general.json
                       the same outputs for triangles in general position,
                       where every operation rounds.  The expected bits come
                       from an exact-rational emulation of IEEE float32
                       (round to nearest, ties to even) applied to mgl32's
                       published vector operations at
                       go-gl/mathgl@5ab0e04e1f55a10f3a71b1fccc60ca2ee9a0bc02:
                       Sub (componentwise), Cross (v1[1]*v2[2] - v1[2]*v2[1],
                       ...), Len = float32(math.Sqrt(float64(Dot(v, v)))),
                       Normalize = v * (1 / Len), Dot = x*x' + y*y' + z*z'
                       (left to right), with no fused multiply-add (Go on
                       amd64 at its default GOAMD64=v1).  The triangles are
                       chosen so that each plausible mistake changes at least
                       one bit: dividing by Len instead of multiplying by its
                       reciprocal, summing Len's squares in float64, another
                       association of Dot or of an edge sum, or fusing a
                       multiply-add.

Run:  python tests/golden/make_golden.py
"""
import json
import math
import os
import struct
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


NEG0 = -0.0

# Each entry: vertices, id, and the expected record worked out by hand.
#  N = normalize((p2-p1) x (p3-p1)); drop = argmax |N| (strict >, lowest wins)
#  c0 = (drop+1)%3, c1 = (drop+2)%3; D = N.p1
#  edge(P,Q;I): nx = P[c1]-Q[c1], ny = Q[c0]-P[c0], d = -(nx P[c0] + ny P[c1]),
#               tr = I[c0] nx + I[c1] ny + d, negate all if tr < 0, divide by tr
TRIS = [
    {  # unit right triangle in z=0: E1 = (0, 1, -0) since d = -(0*0 + 1*0) = -0
        "v": [0, 0, 0, 1, 0, 0, 0, 1, 0], "id": 0x02000000,
        "N": [0.0, 0.0, 1.0], "D": 0.0, "c0": 0, "c1": 1,
        "E": [0.0, 1.0, NEG0, -1.0, -1.0, 1.0],
    },
    {  # x=0 plane, legs 2: N = (4,0,0)/4; E1 = (0, 2, -0)/4, E2 = (-2,-2,4)/4
        "v": [0, 0, 0, 0, 2, 0, 0, 0, 2], "id": 0x02000001,
        "N": [1.0, 0.0, 0.0], "D": 0.0, "c0": 1, "c1": 2,
        "E": [0.0, 0.5, NEG0, -0.5, -0.5, 1.0],
    },
    {  # z=3, clockwise: N = (0,0,-16)/16, D = -3; E1 flipped: (4, -0, +0)/16,
       # E2 flipped: (-4, -4, 16)/16
        "v": [0, 0, 3, 0, 4, 3, 4, 0, 3], "id": 0x01000000,
        "N": [0.0, 0.0, -1.0], "D": -3.0, "c0": 0, "c1": 1,
        "E": [0.25, NEG0, 0.0, -0.25, -0.25, 1.0],
    },
    {  # y=8 plane: e1=(0,0,2), e2=(2,0,0): N = (0*0-2*0, 2*2-0*0, 0*0-0*2) = (0,4,0)/4
       # drop=1, c0=2, c1=0, D = 8.  E1 = edge(p1,p2;p3) on (z,x):
       # nx = p1.x-p2.x = 0, ny = p2.z-p1.z = 2, d = -(0*0 + 2*0) = -0,
       # tr = p3.z*0 + p3.x*2 - 0 = 4 -> (0, 0.5, -0)
       # E2 = edge(p2,p3;p1): nx = p2.x-p3.x = -2, ny = p3.z-p2.z = -2,
       # d = -((-2)*2 + (-2)*0) = 4, tr = 0*(-2) + 0*(-2) + 4 = 4 -> (-0.5,-0.5,1)
        "v": [0, 8, 0, 0, 8, 2, 2, 8, 0], "id": 7,
        "N": [0.0, 1.0, 0.0], "D": 8.0, "c0": 2, "c1": 0,
        "E": [0.0, 0.5, NEG0, -0.5, -0.5, 1.0],
    },
]


def box_rays():
    """Rays inside the box (0,0,0)-(1024,1024,256).  Triangle order of
    AddAxisAlignedRectangularSolid: quads far(z=max) 0-1, near(z=min) 2-3,
    left(x=min) 4-5, right(x=max) 6-7, top(y=max) 8-9, bottom(y=min) 10-11;
    in each quad the first triangle is (v1,v2,v3), the second (v1,v3,v4)."""
    rays = []
    # +z from (100,200,10): far face, x+y=300 < 1024 -> (v1,v3,v4) = tri 1, t=246
    rays.append({"o": [100, 200, 10], "d": [0, 0, 1], "hit": 1, "t": 246.0})
    # +z from (900,900,10): x+y = 1800 > 1024 -> tri 0
    rays.append({"o": [900, 900, 10], "d": [0, 0, 1], "hit": 0, "t": 246.0})
    # -z from (100,200,250): near face z=0; same split as the far face -> tri 3, t=250
    rays.append({"o": [100, 200, 250], "d": [0, 0, -1], "hit": 3, "t": 250.0})
    # +x from (512,512,128): right face at x=1024, exactly on the diagonal
    # v1=(1024,1024,256)-v3=(1024,0,0): both 6 and 7 hit at t=512 -> lower index 6
    rays.append({"o": [512, 512, 128], "d": [1, 0, 0], "hit": 6, "t": 512.0})
    # -x from (512,1000,20): left face x=0, vertices v1=(0,1024,256),
    # v2=(0,1024,0), v3=(0,0,0), v4=(0,0,256); the diagonal v1-v3 is z = y/4;
    # z=20 < 1000/4 -> below the diagonal, i.e. on the side of v2 -> tri 4
    rays.append({"o": [512, 1000, 20], "d": [-1, 0, 0], "hit": 4, "t": 512.0})
    # +y from (10,10,10): top face y=1024; v1=(0,1024,256) v2=(1024,1024,256)
    # v3=(1024,1024,0) v4=(0,1024,0); diagonal v1-v3: z = 256 - x/4;
    # z=10 < 256-2.5 -> side of v4 -> tri 9, t=1014
    rays.append({"o": [10, 10, 10], "d": [0, 1, 0], "hit": 9, "t": 1014.0})
    # tmax = 246 exactly: hits at t == tmax are excluded -> miss
    rays.append({"o": [100, 200, 10], "d": [0, 0, 1], "tmax": 246.0, "hit": -1, "t": 0.0})
    # tmin beyond every wall -> miss
    rays.append({"o": [100, 200, 10], "d": [0, 0, 1], "tmin": 2000.0, "hit": -1, "t": 0.0})
    return rays


# ---- exact-rational IEEE float32 / float64 emulation --------------------------
def _round_bin(q, mant_bits, emin):
    """Nearest binary float to the rational q (ties to even) with mant_bits
    significand bits and minimum normal exponent emin; returns a Fraction
    (infinite results do not occur for the inputs used here)."""
    if q == 0:
        return Fraction(0)
    sgn = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, emin)
    ulp = Fraction(2) ** (e - (mant_bits - 1))
    m = a / ulp
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sgn * fl * ulp


class F32:
    """A float32 value with IEEE signed zero; arithmetic rounds exactly once."""

    def __init__(self, q, neg_zero=False):
        self.q = Fraction(q)
        self.nz = bool(neg_zero) and self.q == 0

    @staticmethod
    def of(x):
        x = struct.unpack("<f", struct.pack("<f", x))[0]
        return F32(Fraction(x), math.copysign(1.0, x) < 0 and x == 0)

    def sign(self):
        return -1 if (self.q < 0 or self.nz) else 1

    def _mk(self, q, zero_sign):
        r = _round_bin(q, 24, -126)
        return F32(r, r == 0 and zero_sign < 0)

    def __add__(self, o):
        q = self.q + o.q
        zs = -1 if (q == 0 and self.q == 0 and o.q == 0 and self.nz and o.nz) else 1
        return self._mk(q, zs)

    def __neg__(self):
        return F32(-self.q, self.q == 0 and not self.nz)

    def __sub__(self, o):
        return self + (-o)

    def __mul__(self, o):
        return self._mk(self.q * o.q, self.sign() * o.sign())

    def __truediv__(self, o):
        return self._mk(self.q / o.q, self.sign() * o.sign())

    def __lt__(self, o):
        return self.q < o.q

    def value(self):
        return -0.0 if self.nz else float(self.q)

    def bits(self):
        return bits(self.value())


def _sqrt64(q):
    """Correctly rounded float64 square root of the rational q >= 0."""
    if q == 0:
        return Fraction(0)
    # scale by 4^k so the integer square root carries > 60 significant bits
    k = 0
    while q * Fraction(4) ** k < Fraction(2) ** 130:
        k += 1
    while q * Fraction(4) ** k > Fraction(2) ** 140:
        k -= 1
    x = q * Fraction(4) ** k
    r = math.isqrt(x.numerator // x.denominator)
    exact = Fraction(r * r) == x
    s = Fraction(r) / Fraction(2) ** k
    # r is floor(sqrt(x)) (to within the truncated fraction); the tail below
    # one unit in 2^-k is far below float64's ulp, so nudging an inexact root
    # up by a quarter unit places it correctly for rounding
    if not exact:
        s += Fraction(1, 4) / Fraction(2) ** k
    return _round_bin(s, 53, -1022)


def mgl_precompute(v, fused=False, recip=True, len2_f64=False, dot_assoc_left=True,
                   edge_assoc_left=True):
    """ChangeIntoIntersectionFormat with mgl32's operations in F32 arithmetic.
    The keyword switches make the mistaken variants the fixture must reject."""
    P = [[F32.of(v[3 * i + c]) for c in range(3)] for i in range(3)]
    p1, p2, p3 = P
    e1 = [p2[c] - p1[c] for c in range(3)]
    e2 = [p3[c] - p1[c] for c in range(3)]

    def fms(a, b, c, d):  # a*b - c*d
        if fused:
            return F32(0)._mk(a.q * b.q - (c * d).q, 1)
        return a * b - c * d
    N = [fms(e1[1], e2[2], e1[2], e2[1]), fms(e1[2], e2[0], e1[0], e2[2]),
         fms(e1[0], e2[1], e1[1], e2[0])]
    if len2_f64:
        l2 = _round_bin(_round_bin(N[0].q ** 2, 53, -1022) + _round_bin(N[1].q ** 2, 53, -1022),
                        53, -1022)
        l2 = _round_bin(l2 + _round_bin(N[2].q ** 2, 53, -1022), 53, -1022)
    else:
        l2 = ((N[0] * N[0] + N[1] * N[1]) + N[2] * N[2]).q
    ln = F32(_round_bin(_sqrt64(l2), 24, -126))
    if recip:
        inv = F32(1) / ln
        N = [N[c] * inv for c in range(3)]
    else:
        N = [N[c] / ln for c in range(3)]
    drop = 0
    for c in (1, 2):
        if abs(N[c].q) > abs(N[drop].q):
            drop = c
    t = [N[c] * p1[c] for c in range(3)]
    D = (t[0] + t[1]) + t[2] if dot_assoc_left else t[0] + (t[1] + t[2])
    c0, c1 = (drop + 1) % 3, (drop + 2) % 3

    def edge(a, b, ins):
        nx = a[c1] - b[c1]
        ny = b[c0] - a[c0]
        d = -(nx * a[c0] + ny * a[c1])
        if edge_assoc_left:
            tr = (ins[c0] * nx + ins[c1] * ny) + d
        else:
            tr = ins[c0] * nx + (ins[c1] * ny + d)
        if tr < F32(0):
            nx, ny, d, tr = -nx, -ny, -d, -tr
        return [nx / tr, ny / tr, d / tr]
    E = edge(p1, p2, p3) + edge(p2, p3, p1)
    return {"N_bits": [x.bits() for x in N], "D_bits": D.bits(), "c0": c0, "c1": c1,
            "E_bits": [x.bits() for x in E]}


VARIANTS = {"fused multiply-add in Cross": {"fused": True},
            "N / Len instead of N * (1 / Len)": {"recip": False},
            "Len's squares summed in float64": {"len2_f64": True},
            "Dot associated right": {"dot_assoc_left": False},
            "edge trial sum associated right": {"edge_assoc_left": False}}


def general_triangles():
    """12 seeded float32 triangles in general position: for each mistaken
    variant the first triangle it gets wrong, then the next triangles that
    some variant gets wrong."""
    import random
    rng = random.Random(20261203)
    cands = []
    for _ in range(2000):
        v = [struct.unpack("<f", struct.pack("<f", rng.uniform(-700.0, 700.0)))[0]
             for _ in range(9)]
        ref = mgl_precompute(v)
        new = {name for name, kw in VARIANTS.items() if mgl_precompute(v, **kw) != ref}
        if new:
            cands.append((v, ref, new))
        if len(cands) >= 12 and set().union(*(c[2] for c in cands)) == set(VARIANTS):
            break
    pick = []
    for name in VARIANTS:
        k = next(i for i, c in enumerate(cands) if name in c[2])
        if k not in pick:
            pick.append(k)
    pick += [i for i in range(len(cands)) if i not in pick][:12 - len(pick)]
    out = []
    for i in sorted(pick):
        v, ref, new = cands[i]
        out.append({"v": v, "id": 0x02000000 | len(out), **ref, "catches": sorted(new)})
    return out


def main():
    with open(os.path.join(HERE, "tri_precompute_general.json"), "w") as f:
        json.dump({"source": "exact-rational float32 emulation of mgl32, see make_golden.py",
                   "triangles": general_triangles()}, f, indent=1)
    tri = []
    for t in TRIS:
        rec = dict(t)
        rec["N_bits"] = [bits(x) for x in t["N"]]
        rec["D_bits"] = bits(t["D"])
        rec["E_bits"] = [bits(x) for x in t["E"]]
        tri.append(rec)
    with open(os.path.join(HERE, "tri_precompute.json"), "w") as f:
        json.dump({"source": "hand-derived, see make_golden.py", "triangles": tri}, f, indent=1)
    with open(os.path.join(HERE, "analytic_rays.json"), "w") as f:
        json.dump({"box_min": [0, 0, 0], "box_max": [1024, 1024, 256], "id": 0x02000000,
                   "rays": box_rays()}, f, indent=1)


if __name__ == "__main__":
    main()
