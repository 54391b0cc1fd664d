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

Run:  python tests/golden/make_golden.py
"""
import json
import os
import struct

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


def main():
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
