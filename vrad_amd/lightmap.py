"""rad/lightmap on the MI355X path: GatherSampleLight and BounceLight, plus
the light-parameter helpers they consume (rad/lightmap/lights.go)."""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, fptr, f32, LIGHT_DTYPE

DIRECT_SCALE = 100.0 * 100.0  # rad/lightmap/lights.go:23
EQUAL_EPSILON = 0.001         # vmath/constants.go:9
NUM_BOUNCE_DEFAULT = 2        # rad/patches/subdivide.go:19


def make_lights(n):
    a = np.zeros(n, LIGHT_DTYPE)
    a["end_fade"] = -1.0
    a["cap_distance"] = 1.0e22
    return a


def point_light(origin, intensity, constant=0.0, linear=0.0, quadratic=0.0):
    """ParseLightPoint + SetLightFalloffParams without _fifty_percent_distance
    (rad/lightmap/lights.go:311-340): attenuations below EQUAL_EPSILON clamp to
    0, all-zero becomes constant 1, intensity scales by c + 100 l + 100^2 q."""
    L = make_lights(1)
    L["type"] = _lib.VR_EMIT_POINT
    L["origin"] = origin
    c, l, q = [0.0 if v < EQUAL_EPSILON else v for v in (constant, linear, quadratic)]
    if c < EQUAL_EPSILON and l < EQUAL_EPSILON and q < EQUAL_EPSILON:
        c = 1.0
    L["constant_attn"], L["linear_attn"], L["quadratic_attn"] = c, l, q
    ratio = np.float32(c) + np.float32(100) * np.float32(l) + np.float32(10000) * np.float32(q)
    inten = f32(intensity, 3)
    if ratio > 0:
        inten = inten * np.float32(ratio)
    L["intensity"] = inten
    return L


def spot_light(origin, direction, intensity, inner_deg, outer_deg, exponent=0.0,
               constant=0.0, linear=0.0, quadratic=0.0):
    """ParseLightSpot (rad/lightmap/lights.go:210-257): cone angles clamp to 90,
    stored as cosines; 180/180 degenerates to a point light."""
    L = point_light(origin, intensity, constant, linear, quadratic)
    inner = inner_deg if inner_deg != 0 else 10.0
    outer = outer_deg if outer_deg != 0 else inner
    outer = max(outer, inner)
    if inner == 180 and outer == 180:
        return L
    inner, outer = min(inner, 90.0), min(outer, 90.0)
    L["type"] = _lib.VR_EMIT_SPOTLIGHT
    d = f32(direction, 3)
    L["normal"] = d / np.float32(np.linalg.norm(d))
    L["stopdot"] = math.cos(inner / 180.0 * math.pi)
    L["stopdot2"] = math.cos(outer / 180.0 * math.pi)
    L["exponent"] = exponent
    return L


def sun_light(direction, intensity):
    L = make_lights(1)
    L["type"] = _lib.VR_EMIT_SKYLIGHT
    d = np.asarray(direction, np.float64)
    L["normal"] = d / np.linalg.norm(d)
    L["intensity"] = intensity
    return L


def solve_inverse_quadratic(x1, y1, x2, y2, x3, y3):
    """vmath/quadratic/solver.go:48-63: a, b, c with a x^2 + b x + c = y."""
    f = np.float32
    x1, y1, x2, y2, x3, y3 = map(f, (x1, y1, x2, y2, x3, y3))
    det = (x1 - x2) * (x1 - x3) * (x2 - x3)
    if det == 0:
        return None
    a = (x3 * (-y1 + y2) + x2 * (y1 - y3) + x1 * (-y2 + y3)) / det
    b = (x3 * x3 * (y1 - y2) + x1 * x1 * (y2 - y3) + x2 * x2 * (-y1 + y3)) / det
    c = (x1 * x3 * (-x1 + x3) * y2 + x2 * x2 * (x3 * y1 - x1 * y3) +
         x2 * (-(x3 * x3 * y1) + x1 * x1 * y3)) / det
    return f(a), f(b), f(c)


class _Samples:
    def __init__(self, pos, normal):
        p = f32(pos).reshape(-1, 3)
        nrm = f32(normal).reshape(-1, 3)
        self.cols = [np.ascontiguousarray(p[:, k]) for k in range(3)] + \
                    [np.ascontiguousarray(nrm[:, k]) for k in range(3)]
        self.s = _lib.Samples(*[fptr(c) for c in self.cols])
        self.n = p.shape[0]


def GatherSampleLight(env, positions, normals, lights, normal_offset=0.1, skip_id=-1):
    """Direct light at each sample: rgb (n,3) float32."""
    s = _Samples(positions, normals)
    L = np.ascontiguousarray(lights, LIGHT_DTYPE)
    rgb = np.zeros((s.n, 3), np.float32)
    check(lib.vr_gather_sample_light(env.handle, C.byref(s.s), s.n, ptr(L), len(L),
                                     float(normal_offset), int(skip_id), ptr(rgb)))
    return rgb


def BounceLight(transfers, emit0, refl, area, total=None, max_bounces=100, stop_frac=1e-3):
    """Jacobi bounce.  Returns (total (n,3), energy per bounce (k,), k)."""
    e0 = f32(emit0).reshape(-1, 3)
    n = e0.shape[0]
    rf = f32(refl).reshape(n, 3)
    ar = f32(area).reshape(n)
    tot = np.zeros((n, 3), np.float32) if total is None else f32(total).reshape(n, 3).copy()
    energy = np.zeros(max(max_bounces, 1), np.float64)
    nb = C.c_int32(0)
    prm = _lib.BounceParams(int(max_bounces), float(stop_frac))
    check(lib.vr_bounce_light(transfers.handle, ptr(e0), ptr(rf), ptr(ar), ptr(tot),
                              C.byref(prm), ptr(energy), C.byref(nb)))
    return tot, energy[:nb.value], nb.value


def BounceLightBatch(jobs, max_bounces=100, stop_frac=1e-3, use_graph=True):
    """Batched BounceLight over independent transfer sets in one launch per
    bounce (optionally a replayed hipGraph of 8 bounces).  `jobs` is a list of
    (transfers, emit0, refl, area, total_or_None); returns a list of
    (total (n,3), energy per bounce (k,), k), equal per job to BounceLight."""
    keep, arr = [], (_lib.BounceJob * max(len(jobs), 1))()
    for i, (xf, emit0, refl, area, total) in enumerate(jobs):
        e0 = f32(emit0).reshape(-1, 3)
        n = e0.shape[0]
        rf = f32(refl).reshape(n, 3)
        ar = f32(area).reshape(n)
        tot = np.zeros((n, 3), np.float32) if total is None else f32(total).reshape(n, 3).copy()
        energy = np.zeros(max(max_bounces, 1), np.float64)
        keep.append((e0, rf, ar, tot, energy))
        h = xf.handle
        arr[i] = _lib.BounceJob(h.value if isinstance(h, C.c_void_p) else h, _lib.fptr(e0), _lib.fptr(rf), _lib.fptr(ar),
                                _lib.fptr(tot), energy.ctypes.data_as(C.POINTER(C.c_double)), 0)
    prm = _lib.BounceParams(int(max_bounces), float(stop_frac))
    check(lib.vr_bounce_light_batch(arr, len(jobs), C.byref(prm), 1 if use_graph else 0))
    return [(k[3], k[4][:arr[i].n_bounces], arr[i].n_bounces) for i, k in enumerate(keep)]
