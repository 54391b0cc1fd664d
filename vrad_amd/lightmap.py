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


def _falloff_fifty_percent(d50, d0, hardfalloff):
    """The _fifty_percent_distance branch of SetLightFalloffParams
    (rad/lightmap/lights.go:262-310), in float32 as the Go code computes it:
    fit 1/(c + b d + a d^2) through 1 at 0, 1/2 at d50 and 1/256 at d0 with
    SolveInverseQuadraticMonotonic, rescale so the d50 value is exact, then
    set the fade (hard falloff) or the cap where the quadratic turns.
    Returns (c, l, q, start_fade, end_fade, cap_distance)."""
    f = np.float32
    d50, d0 = f(d50), f(d0)
    start, end, cap = f(0.0), f(-1.0), f(1.0e22)
    if d0 < d50:
        d0 = f(2.0) * d50
    a, b, c = f(0.0), f(1.0), f(0.0)
    sol = solve_inverse_quadratic_monotonic(0.0, 1.0, d50, 2.0, d0, 256.0)
    if sol is not None:
        a, b, c = sol
    v50 = c + d50 * (b + d50 * a)
    scale = f(2.0) / v50
    a, b, c = a * scale, b * scale, c * scale
    if hardfalloff:
        end = d0
        start = f(0.75) * d0 + f(0.25) * d50
    elif abs(float(a)) > 0.0:
        fl_max = b / (f(-2.0) * a)
        if fl_max > 0.0:
            cap, start, end = fl_max, fl_max, f(10.0) * fl_max
    return c, b, a, start, end, cap


def point_light(origin, intensity, constant=0.0, linear=0.0, quadratic=0.0,
                fifty_percent_distance=0.0, zero_percent_distance=0.0, hardfalloff=False):
    """ParseLightPoint + SetLightFalloffParams (rad/lightmap/lights.go:259-341,
    417-426).  With fifty_percent_distance == 0: attenuations below
    EQUAL_EPSILON clamp to 0, all-zero becomes constant 1, and intensity
    scales by c + 100 l + 100^2 q.  Otherwise the attenuations, fade and cap
    come from the fifty/zero percent distances and intensity is unscaled."""
    L = make_lights(1)
    L["type"] = _lib.VR_EMIT_POINT
    L["origin"] = origin
    if fifty_percent_distance != 0:
        c, l, q, start, end, cap = _falloff_fifty_percent(
            fifty_percent_distance, zero_percent_distance, hardfalloff)
        L["constant_attn"], L["linear_attn"], L["quadratic_attn"] = c, l, q
        L["start_fade"], L["end_fade"], L["cap_distance"] = start, end, cap
        L["intensity"] = f32(intensity, 3)
        return L
    c, l, q = [np.float32(v) for v in (constant, linear, quadratic)]
    c, l, q = [np.float32(0.0) if v < np.float32(EQUAL_EPSILON) else v for v in (c, l, q)]
    if c < np.float32(EQUAL_EPSILON) and l < np.float32(EQUAL_EPSILON) and \
            q < np.float32(EQUAL_EPSILON):
        c = np.float32(1.0)
    L["constant_attn"], L["linear_attn"], L["quadratic_attn"] = c, l, q
    ratio = c + np.float32(100) * l + np.float32(10000) * q
    inten = f32(intensity, 3)
    if ratio > 0:
        inten = inten * np.float32(ratio)
    L["intensity"] = inten
    return L


def spot_cosine(deg):
    """float32(math.Cos(float64(deg / 180 * vmath.PI))) with deg float32 and
    the division and product in float32 (rad/lightmap/lights.go:249-250;
    vmath.PI, an untyped constant, becomes float32 there)."""
    x = np.float32(deg) / np.float32(180.0) * np.float32(math.pi)
    return np.float32(math.cos(float(x)))


def spot_light(origin, direction, intensity, inner_deg, outer_deg, exponent=0.0,
               constant=0.0, linear=0.0, quadratic=0.0, fifty_percent_distance=0.0,
               zero_percent_distance=0.0, hardfalloff=False):
    """ParseLightSpot (rad/lightmap/lights.go:210-257): cone angles clamp to 90,
    stored as float32 cosines; 180/180 degenerates to a point light."""
    L = point_light(origin, intensity, constant, linear, quadratic, fifty_percent_distance,
                    zero_percent_distance, hardfalloff)
    inner = np.float32(inner_deg) if inner_deg != 0 else np.float32(10.0)
    outer = np.float32(outer_deg) if outer_deg != 0 else inner
    outer = max(outer, inner)
    if inner == 180 and outer == 180:
        return L
    inner, outer = min(inner, np.float32(90.0)), min(outer, np.float32(90.0))
    L["type"] = _lib.VR_EMIT_SPOTLIGHT
    d = f32(direction, 3)
    L["normal"] = d / np.float32(np.linalg.norm(d))
    L["stopdot"] = spot_cosine(inner)
    L["stopdot2"] = spot_cosine(outer)
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
    """vmath/quadratic/solver.go:48-63: a, b, c with a x^2 + b x + c = y, in
    float32 with Go's left-to-right evaluation; None when det == 0."""
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


def _flerp(f1, f2, i1, i2, x):
    """vmath/quadratic/lerp.go:14-16, float32 arithmetic widened at the end."""
    f = np.float32
    f1, f2, i1, i2, x = map(f, (f1, f2, i1, i2, x))
    return float(f1 + (f2 - f1) * (x - i1) / (i2 - i1))


def solve_inverse_quadratic_monotonic(x1, y1, x2, y2, x3, y3):
    """vmath/quadratic/solver.go:4-45.  The points are sorted by x first (the
    intent: the Go vSwap at :65-69 swaps only its local pointers).  While the
    fit's derivative term 2a + b has the wrong sign for monotonic y, the middle
    y is blended toward the chord in steps of 0.05 (float64 counter).
    Returns (a, b, c) or None when the points do not determine a quadratic."""
    f = np.float32
    x1, y1, x2, y2, x3, y3 = map(f, (x1, y1, x2, y2, x3, y3))
    if x1 > x2:
        x1, x2, y1, y2 = x2, x1, y2, y1
    if x2 > x3:
        x2, x3, y2, y3 = x3, x2, y3, y2
    if x1 > x2:
        x1, x2, y1, y2 = x2, x1, y2, y1
    sol = None
    blend = 0.0
    while blend <= 1.0:
        tempy2 = (1.0 - blend) * float(y2) + blend * _flerp(y1, y3, x1, x3, x2)
        sol = solve_inverse_quadratic(x1, y1, x2, f(tempy2), x3, y3)
        if sol is None:
            return None
        a, b, c = sol
        deriv = f(2.0) * a + b
        if y1 < y2 < y3:
            if deriv >= 0.0:
                return sol
        elif y1 > y2 > y3:
            if deriv <= 0.0:
                return sol
        else:
            return sol
        blend += 0.05
    return sol


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


def GatherSampleLightBatch(items, normal_offset=0.1, skip_id=-1):
    """GatherSampleLight for many scenes in one launch: items is a list of
    (env, positions, normals, lights); returns one rgb (n,3) per item, each
    equal to GatherSampleLight on that item alone."""
    k = len(items)
    if k == 0:
        return []
    smp, keep, outs = (_lib.Samples * k)(), [], []
    hs = (C.c_void_p * k)()
    lp = (C.c_void_p * k)()
    rp = (C.c_void_p * k)()
    n = np.zeros(k, np.int64)
    nl = np.zeros(k, np.int32)
    for q, (env, pos, nrm, lights) in enumerate(items):
        s = _Samples(pos, nrm)
        L = np.ascontiguousarray(lights, LIGHT_DTYPE)
        rgb = np.zeros((s.n, 3), np.float32)
        keep.append((s, L))
        outs.append(rgb)
        smp[q] = s.s
        n[q], nl[q] = s.n, len(L)
        hs[q] = env.handle.value if isinstance(env.handle, C.c_void_p) else env.handle
        lp[q] = L.ctypes.data if len(L) else None
        rp[q] = rgb.ctypes.data
    check(lib.vr_gather_sample_light_batch(hs, smp, ptr(n), lp, ptr(nl), k, float(normal_offset),
                                           int(skip_id), rp))
    return outs


def CanLeafTraceToSky(env, mins, maxs):
    """CanLeafTraceToSky (rad/lightmap/lightmap.go:425-451) for many leaves at
    once: mins, maxs are (n, 3) int16 leaf bounds; returns bool (n,)."""
    mn = np.ascontiguousarray(mins, np.int16).reshape(-1, 3)
    mx = np.ascontiguousarray(maxs, np.int16).reshape(-1, 3)
    if mn.shape != mx.shape:
        raise ValueError("mins and maxs differ in shape")
    out = np.zeros(len(mn), np.uint8)
    check(lib.vr_can_leaf_trace_to_sky(env.handle, ptr(mn), ptr(mx), len(mn), ptr(out)))
    return out.astype(bool)


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
        arr[i] = _lib.BounceJob(h.value if isinstance(h, C.c_void_p) else h, _lib.fptr(e0),
                                _lib.fptr(rf), _lib.fptr(ar), _lib.fptr(tot),
                                energy.ctypes.data_as(C.POINTER(C.c_double)), 0)
    prm = _lib.BounceParams(int(max_bounces), float(stop_frac))
    check(lib.vr_bounce_light_batch(arr, len(jobs), C.byref(prm), 1 if use_graph else 0))
    return [(k[3], k[4][:arr[i].n_bounces], arr[i].n_bounces) for i, k in enumerate(keep)]
