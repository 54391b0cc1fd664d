"""raytracer/trace on the MI355X path: TestLineDoesHitSky
(raytracer/trace/testline.go:18-94) for a FourVectors packet, and the batched
form.  The 3D-skybox recursion (:60-89) and texture shadows (:53-55) are off
by default in the reference and are not provided."""
import numpy as np

from ._lib import f32
from .raytracer import GetEnvironment, TRACE_ID_STATICPROP


def TestLineDoesHitSky(start, stop, canRecurse=False, staticPropToSkip=0, doDebug=False,
                       env=None):
    """start/stop: FourVectors as (3, 4) arrays.  Returns fractionVisible (4,)
    = 1 - occlusion, occlusion = closest hit in [0, len) is not sky."""
    env = env or GetEnvironment()
    s = f32(start, (3, 4)).T
    e = f32(stop, (3, 4)).T
    vis, _ = env.TestLines(s, e, skip_id=TRACE_ID_STATICPROP | int(staticPropToSkip))
    return vis.astype(np.float32)


def TestLines(env, starts, ends, skip_id=-1):
    return env.TestLines(starts, ends, skip_id)
