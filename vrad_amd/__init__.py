"""MI355X-native radiosity bake path (VRAD's raytracer / transfers / bounce)
over hand-written gfx950 HIP kernels behind the C-ABI in include/vrad_hip.h."""
from . import _lib
from ._lib import VrError
from .raytracer import Environment, GetEnvironment, FourRays, RayTracingResult
from .trace import TestLineDoesHitSky
from .patches import MakeTransfers, Transfers
from .lightmap import GatherSampleLight, BounceLight

__all__ = ["Environment", "GetEnvironment", "FourRays", "RayTracingResult",
           "TestLineDoesHitSky", "MakeTransfers", "Transfers", "GatherSampleLight",
           "BounceLight", "VrError"]
