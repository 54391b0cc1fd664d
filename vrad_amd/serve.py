"""Re-bake serving against resident scenes (no reference counterpart).  A
Server keeps each scene's kd-tree, patches and transfers on the device; a
request moves some of the scene's lights, re-gathers the direct light and
runs a fixed number of bounces.  The batched mode admits requests by free
slots and device memory and runs one bounce of every admitted request per
launch (continuous batching); the plain mode runs one request at a time on
the plain kernels."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, f32, lib, ptr

BATCHED, PLAIN = _lib.VR_SERVE_BATCHED, _lib.VR_SERVE_PLAIN


class Server:
    def __init__(self, mode=BATCHED, max_active=8, hbm_budget=0, bounces=8, normal_offset=0.1):
        self._h = C.c_void_p()
        prm = _lib.ServerParams(int(mode), int(max_active), int(hbm_budget), int(bounces),
                                float(normal_offset))
        check(lib.vr_server_create(C.byref(self._h), C.byref(prm)))
        self.bounces = int(bounces)
        self._n = []
        self._keep = []

    def add_scene(self, env, transfers, positions, normals, refl, area, lights):
        """Register a built Environment with its Transfers, patch sample points
        and normals (n,3), reflectivity (n,3), area (n,) and light records."""
        pos = f32(positions).reshape(-1, 3)
        nrm = f32(normals).reshape(-1, 3)
        cols = [np.ascontiguousarray(pos[:, k]) for k in range(3)] + \
               [np.ascontiguousarray(nrm[:, k]) for k in range(3)]
        smp = _lib.Samples(*[_lib.fptr(c) for c in cols])
        rf = f32(refl).reshape(-1, 3)
        ar = f32(area).reshape(-1)
        lt = np.ascontiguousarray(lights)
        sid = C.c_int32()
        check(lib.vr_server_add_scene(self._h, env.handle, transfers.handle, C.byref(smp),
                                      ptr(rf), ptr(ar), ptr(lt), len(lt), C.byref(sid)))
        self._n.append(len(ar))
        return sid.value

    def submit(self, scene_id, moves=()):
        """moves: iterable of (light index, (dx, dy, dz)).  Returns a ticket."""
        idx = np.array([m[0] for m in moves], np.int32)
        off = np.array([m[1] for m in moves], np.float32).reshape(-1, 3)
        t = C.c_int64()
        check(lib.vr_server_submit(self._h, int(scene_id), ptr(idx), ptr(off), len(idx),
                                   C.byref(t)))
        return t.value, int(scene_id)

    def wait(self, ticket, timeout_ms=-1):
        """(total (n,3), energy per bounce, latency ms) of a submitted request."""
        t, sid = ticket
        tot = np.zeros((self._n[sid], 3), np.float32)
        en = np.zeros(max(self.bounces, 1), np.float64)
        nb = C.c_int32()
        lat = C.c_double()
        check(lib.vr_server_wait(self._h, t, int(timeout_ms), ptr(tot), ptr(en), C.byref(nb),
                                 C.byref(lat)))
        return tot, en[:nb.value], lat.value

    def stats(self):
        s = _lib.ServerStats()
        check(lib.vr_server_get_stats(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in s._fields_}

    def close(self):
        if self._h:
            lib.vr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
