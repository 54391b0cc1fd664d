"""Autotuner for the kernel-shape knobs (vr_set_tuning).

No reference counterpart: the reference has no GPU path and nothing to tune.
It sweeps the knobs that change only a kernel's schedule, times each setting
on the caller's own workload with the library's HIP-event kernel timers, and
keeps the fastest:

  visibility  traversal-stack entries in LDS ("trace_lds_entries") x rays per
              wavefront ("trace_rays_per_lane", 64 x 1/2/4/8 rays per wave)
  bounce      gather tile width, kernel and prefetch depth ("bounce_kernel",
              "bounce_tile_cols", "bounce_group", "bounce_passes",
              "bounce_prefetch"), optionally the row placement
              ("bounce_placement")

Visibility results never depend on these knobs.  For BounceLight the kernel
and tile width fix the FP32 summation order (each order is reproduced bit for
bit by the CPU restatement), so tune_bounce() by default varies only what
keeps the current order (the prefetch depth); keep_order=False searches every
compiled shape.

Chosen settings are cached per (device, workload key), in memory and, when a
path is given, in a JSON file, so a service tunes a workload class once.
"""
import json
import os
import time

from . import _lib

VIS_STACKS = (2, 4, 6, 8, 10, 12, 16, 24)
VIS_RAYS = (1, 2, 4, 8)
BOUNCE_KNOBS = ("bounce_kernel", "bounce_tile_cols", "bounce_group", "bounce_passes",
                "bounce_prefetch")
# every compiled gather shape: (kernel, tile cols, group, passes, prefetch)
BOUNCE_SHAPES = ((2, 4096, 0, 0, 1), (2, 4096, 0, 0, 2), (2, 2048, 0, 0, 1),
                 (0, 4096, 16, 1, 1), (0, 4096, 16, 1, 3), (0, 4096, 16, 1, 7),
                 (0, 4096, 16, 2, 1), (0, 4096, 16, 2, 3), (0, 4096, 8, 2, 1), (0, 4096, 8, 2, 3),
                 (0, 4096, 8, 1, 3), (0, 4096, 8, 1, 7), (0, 2048, 8, 2, 1), (0, 2048, 16, 2, 1),
                 (1, 4096, 8, 2, 1))


def get_knob(name, lib=None):
    lib = lib or _lib.lib
    v = _lib.C.c_int()
    _lib.check(lib.vr_get_tuning(name.encode(), _lib.C.byref(v)))
    return v.value


def set_knob(name, value, lib=None):
    _lib.check((lib or _lib.lib).vr_set_tuning(name.encode(), int(value)))


def kernel_ms(family, lib=None):
    """Mean launch time (ms) of a kernel family since the last profile reset."""
    lib = lib or _lib.lib
    ms, n = _lib.C.c_double(), _lib.C.c_int64()
    _lib.check(lib.vr_profile_get(family.encode(), _lib.C.byref(ms), _lib.C.byref(n)))
    if n.value == 0:
        raise RuntimeError(f"no {family} launch was timed")
    return ms.value / n.value


def timed(run, family, reps=3, lib=None):
    """Median over reps of the mean `family` launch time of one run() call
    (after one untimed call for set-up and clocks)."""
    lib = lib or _lib.lib
    run()
    out = []
    for _ in range(reps):
        _lib.check(lib.vr_profile_reset())
        _lib.check(lib.vr_profile_enable(1))
        try:
            run()
        finally:
            _lib.check(lib.vr_profile_enable(0))
        out.append(kernel_ms(family, lib))
    out.sort()
    return out[len(out) // 2]


class Autotuner:
    """Sweeps knob settings over a workload and keeps the fastest.

    measure(knobs) -> ms is the only device-facing step; the default applies
    the knobs with vr_set_tuning and times run() with the kernel timers.  A
    caller may pass its own measure (the tests do, without a GPU)."""

    def __init__(self, device="gpu0", cache_path=None, reps=3, lib=None, setter=None,
                 final_rounds=0, finalists=3):
        self.device = device
        self.cache_path = cache_path
        self.reps = reps
        # final_rounds > 0: after the sweep, time the `finalists` fastest
        # candidates again, in turn, that many rounds, and keep each one's
        # best time, so a choice between near-equal settings does not rest
        # on one sample taken while the clocks drifted
        self.final_rounds = final_rounds
        self.finalists = finalists
        self.lib = lib
        self.setter = setter or (lambda k, v: set_knob(k, v, self.lib))
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def _key(self, kind, key):
        return f"{self.device}|{kind}|{key}"

    def _save(self):
        if not self.cache_path:
            return
        d = os.path.dirname(os.path.abspath(self.cache_path))
        os.makedirs(d, exist_ok=True)
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)

    def apply(self, knobs):
        for k, v in knobs.items():
            self.setter(k, v)

    def sweep(self, kind, key, candidates, measure, force=False):
        """Times every candidate knob set (dicts) with measure(knobs) and
        applies the fastest; a cached choice for (device, kind, key) is
        applied without timing unless force.  Returns the result record."""
        ck = self._key(kind, key)
        if not force and ck in self.cache:
            rec = self.cache[ck]
            self.apply(rec["best"])
            return dict(rec, cached=True)
        table = []
        for knobs in candidates:
            self.apply(knobs)
            table.append((measure(knobs), knobs))
        ms, best = min(table, key=lambda x: x[0])
        final = None
        if self.final_rounds > 0 and len(table) > 1:
            final = [[m, k] for m, k in sorted(table, key=lambda x: x[0])[:self.finalists]]
            for _ in range(self.final_rounds):
                for f in final:
                    self.apply(f[1])
                    f[0] = min(f[0], measure(f[1]))
            ms, best = min(final, key=lambda x: x[0])
        self.apply(best)
        rec = {"best": best, "best_ms": ms, "table": [{"ms": m, "knobs": k} for m, k in table],
               "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
        if final is not None:
            rec["final"] = [{"ms": m, "knobs": k} for m, k in final]
        self.cache[ck] = rec
        self._save()
        return dict(rec, cached=False)

    # ---- workloads -----------------------------------------------------------
    def tune_visibility(self, run, key, stacks=VIS_STACKS, rays=VIS_RAYS, family="test_line",
                        force=False, measure=None):
        """run() launches the caller's visibility batch (TestLine or
        TraceRays); key names the workload class (e.g. scene and batch size)."""
        cands = [{"trace_lds_entries": s, "trace_rays_per_lane": r} for s in stacks for r in rays]
        measure = measure or (lambda knobs: timed(run, family, self.reps, self.lib))
        return self.sweep("visibility", key, cands, measure, force)

    def tune_bounce(self, run, key, keep_order=True, shapes=BOUNCE_SHAPES, force=False,
                    measure=None, current=None, placements=None):
        """run() performs the caller's bounces (BounceLight or a batch).  With
        keep_order, only shapes with the current kernel and tile width (the
        FP32 summation order) are tried.  placements (e.g. (0, 1, 2)) also
        sweeps "bounce_placement", the dealing of rows to the gather's chunks
        (part of the stream gather's order, so keep_order holds it at the
        current value for that kernel); None leaves it as it is."""
        if keep_order:
            cur = current or (get_knob("bounce_kernel", self.lib),
                              get_knob("bounce_tile_cols", self.lib))
            shapes = [s for s in shapes if (s[0], s[1]) == tuple(cur[:2])]
        cands = [dict(zip(BOUNCE_KNOBS, s)) for s in shapes]
        if placements is not None:
            cur_p = current[2] if current and len(current) > 2 else \
                get_knob("bounce_placement", self.lib)
            cands = [dict(c, bounce_placement=p) for c in cands for p in placements
                     if not (keep_order and c["bounce_kernel"] == 2 and p != cur_p)]
        measure = measure or (lambda knobs: timed(run, "bounce_gather", self.reps, self.lib))
        return self.sweep(f"bounce-{'order' if keep_order else 'any'}", key, cands, measure,
                          force)
