"""The autotuner's selection and cache (no GPU: injected knob setter and
timings)."""
import json

from vrad_amd import autotune


class Knobs:
    def __init__(self):
        self.v = {}

    def set(self, k, val):
        self.v[k] = val


def test_visibility_sweep_picks_fastest_and_applies_it(tmp_path):
    kn = Knobs()
    tu = autotune.Autotuner(device="dev", cache_path=str(tmp_path / "t.json"), setter=kn.set)
    seen = []

    def measure(k):  # fastest at 12 entries, 2 rays per lane
        seen.append(dict(k))
        return abs(k["trace_lds_entries"] - 12) + abs(k["trace_rays_per_lane"] - 2) + 1.0

    rec = tu.tune_visibility(run=None, key="level:4M", measure=measure)
    assert len(seen) == len(autotune.VIS_STACKS) * len(autotune.VIS_RAYS)
    assert rec["best"] == {"trace_lds_entries": 12, "trace_rays_per_lane": 2}
    assert rec["best_ms"] == 1.0 and not rec["cached"]
    assert kn.v == rec["best"]                       # the winner is left applied
    saved = json.load(open(tmp_path / "t.json"))
    assert saved["dev|visibility|level:4M"]["best"] == rec["best"]
    # a second tuner with the same cache applies the choice without timing
    kn2 = Knobs()
    tu2 = autotune.Autotuner(device="dev", cache_path=str(tmp_path / "t.json"), setter=kn2.set)
    rec2 = tu2.tune_visibility(run=None, key="level:4M",
                               measure=lambda k: (_ for _ in ()).throw(AssertionError("timed")))
    assert rec2["cached"] and kn2.v == rec["best"]
    # another device or workload key is tuned afresh
    rec3 = autotune.Autotuner(device="other", setter=kn2.set).tune_visibility(
        run=None, key="level:4M", measure=measure)
    assert not rec3["cached"]


def test_bounce_sweep_keeps_the_summation_order():
    kn = Knobs()
    tu = autotune.Autotuner(setter=kn.set)
    tried = []

    def measure(k):
        tried.append((k["bounce_kernel"], k["bounce_tile_cols"], k["bounce_prefetch"]))
        return 1.0 / k["bounce_prefetch"]

    rec = tu.tune_bounce(run=None, key="250k", measure=measure, current=(2, 4096))
    # only shapes with the current kernel and tile width (same FP32 order)
    assert tried and all(t[:2] == (2, 4096) for t in tried)
    assert rec["best"]["bounce_prefetch"] == max(t[2] for t in tried)
    tried.clear()
    rec = tu.tune_bounce(run=None, key="250k", keep_order=False, measure=measure)
    assert len(tried) == len(autotune.BOUNCE_SHAPES)
    assert {t[0] for t in tried} == {0, 1, 2}


def test_bounce_sweep_of_row_placements():
    kn = Knobs()
    tu = autotune.Autotuner(setter=kn.set)
    seen = []

    def measure(k):
        seen.append((k["bounce_kernel"], k["bounce_placement"]))
        return 1.0 + k["bounce_placement"] * (-0.1 if k["bounce_kernel"] == 0 else 0.1)

    # keeping the order: the stream gather stays at the current placement,
    # the tiled gather (placement-independent order) tries every placement
    rec = tu.tune_bounce(run=None, key="p", measure=measure, current=(2, 4096, 0),
                         placements=(0, 1, 2))
    assert seen and {p for k, p in seen} == {0}
    assert rec["best"]["bounce_placement"] == 0
    seen.clear()
    rec = tu.tune_bounce(run=None, key="p-tiled", measure=measure, current=(0, 4096, 0),
                         placements=(0, 1, 2))
    assert {p for k, p in seen} == {0, 1, 2} and all(k == 0 for k, _ in seen)
    assert rec["best"]["bounce_placement"] == 2
    seen.clear()
    rec = tu.tune_bounce(run=None, key="q", keep_order=False, measure=measure,
                         placements=(0, 1, 2))
    assert len(seen) == 3 * len(autotune.BOUNCE_SHAPES)


def test_final_rounds_retime_the_fastest_candidates(tmp_path):
    # the first pass catches S10 in a slow sample and sees S6 fastest; the
    # finalists' later samples show S10 fastest, and the choice follows each
    # one's best time.  Without final rounds the first pass decides.
    def scripted():
        first = {(6, 1): 0.611, (8, 1): 0.612, (10, 1): 0.640}
        later = {(6, 1): 0.618, (8, 1): 0.612, (10, 1): 0.605}
        seen = {}

        def measure(knobs):
            k = (knobs["trace_lds_entries"], knobs["trace_rays_per_lane"])
            seen[k] = seen.get(k, 0) + 1
            return (first if seen[k] == 1 else later).get(k, 1.0)
        return measure, seen

    kn = Knobs()
    measure, seen = scripted()
    rec = autotune.Autotuner(setter=kn.set, final_rounds=2).tune_visibility(
        lambda: None, key="k", measure=measure)
    assert rec["best"] == {"trace_lds_entries": 10, "trace_rays_per_lane": 1}
    assert kn.v["trace_lds_entries"] == 10
    assert [f["knobs"]["trace_lds_entries"] for f in rec["final"]] == [6, 8, 10]
    assert seen[(8, 1)] == seen[(10, 1)] == seen[(6, 1)] == 3
    assert all(n == 1 for k, n in seen.items() if k not in ((8, 1), (10, 1), (6, 1)))
    kn2 = Knobs()
    measure2, _ = scripted()
    rec2 = autotune.Autotuner(setter=kn2.set).tune_visibility(lambda: None, key="k",
                                                               measure=measure2)
    assert rec2["best"] == {"trace_lds_entries": 6, "trace_rays_per_lane": 1}
    assert "final" not in rec2
