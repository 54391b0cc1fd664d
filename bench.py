"""Benchmark of the MI355X radiosity-bake path.

Headline (value): visibility throughput in Mrays/s on the "plain GPU path at
scale" workload: a procedural 8x8-room level (~175k triangles, chop 32,
16 point lights + sun) and batches of 4M patch-pair TestLine segments in
MakeTransfers order (rows of the upper triangle that pass the facing cull),
resident in HBM.  One step = one 4M-segment batch through the visibility
kernel.

Also reported (same run): the traversal-stack autotune sweep, the plain path
and, with --random-pairs, an incoherent batch; BounceLight gather bandwidth on the
precision-form transfer shape (250k rows, lognormal lengths, mean 2000, FP32
and FP16) after a kernel-shape sweep; and the CPU restatement timed on the
same segments.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SEG_BYTES = 26  # 24 B segment in + 1 B visible + 1 B sky out
LEVEL = dict(seed=2, rooms_x=8, rooms_y=8, chop=32.0, tess=64.0, n_lights=16, sun=True)


def pair_segments(o, nrm, i, j, eps=0.1):
    e = np.float32(eps)
    return (o[i] + e * nrm[i]).astype(np.float32), (o[j] + e * nrm[j]).astype(np.float32)


def row_batch(s, n_seg, i0):
    """One MakeTransfers visibility batch: rows i = i0, i0+1, ... with every
    j > i that passes the facing cull, in the order the pair kernel emits."""
    o, nrm = s.p_origin, s.p_normal
    n = len(o)
    out_i, out_j, have, i = [], [], 0, i0
    while have < n_seg and i < n - 1:
        j = np.arange(i + 1, n)
        dv = o[j] - o[i]
        di = dv @ nrm[i]
        dj = -np.einsum("ij,ij->i", nrm[j], dv)
        j = j[(di > 0) & (dj > 0)]
        out_i.append(np.full(len(j), i))
        out_j.append(j)
        have += len(j)
        i += 1
    i = np.concatenate(out_i)[:n_seg]
    j = np.concatenate(out_j)[:n_seg]
    return pair_segments(o, nrm, i, j)


def random_batch(s, n_seg, seed):
    """Random patch pairs passing the facing cull (no ray coherence)."""
    rng = np.random.default_rng(seed)
    o, nrm = s.p_origin, s.p_normal
    n = len(o)
    out_i, out_j, have = [], [], 0
    while have < n_seg:
        i = rng.integers(0, n, 2 * n_seg)
        j = rng.integers(0, n, 2 * n_seg)
        lo, hi = np.minimum(i, j), np.maximum(i, j)
        keep = lo != hi
        lo, hi = lo[keep], hi[keep]
        dv = o[hi] - o[lo]
        di = np.einsum("ij,ij->i", nrm[lo], dv)
        dj = -np.einsum("ij,ij->i", nrm[hi], dv)
        keep = (di > 0) & (dj > 0)
        out_i.append(lo[keep])
        out_j.append(hi[keep])
        have += int(keep.sum())
    return pair_segments(o, nrm, np.concatenate(out_i)[:n_seg], np.concatenate(out_j)[:n_seg])


class Ctx:
    def __init__(self):
        import torch
        from vrad_amd import _lib
        self.torch = torch
        self.lib, self.check, self.L = _lib.lib, _lib.check, _lib
        self.check(self.lib.vr_init(0))
        self.dev = torch.device("cuda:0")
        self.stream = torch.cuda.current_stream(self.dev)
        self.sptr = _lib.C.c_void_p(self.stream.cuda_stream)

    def kernel_ms(self, name):
        C = self.L.C
        ms, n = C.c_double(), C.c_int64()
        self.check(self.lib.vr_profile_get(name, C.byref(ms), C.byref(n)))
        return ms.value / max(n.value, 1)


def run_visibility(cx, args):
    from vrad_amd import scenes
    from vrad_amd.raytracer import Environment
    torch, lib, check, L = cx.torch, cx.lib, cx.check, cx.L
    s = scenes.level(**LEVEL)
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    tb = time.time()
    env.SetupAccelerationStructure()
    build_s = time.time() - tb
    info = env.info()
    st, en = row_batch(s, args.segments, i0=len(s.p_area) // 2)
    n = len(st)

    class Batch:
        def __init__(self, a, b):
            self.cols = [torch.from_numpy(np.ascontiguousarray(x[:, k])).to(cx.dev)
                         for x in (a, b) for k in range(3)]
            self.n = len(a)
            self.vis = torch.empty(self.n, dtype=torch.uint8, device=cx.dev)
            self.sky = torch.empty(self.n, dtype=torch.uint8, device=cx.dev)
            self.segs = L.Segs(*[L.C.cast(c.data_ptr(), L.FP) for c in self.cols])

        def step(self):
            check(lib.vr_test_line_dev(env.handle, L.C.byref(self.segs), self.n, -1,
                                       L.C.c_void_p(self.vis.data_ptr()),
                                       L.C.c_void_p(self.sky.data_ptr()), cx.sptr))

        def run(self, steps, warmup):
            for _ in range(warmup):
                self.step()
            torch.cuda.synchronize()
            check(lib.vr_profile_reset())
            check(lib.vr_profile_enable(1))
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(steps):
                self.step()
            torch.cuda.synchronize()
            el = time.perf_counter() - t
            check(lib.vr_profile_enable(0))
            return el, cx.kernel_ms(b"test_line")

    batch = Batch(st, en)
    # autotune the LDS share of the traversal stack on this batch
    sweep = {}
    for S in (4, 8, 12, 16, 24):
        check(lib.vr_set_tuning(b"trace_lds_entries", S))
        _, km = batch.run(3, 1)
        sweep[S] = round(n / (km * 1e-3) / 1e6, 1)
    best_s = max(sweep, key=sweep.get)
    check(lib.vr_set_tuning(b"trace_lds_entries", best_s))
    elapsed, kern_ms = batch.run(args.steps, args.warmup)
    vis = batch.vis.cpu().numpy()
    # the plain path (scratch stacks, no XCD remap) and an incoherent batch
    check(lib.vr_set_path(L.VR_PATH_PLAIN))
    _, plain_ms = batch.run(3, 1)
    check(lib.vr_set_path(L.VR_PATH_OPT))
    extra = {"lds_entries_sweep_Mrays": sweep, "lds_entries": best_s,
             "plain_path_Mrays": round(n / (plain_ms * 1e-3) / 1e6, 1)}
    if args.random_pairs:
        rb = Batch(*random_batch(s, args.segments, seed=20))
        _, rand_ms = rb.run(3, 1)
        del rb
        extra["random_pairs_Mrays"] = round(n / (rand_ms * 1e-3) / 1e6, 1)
    return dict(s=s, st=st, en=en, n=n, vis=vis, elapsed=elapsed, kern_ms=kern_ms,
                info=info, build_s=build_s, extra=extra)


def run_cpu(v, args):
    """The CPU restatement on the same batch, repeated for ~cpu_seconds."""
    import oracle
    s, st, en, n = v["s"], v["st"], v["en"], v["n"]
    threads = min(16, os.cpu_count() or 1)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    passes, cpu_s = 0, 0.0
    while passes < 512 and (passes == 0 or cpu_s < args.cpu_seconds):
        c0 = time.perf_counter()
        cv, _ = oracle.testline(kd, tri, st, en, threads=threads)
        cpu_s += time.perf_counter() - c0
        passes += 1
    return {"value": round(n * passes / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"the {n}-segment batch x {passes} passes, kd-tree TestLine "
                      f"(oracle/oracle.c), {cpu_s:.1f} s",
            "bitexact_vs_gpu": bool(np.array_equal(cv, v["vis"]))}


def run_bounce(cx, args):
    from vrad_amd import lightmap
    from vrad_amd.patches import Transfers
    lib, check, L = cx.lib, cx.check, cx.L
    N = 250_000
    rng = np.random.default_rng(8)
    e0 = rng.random((N, 3)).astype(np.float32)
    refl = rng.uniform(0.2, 0.8, (N, 3)).astype(np.float32)
    area = rng.uniform(16, 1024, N).astype(np.float32)
    xf = Transfers.synthetic(seed=3, n_rows=N, mean_len=2000.0, sigma=1.0, max_len=40000)
    T = xf.info().nnz

    knobs = (b"bounce_kernel", b"bounce_tile_cols", b"bounce_group", b"bounce_passes",
             b"bounce_prefetch")

    def set_bounce(*shape):
        for k, val in zip(knobs, shape):
            check(lib.vr_set_tuning(k, val))

    def bounce_ms(nb):
        check(lib.vr_profile_reset())
        check(lib.vr_profile_enable(1))
        lightmap.BounceLight(xf, e0, refl, area, max_bounces=nb, stop_frac=0.0)
        check(lib.vr_profile_enable(0))
        return cx.kernel_ms(b"bounce_gather")

    # (kernel, tile cols, group, passes loaded ahead of use, steps loads run ahead)
    shapes = [(0, 4096, 16, 1, 1), (0, 4096, 16, 1, 3), (0, 4096, 16, 2, 1), (0, 4096, 8, 2, 1),
              (0, 4096, 8, 2, 3), (0, 4096, 8, 1, 3), (0, 2048, 8, 2, 1), (0, 2048, 16, 2, 1),
              (1, 4096, 8, 2, 1)]

    def shape_name(shp):
        return "tiled%dx%d-p%d-d%d" % shp[1:] if shp[0] == 0 else "wave"

    # best shapes of earlier sweeps, used when the sweep is skipped
    default_shape = {"fp32": (0, 4096, 16, 1, 3), "fp16": (0, 4096, 8, 2, 3)}

    out = {}
    for name, prec, cb, sb in (("fp32", L.VR_COEF_F32, 4, 0), ("fp16", L.VR_COEF_F16, 2, 4)):
        xf.set_precision(prec)
        # stored layout read by the tiled gather: 16-bit column (LDS offset)
        # + coefficient per transfer; per patch row_ptr 8, emit read 12,
        # reflectivity 12, emit write 12, total read+write 24, scale
        B = T * (2 + cb) + N * (8 + 12 + 12 + 12 + 24 + sb)
        B_i32 = T * (4 + cb) + N * (8 + 12 + 12 + 12 + 24 + sb)
        sweep = {}
        for shp in (shapes if args.bounce_sweep else [default_shape[name]]):
            set_bounce(*shp)
            key = shape_name(shp)
            sweep[key] = round(bounce_ms(3), 4)
        best = min(sweep, key=sweep.get)
        set_bounce(*{shape_name(x): x for x in shapes}[best])
        per = bounce_ms(8)
        gbps = B / (per * 1e-3) / 1e9
        out[name] = {"ms_per_bounce": round(per, 4), "GBps": round(gbps, 1),
                     "frac": round(gbps / HBM_PEAK_GBPS, 4), "bytes_per_bounce": B,
                     "bytes_per_bounce_int32_index": B_i32,
                     "transfers": T, "rows": N, "kernel": best, "sweep_ms": sweep,
                     "traffic": measured_traffic("bounce_gather_" + name)[0]}
    set_bounce(*shapes[0])
    xf.close()
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/traffic.json, written from the PMC runs and
    scaled by the copy-kernel calibration), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
        return t[kernel]["bytes_per_launch"], t[kernel]["source"]
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--segments", type=int, default=4 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="repeat the CPU pass over the batch until this much time has run")
    ap.add_argument("--no-bounce", action="store_true")
    ap.add_argument("--no-trace", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--random-pairs", action="store_true",
                    help="also time an incoherent batch of random patch pairs")
    ap.add_argument("--bounce-sweep", type=int, default=1,
                    help="sweep the bounce kernel shapes (autotune) before timing")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="finish with a device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; runs on more than one GPU are not part of "
              "this benchmark (the multi-GPU sharding is a design, not built)", file=sys.stderr)
        sys.exit(2)

    cx = Ctx()
    v = None if args.no_trace else run_visibility(cx, args)
    cpu = None if (v is None or args.no_cpu) else run_cpu(v, args)
    bounce = None if args.no_bounce else run_bounce(cx, args)

    if args.calib_copy_mb > 0:
        torch = cx.torch
        a = torch.empty(args.calib_copy_mb << 20, dtype=torch.uint8, device=cx.dev).fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()

    out = {
        "metric": "Mrays/s visibility; bounce HBM GB/s; re-bake p50/p99 ms vs plain GPU",
        "value": None, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "plain GPU path at scale: 8x8-room procedural level, "
                               "4M MakeTransfers TestLine segments per batch (row order)"},
        "roofline": None, "cpu_baseline": cpu, "visibility": None, "bounce": bounce,
    }
    if v is not None:
        n, info = v["n"], v["info"]
        achieved = SEG_BYTES * n / (v["kern_ms"] * 1e-3) / 1e9
        out.update({
            "value": round(n * args.steps / v["elapsed"] / 1e6, 2),
            "ms_per_step": round(v["elapsed"] / args.steps * 1e3, 4),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
                         "traffic": measured_traffic("test_line")[0], "kernel": "test_line",
                         "kernel_ms": round(v["kern_ms"], 4),
                         "algorithmic_bytes_per_launch": SEG_BYTES * n},
            "visibility": v["extra"],
        })
        out["config"].update({"triangles": int(info.n_tri), "patches": int(len(v["s"].p_area)),
                              "kd_nodes": int(info.n_nodes), "kd_leaf_refs": int(info.n_leaf_refs),
                              "kd_build_s": round(v["build_s"], 2), "segments_per_step": n,
                              "visible_fraction": round(float(v["vis"].mean()), 4)})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
