"""Benchmark of the MI355X radiosity-bake path.

Headline (value): visibility throughput in Mrays/s on the "plain GPU path at
scale" workload: a procedural 8x8-room level (~175k triangles, chop 32,
16 point lights + sun) and batches of 4M patch-pair TestLine segments in
MakeTransfers order (rows of the upper triangle that pass the facing cull),
resident in HBM.  One step = one 4M-segment batch through the visibility
kernel.

Also reported (same run):
  visibility   the autotuner's sweep (traversal-stack entries in LDS x rays
               per wavefront), the plain path and, with --random-pairs, an
               incoherent batch;
  cpu_baseline the CPU restatement on the same segments (host cores);
  bounce       BounceLight gather bandwidth on the precision-form transfer
               shape (250k rows, lognormal lengths, mean 2000; FP32 and FP16)
               after the autotuner's sweep of the gather shapes (kernel, tile
               width, prefetch), beside a plain streaming-read probe;
  room_bake    the CPU-runnable closed room baked end to end on the GPU and
               by the CPU restatement, compared bit for bit;
  batch_bake   64 procedural scenes of uneven size: bounces of all of them in
               one launch per bounce (replayed hipGraph) vs one by one and vs
               the plain path;
  serving      re-bake requests (Poisson arrivals) against 32 resident scenes:
               p50/p99 latency and throughput of continuous batching vs the
               plain sequential re-bake, alone-vs-mix equality, FP16 luxel
               error vs the FP64 recurrence;
  pvs_culling  MakeTransfers with PVS culling of candidate pairs (room-grid
               BSP, adjacency PVS) on the 4x4-room level against the unculled
               run, and on the 8x8-room level;
  luxel_stage  patch generation, lightmap extents and luxels of the 8x8-room
               level's faces against the CPU restatement, and a whole bake
               from faces to luxels on a 3x3-room level.

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
# row placement of the gather in the level bake (rows clustered by room)
LEVEL_BAKE_PLACEMENT = 2
# the largest generated level: ~1.55 M triangles, ~264k patches
LEVEL_16X16 = dict(seed=3, rooms_x=16, rooms_y=16, chop=72.0, tess=40.0, n_lights=16, sun=True)


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def pvs_groups(s):
    """Patch clusters of the level's room grid, the one-doorway PVS (either
    direction), and the patches of each room."""
    from vrad_amd import pvs, scenes
    bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y, pvs_hops=1)
    cl = pvs.ClusterFromPoint(bsp, s.p_origin).astype(np.int64)
    rows, _ = pvs.DecompressVis(vis)
    nc = vis.n_clusters
    rv = np.unpackbits(rows, axis=1, bitorder="little")[:, :nc].astype(bool)
    return cl, rv | rv.T, [np.flatnonzero(cl == c) for c in range(nc)], rows


def pvs_pair_batch(s, n_seg, first_room):
    """A PVS-culled MakeTransfers visibility batch in the order
    vr_make_transfers_pvs forms its candidates: patches grouped by room
    (cluster), then for each room pair a <= b whose PVS bits pass (the room
    grid's PVS of one doorway), the rows of a against the columns of b (j > i
    within one room), facing cull; starting at room first_room.  About
    two thirds of these segments are visible, and a visible segment walks its
    whole length."""
    cl, rv, groups, _ = pvs_groups(s)
    nc = len(groups)
    o, nrm = s.p_origin, s.p_normal
    out_i, out_j, have = [], [], 0
    for a in list(range(first_room, nc)) + list(range(first_room)):
        for b in range(a, nc):
            if not rv[a, b] or have >= n_seg:
                continue
            for i in groups[a]:
                j = groups[b][groups[b] > i] if a == b else groups[b]
                dv = o[j] - o[i]
                j = j[(dv @ nrm[i] > 0) & (-np.einsum("ij,ij->i", nrm[j], dv) > 0)]
                out_i.append(np.full(len(j), i))
                out_j.append(j)
                have += len(j)
                if have >= n_seg:
                    break
        if have >= n_seg:
            break
    i = np.concatenate(out_i)[:n_seg]
    j = np.concatenate(out_j)[:n_seg]
    return pair_segments(o, nrm, i, j)


def hemisphere_rays(s, n_ray, seed):
    """Closest-hit rays (Trace4Rays' work): from patch centres moved 0.1 along
    the normal, cosine-distributed about the normal, tmax = MAX_TRACE_LENGTH."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(s.p_area), n_ray)
    nrm = s.p_normal[k].astype(np.float64)
    u1, u2 = rng.random(n_ray), rng.random(n_ray)
    r, ph = np.sqrt(u1), 2 * np.pi * u2
    a = np.where(np.abs(nrm[:, 0:1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]])
    t1 = np.cross(nrm, a)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(nrm, t1)
    d = (r * np.cos(ph))[:, None] * t1 + (r * np.sin(ph))[:, None] * t2 + \
        np.sqrt(1 - u1)[:, None] * nrm
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = (s.p_origin[k] + np.float32(0.1) * s.p_normal[k]).astype(np.float32)
    return o, d.astype(np.float32)


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


def kd_build_compare(s, env):
    """Scene build (triangle records, both kd-trees, upload) with the tree built
    on the GPU and on the host threads; both must give env's tree."""
    from vrad_amd import autotune
    from vrad_amd.raytracer import Environment
    out = {}
    old = autotune.get_knob("kd_build")
    try:
        for name, mode in (("device", 1), ("host", 0)):
            autotune.set_knob("kd_build", mode)
            e = Environment()
            e.AddTriangles(s.verts9, s.ids)
            t0 = time.perf_counter()
            e.SetupAccelerationStructure(n_threads=16)
            out[f"{name}_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
            same = (np.array_equal(e.OptimizedKDTree[0], env.OptimizedKDTree[0])
                    and np.array_equal(e.OptimizedKDTree[1].view(np.uint32),
                                       env.OptimizedKDTree[1].view(np.uint32))
                    and np.array_equal(e.TriangleIndexList, env.TriangleIndexList))
            out[f"{name}_same_tree"] = bool(same)
            e.close()
    finally:
        autotune.set_knob("kd_build", old)
    out["host_threads"] = 16
    out["note"] = "whole vr_scene_build: triangle records (host), kd-trees, upload"
    return out


class Batch:
    """A TestLine batch resident on the device (vr_test_line_dev)."""

    def __init__(self, cx, env, a, b):
        self.cx, self.env = cx, env
        torch, L = cx.torch, cx.L
        self.cols = [torch.from_numpy(np.ascontiguousarray(x[:, k])).to(cx.dev)
                     for x in (a, b) for k in range(3)]
        self.n = len(a)
        self.vis = torch.empty(self.n, dtype=torch.uint8, device=cx.dev)
        self.sky = torch.empty(self.n, dtype=torch.uint8, device=cx.dev)
        self.segs = L.Segs(*[L.C.cast(c.data_ptr(), L.FP) for c in self.cols])

    def step(self):
        cx, L = self.cx, self.cx.L
        cx.check(cx.lib.vr_test_line_dev(self.env.handle, L.C.byref(self.segs), self.n, -1,
                                         L.C.c_void_p(self.vis.data_ptr()),
                                         L.C.c_void_p(self.sky.data_ptr()), cx.sptr))

    def run(self, steps, warmup):
        cx, torch, check, lib = self.cx, self.cx.torch, self.cx.check, self.cx.lib
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


class RayBatch:
    """A closest-hit batch resident on the device (vr_trace_closest_dev)."""

    def __init__(self, cx, env, o, d):
        self.cx, self.env = cx, env
        torch, L = cx.torch, cx.L
        cols = [o[:, k] for k in range(3)] + [d[:, k] for k in range(3)]
        cols += [np.zeros(len(o), np.float32),
                 np.full(len(o), 1.732050807569 * 32768.0, np.float32)]
        self.cols = [torch.from_numpy(np.ascontiguousarray(c)).to(cx.dev) for c in cols]
        self.n = len(o)
        self.hit = torch.empty(self.n, dtype=torch.int32, device=cx.dev)
        self.t = torch.empty(self.n, dtype=torch.float32, device=cx.dev)
        self.rays = L.Rays(*[L.C.cast(c.data_ptr(), L.FP) for c in self.cols])

    def step(self):
        cx, L = self.cx, self.cx.L
        cx.check(cx.lib.vr_trace_closest_dev(self.env.handle, L.C.byref(self.rays), self.n,
                                             -1, L.C.c_void_p(self.hit.data_ptr()),
                                             L.C.c_void_p(self.t.data_ptr()), None, cx.sptr))

    def run(self, steps, warmup):
        cx, torch, check, lib = self.cx, self.cx.torch, self.cx.check, self.cx.lib
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        check(lib.vr_profile_reset())
        check(lib.vr_profile_enable(1))
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        check(lib.vr_profile_enable(0))
        return None, cx.kernel_ms(b"trace_closest")


def run_visibility(cx, args):
    progress("visibility")
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
    kd_build = kd_build_compare(s, env)
    st, en = row_batch(s, args.segments, i0=len(s.p_area) // 2)
    n = len(st)

    batch = Batch(cx, env, st, en)
    # autotune traversal-stack entries in LDS x rays per wavefront on this batch
    from vrad_amd import autotune
    rec = autotune.Autotuner(reps=3, final_rounds=2).tune_visibility(batch.step,
                                                                     key=f"level:{n}")
    sweep = {"S%d-R%d" % (r["knobs"]["trace_lds_entries"], r["knobs"]["trace_rays_per_lane"]):
             round(n / (r["ms"] * 1e-3) / 1e6, 1) for r in rec["table"]}
    best_s = rec["best"]
    elapsed, kern_ms = batch.run(args.steps, args.warmup)
    vis = batch.vis.cpu().numpy()
    # the plain path (one thread per segment, scratch-memory stacks) and an incoherent batch
    check(lib.vr_set_path(L.VR_PATH_PLAIN))
    _, plain_ms = batch.run(3, 1)
    check(lib.vr_set_path(L.VR_PATH_OPT))
    extra = {"autotune_Mrays": sweep, "tuned": best_s,
             "plain_path_Mrays": round(n / (plain_ms * 1e-3) / 1e6, 1), "kd_build": kd_build}
    # the visible-heavy batch: PVS-culled MakeTransfers rows of the same level
    progress("visibility: PVS-culled (visible-heavy) batch")
    pst, pen = pvs_pair_batch(s, args.segments, first_room=27)
    pb = Batch(cx, env, pst, pen)
    prec = autotune.Autotuner(reps=3, final_rounds=2).tune_visibility(
        pb.step, key=f"level-pvs:{pb.n}", stacks=(4, 6, 8, 10, 12), rays=(1,))
    _, pvs_ms = pb.run(args.steps, args.warmup)
    extra["pvs_batch"] = {
        "segments": pb.n, "visible_fraction": round(float(pb.vis.float().mean().item()), 4),
        "Mrays": round(pb.n / (pvs_ms * 1e-3) / 1e6, 1), "kernel_ms": round(pvs_ms, 4),
        "tuned": prec["best"],
        "autotune_Mrays": {"S%d" % r["knobs"]["trace_lds_entries"]:
                           round(pb.n / (r["ms"] * 1e-3) / 1e6, 1) for r in prec["table"]},
        "note": "the 8x8-room level's PVS-culled MakeTransfers candidates (room-pair "
                "order of vr_make_transfers_pvs, rooms from 27 on, one-doorway PVS, facing cull)"}
    del pb
    # closest hit (Trace4Rays' work at batch scale)
    progress("visibility: closest-hit rays")
    ro, rd = hemisphere_rays(s, args.segments, seed=21)
    rb = RayBatch(cx, env, ro, rd)
    crec = autotune.Autotuner(reps=3, final_rounds=2).tune_visibility(
        rb.step, key=f"level-closest:{rb.n}", stacks=(4, 6, 8, 10, 12), rays=(1,),
        family="trace_closest")
    _, cl_ms = rb.run(args.steps, args.warmup)
    extra["closest_hit"] = {
        "rays": rb.n, "Mrays": round(rb.n / (cl_ms * 1e-3) / 1e6, 1),
        "kernel_ms": round(cl_ms, 4), "tuned": crec["best"],
        "hit_fraction": round(float((rb.hit >= 0).float().mean().item()), 4),
        "note": "patch-centre rays, cosine-distributed about the normal, tmax = MAX_TRACE_LENGTH; "
                "40 B per ray algorithmic"}
    del rb
    autotune.Autotuner().apply(best_s)
    if args.random_pairs:
        rb = Batch(cx, env, *random_batch(s, args.segments, seed=20))
        _, rand_ms = rb.run(3, 1)
        del rb
        extra["random_pairs_Mrays"] = round(n / (rand_ms * 1e-3) / 1e6, 1)
    return dict(s=s, st=st, en=en, n=n, vis=vis, elapsed=elapsed, kern_ms=kern_ms,
                info=info, build_s=build_s, extra=extra)


def run_level16x16(cx, args):
    """Visibility on the largest generated level: 16 x 16 rooms,
    ~1.55 M triangles / ~264k patches.  Scene build with the tree
    on the GPU and on the host (same tree), then a 4 M-segment MakeTransfers
    batch in row order and 4 M closest-hit rays, each at its autotuned shape."""
    progress("16x16-room level: build")
    from vrad_amd import autotune, scenes
    from vrad_amd.raytracer import Environment
    s = scenes.level(**LEVEL_16X16)
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    t0 = time.perf_counter()
    env.SetupAccelerationStructure(n_threads=16)
    build_ms = (time.perf_counter() - t0) * 1e3
    info = env.info()
    kd = kd_build_compare(s, env)
    out = {"generator": dict(LEVEL_16X16), "triangles": int(info.n_tri),
           "patches": int(len(s.p_area)), "kd_nodes": int(info.n_nodes),
           "kd_leaf_refs": int(info.n_leaf_refs), "scene_build_ms": round(build_ms, 1),
           "kd_build": kd}
    progress("16x16-room level: TestLine batch")
    st, en = row_batch(s, args.segments, i0=len(s.p_area) // 2)
    b = Batch(cx, env, st, en)
    rec = autotune.Autotuner(reps=3, final_rounds=2).tune_visibility(
        b.step, key=f"level16x16:{b.n}", stacks=(6, 8, 10, 12), rays=(1,))
    _, ms = b.run(args.steps, args.warmup)
    out["test_line"] = {"segments": b.n, "Mrays": round(b.n / (ms * 1e-3) / 1e6, 1),
                        "kernel_ms": round(ms, 4), "tuned": rec["best"],
                        "visible_fraction": round(float(b.vis.float().mean().item()), 4),
                        "note": "MakeTransfers rows from the middle patch on, facing cull"}
    del b
    progress("16x16-room level: closest-hit rays")
    ro, rd = hemisphere_rays(s, args.segments, seed=33)
    rb = RayBatch(cx, env, ro, rd)
    rec = autotune.Autotuner(reps=3, final_rounds=2).tune_visibility(
        rb.step, key=f"level16x16-closest:{rb.n}", stacks=(6, 8, 10, 12), rays=(1,),
        family="trace_closest")
    _, ms = rb.run(args.steps, args.warmup)
    out["closest_hit"] = {"rays": rb.n, "Mrays": round(rb.n / (ms * 1e-3) / 1e6, 1),
                          "kernel_ms": round(ms, 4), "tuned": rec["best"],
                          "hit_fraction": round(float((rb.hit >= 0).float().mean().item()), 4)}
    del rb
    env.close()
    return out


def run_cpu(v, args):
    """The CPU restatement on the same batch, repeated for ~cpu_seconds."""
    import oracle
    s, st, en, n = v["s"], v["st"], v["en"], v["n"]
    threads = min(16, os.cpu_count() or 1)
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    passes, cpu_s, per = 0, 0.0, []
    while passes < 512 and (passes < 5 or cpu_s < args.cpu_seconds):
        c0 = time.perf_counter()
        cv, _ = oracle.testline(kd, tri, st, en, threads=threads)
        per.append(time.perf_counter() - c0)
        cpu_s += per[-1]
        passes += 1
    # one thread, on the batch's first 1 M segments (at least 5 passes)
    m1 = min(n, 1 << 20)
    one = []
    while len(one) < 5 or sum(one) < 2.0:
        c0 = time.perf_counter()
        oracle.testline(kd, tri, st[:m1], en[:m1], threads=1)
        one.append(time.perf_counter() - c0)
    # the traversal's work on this batch (closest-hit rule, untimed pass)
    wv, visits, tests = oracle.testline_stats(kd, tri, st, en, threads=threads)
    v["work"] = {"kd_node_visits_per_segment": round(visits / n, 2),
                 "triangle_tests_per_segment": round(tests / n, 2),
                 "gpu_giga_node_visits_per_s": round(visits / (v["kern_ms"] * 1e-3) / 1e9, 2),
                 "counted_by": "oracle kd traversal (closest-hit order)",
                 "visible_matches": bool(np.array_equal(wv, cv))}
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                         "")
    except OSError:
        pass
    return {"value": round(n / float(np.median(per)) / 1e6, 3), "unit": "Mrays/s",
            "cores": threads, "kind": "port",
            "sample": f"the {n}-segment batch x {passes} passes (median pass), kd-tree TestLine "
                      f"(oracle/oracle.c), {cpu_s:.1f} s",
            "value_1_thread": round(m1 / float(np.median(one)) / 1e6, 3),
            "sample_1_thread": f"the batch's first {m1} segments x {len(one)} passes (median)",
            "nproc": os.cpu_count(), "cpu_model": model,
            "bitexact_vs_gpu": bool(np.array_equal(cv, v["vis"]))}


def run_bounce(cx, args):
    progress("bounce gather sweep")
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

    def bounce_ms(nb, reps=1, warm_s=0.0):
        # one untimed call (tile layout for the shape), then untimed calls for
        # warm_s seconds (the GPU reaches steady clocks only after some
        # tenths of a second of this work), then the median over reps calls
        # of the gather's mean launch time
        lightmap.BounceLight(xf, e0, refl, area, max_bounces=1, stop_frac=0.0)
        t_end = time.perf_counter() + warm_s
        while time.perf_counter() < t_end:
            lightmap.BounceLight(xf, e0, refl, area, max_bounces=nb, stop_frac=0.0)
        out = []
        for _ in range(reps):
            check(lib.vr_profile_reset())
            check(lib.vr_profile_enable(1))
            lightmap.BounceLight(xf, e0, refl, area, max_bounces=nb, stop_frac=0.0)
            check(lib.vr_profile_enable(0))
            out.append(cx.kernel_ms(b"bounce_gather"))
        return float(np.median(out))

    # (kernel, tile cols, group, passes loaded ahead of use, steps loads run ahead)
    from vrad_amd import autotune
    shapes = list(autotune.BOUNCE_SHAPES)

    def shape_name(shp):
        if shp[0] == 2:
            return "stream%d-d%d" % (shp[1], shp[4])
        return "tiled%dx%d-p%d-d%d" % shp[1:] if shp[0] == 0 else "wave"

    # best shapes of earlier sweeps, used when the sweep is skipped
    default_shape = {"fp32": (2, 4096, 0, 0, 1), "fp16": (2, 4096, 0, 0, 1),
                     "bf16": (2, 4096, 0, 0, 1)}

    out = {}
    for name, prec, cb, sb in (("fp32", L.VR_COEF_F32, 4, 0), ("fp16", L.VR_COEF_F16, 2, 4),
                               ("bf16", L.VR_COEF_BF16, 2, 4)):
        xf.set_precision(prec)
        # stored layout read by the tiled gather: 16-bit column (LDS offset)
        # + coefficient per transfer; per patch row_ptr 8, emit read 12,
        # reflectivity 12, emit write 12, total read+write 24, scale
        B = T * (2 + cb) + N * (8 + 12 + 12 + 12 + 24 + sb)
        B_i32 = T * (4 + cb) + N * (8 + 12 + 12 + 12 + 24 + sb)
        # autotune every compiled gather shape (each is its own restated order;
        # the median of 3 timings per shape)
        tu = autotune.Autotuner(reps=3)
        rec = tu.tune_bounce(
            lambda: lightmap.BounceLight(xf, e0, refl, area, max_bounces=3, stop_frac=0.0),
            key=f"{name}:{N}:{T}", keep_order=False,
            shapes=shapes if args.bounce_sweep else [default_shape[name]])
        sweep = {shape_name(tuple(r["knobs"][k] for k in autotune.BOUNCE_KNOBS)): round(r["ms"], 4)
                 for r in rec["table"]}
        best = shape_name(tuple(rec["best"][k] for k in autotune.BOUNCE_KNOBS))
        per = bounce_ms(16, reps=5, warm_s=0.3)
        gbps = B / (per * 1e-3) / 1e9
        out[name] = {"ms_per_bounce": round(per, 4), "GBps": round(gbps, 1),
                     "frac": round(gbps / HBM_PEAK_GBPS, 4), "bytes_per_bounce": B,
                     "bytes_per_bounce_int32_index": B_i32,
                     "transfers": T, "rows": N, "kernel": best, "sweep_ms": sweep,
                     "traffic": measured_traffic("bounce_gather_" + name)[0]}
    set_bounce(2, 4096, 0, 0, 0)  # the library defaults, which the later sections use
    # what a plain streaming read reaches on this device (16-byte loads over
    # 4 GiB, best of 5): the practical ceiling beside the 8 TB/s peak
    g = L.C.c_double()
    check(lib.vr_measure_read_bandwidth(4 << 30, 5, L.C.byref(g)))
    out["read_probe_GBps"] = round(g.value, 1)
    xf.close()
    return out


def run_room_bake(cx, args):
    """The CPU-runnable case end to end: closed room (40 pillars), patches
    chopped at 32 units, one point light; MakeTransfers, direct light and
    FP32 bounces to the energy stop rule (1e-3, at most 100), on the GPU and
    with the CPU restatement on the host cores; results compared bit for bit."""
    import oracle
    from vrad_amd import lightmap, scenes
    from vrad_amd.patches import MakeTransfers
    from vrad_amd.raytracer import Environment
    progress("closed-room full bake")
    s = scenes.room(seed=1, n_pillars=40, chop=32.0, tess=32.0)

    def gpu_bake():
        t = {}
        t0 = time.perf_counter()
        env = Environment()
        env.AddTriangles(s.verts9, s.ids)
        env.SetupAccelerationStructure()
        t["tree_ms"] = (time.perf_counter() - t0) * 1e3
        t1 = time.perf_counter()
        xf = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area)
        t["make_transfers_ms"] = (time.perf_counter() - t1) * 1e3
        t2 = time.perf_counter()
        direct = lightmap.GatherSampleLight(env, s.p_origin, s.p_normal, s.lights)
        emit0 = (direct * s.p_refl).astype(np.float32)
        t["direct_ms"] = (time.perf_counter() - t2) * 1e3
        t3 = time.perf_counter()
        tot, en, k = lightmap.BounceLight(xf, emit0, s.p_refl, s.p_area, total=direct)
        t["bounce_ms"] = (time.perf_counter() - t3) * 1e3
        t["total_ms"] = (time.perf_counter() - t0) * 1e3
        nnz = xf.info().nnz
        xf.close()
        return t, tot, k, nnz

    gpu_bake()                       # first call pays one-time set-up
    tg, tot, k, nnz = gpu_bake()
    threads = min(16, os.cpu_count() or 1)
    tc = {}
    c0 = time.perf_counter()
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    tc["tree_ms"] = (time.perf_counter() - c0) * 1e3
    c1 = time.perf_counter()
    rp, col, coef, _ = oracle.make_transfers(kd, tri, s.p_origin, s.p_normal, s.p_area,
                                             threads=threads)
    tc["make_transfers_ms"] = (time.perf_counter() - c1) * 1e3
    c2 = time.perf_counter()
    direct = oracle.gather_light(kd, tri, s.p_origin, s.p_normal, s.lights, threads=threads)
    emit0 = (direct * s.p_refl).astype(np.float32)
    tc["direct_ms"] = (time.perf_counter() - c2) * 1e3
    c3 = time.perf_counter()
    otot, oen, ok = oracle.bounce_f32(rp, col, coef, emit0, s.p_refl, s.p_area, total=direct,
                                      threads=threads)
    tc["bounce_ms"] = (time.perf_counter() - c3) * 1e3
    tc["total_ms"] = (time.perf_counter() - c0) * 1e3
    r = lambda d: {kk: round(v, 2) for kk, v in d.items()}
    return {"triangles": int(len(s.ids)), "patches": int(len(s.p_area)), "transfers": int(nnz),
            "bounces": int(k), "gpu": r(tg), "cpu": r(tc), "cpu_threads": threads,
            "speedup_total": round(tc["total_ms"] / tg["total_ms"], 1),
            "bitexact_vs_cpu": bool(k == ok and np.array_equal(tot.view(np.uint32),
                                                                otot.view(np.uint32)))}


def run_pvs(cx, args):
    """PVS culling of MakeTransfers candidate pairs on generated levels (room
    grid BSP, each room's PVS = itself and the rooms one doorway away): the
    4x4-room level unculled and culled, and the 8x8-room level culled."""
    from vrad_amd import pvs, scenes
    from vrad_amd.patches import MakeTransfers
    from vrad_amd.raytracer import Environment

    def level_case(rx, ry, unculled, bounce=False):
        progress(f"PVS: {rx}x{ry}-room level")
        s = scenes.level(seed=2, rooms_x=rx, rooms_y=ry, chop=32.0, tess=64.0)
        env = Environment()
        env.AddTriangles(s.verts9, s.ids)
        env.SetupAccelerationStructure()
        bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y, pvs_hops=1)
        pvs.ClusterFromPoint(bsp, s.p_origin[:1])  # set-up outside the timing
        t0 = time.perf_counter()
        cl = pvs.ClusterFromPoint(bsp, s.p_origin).astype(np.int32)
        t1 = time.perf_counter()
        rows, _ = pvs.DecompressVis(vis)
        t2 = time.perf_counter()
        r = {"patches": int(len(s.p_area)), "clusters": int(vis.n_clusters),
             "cluster_from_point_ms": round((t1 - t0) * 1e3, 2),
             "decompress_vis_ms": round((t2 - t1) * 1e3, 2),
             "patches_without_cluster": int((cl < 0).sum())}
        t3 = time.perf_counter()
        xp = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=cl, pvs_rows=rows,
                           normalize=False)
        ip = xp.info()
        r["pvs"] = {"make_transfers_s": round(time.perf_counter() - t3, 3),
                    "pairs_considered": int(ip.pairs_considered),
                    "segments_traced": int(ip.segments_traced), "transfers": int(ip.nnz)}
        r["pvs"]["segments_per_s_G"] = round(ip.segments_traced / r["pvs"]["make_transfers_s"]
                                             / 1e9, 3)
        if bounce:
            r["bounce_on_this_csr"] = real_csr_bounce(cx, xp, s)
        if unculled:
            t4 = time.perf_counter()
            xf = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, normalize=False)
            inf = xf.info()
            r["unculled"] = {"make_transfers_s": round(time.perf_counter() - t4, 3),
                             "pairs_considered": int(inf.pairs_considered),
                             "segments_traced": int(inf.segments_traced),
                             "transfers": int(inf.nnz)}
            rp, col, coef = xf.download()
            prp, pcol, pcoef = xp.download()
            row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
            a, b = cl[row], cl[col]
            rv = np.unpackbits(rows, axis=1, bitorder="little")[:, :vis.n_clusters].astype(bool)
            keep = rv[a, b] | rv[b, a]
            r["pvs_result_equals_filtered_unculled"] = bool(
                np.array_equal(pcol, col[keep]) and
                np.array_equal(pcoef.view(np.uint32), coef[keep].view(np.uint32)))
            r["transfers_outside_pvs"] = int((~keep).sum())
            r["speedup"] = round(r["unculled"]["make_transfers_s"] / r["pvs"]["make_transfers_s"], 2)
            xf.close()
        xp.close()
        env.close()
        return r

    return {"note": "PVS = each room and its grid neighbours (one doorway); normalize off so "
                    "culled rows are the unculled rows minus the culled pairs",
            "level_4x4": level_case(4, 4, True)}


def run_level_bake(cx, args):
    """The plain GPU path at scale, end to end: the 8x8-room level (16 point
    lights + sun), scene build, clusters and PVS, PVS-culled MakeTransfers
    (row-normalised), direct light, 100 FP32 bounces.  Checked on the CPU
    (oracle) on a sampled subset: the direct light of 4096 patches, and the
    transfer columns of 64 rows against TestLine over every PVS candidate of
    those rows.  Then the bounce rate on this CSR in FP32 and FP16."""
    import oracle
    from vrad_amd import lightmap, pvs, scenes
    from vrad_amd.patches import MakeTransfers
    from vrad_amd.raytracer import Environment
    progress("level bake: 8x8-room level end to end")
    s = scenes.level(**LEVEL)
    np_ = len(s.p_area)
    times = {}
    t0 = time.perf_counter()
    env = Environment()
    env.AddTriangles(s.verts9, s.ids)
    env.SetupAccelerationStructure()
    times["scene_build_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y, pvs_hops=1)
    cl = pvs.ClusterFromPoint(bsp, s.p_origin).astype(np.int32)
    rows, _ = pvs.DecompressVis(vis)
    times["clusters_and_pvs_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    xf = MakeTransfers(env, s.p_origin, s.p_normal, s.p_area, cluster=cl, pvs_rows=rows)
    times["make_transfers_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    direct = lightmap.GatherSampleLight(env, s.p_origin, s.p_normal, s.lights)
    times["direct_light_ms"] = (time.perf_counter() - t0) * 1e3
    emit0 = (direct * s.p_refl).astype(np.float32)
    # MakeTransfers rows cluster by room: the gather deals each chunk's rows
    # to its waves by row index (bounce_placement 2, its own restated order);
    # the tile layout for it is built inside the timed bounce call
    from vrad_amd import autotune
    old_place = autotune.get_knob("bounce_placement")
    autotune.set_knob("bounce_placement", LEVEL_BAKE_PLACEMENT)
    try:
        t0 = time.perf_counter()
        tot, en, k = lightmap.BounceLight(xf, emit0, s.p_refl, s.p_area, total=direct,
                                          max_bounces=100, stop_frac=0.0)
        times["bounces_ms"] = (time.perf_counter() - t0) * 1e3
    finally:
        autotune.set_knob("bounce_placement", old_place)
    times["total_ms"] = sum(times.values())
    times = {a: round(b, 1) for a, b in times.items()}
    inf = xf.info()
    progress("level bake: CPU checks on a sample")
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    rng = np.random.default_rng(77)
    pick = rng.choice(np_, 4096, replace=False)
    want = oracle.gather_light(kd, tri, s.p_origin[pick], s.p_normal[pick], s.lights, threads=16)
    direct_ok = bool(np.array_equal(want.view(np.uint32), direct[pick].view(np.uint32)))
    rp, col, _ = xf.download()
    cl64, rv, groups, _ = pvs_groups(s)
    o, nrm = s.p_origin, s.p_normal
    rows_ok, seg_checked = True, 0
    for i in rng.choice(np_, 64, replace=False):
        j = np.arange(np_)
        ok = rv[cl64[i], np.maximum(cl64, 0)] | (cl64 < 0) | (cl64[i] < 0)
        j = j[ok & (j != i)]
        # the kernels cull and trace a pair in ascending patch order, with
        # one float32 rounding per operation
        lo, hi = np.minimum(i, j), np.maximum(i, j)
        d = [o[hi, c] - o[lo, c] for c in range(3)]
        di = (nrm[lo, 0] * d[0] + nrm[lo, 1] * d[1]) + nrm[lo, 2] * d[2]
        dj = -((nrm[hi, 0] * d[0] + nrm[hi, 1] * d[1]) + nrm[hi, 2] * d[2])
        keep = (di > 0) & (dj > 0)
        j, lo, hi = j[keep], lo[keep], hi[keep]
        a, b = pair_segments(o, nrm, lo, hi)
        v, _ = oracle.testline(kd, tri, a, b, threads=16)
        seg_checked += len(j)
        got = col[rp[i]:rp[i + 1]]
        got = got[got != i]             # padding entries (col = row, coef = 0)
        rows_ok = rows_ok and np.array_equal(np.sort(got), np.sort(j[v.astype(bool)]))
    del rp, col
    out = {"triangles": int(len(s.ids)), "patches": int(np_),
           "lights": int(len(s.lights)), "sun": bool((s.lights["type"] == 3).any()),
           "transfers": int(inf.nnz), "segments_traced": int(inf.segments_traced),
           "bounces": int(k), "bounce_placement": LEVEL_BAKE_PLACEMENT,
           "energy_first_last": [float(en[0]), float(en[-1])],
           "times": times,
           "cpu_check": {"direct_light_4096_patches_bitexact": direct_ok,
                         "rows_checked": 64, "segments_checked": int(seg_checked),
                         "transfer_columns_equal_oracle_visibility": bool(rows_ok)}}
    out["bounce_rate_on_this_csr"] = real_csr_bounce(cx, xf, s)
    xf.close()
    env.close()
    return out


def real_csr_bounce(cx, xf, s):
    """BounceLight GB/s on a transfer set MakeTransfers built (clustered
    columns, rows of every length), FP32 and FP16, with each row placement:
    8 bounces per call, the median of 3 calls after a warm-up call that
    builds the tile layout and 0.2 s of untimed calls."""
    from vrad_amd import autotune, lightmap
    lib, check, L = cx.lib, cx.check, cx.L
    inf = xf.info()
    N, T = int(inf.n_rows), int(inf.stored_nnz)
    rng = np.random.default_rng(9)
    e0 = rng.random((N, 3)).astype(np.float32)
    out = {"rows": N, "transfers": int(inf.nnz), "stored_entries": T}
    old = autotune.get_knob("bounce_placement")
    try:
        for placement, pname in ((0, "strided"), (1, "row_runs"), (2, "strided_by_index")):
            autotune.set_knob("bounce_placement", placement)
            for name, prec, cb, sb in (("fp32", L.VR_COEF_F32, 4, 0),
                                       ("fp16", L.VR_COEF_F16, 2, 4)):
                xf.set_precision(prec)
                lightmap.BounceLight(xf, e0, s.p_refl, s.p_area, max_bounces=1, stop_frac=0.0)
                t_end = time.perf_counter() + 0.2   # steady clocks before timing
                while time.perf_counter() < t_end:
                    lightmap.BounceLight(xf, e0, s.p_refl, s.p_area, max_bounces=8,
                                         stop_frac=0.0)
                ms = []
                for _ in range(3):
                    check(lib.vr_profile_reset())
                    check(lib.vr_profile_enable(1))
                    lightmap.BounceLight(xf, e0, s.p_refl, s.p_area, max_bounces=8,
                                         stop_frac=0.0)
                    check(lib.vr_profile_enable(0))
                    ms.append(cx.kernel_ms(b"bounce_gather"))
                per = float(np.median(ms))
                B = T * (2 + cb) + N * (8 + 12 + 12 + 12 + 24 + sb)
                out.setdefault(pname, {})[name] = {
                    "ms_per_bounce": round(per, 4), "GBps": round(B / (per * 1e-3) / 1e9, 1),
                    "frac": round(B / (per * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                    "bytes_per_bounce": B}
    finally:
        autotune.set_knob("bounce_placement", old)
        xf.set_precision(L.VR_COEF_F32)
    out["note"] = ("strided: each gather chunk takes a stride through all rows longest-first; "
                   "row_runs: a run of consecutive rows (bounce_placement = 1); "
                   "strided_by_index: strided chunks, rows dealt to waves by index (2)")
    return out


def run_luxel_stage(cx, args):
    """The luxel stage on the 8x8-room level's faces (lightmap vectors of 16
    units per luxel): patch generation, lightmap extents and the luxels of
    every face, against the CPU restatement (one thread); then a whole bake
    from faces to luxels on a 3x3-room level: patches, clusters, PVS,
    MakeTransfers, direct light, bounces, FinalLightFace."""
    import oracle
    from vrad_amd import _lib, lightmap, luxels, pvs, scenes
    from vrad_amd.patches import MakePatches, MakeTransfers
    from vrad_amd.raytracer import Environment
    progress("luxel stage: 8x8-room level faces")
    faces = scenes.faces_of("level", **LEVEL)
    MakePatches(faces[:1])  # set-up outside the timing
    t0 = time.perf_counter()
    patches = MakePatches(faces)
    t1 = time.perf_counter()
    lm, n_lux = luxels.CalcFaceExtents(faces)
    t2 = time.perf_counter()
    val = np.random.default_rng(0).uniform(0, 1, (len(patches), 3)).astype(np.float32)
    lx = luxels.FaceLuxels(faces, lm, patches, val)
    t3 = time.perf_counter()
    c0 = time.perf_counter()
    olm, _ = oracle.calc_face_extents(faces)
    opos, opat = oracle.face_luxels(faces, olm, patches)
    c1 = time.perf_counter()
    out = {"faces": int(len(faces)), "patches": int(len(patches)),
           "leaf_patches": int((patches["child1"] == -1).sum()), "luxels": int(n_lux),
           "gpu_ms": {"make_patches": round((t1 - t0) * 1e3, 2),
                      "face_extents": round((t2 - t1) * 1e3, 2),
                      "face_luxels": round((t3 - t2) * 1e3, 2)},
           "cpu_restatement_ms": round((c1 - c0) * 1e3, 1), "cpu_threads": 1,
           "bitexact_vs_cpu": bool(lm.tobytes() == olm.tobytes()
                                   and np.array_equal(lx["patch"], opat)
                                   and np.array_equal(lx["pos"].view(np.uint32),
                                                      opos.view(np.uint32))),
           "note": "GPU times are whole calls from host arrays (upload, kernels, download)"}

    progress("bake from faces to luxels: 3x3-room level")
    kw = dict(seed=5, rooms_x=3, rooms_y=3, chop=32.0, tess=64.0, n_lights=4, sun=True)
    s = scenes.level(**kw)
    faces = scenes.faces_of("level", **kw)

    def bake(errors=False):
        T = {}
        tb = time.perf_counter()
        env = Environment()
        env.AddTriangles(s.verts9, s.ids)
        env.SetupAccelerationStructure()
        T["scene_build"] = time.perf_counter() - tb
        t = time.perf_counter()
        patches = MakePatches(faces)
        leaf = np.flatnonzero(patches["child1"] == -1)
        T["make_patches"] = time.perf_counter() - t
        t = time.perf_counter()
        bsp, vis = scenes.grid_bsp(s.grid_x, s.grid_y, pvs_hops=1)
        cl = pvs.PatchClusters(bsp, patches[leaf])
        rows, _ = pvs.DecompressVis(vis)
        T["clusters_and_pvs"] = time.perf_counter() - t
        o, nrm, area = patches["origin"][leaf], patches["normal"][leaf], patches["area"][leaf]
        t = time.perf_counter()
        xf = MakeTransfers(env, o, nrm, area, cluster=cl, pvs_rows=rows)
        T["make_transfers"] = time.perf_counter() - t
        refl = np.full((len(leaf), 3), 0.5, np.float32)
        t = time.perf_counter()
        direct = lightmap.GatherSampleLight(env, o, nrm, s.lights)
        T["direct"] = time.perf_counter() - t
        t = time.perf_counter()
        tot, _, k = lightmap.BounceLight(xf, (direct * refl).astype(np.float32), refl, area,
                                         total=direct)
        T["bounce"] = time.perf_counter() - t
        pt = np.zeros((len(patches), 3), np.float32)
        pd = np.zeros((len(patches), 3), np.float32)
        pt[leaf], pd[leaf] = tot, direct
        t = time.perf_counter()
        _, light, lx = luxels.FinalLightFace(env, faces, patches, pt, pd, s.lights)
        T["final_light_face"] = time.perf_counter() - t
        T["total"] = time.perf_counter() - tb
        r = {"faces": int(len(faces)), "leaf_patches": int(len(leaf)),
             "transfers": int(xf.info().nnz), "bounces": int(k), "luxels": int(len(light)),
             "ms": {a: round(b * 1e3, 2) for a, b in T.items()}}
        if errors:
            # luxel error against the FP64 recurrence of the same k bounces
            # (same transfers, same direct light), FP32 and FP16 coefficients
            emit0 = (direct * refl).astype(np.float32)
            rp_, col_, coef_ = xf.download()
            t64, _, _ = oracle.bounce_f64(rp_, col_, coef_, emit0, refl, area,
                                          total=direct.astype(np.float64), max_bounces=k,
                                          stop_frac=0.0)
            dlux = lightmap.GatherSampleLight(env, lx["pos"], lx["normal"], s.lights)
            b64 = np.zeros((len(patches), 3))
            b64[leaf] = t64 - direct
            l64 = dlux.astype(np.float64) + b64[lx["patch"]]
            nz = l64 != 0

            def rel(l):
                return float(np.max(np.abs(l.astype(np.float64)[nz] - l64[nz]) / np.abs(l64[nz])))
            xf.set_precision(_lib.VR_COEF_F16)
            t16, _, _ = lightmap.BounceLight(xf, emit0, refl, area, total=direct, max_bounces=k,
                                             stop_frac=0.0)
            b16 = np.zeros((len(patches), 3), np.float32)
            b16[leaf] = t16 - direct
            r["luxel_rel_err_vs_fp64"] = {"fp32": rel(light),
                                          "fp16": rel((dlux + b16[lx["patch"]]).astype(np.float32)),
                                          "luxels_compared": int(nz.any(axis=1).sum()),
                                          "zero_where_fp64_zero": bool(np.all(light[~nz] == 0))}
        xf.close()
        env.close()
        return r, patches, leaf, cl, rows, light, lx, k

    first, *_ = bake()
    r, patches, leaf, cl, rows, light, lx, k = bake(errors=True)
    r["first_call_total_ms"] = first["ms"]["total"]
    # the same bake by the CPU restatement, from the same patches
    threads = min(16, os.cpu_count() or 1)
    c0 = time.perf_counter()
    tri = oracle.precompute(s.verts9, s.ids)
    kd = oracle.KdTree(s.verts9)
    o, nrm, area = patches["origin"][leaf], patches["normal"][leaf], patches["area"][leaf]
    rp, col, coef, _ = oracle.make_transfers(kd, tri, o, nrm, area, threads=threads, cluster=cl,
                                             pvs_rows=rows)
    direct = oracle.gather_light(kd, tri, o, nrm, s.lights, threads=threads)
    refl = np.full((len(leaf), 3), 0.5, np.float32)
    otot, _, ok = oracle.bounce_f32(rp, col, coef, (direct * refl).astype(np.float32), refl, area,
                                    total=direct, threads=threads)
    olm, _ = oracle.calc_face_extents(faces)
    opos, opat = oracle.face_luxels(faces, olm, patches)
    bounced = np.zeros((len(patches), 3), np.float32)
    bounced[leaf] = otot - direct
    onrm = faces["normal"][np.repeat(np.arange(len(faces)), olm["n_luxels"])]
    olight = (oracle.gather_light(kd, tri, opos, onrm, s.lights, threads=threads)
              + bounced[opat]).astype(np.float32)
    r["cpu_restatement_ms"] = round((time.perf_counter() - c0) * 1e3, 1)
    r["cpu_threads"] = threads
    r["bitexact_vs_cpu"] = bool(ok == k and np.array_equal(light.view(np.uint32),
                                                           olight.view(np.uint32)))
    out["bake_3x3"] = r
    return out


def batching_scenes(n_scenes, seed=64):
    """Scenes for the many-scene bake: triangle and patch targets drawn
    log-uniformly from [2k, 200k] and [1k, 60k], 1-16 lights, sun on half.
    Rooms follow the patch target (about 3000 patches per room at chop 32),
    then the wall tessellation is solved for the triangle target and the chop
    for the patch target from one calibration scene (triangles ~ 1/tess^2,
    patches ~ 1/chop^2)."""
    from vrad_amd import scenes
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        t_tri = float(np.exp(rng.uniform(np.log(2000.0), np.log(200000.0))))
        t_pat = float(np.exp(rng.uniform(np.log(1000.0), np.log(60000.0))))
        nl, sun = int(rng.integers(1, 17)), bool(rng.integers(0, 2))
        rooms = int(np.clip(round(t_pat / 3000.0), 1, 16))
        rx = int(np.clip(round(np.sqrt(rooms)), 1, 4))
        ry = int(np.clip(round(rooms / rx), 1, 4))
        c = scenes.level(seed=500 + i, rooms_x=rx, rooms_y=ry, chop=64.0, tess=32.0,
                         n_lights=nl, sun=sun)
        tess = float(np.clip(32.0 * np.sqrt(len(c.ids) / t_tri), 8.0, 512.0))
        chop = float(np.clip(64.0 * np.sqrt(len(c.p_area) / t_pat), 12.0, 512.0))
        out.append(scenes.level(seed=500 + i, rooms_x=rx, rooms_y=ry, chop=chop, tess=tess,
                                n_lights=nl, sun=sun))
    return out


def run_batch_bake(cx, args):
    """Many-scene bake: 64 procedural scenes of uneven size baked together.
    Visibility of every scene's MakeTransfers candidate pairs shares one
    launch per round (vr_make_transfers_batch), direct light of every scene
    is one launch (vr_gather_sample_light_batch), and every scene's FP32
    bounces to the stop rule run in one batched launch per bounce (a replayed
    hipGraph of 8 bounces).  Each phase is also run one scene at a time, and
    the bounces on the plain path; results compared bit for bit.  The
    bounces' tile layouts are built before either bounce leg is timed."""
    from vrad_amd import lightmap
    from vrad_amd.patches import MakeTransfers, MakeTransfersBatch
    from vrad_amd.raytracer import Environment
    lib, check, L = cx.lib, cx.check, cx.L
    progress(f"batched bake of {args.batch_scenes} scenes")
    sc = batching_scenes(args.batch_scenes)
    envs = []
    t0 = time.perf_counter()
    for s in sc:
        env = Environment()
        env.AddTriangles(s.verts9, s.ids)
        env.SetupAccelerationStructure()
        envs.append(env)
    t_build = time.perf_counter() - t0
    items = [(env, s.p_origin, s.p_normal, s.p_area) for s, env in zip(sc, envs)]
    gitems = [(env, s.p_origin, s.p_normal, s.lights) for s, env in zip(sc, envs)]
    # warm both paths' kernels on the two smallest scenes
    small = sorted(range(len(sc)), key=lambda k: len(sc[k].p_area))[:2]
    for x in MakeTransfersBatch([items[k] for k in small]):
        x.close()
    lightmap.GatherSampleLightBatch([gitems[k] for k in small])
    # one scene at a time
    progress("batched bake: MakeTransfers and direct light one scene at a time")
    t0 = time.perf_counter()
    single = [MakeTransfers(*it) for it in items]
    t_xf_single = time.perf_counter() - t0
    t0 = time.perf_counter()
    d_single = [lightmap.GatherSampleLight(*g) for g in gitems]
    t_direct_single = time.perf_counter() - t0
    info_single = [(x.info().nnz, x.info().segments_traced) for x in single]
    check_ix = sorted(range(len(sc)), key=lambda k: len(sc[k].p_area))
    check_ix = check_ix[:6] + check_ix[-2:]
    ref_csr = {k: single[k].download() for k in check_ix}
    for x in single:
        x.close()
    # batched
    progress("batched bake: MakeTransfers and direct light batched")
    t0 = time.perf_counter()
    xfs = MakeTransfersBatch(items)
    t_xf_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    direct = lightmap.GatherSampleLightBatch(gitems)
    t_direct_batch = time.perf_counter() - t0
    same_xf = all((x.info().nnz, x.info().segments_traced) == info_single[k]
                  for k, x in enumerate(xfs))
    for k in check_ix:
        a, b = ref_csr[k], xfs[k].download()
        same_xf = same_xf and all(np.array_equal(u.view(np.uint32) if u.dtype == np.float32 else u,
                                                 v.view(np.uint32) if v.dtype == np.float32 else v)
                                  for u, v in zip(a, b))
    same_direct = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                      for a, b in zip(direct, d_single))
    jobs = [(xf, (d * s.p_refl).astype(np.float32), s.p_refl, s.p_area, d)
            for xf, d, s in zip(xfs, direct, sc)]
    # tile layouts of every job, outside both timed legs
    lightmap.BounceLightBatch(jobs, max_bounces=1, stop_frac=0.0, use_graph=False)
    progress("batched bake: bounces")
    t0 = time.perf_counter()
    batch = lightmap.BounceLightBatch(jobs, max_bounces=100, stop_frac=1e-3, use_graph=True)
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    one = [lightmap.BounceLight(xf, e0, rf, ar, total=d) for xf, e0, rf, ar, d in jobs]
    t_single = time.perf_counter() - t0
    check(lib.vr_set_path(L.VR_PATH_PLAIN))
    try:
        t0 = time.perf_counter()
        for xf, e0, rf, ar, d in jobs:
            lightmap.BounceLight(xf, e0, rf, ar, total=d)
        t_plain = time.perf_counter() - t0
    finally:
        check(lib.vr_set_path(L.VR_PATH_OPT))
    same = all(k1 == k2 and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
               for (t1, _, k1), (t2, _, k2) in zip(batch, one))
    n = [len(s.p_area) for s in sc]
    tris = [len(s.ids) for s in sc]
    nnz = [x.info().nnz for x in xfs]
    segs = sum(x.info().segments_traced for x in xfs)
    ks = [b[2] for b in batch]
    for x in xfs:
        x.close()
    for e in envs:
        e.close()
    return {"scenes": len(sc), "triangles": [int(min(tris)), int(max(tris))],
            "triangles_total": int(sum(tris)),
            "patches": [int(min(n)), int(max(n))], "patches_total": int(sum(n)),
            "lights": [int(min(len(s.lights) for s in sc)), int(max(len(s.lights) for s in sc))],
            "transfers_total": int(sum(nnz)), "segments_traced": int(segs),
            "bounces": [int(min(ks)), int(max(ks))],
            "scene_build_s": round(t_build, 3),
            "make_transfers_batched_s": round(t_xf_batch, 3),
            "make_transfers_one_by_one_s": round(t_xf_single, 3),
            "make_transfers_equal": bool(same_xf),
            "direct_batched_s": round(t_direct_batch, 4),
            "direct_one_by_one_s": round(t_direct_single, 4),
            "direct_equal": bool(same_direct),
            "bounce_batched_graph_s": round(t_batch, 4), "bounce_one_by_one_s": round(t_single, 4),
            "bounce_plain_s": round(t_plain, 4),
            "batched_vs_plain": round(t_plain / t_batch, 1),
            "batched_equals_one_by_one": bool(same),
            "note": "tile layouts built before both timed bounce legs; MakeTransfers equality "
                    "by nnz and segments for all scenes and full CSR for the 6 smallest and "
                    "2 largest"}


def run_serving(cx, args):
    """Serving: resident scenes from the level generator (mixed sizes,
    synthetic lognormal transfers, FP16 with per-row scale), a seeded
    Poisson stream of re-bake requests (1-4 lights moved by
    U[-64,64]^3, direct re-gather + 8 bounces); p50/p99 latency (completion -
    arrival) and throughput of the batched server and of the plain sequential
    server on the same arrival trace; a sample of requests re-run alone must
    equal its results in the mix."""
    from vrad_amd import scenes
    from vrad_amd.patches import Transfers
    from vrad_amd.raytracer import Environment
    from vrad_amd.serve import BATCHED, PLAIN, Server
    L = cx.L
    progress(f"serving: building {args.serve_scenes} resident scenes")
    rng = np.random.default_rng(2026)
    resident = []
    for i in range(args.serve_scenes):
        rx, ry = (int(v) for v in rng.integers(2, 7, 2))
        s = scenes.level(seed=900 + i, rooms_x=rx, rooms_y=ry, chop=32.0, tess=64.0,
                         n_lights=int(rng.integers(4, 17)), sun=True)
        env = Environment()
        env.AddTriangles(s.verts9, s.ids)
        env.SetupAccelerationStructure()
        n = len(s.p_area)
        xf = Transfers.synthetic(seed=7000 + i, n_rows=n, mean_len=2000.0, sigma=1.0,
                                 max_len=min(40000, n - 1), precision=L.VR_COEF_F16)
        resident.append((s, env, xf))
    # the FP32 coefficients of the smallest scene, for the FP64 error check
    # (read before any server holds the scenes' tile-blocked copies)
    small = int(np.argmin([len(r[0].p_area) for r in resident]))
    xs = resident[small][2]
    xs.set_precision(L.VR_COEF_F32)
    csr32 = xs.download()
    xs.set_precision(L.VR_COEF_F16)
    def make_trace(n_req, rate, seed):
        g = np.random.default_rng(seed)
        out, t = [], 0.0
        for _ in range(n_req):
            t += float(g.exponential(1.0 / rate))
            sid = int(g.integers(len(resident)))
            nl = len(resident[sid][0].lights)
            k = int(g.integers(1, min(4, nl) + 1))
            moves = [(int(li), g.uniform(-64, 64, 3).astype(np.float32))
                     for li in g.choice(nl, size=k, replace=False)]
            out.append((t, sid, moves))
        return out

    trace = make_trace(args.serve_requests, args.serve_rate, 2027)

    def serve(mode, max_active, trace=trace, keep=None):
        sv = keep or Server(mode=mode, max_active=max_active, hbm_budget=16 << 30, bounces=8)
        if keep is None:
            for s, env, xf in resident:
                sv.add_scene(env, xf, s.p_origin, s.p_normal, s.p_refl, s.p_area, s.lights)
            sv.wait(sv.submit(0, []))                   # first request pays set-up
        st0 = sv.stats()
        progress(f"serving trace, mode {mode}: {len(trace)} requests")
        tickets, t0 = [], time.perf_counter()
        for i, (at, sid, moves) in enumerate(trace):
            if i % 200 == 199:
                progress(f"  {i + 1} submitted, {sv.stats()['completed']} done")
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(sv.submit(sid, moves))
        res = [sv.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        st = sv.stats()
        lat = np.array([r[2] for r in res])
        it = int(st["iterations"] - st0["iterations"])
        out = {"p50_ms": round(float(np.percentile(lat, 50)), 2),
               "p99_ms": round(float(np.percentile(lat, 99)), 2),
               "mean_ms": round(float(lat.mean()), 2),
               "throughput_rps": round(len(res) / wall, 2), "wall_s": round(wall, 2),
               "max_active_seen": int(st["max_active_seen"]), "iterations": it}
        if mode == BATCHED and it > 0:
            # request-bounces per batched launch (8 bounces per request)
            out["mean_requests_per_launch"] = round(8 * len(res) / it, 2)
        return sv, out, res

    def service_times(sv, n_req):
        """Closed loop: each request submitted after the previous finished,
        so its latency is its service time (no queueing)."""
        lat = []
        for _, sid, moves in make_trace(n_req, 1.0, 4242):
            lat.append(sv.wait(sv.submit(sid, moves))[2])
        lat = np.array(lat)
        return {"requests": n_req, "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "mean_ms": round(float(lat.mean()), 2)}

    sv, batched, res_b = serve(BATCHED, 16)
    # alone vs mix: a sample of the trace's requests re-run one at a time
    same = True
    for j in range(0, len(trace), max(1, len(trace) // 16)):
        _, sid, moves = trace[j]
        tot, en, _ = sv.wait(sv.submit(sid, moves))
        same = same and np.array_equal(tot.view(np.uint32), res_b[j][0].view(np.uint32)) \
            and np.array_equal(en, res_b[j][1])
    # heavier arrival rates, where continuous batching engages
    sweep = {}
    for rate in args.serve_sweep:
        _, r, _ = serve(BATCHED, 16, make_trace(args.serve_sweep_requests, rate, 3000 + int(rate)),
                        keep=sv)
        sweep[f"{rate:g}_rps"] = r
    svc_batched = service_times(sv, args.serve_service_requests)
    sv.close()
    # luxel error of a served FP16 re-bake against the FP64 recurrence (same
    # direct light, FP32 coefficients), on the smallest scene
    import oracle
    js = next(j for j, (_, sid, _) in enumerate(trace) if sid == small)
    s_sm = resident[small][0]
    lights = s_sm.lights.copy()
    for li, d in trace[js][2]:
        lights["origin"][li] += d
    direct = oracle.gather_light(oracle.KdTree(s_sm.verts9), oracle.precompute(s_sm.verts9, s_sm.ids),
                                 s_sm.p_origin, s_sm.p_normal, lights, threads=16)
    emit0 = (direct * s_sm.p_refl).astype(np.float32)
    rp, col, coef = csr32
    t64, _, _ = oracle.bounce_f64(rp, col, coef, emit0, s_sm.p_refl, s_sm.p_area,
                                  total=direct.astype(np.float64), max_bounces=8, stop_frac=0.0)
    got = res_b[js][0].astype(np.float64)
    nz = t64 != 0
    lux_err = float(np.max(np.abs(got[nz] - t64[nz]) / np.abs(t64[nz]))) if nz.any() else 0.0
    lux_err_zero_ok = bool(np.all(got[~nz] == 0))
    del res_b
    # the plain server on the head of the same trace: its service time is
    # above the mean arrival gap, so its queue (and p99) grows with the
    # trace's length; the closed-loop service times compare without queueing
    svp, plain, _ = serve(PLAIN, 1, trace[:args.serve_plain_requests])
    svc_plain = service_times(svp, args.serve_service_requests)
    svp.close()
    sizes = [len(r[0].p_area) for r in resident]
    nnz = [r[2].info().nnz for r in resident]
    for _, _, xf in resident:
        xf.close()
    return {"requests": len(trace), "rate_rps": args.serve_rate, "scenes": len(resident),
            "patches": [int(min(sizes)), int(max(sizes))], "transfers_total": int(sum(nnz)),
            "batched": batched, "batched_rate_sweep": sweep,
            "plain_sequential": dict(plain, trace_requests=len(trace[:args.serve_plain_requests]),
                                     note="arrivals exceed its capacity: p99 grows with the "
                                          "trace length"),
            "p99_ratio_plain_over_batched": round(plain["p99_ms"] / batched["p99_ms"], 1),
            "service_time": {"batched": svc_batched, "plain": svc_plain,
                             "p99_ratio_plain_over_batched":
                                 round(svc_plain["p99_ms"] / svc_batched["p99_ms"], 1),
                             "p50_ratio_plain_over_batched":
                                 round(svc_plain["p50_ms"] / svc_batched["p50_ms"], 1)},
            "p99_ratio_plain_service_over_batched_under_load":
                round(svc_plain["p99_ms"] / batched["p99_ms"], 1),
            "alone_equals_mix": bool(same),
            "luxel_rel_err_fp16_vs_fp64": lux_err, "luxel_zero_where_fp64_zero": lux_err_zero_ok,
            "luxel_err_sample": f"request {js} of the trace on the smallest scene "
                                f"({len(s_sm.p_area)} patches), 8 bounces"}


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
    ap.add_argument("--no-room", action="store_true", help="skip the closed-room full bake")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched bake")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving trace")
    ap.add_argument("--no-pvs", action="store_true", help="skip the PVS-culled MakeTransfers")
    ap.add_argument("--no-luxels", action="store_true", help="skip the luxel stage")
    ap.add_argument("--no-big-level", action="store_true",
                    help="skip the 16x16-room level (1.55 M triangles)")
    ap.add_argument("--no-level", action="store_true",
                    help="skip the 8x8-room level bake end to end")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--serve-scenes", type=int, default=32)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=[100.0, 300.0],
                    help="further arrival rates for the batched server")
    ap.add_argument("--serve-sweep-requests", type=int, default=600)
    ap.add_argument("--serve-plain-requests", type=int, default=300)
    ap.add_argument("--serve-service-requests", type=int, default=60)
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
    big = None if (args.no_trace or args.no_big_level) else run_level16x16(cx, args)
    bounce = None if args.no_bounce else run_bounce(cx, args)
    room = None if args.no_room else run_room_bake(cx, args)
    batch = None if args.no_batch else run_batch_bake(cx, args)
    serving = None if args.no_serve else run_serving(cx, args)
    pvs_out = None if args.no_pvs else run_pvs(cx, args)
    level = None if args.no_level else run_level_bake(cx, args)
    lux = None if args.no_luxels else run_luxel_stage(cx, args)

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
        "room_bake": room, "batch_bake": batch, "serving": serving, "pvs_culling": pvs_out,
        "level_bake": level, "level_16x16": big,
        "luxel_stage": lux,
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
            "visibility": dict(v["extra"], **({"work": v["work"]} if "work" in v else {})),
        })
        out["config"].update({"triangles": int(info.n_tri), "patches": int(len(v["s"].p_area)),
                              "kd_nodes": int(info.n_nodes), "kd_leaf_refs": int(info.n_leaf_refs),
                              "kd_build_s": round(v["build_s"], 2), "segments_per_step": n,
                              "visible_fraction": round(float(v["vis"].mean()), 4)})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
