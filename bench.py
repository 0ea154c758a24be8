#!/usr/bin/env python3
"""Benchmark of the batch pipeline on one MI355X.

Workload: 4096 x 4096 image, 200,000 distinct uniform seeds (four seed sets,
default_rng 2, 12, 22 and 32, rotated across the steps), default border
padding 16, full pipeline per step -- seed ordering, the padded
Voronoi flood fill, triangle extraction with first-registration numbering,
per-pixel triangle assignment and the (H, W, 3) int32 output -- with the seeds
already resident in HBM and the output left in HBM
(dl_triangulate_device).  Every step's output is bit-exact with the reference
NumPy path (tests/test_gpu_parity.py and tests/test_gpu_large.py).

Prints one JSON line (rank 0, N = 1): value = Mpixel/s over the timed steps; a
`roofline` object for the dominant kernel from HIP-event kernel timings over a
second run of the same steps (events around every kernel on its stream); a
`cpu_baseline` object timing the CPU oracle (the C restatement of the
reference path) on a bounded sample on this host; and the batched, incremental
and serving workloads.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4   # wave64 VALU instructions, G/s
KINDS = ["fill", "seeds", "detect", "register", "bin", "assign_output", "init"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def graph_stats(lib, ctx):
    from integer_delauney_amd import _lib
    c, r = ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.check(lib.dl_get_graph_stats(ctx.handle, ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


def kernel_stats(lib, ctx):
    counts = (ctypes.c_int64 * 7)()
    ms = (ctypes.c_double * 7)()
    lib.dl_get_kernel_stats(ctx.handle, counts, ms)
    return {KINDS[i]: (int(counts[i]), float(ms[i])) for i in range(7)}


def _oracle_rate(side: int, n: int, rng: int, threads: int, min_seconds: float):
    """Mpixel/s of the C oracle's full pipeline on a side x side image with n
    uniform seeds, repeated until min_seconds of work."""
    from integer_delauney_amd.workloads import uniform_seeds
    from oracle import oracle as O

    seeds = uniform_seeds(side, side, n, rng)
    used = O.set_threads(threads)
    reps, t = 0, 0.0
    while t < min_seconds or reps == 0:
        t0 = time.perf_counter()
        O.triangulate(side, side, seeds, None, assign_method="raster")
        t += time.perf_counter() - t0
        reps += 1
    O.set_threads(0)
    return side * side * reps / t / 1e6, used, reps, t


def cpu_baselines():
    """The C oracle (the CPU restatement of the reference path) on this host
    at the 4096^2 / 200k workload and at the 256^2 / 1,000-seed one, with one
    thread and with every host core (OpenMP: flood-fill passes split by rows /
    column strips, raster by row bands; detection stays sequential, as the
    registration order requires), plus the reference NumPy path's own time on
    the 256^2 workload, measured in the build container (the reference does
    not travel to this host)."""
    import json as _json

    out = []
    for side, n, rng, label in ((4096, 200_000, 2, "4096x4096, 200000 uniform seeds"),
                                (256, 1000, 1, "256x256, 1000 uniform seeds")):
        for threads in (0, 1):
            rate, used, reps, t = _oracle_rate(side, n, rng, threads, 3.0 if side > 256 else 2.0)
            out.append({"value": rate, "unit": "Mpixel/s", "cores": used, "kind": "port",
                        "sample": f"{label} (default_rng({rng})), full pipeline in the C oracle: "
                                  f"unpadded + padded flood fill, detection, exact raster; "
                                  f"{reps} run(s), {t:.2f} s",
                        "seconds": t, "host_cpu_count": os.cpu_count()})
    path = os.path.join(ROOT, "tests", "golden", "c1_timing.json")
    if os.path.exists(path):
        with open(path) as f:
            c1 = _json.load(f)
        out.append({"value": c1["mpixel_per_s"], "unit": "Mpixel/s", "cores": c1["cores"],
                    "kind": "reference",
                    "sample": f"the reference NumPy path (delauney.reference Voronoi + "
                              f"GridTriangulation) on {c1['workload']}, {c1['total_s']:.1f} s, "
                              f"timed in the build container ({c1['cpu']}, NumPy {c1['numpy']}), "
                              f"not on this host",
                    "seconds": c1["total_s"]})
    return out


def bench_batch(lib, ctx, n_images: int, steps: int):
    """Batched uneven images (512 by default: sides uniform in [128, 2048],
    0.5-2 % seed density, padding 16): the whole batch per step through
    dl_triangulate_batch_device, against the plain single-image path
    (dl_triangulate_device once per image) on the same device-resident
    inputs.  The batched output is compared with the single-image outputs
    bit for bit on the device."""
    import numpy as np
    import torch

    from integer_delauney_amd import _lib
    from integer_delauney_amd.workloads import batch_images

    log(f"generating {n_images} batch images")
    imgs = batch_images(n_images, 4)
    k = len(imgs)
    widths = np.array([w for w, _, _ in imgs], dtype=np.int32)
    heights = np.array([h for _, h, _ in imgs], dtype=np.int32)
    counts = np.array([s.shape[0] for _, _, s in imgs], dtype=np.int64)
    pads = np.full(k, 16, dtype=np.int32)
    xy = np.ascontiguousarray(np.concatenate([s for _, _, s in imgs]))
    px = widths.astype(np.int64) * heights.astype(np.int64)
    total_px = int(px.sum())
    d_xy = torch.from_numpy(xy).to("cuda")
    d_out = torch.empty(3 * total_px, dtype=torch.int32, device="cuda")
    ntri = np.zeros(k, dtype=np.int64)
    I64P = ctypes.POINTER(ctypes.c_int64)

    def batch_step():
        _lib.check(lib.dl_triangulate_batch_device(
            ctx.handle, k, _lib.i32p(widths), _lib.i32p(heights), counts.ctypes.data_as(I64P),
            ctypes.c_void_p(d_xy.data_ptr()), _lib.i32p(pads), ctypes.c_void_p(d_out.data_ptr()),
            None, ntri.ctypes.data_as(I64P)))

    # Warm-up: first call (flood-fill shape tuning on a new canvas class),
    # then predicted fills; the default graph mode (auto: uncaptured at this
    # size).
    for _ in range(3):
        batch_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch_step()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    log("batched steps done")
    ref = d_out.clone()
    ntri_batch = ntri.copy()
    # The same batches as hipGraph replays (captured on the third call).
    _lib.check(lib.dl_set_graphs(ctx.handle, 1))
    for _ in range(3):
        batch_step()
    torch.cuda.synchronize()
    g0 = graph_stats(lib, ctx)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch_step()
    torch.cuda.synchronize()
    t_replayed = (time.perf_counter() - t0) / steps
    replays = graph_stats(lib, ctx)[1] - g0[1]
    _lib.check(lib.dl_set_graphs(ctx.handle, 2))
    same_replayed = bool(torch.equal(ref, d_out))
    log("replayed batched steps done")

    seed_off = np.concatenate([[0], np.cumsum(counts)])
    pix_off = np.concatenate([[0], np.cumsum(px)])
    nt = ctypes.c_int64(0)
    ntri_single = np.zeros(k, dtype=np.int64)

    def single_pass():
        for i in range(k):
            _lib.check(lib.dl_triangulate_device(
                ctx.handle, int(widths[i]), int(heights[i]),
                ctypes.c_void_p(d_xy.data_ptr() + 8 * int(seed_off[i])), int(counts[i]), 16,
                ctypes.c_void_p(d_out.data_ptr() + 12 * int(pix_off[i])), None, ctypes.byref(nt)))
            ntri_single[i] = nt.value

    d_out.fill_(-7)
    single_pass()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    single_pass()
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t1
    log("single-image pass done")
    same = bool(torch.equal(ref, d_out)) and bool(np.array_equal(ntri_batch, ntri_single))
    del ref
    return {"workload": f"{k} images, sides uniform in [128, 2048], seed density uniform in "
                        f"0.5-2 %, padding 16, full pipeline, inputs and outputs in HBM",
            "images": k, "mpixel": total_px / 1e6, "seeds": int(counts.sum()),
            "triangles": int(ntri_batch.sum()), "steps": steps,
            "ms_per_batch": t_batch * 1e3, "batched_mpixel_per_s": total_px / t_batch / 1e6,
            "single_image_path_ms": t_single * 1e3,
            "single_image_path_mpixel_per_s": total_px / t_single / 1e6,
            "speedup_vs_single_image_path": t_single / t_batch,
            "batched_equals_single_bitwise": same,
            "graph_replays_in_timed_steps": replays,
            "replayed_ms_per_batch": t_replayed * 1e3,
            "replayed_equals_uncaptured_bitwise": same_replayed}


def bench_small_images(lib, ctx, steps: int):
    """Small single images through dl_triangulate_device, the regime where the
    host's issue of ~30 launches per call is comparable to the GPU work:
    256 x 256 with 1,000 uniform seeds (the reference's CPU-runnable
    configuration, default_rng(1)) and 1024 x 1024 at the 4096^2 seed density.
    Calls per second with hipGraph replay and with every launch issued from
    the host; the two outputs are compared bit for bit."""
    import torch

    from integer_delauney_amd import _lib
    from integer_delauney_amd.workloads import uniform_seeds

    res = []
    for side, n, rng in ((256, 1000, 1), (1024, 12500, 2)):
        seeds = uniform_seeds(side, side, n, rng)
        d_xy = torch.from_numpy(seeds).to("cuda")
        d_out = torch.empty((side, side, 3), dtype=torch.int32, device="cuda")
        nt = ctypes.c_int64(0)

        def call():
            _lib.check(lib.dl_triangulate_device(ctx.handle, side, side,
                                                 ctypes.c_void_p(d_xy.data_ptr()), n, 16,
                                                 ctypes.c_void_p(d_out.data_ptr()), None,
                                                 ctypes.byref(nt)))

        out = {}
        for graphs in (1, 0):
            _lib.check(lib.dl_set_graphs(ctx.handle, graphs))
            for _ in range(4):
                call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                call()
            torch.cuda.synchronize()
            out[graphs] = ((time.perf_counter() - t0) / steps * 1e3, d_out.clone())
        _lib.check(lib.dl_set_graphs(ctx.handle, 2))
        res.append({"side": side, "seeds": n, "triangles": int(nt.value), "calls": steps,
                    "graph_ms_per_call": out[1][0], "uncaptured_ms_per_call": out[0][0],
                    "graph_mpixel_per_s": side * side / out[1][0] / 1e3,
                    "uncaptured_mpixel_per_s": side * side / out[0][0] / 1e3,
                    "equal_bitwise": bool(torch.equal(out[1][1], out[0][1]))})
        log(f"small images {side}^2 done")
    return res


def bench_incremental(lib, ctx, side: int, n0: int, rounds: int, n_ins: int, n_del: int,
                      record: bool = True):
    """Clustered seeds + incremental updates: a side x side session with n0
    seeds from a 64-component Gaussian mixture (sigma 200 px), then `rounds`
    rounds of n_ins new mixture samples and n_del deletions drawn uniformly
    from the live set, each round finalised on the device (seed set, grid and
    rows stay in HBM).  The final grid is compared bit for bit with a fresh
    batch run on the final seed set."""
    import numpy as np
    import torch

    from integer_delauney_amd import _lib
    from integer_delauney_amd.workloads import clustered_seeds

    log(f"generating {n0} clustered seeds on {side}x{side}")
    rng = np.random.default_rng(3)
    seeds0 = clustered_seeds(side, side, n0, 3)
    centres = np.random.default_rng(3).uniform([0, 0], [side, side], size=(64, 2))
    live = seeds0.copy()
    taken = np.zeros(side * side, dtype=bool)
    taken[live[:, 0].astype(np.int64) * side + live[:, 1]] = True
    plan = []
    for r in range(rounds):
        ins = []
        while len(ins) < n_ins:
            k = 2 * (n_ins - len(ins)) + 64
            comp = rng.integers(0, 64, size=k)
            pts = np.rint(centres[comp] + rng.normal(0.0, 200.0, size=(k, 2))).astype(np.int64)
            ok = (pts[:, 0] >= 0) & (pts[:, 0] < side) & (pts[:, 1] >= 0) & (pts[:, 1] < side)
            for x, y in pts[ok].tolist():
                key = x * side + y
                if not taken[key]:
                    taken[key] = True
                    ins.append((x, y))
                    if len(ins) == n_ins:
                        break
        ins = np.asarray(ins, dtype=np.int32)
        live = np.concatenate([live, ins])
        drop = rng.choice(live.shape[0], size=n_del, replace=False)
        dele = np.ascontiguousarray(live[drop])
        taken[dele[:, 0].astype(np.int64) * side + dele[:, 1]] = False
        keep = np.ones(live.shape[0], dtype=bool)
        keep[drop] = False
        live = live[keep]
        plan.append((np.ascontiguousarray(ins), dele))
    log("update rounds generated")

    h = ctypes.c_void_p()
    _lib.check(lib.dl_session_create(ctx.handle, side, side, n0 + rounds * n_ins, 16,
                                     ctypes.byref(h)))
    if record:
        # keep each finalise's fill trajectory: the next finalise recomputes
        # only the tiles the round's inserts and deletes reach (exact)
        _lib.check(lib.dl_session_set_fill_record(h, 1))
    nt = ctypes.c_int64(0)
    try:
        t0 = time.perf_counter()
        _lib.check(lib.dl_session_insert(ctx.handle, h, _lib.i32p(seeds0), None, seeds0.shape[0]))
        _lib.check(lib.dl_session_finalise(ctx.handle, h, ctypes.byref(nt)))
        t_build = time.perf_counter() - t0
        rs0 = [ctypes.c_int64(0) for _ in range(3)]
        _lib.check(lib.dl_get_record_stats(ctx.handle, *[ctypes.byref(v) for v in rs0]))
        per_round = []
        fill_rounds = []   # per round: fill launches, launches to convergence, tile visits
        for ins, dele in plan:
            t1 = time.perf_counter()
            _lib.check(lib.dl_session_insert(ctx.handle, h, _lib.i32p(ins), None, ins.shape[0]))
            _lib.check(lib.dl_session_remove(ctx.handle, h, _lib.i32p(dele), dele.shape[0]))
            _lib.check(lib.dl_session_finalise(ctx.handle, h, ctypes.byref(nt)))
            per_round.append(time.perf_counter() - t1)
            la, cv, vi = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
            _lib.check(lib.dl_get_fill_stats(ctx.handle, ctypes.byref(la), ctypes.byref(cv),
                                             ctypes.byref(vi), None, None))
            rv, re_, rc = ctypes.c_int(0), ctypes.c_int64(0), ctypes.c_int64(0)
            _lib.check(lib.dl_session_fill_record_stats(h, ctypes.byref(rv), ctypes.byref(re_),
                                                        ctypes.byref(rc)))
            fill_rounds.append((la.value, cv.value, vi.value, rc.value, rv.value))
        log("update rounds done")
        rs1 = [ctypes.c_int64(0) for _ in range(3)]
        _lib.check(lib.dl_get_record_stats(ctx.handle, *[ctypes.byref(v) for v in rs1]))
        retained_rounds = int(rs1[2].value - rs0[2].value)
        # bit-exact check against a fresh batch run on the final seed set
        n_live = ctypes.c_int64(0)
        _lib.check(lib.dl_session_seed_count(h, ctypes.byref(n_live)))
        d_xy = torch.from_numpy(live).to("cuda")
        d_ref = torch.empty((side, side, 3), dtype=torch.int32, device="cuda")
        nt2 = ctypes.c_int64(0)
        # (with stage timing on, which also counts the fill's tile visits; every
        # round's fill is this same full fill of its live set)
        _lib.check(lib.dl_set_timing(ctx.handle, 1))
        _lib.check(lib.dl_triangulate_device(ctx.handle, side, side, ctypes.c_void_p(d_xy.data_ptr()),
                                             live.shape[0], 16, ctypes.c_void_p(d_ref.data_ptr()),
                                             None, ctypes.byref(nt2)))
        torch.cuda.synchronize()
        la, cv, vi = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.check(lib.dl_get_fill_stats(ctx.handle, ctypes.byref(la), ctypes.byref(cv),
                                         ctypes.byref(vi), None, None))
        _lib.check(lib.dl_set_timing(ctx.handle, 0))
        batch_fill = (la.value, cv.value, vi.value)
        # the session's grid through the host API (12 B/px device-to-host)
        host = np.empty((side, side, 3), dtype=np.int32)
        _lib.check(lib.dl_session_copy_result(h, _lib.i32p(host), None))
        same = bool(torch.equal(torch.from_numpy(host).to("cuda"), d_ref)) and nt.value == nt2.value \
            and n_live.value == live.shape[0]
    finally:
        lib.dl_session_destroy(h)
    pr = np.array(per_round)
    return {"workload": f"{side}x{side}, {n0} seeds from a 64-component Gaussian mixture "
                        f"(sigma 200 px), {rounds} rounds of {n_ins} inserts + {n_del} deletes, "
                        f"each round finalised on the device, padding 16",
            "initial_build_ms": t_build * 1e3, "rounds": rounds,
            "round_ms_mean": float(pr.mean() * 1e3), "round_ms_p50": float(np.percentile(pr, 50) * 1e3),
            "round_ms_p99": float(np.percentile(pr, 99) * 1e3),
            "effective_mpixel_per_s": side * side / pr.mean() / 1e6,
            "final_seeds": int(live.shape[0]), "final_triangles": int(nt.value),
            "fill_launches_per_round": float(np.mean([f[0] for f in fill_rounds])),
            "fill_launches_to_convergence_per_round": float(np.mean([f[1] for f in fill_rounds])),
            "fill_record": bool(record),
            "fill_tile_computations_per_round": float(np.mean([f[3] for f in fill_rounds]))
            if record else None,
            "rounds_with_a_record": int(sum(f[4] for f in fill_rounds)) if record else None,
            "rounds_replayed_from_the_retained_plane": retained_rounds if record else None,
            "fill_tile_visits_final_batch": batch_fill[2],
            "fill_launches_final_batch": batch_fill[0],
            "final_equals_batch_bitwise": same}


def bench_published(reps: int):
    """The inputs of the reference's own published measurements (its CUDA path
    on an unspecified NVIDIA GPU, API_REFERENCE.md:36 and 294-300), timed here
    through the same Python API with host arrays in and out: triangulate() and
    the Voronoi + GridTriangulation pair at 1510 x 1018 with 30,000 seeds, and
    the incremental workload on 1240 x 1240 (22,000 initial seeds, then five
    scattered batches of 1,800): insert() per batch (six calls), six
    insert_deferred() calls and one finalise(), and one stateless rebuild of
    the final set.  Seeds are uniform (the reference does not publish its
    draws); padding is the default."""
    import numpy as np

    import integer_delauney_amd as dl
    from integer_delauney_amd.workloads import uniform_seeds

    def best(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts) * 1e3)

    W, H = 1510, 1018
    s = uniform_seeds(W, H, 30_000, 5)
    tri_ms = best(lambda: dl.triangulate(W, H, s))
    tri_arrays_ms = best(lambda: dl.triangulate(W, H, s, as_arrays=True))
    pair_ms = best(lambda: dl.GridTriangulation().compute(dl.Voronoi().compute(W, H, s), s))
    S = 1240
    allp = uniform_seeds(S, S, 22_000 + 5 * 1_800, 6)
    batches = [allp[:22_000]] + [allp[22_000 + 1_800 * k: 22_000 + 1_800 * (k + 1)] for k in range(5)]

    def per_round():
        m = dl.Delaunay(S, S, allp.shape[0])
        for b in batches:
            m.insert(b, as_arrays=True)

    def deferred():
        m = dl.Delaunay(S, S, allp.shape[0])
        for b in batches:
            m.insert_deferred(b)
        m.finalise(as_arrays=True)

    ins_ms = best(per_round)
    def_ms = best(deferred)
    rebuild_ms = best(lambda: dl.triangulate(S, S, allp, as_arrays=True))
    log("published-workload timings done")
    return {"note": "inputs of the reference's published CUDA measurements (unspecified NVIDIA "
                    "GPU, API_REFERENCE.md:36, 294-300); here on one MI355X through the Python "
                    "API, host arrays in and out, median of %d" % reps,
            "triangulate_1510x1018_30k_ms": tri_ms, "reference_ms": 116.8,
            "triangulate_1510x1018_30k_as_arrays_ms": tri_arrays_ms,
            "voronoi_plus_grid_triangulation_ms": pair_ms, "reference_pair_ms": 157.0,
            "incremental_1240_insert_x6_ms": ins_ms, "reference_insert_x6_ms": 181.0,
            "incremental_1240_deferred_x6_finalise_ms": def_ms, "reference_deferred_ms": 76.0,
            "incremental_1240_stateless_rebuild_ms": rebuild_ms, "reference_rebuild_ms": 103.0}


def bench_queries(side: int, n_seeds: int, n_points: int, reps: int):
    """Session queries on one image: locate / in_circumsphere over n_points
    host points, the edge list, edge scores, best-k midpoint selection and
    finalise_device, each through the Python API (host arrays in and out)."""
    import numpy as np

    import integer_delauney_amd as dl
    from integer_delauney_amd.workloads import uniform_seeds

    log(f"queries: {side}x{side}, {n_seeds} seeds, {n_points} points")
    rng = np.random.default_rng(7)
    seeds = uniform_seeds(side, side, n_seeds, 2)
    vals = rng.random(n_seeds).astype(np.float32)
    mesh = dl.Delaunay(side, side, n_seeds, 16)
    mesh.insert_deferred(seeds, vals)
    mesh.finalise_device()
    q = rng.integers(0, side, size=(n_points, 2)).astype(np.int32)
    t = rng.random(n_points) * 4.0

    def timed(fn):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        return (time.perf_counter() - t0) / reps * 1e3, r

    out = {"side": side, "seeds": n_seeds, "points": n_points}
    ms, _ = timed(lambda: mesh.locate(q))
    out["locate_ms"] = ms
    out["locate_mpoints_per_s"] = n_points / ms / 1e3
    ms, _ = timed(lambda: mesh.in_circumsphere(q, t))
    out["in_circumsphere_ms"] = ms
    out["in_circumsphere_mpoints_per_s"] = n_points / ms / 1e3

    # The edge list is built on first use after a change and cached: time the
    # build (first call) and a cached call.
    t0 = time.perf_counter()
    e = mesh.get_edges()
    out["get_edges_build_ms"] = (time.perf_counter() - t0) * 1e3
    out["edges"] = int(e.shape[0])
    ms, _ = timed(lambda: mesh.get_edges())
    out["get_edges_cached_ms"] = ms
    ms, _ = timed(lambda: mesh.edge_scores(2.0))
    out["edge_scores_ms"] = ms
    for k in (10_000, 100_000):
        ms, mid = timed(lambda: mesh.select_midpoints(2.0, k))
        out[f"select_{k}_ms"] = ms
        out[f"select_{k}_returned"] = int(mid.shape[0])
    ms, _ = timed(lambda: mesh.finalise_device())
    out["finalise_device_split_ms"] = ms
    mesh.close()
    log("queries done")
    return out


def bench_serving(n_requests: int, n_sessions: int, slots: int = 3, loads=(0.3, 0.6, 0.9)):
    """Serving loop (integer_delauney_amd.serving): requests with sides
    log-uniform in [64, 8192] (width and height independent) at 1 % seed
    density, 20 % of them insert/delete batches of 1-5 % of a resident
    session's seeds (half inserts, half deletes).  First the saturated
    throughput (all requests queued at once), then Poisson arrivals at three
    offered loads (fractions of that throughput) with p50/p99 latency from
    arrival to completion."""
    import numpy as np

    from integer_delauney_amd.serving import Server, TriangulateRequest, UpdateRequest
    from integer_delauney_amd.workloads import _uniform_fast

    rng = np.random.default_rng(5)

    def side():
        return int(round(np.exp(rng.uniform(np.log(64), np.log(8192)))))

    srv = Server(hbm_budget=160 << 30, slots=slots)
    model = {}
    for _ in range(n_sessions):
        w, h = side(), side()
        n0 = max(3, int(0.01 * w * h))
        s0 = _uniform_fast(w, h, n0, rng)
        sid = srv.create_session(w, h, 2 * n0 + 16, seeds=s0)
        taken = np.zeros(w * h, dtype=bool)
        taken[s0[:, 0].astype(np.int64) * h + s0[:, 1]] = True
        model[sid] = {"w": w, "h": h, "live": s0, "taken": taken, "cap": 2 * n0 + 16}
    log(f"serving: {n_sessions} resident sessions created")

    def next_request():
        if rng.random() < 0.2:
            sid = int(rng.integers(0, n_sessions))
            m = model[sid]
            live = m["live"]
            k = max(2, int(round(rng.uniform(0.01, 0.05) * live.shape[0])))
            k_del = min(k // 2, live.shape[0] - 3)
            k_ins = min(k - k // 2, m["cap"] - live.shape[0] + k_del)
            idx = rng.choice(live.shape[0], size=k_del, replace=False)
            drop = np.ascontiguousarray(live[idx])
            m["taken"][drop[:, 0].astype(np.int64) * m["h"] + drop[:, 1]] = False
            add = []
            while len(add) < k_ins:
                x, y = int(rng.integers(0, m["w"])), int(rng.integers(0, m["h"]))
                if not m["taken"][x * m["h"] + y]:
                    m["taken"][x * m["h"] + y] = True
                    add.append((x, y))
            add = np.asarray(add, dtype=np.int32).reshape(-1, 2)
            keep = np.ones(live.shape[0], dtype=bool)
            keep[idx] = False
            m["live"] = np.concatenate([live[keep], add])
            return UpdateRequest(sid, insert=add, remove=drop), m["w"] * m["h"]
        w, h = side(), side()
        n = max(3, int(0.01 * w * h))
        return TriangulateRequest(w, h, _uniform_fast(w, h, n, rng)), w * h

    def run(load_rate=None):
        reqs = [next_request() for _ in range(n_requests)]
        px = sum(p for _, p in reqs)
        futs = []
        b0 = len(srv.batches)
        t0 = time.perf_counter()
        if load_rate is None:
            arrivals = np.zeros(n_requests)
        else:
            arrivals = np.cumsum(rng.exponential(1.0 / load_rate, size=n_requests))
        for (req, _), a in zip(reqs, arrivals):
            while time.perf_counter() - t0 < a:
                time.sleep(min(0.0005, max(0.0, a - (time.perf_counter() - t0))))
            f = srv.submit(req, keep=False)
            f.t_arrival = t0 + a
            futs.append(f)
        lat = []
        for f in futs:
            f.result()
            lat.append(f.t_done - f.t_arrival)
        makespan = max(f.t_done for f in futs) - t0
        del futs
        time.sleep(0.05)   # the last workers file their batch records
        recs = srv.batches[b0:]
        dev = sum(b.gpu_ms for b in recs)
        cpy = sum(b.copy_ms for b in recs)
        lat = np.array(lat) * 1e3
        return {"requests": n_requests, "mpixel": px / 1e6, "makespan_s": makespan,
                "throughput_mpixel_per_s": px / makespan / 1e6,
                "latency_ms_p50": float(np.percentile(lat, 50)),
                "latency_ms_p99": float(np.percentile(lat, 99)),
                "latency_ms_max": float(lat.max()),
                "batches": len(recs),
                "pipeline_device_ms_sum": dev, "copy_back_ms_sum": cpy,
                "copy_back_share": cpy / (dev + cpy) if dev + cpy else None,
                "rejected": srv.rejected}

    def run_plain():
        """The same kind of stream, one request at a time through the plain
        host API (triangulate per image; session updates and finalise per
        request), no batching, no overlap."""
        import ctypes as ct

        from integer_delauney_amd import _lib, triangulate
        lib = _lib.load()
        ctx = _lib.context()
        reqs = [next_request() for _ in range(n_requests)]
        px = sum(p for _, p in reqs)
        t0 = time.perf_counter()
        for req, _ in reqs:
            if isinstance(req, UpdateRequest):
                h = srv._sessions[req.session_id].handle
                _lib.check(lib.dl_session_update(ctx.handle, h, _lib.i32p(req.remove),
                                                 req.remove.shape[0], _lib.i32p(req.insert),
                                                 None, req.insert.shape[0]))
                n = ct.c_int64()
                _lib.check(lib.dl_session_finalise(ctx.handle, h, ct.byref(n)))
                m = model[req.session_id]
                g = np.empty((m["h"], m["w"], 3), dtype=np.int32)
                _lib.check(lib.dl_session_copy_result(h, _lib.i32p(g), None))
            else:
                triangulate(req.width, req.height, req.seeds, as_arrays=True)
        dt = time.perf_counter() - t0
        return {"requests": n_requests, "mpixel": px / 1e6, "seconds": dt,
                "throughput_mpixel_per_s": px / dt / 1e6}

    try:
        run(None)   # warm-up (workspaces grow to their working size)
        sat = run(None)
        log(f"serving: saturated {sat['throughput_mpixel_per_s']:.0f} Mpx/s")
        mean_px = sat["mpixel"] * 1e6 / n_requests
        out = {"workload": f"{n_requests} requests per phase, sides log-uniform in [64, 8192], "
                           f"1 % seed density, 20 % updates (1-5 % of the seeds of one of "
                           f"{n_sessions} resident sessions), padding 16, {slots} worker slots, "
                           f"HBM budget 160 GiB",
               "saturated": sat, "offered": []}
        for f in loads:
            rate = f * sat["throughput_mpixel_per_s"] * 1e6 / mean_px   # requests / s
            r = run(rate)
            r["offered_fraction"] = f
            r["offered_requests_per_s"] = rate
            out["offered"].append(r)
            log(f"serving: load {f}: p50 {r['latency_ms_p50']:.1f} ms p99 {r['latency_ms_p99']:.1f} ms")
        plain = run_plain()
        out["plain_single_image_path"] = plain
        out["saturated_vs_plain"] = sat["throughput_mpixel_per_s"] / plain["throughput_mpixel_per_s"]
        log(f"serving: plain path {plain['throughput_mpixel_per_s']:.0f} Mpx/s")
        out["batches"] = len(srv.batches)
        out["max_batch"] = max(b.requests for b in srv.batches)
        return out
    finally:
        srv.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--bfs-config", type=str, default="", help="waves,rows,iters")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--e2e-steps", type=int, default=3)
    ap.add_argument("--batch-images", type=int, default=512,
                    help="images of the batched workload (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--small-steps", type=int, default=200,
                    help="calls per small-image case, graphs on and off (0: skip)")
    ap.add_argument("--inc-record", type=int, default=1,
                    help="1: the incremental session keeps its fill record (replays only what "
                         "each round's changes reach); 0: every round recomputes the fill")
    ap.add_argument("--inc-rounds", type=int, default=100,
                    help="update rounds of the incremental workload (0: skip it)")
    ap.add_argument("--inc-side", type=int, default=16384)
    ap.add_argument("--inc-seeds", type=int, default=2_000_000)
    ap.add_argument("--serve-requests", type=int, default=300,
                    help="requests per phase of the serving workload (0: skip it)")
    ap.add_argument("--serve-sessions", type=int, default=8)
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--autotune", type=int, default=1,
                    help="0: fixed default flood-fill shape (no first-use timing of the candidates)")
    ap.add_argument("--published-reps", type=int, default=5,
                    help="repetitions of the reference's published workloads (0: skip)")
    ap.add_argument("--query-points", type=int, default=1_000_000,
                    help="points per locate / in_circumsphere call in the query section (0: skip)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mb:
        import torch
        n = args.copy_probe_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "dispatches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (a batch would "
              f"shard by image, one large image by row strips; neither is built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from integer_delauney_amd import _lib
    from integer_delauney_amd.workloads import uniform_seeds

    W = H = 4096
    N = 200_000
    P = 16
    SETS = (2, 12, 22, 32)   # default_rng seeds of the rotated seed sets
    log("generating seeds")
    seed_sets = [uniform_seeds(W, H, N, r) for r in SETS]
    seeds = seed_sets[0]
    lib = _lib.load()
    ctx = _lib.context()
    _lib.check(lib.dl_set_autotune(ctx.handle, 1 if args.autotune else 0))
    if args.bfs_config:
        w, r, k = (int(v) for v in args.bfs_config.split(","))
        _lib.check(lib.dl_set_bfs_config(ctx.handle, w, r, k))
    # Four distinct seed sets, resident in HBM before timing; step i uses set
    # i mod 4, so the fill's launch-count prediction meets inputs it has not
    # seen in the previous call.
    d_sets = [torch.from_numpy(x).to("cuda") for x in seed_sets]
    d_out = torch.empty((H, W, 3), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    n_tri = ctypes.c_int64(0)
    tri_counts = {}

    def step(i, d_xy=None):
        d = d_sets[i % len(d_sets)] if d_xy is None else d_xy
        _lib.check(lib.dl_triangulate_device(ctx.handle, W, H, ctypes.c_void_p(d.data_ptr()),
                                             N, P, ctypes.c_void_p(d_out.data_ptr()), None,
                                             ctypes.byref(n_tri)))
        tri_counts[i % len(d_sets)] = int(n_tri.value)

    def shortfalls():
        sf = ctypes.c_int64(0)
        _lib.check(lib.dl_get_fill_stats(ctx.handle, None, None, None, None, ctypes.byref(sf)))
        return int(sf.value)

    for i in range(args.warmup):
        step(i)
        log(f"warmup {i + 1}/{args.warmup} done")
    # Timed region: the pipeline alone, in the library's default graph mode
    # (auto: at this size every launch is issued from the host).
    g_main = graph_stats(lib, ctx)
    sf0 = shortfalls()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"timed {args.steps} steps")
    timed_shortfalls = shortfalls() - sf0
    default_replays = graph_stats(lib, ctx)[1] - g_main[1]
    ms_per_step = (t1 - t0) / args.steps * 1e3
    value = W * H / (ms_per_step / 1e3) / 1e6
    # The same steps as hipGraph replays (graphs on: the third call of the
    # signature is captured, the timed ones replay it).  A graph holds the
    # seed buffer's address, so here each step first copies its seed set
    # into one fixed buffer (a 1.6 MB device copy, inside the timed loop).
    d_fixed = torch.empty_like(d_sets[0])
    _lib.check(lib.dl_set_graphs(ctx.handle, 1))
    for i in range(3):
        d_fixed.copy_(d_sets[i % len(d_sets)])
        step(i, d_fixed)
    g_before = graph_stats(lib, ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        d_fixed.copy_(d_sets[i % len(d_sets)])
        step(i, d_fixed)
    torch.cuda.synchronize()
    replayed_ms = (time.perf_counter() - t0) / args.steps * 1e3
    g_after = graph_stats(lib, ctx)
    _lib.check(lib.dl_set_graphs(ctx.handle, 2))
    log("replayed steps done")
    graphs = {"default_mode": "auto (capture below 2^22 canvas cells)",
              "replays_in_default_steps": default_replays,
              "captures": g_after[0], "replays_in_timed_steps": g_after[1] - g_before[1],
              "replayed_ms_per_step": replayed_ms,
              "replayed_mpixel_per_s": W * H / (replayed_ms / 1e3) / 1e6}
    # Kernel-timing region: the same steps again with a pair of HIP events
    # around every kernel, on the stream the kernels run on (the events add
    # host work, so this region is not the one `value` comes from).
    _lib.check(lib.dl_reset_kernel_stats(ctx.handle))
    _lib.check(lib.dl_set_kernel_profiling(ctx.handle, 1))
    torch.cuda.synchronize()
    fill_launches = fill_visits = fill_converged = 0
    t2 = time.perf_counter()
    for i in range(args.steps):
        step(i)
        la, cv, vi = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.check(lib.dl_get_fill_stats(ctx.handle, ctypes.byref(la), ctypes.byref(cv),
                                         ctypes.byref(vi), None, None))
        fill_launches += la.value
        fill_converged += cv.value
        fill_visits += vi.value
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    _lib.check(lib.dl_set_kernel_profiling(ctx.handle, 0))
    profiled_ms_per_step = (t3 - t2) / args.steps * 1e3
    stats = kernel_stats(lib, ctx)

    # Algorithmic bytes per launch of each kernel kind.
    Wp, Hp = W + 2 * P, H + 2 * P
    T = int(n_tri.value)
    cw, cr, ck = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(lib.dl_get_bfs_config(ctx.handle, ctypes.byref(cw), ctypes.byref(cr), ctypes.byref(ck)))
    cfg_w, cfg_r, cfg_k = cw.value, cr.value, ck.value
    tile_cells = (128 - 2 * cfg_k) * (cfg_w * cfg_r - 2 * cfg_k)
    alg = {
        # 4 B read + 4 B written per interior cell of every visited tile,
        # per launch (the kernel events count every launch issued)
        "fill": 8.0 * tile_cells * fill_visits / max(stats["fill"][0], 1),
        # labels read once (4 B/cell) + candidate records written/read (20 B each)
        "detect": 4.0 * Wp * Hp,
        # fused assignment + output: (H, W, 3) int32 written once, the label
        # plane read once, triangle rows and vertex labels read (32 B each)
        "assign_output": (12.0 + 4.0) * W * H + 32.0 * T,
    }
    per_kernel = {}
    for k, (c, t) in stats.items():
        if c:
            per_kernel[k] = {"launches_per_step": c / args.steps, "ms_per_step": t / args.steps,
                             "avg_launch_us": t / c * 1e3}
    dominant = max(per_kernel, key=lambda k: per_kernel[k]["ms_per_step"])
    dom = per_kernel[dominant]
    achieved = alg.get(dominant, 0.0) / (dom["avg_launch_us"] * 1e-6) / 1e9
    # HBM bytes per launch of the dominant kernel, from the FETCH_SIZE and
    # WRITE_SIZE counter runs summarised under profiles/ (valid for the
    # flood-fill configuration they were taken with).
    traffic = None
    traffic_src = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        ent = tj.get(dominant)
        if ent and (dominant != "fill" or ent.get("config") == [cfg_w, cfg_r, cfg_k]):
            traffic = ent["hbm_bytes_per_launch"]
            traffic_src = tj.get("source")
    # VALU issue of the dominant kernel: wave instructions per launch from
    # the SQ counter runs summarised under profiles/, over the live launch
    # time; peak = 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU
    # instruction (16 lanes per SIMD) = 614.4 G instructions/s.
    valu = None
    vpath = os.path.join(ROOT, "profiles", "valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            vj = json.load(f)
        ent = vj.get(dominant)
        if ent and (dominant != "fill" or ent.get("config") == [cfg_w, cfg_r, cfg_k]):
            ach = ent["valu_insts_per_launch"] / (dom["avg_launch_us"] * 1e-6) / 1e9
            valu = {"kernel": dominant, "bound": "valu", "achieved": ach,
                    "peak": VALU_PEAK_GINST, "unit": "G wave-instructions/s",
                    "frac": ach / VALU_PEAK_GINST,
                    "valu_insts_per_launch": ent["valu_insts_per_launch"],
                    "source": vj.get("source")}
    b_alg_image = 12.0 * W * H + 8.0 * Wp * Hp + 8.0 * N + 32.0 * T

    # End to end through the host API: H2D of the seeds, D2H of the grid.
    e2e = None
    if args.e2e_steps > 0:
        from integer_delauney_amd import triangulate
        triangulate(W, H, seeds, P, as_arrays=True)
        t2 = time.perf_counter()
        for i in range(args.e2e_steps):
            triangulate(W, H, seed_sets[i % len(seed_sets)], P, as_arrays=True)
        t3 = time.perf_counter()
        e2e = W * H / ((t3 - t2) / args.e2e_steps) / 1e6
        log("end-to-end steps done")

    small = bench_small_images(lib, ctx, args.small_steps) if args.small_steps > 0 else None

    batch = None
    if args.batch_images > 0:
        batch = bench_batch(lib, ctx, args.batch_images, args.batch_steps)

    incremental = None
    if args.inc_rounds > 0:
        incremental = bench_incremental(lib, ctx, args.inc_side, args.inc_seeds, args.inc_rounds,
                                        10_000, 5_000, record=args.inc_record != 0)

    serving = None
    if args.serve_requests > 0:
        serving = bench_serving(args.serve_requests, args.serve_sessions, args.serve_slots)

    published = bench_published(args.published_reps) if args.published_reps > 0 else None
    queries = None
    if args.query_points > 0:
        queries = bench_queries(W, N, args.query_points, 5)

    cpu = cpus = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle samples")
        cpus = cpu_baselines()
        cpu = cpus[0]   # the 4096^2 workload on every host core

    line = {
        "metric": "Mpixel/s full pipeline (bit-exact) + % HBM roofline; p50/p99 latency served",
        "value": value,
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "single image 4096x4096, 200000 distinct uniform seeds, "
                               "border padding 16, full pipeline (Voronoi + Delaunay "
                               "extraction + pixel-to-triangle assignment), inputs and outputs "
                               "resident in HBM; four seed sets (default_rng 2, 12, 22, 32) "
                               "rotated across the steps",
                   "width": W, "height": H, "seeds": N, "border_padding": P,
                   "seed_sets": list(SETS), "triangles": T,
                   "triangles_per_set": [tri_counts.get(k) for k in range(len(SETS))],
                   "parallelism": "single GPU"},
        "fill_prediction_shortfalls_in_timed_steps": timed_shortfalls,
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "alg_bytes_per_launch": alg.get(dominant),
                     "avg_launch_us": dom["avg_launch_us"]},
        "roofline_valu": valu,
        "graphs": graphs,
        "kernel_timing_region_ms_per_step": profiled_ms_per_step,
        "pipeline_alg_bytes_per_image": b_alg_image,
        "pipeline_hbm_frac": b_alg_image / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "kernels": per_kernel,
        "fill": {"launches_per_step": fill_launches / args.steps,
                 "launches_to_convergence_per_step": fill_converged / args.steps,
                 "tile_visits_per_step": fill_visits / args.steps,
                 "config": [cfg_w, cfg_r, cfg_k]},
        "end_to_end_mpixel_per_s": e2e,
        "small_images": small,
        "batched": batch,
        "incremental": incremental,
        "serving": serving,
        "queries": queries,
        "published_workloads": published,
        "cpu_baseline": cpu,
        "cpu_baselines": cpus,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
