"""The clustered incremental workload: seeds from a Gaussian mixture, then
rounds of inserts and deletes against a resident device session.

The reference's session semantics are insert_deferred / finalise
(src/cuda/delaunay.cu:184-372); its incremental test holds each round to the
batch result (tests/test_incremental_cuda.py:232-313).  Here each finalise
must equal the batch pipeline on the live seeds, bit for bit.

* At 1536 x 1536 (16 components, sigma 20 px -- the 16384^2 workload's
  sigma 200 px scaled to the canvas -- 20,000 seeds, ten rounds of 1,000
  inserts and 500 deletes) every round is compared with the oracle.
* At the full size (16384 x 16384, 2,000,000 seeds, 64 components, sigma
  200 px, two rounds of 10,000 inserts and 5,000 deletes; a canvas of
  16416^2 = 269,485,056 cells, above 2^28) the oracle's flood fill would
  take hours, so the fill is checked by what the mathematics fixes: every
  cell is defined, its d^2 is the squared distance to its labelled seed, the
  grid is a fixed point of every one of the four directional passes (the
  reference stops only after an iteration in which no pass changed a cell,
  voronoi.py:110-163), and d^2 is at least the exact nearest-seed distance.
  Detection and assignment are then checked bit for bit with the oracle on
  the GPU's own padded canvas, and the session's result with the batch
  path's.
"""
import ctypes
import time

import numpy as np
import pytest

import integer_delauney_amd as dl
from integer_delauney_amd import _lib
from integer_delauney_amd.workloads import clustered_more, clustered_seeds
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rows(m):
    return np.array([m[i] for i in range(len(m))], dtype=np.int32).reshape(-1, 5)


def _keys(xy, H):
    return xy[:, 0].astype(np.int64) * H + xy[:, 1]


def _round_plan(W, H, live, n_ins, n_del, rng, seed, components, sigma, centres_seed):
    ins = clustered_more(W, H, live, n_ins, seed, components=components, sigma=sigma,
                         centres_seed=centres_seed)
    live = np.concatenate([live, ins])
    drop = rng.choice(live.shape[0], size=n_del, replace=False)
    dele = np.ascontiguousarray(live[drop])
    keep = np.ones(live.shape[0], dtype=bool)
    keep[drop] = False
    return ins, dele, np.ascontiguousarray(live[keep])


def test_clustered_session_rounds_equal_oracle():
    W = H = 1536
    comps, sigma = 16, 20.0
    O.set_threads(0)
    seeds = clustered_seeds(W, H, 20_000, 3, components=comps, sigma=sigma)
    rng = np.random.default_rng(30)
    s = dl.Delaunay(W, H, 40_000, 16)
    s.insert_deferred(seeds)
    live = seeds
    for r in range(10):
        ins, dele, live = _round_plan(W, H, live, 1_000, 500, rng, 100 + r, comps, sigma, 3)
        s.insert_deferred(ins)
        s.remove(dele)
        assert s.seed_count == live.shape[0]
        t0 = time.perf_counter()
        m, grid = s.finalise()
        t_gpu = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref_rows, ref_grid = O.triangulate(W, H, live, 16, return_rows=True)
        _say(f"round {r}: {live.shape[0]} seeds, {len(ref_rows)} triangles, "
             f"gpu {t_gpu * 1e3:.1f} ms, oracle {time.perf_counter() - t0:.1f} s")
        np.testing.assert_array_equal(grid, ref_grid, err_msg=f"round {r}")
        np.testing.assert_array_equal(_rows(m), ref_rows, err_msg=f"round {r}")
        np.testing.assert_array_equal(np.sort(_keys(s.get_seeds(), H)), np.sort(_keys(live, H)))


# --- full size -----------------------------------------------------------------

def _check_fill(ids, d2, sx, sy, off, name, rows_per_chunk=512):
    """ids, d2: (h, w) owner ids and squared distances of a flood fill whose
    seeds sit at (sx + off, sy + off) in canvas coordinates.  Asserts every
    cell is defined, d2 is the distance to the owner, and no cell would change
    in any of the four passes (a candidate from a neighbour is never strictly
    closer, nor equally close with a higher id)."""
    h, w = ids.shape
    n = sx.shape[0]
    X = np.arange(w, dtype=np.int64)[None, :]
    for y0 in range(0, h, rows_per_chunk):
        y1 = min(h, y0 + rows_per_chunk)
        Y = np.arange(y0, y1, dtype=np.int64)[:, None]
        idc = ids[y0:y1]
        assert idc.min() >= 0 and idc.max() < n, f"{name}: undefined cell in rows {y0}-{y1}"
        ox = sx[idc] + off
        oy = sy[idc] + off
        dd = (X - ox) ** 2 + (Y - oy) ** 2
        np.testing.assert_array_equal(d2[y0:y1], dd, err_msg=f"{name}: d^2 rows {y0}-{y1}")
        # neighbours along x (passes 1 and 2)
        for a, b in ((slice(1, None), slice(None, -1)), (slice(None, -1), slice(1, None))):
            nid = idc[:, b]
            cand = (X[:, a] - sx[nid] - off) ** 2 + (Y - sy[nid] - off) ** 2
            cur = dd[:, a]
            bad = (cand < cur) | ((cand == cur) & (nid > idc[:, a]))
            assert not bad.any(), f"{name}: an x pass would change rows {y0}-{y1}"
        # neighbours along y (passes 3 and 4), rows y0-1 and y1 included
        lo, hi = max(0, y0 - 1), min(h, y1 + 1)
        ext = ids[lo:hi]
        for dy in (-1, 1):
            ys = np.arange(y0, y1)
            ok = (ys + dy >= 0) & (ys + dy < h)
            ys = ys[ok]
            nid = ext[ys + dy - lo]
            cur_id = ids[ys]
            Yv = ys.astype(np.int64)[:, None]
            cand = (X - sx[nid] - off) ** 2 + (Yv - sy[nid] - off) ** 2
            cur = dd[ys - y0]
            bad = (cand < cur) | ((cand == cur) & (nid > cur_id))
            assert not bad.any(), f"{name}: a y pass would change rows {y0}-{y1}"


def _triangulate_c_abi(W, H, seeds, P):
    """triangulate through the C ABI: grid, rows and the padded canvas."""
    lib = _lib.load()
    ctx = _lib.context()
    grid = np.empty((H, W, 3), dtype=np.int32)
    n = ctypes.c_int64(0)
    _lib.check(lib.dl_triangulate(ctx.handle, W, H, _lib.i32p(seeds), seeds.shape[0], P,
                                  _lib.i32p(grid), ctypes.byref(n)))
    rows = np.empty((int(n.value), 5), dtype=np.int32)
    _lib.check(lib.dl_copy_triangles(ctx.handle, _lib.i32p(rows)))
    w, h = ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(lib.dl_copy_padded_voronoi(ctx.handle, None, ctypes.byref(w), ctypes.byref(h)))
    pad = np.empty((h.value, w.value, 2), dtype=np.int32)
    _lib.check(lib.dl_copy_padded_voronoi(ctx.handle, _lib.i32p(pad), ctypes.byref(w),
                                          ctypes.byref(h)))
    return grid, rows, pad


@pytest.mark.slow
@pytest.mark.timeout(1200)
def test_clustered_16384_2m_full_size():
    W = H = 16384
    P = 16
    O.set_threads(0)
    t0 = time.perf_counter()
    seeds0 = clustered_seeds(W, H, 2_000_000, 3)
    rng = np.random.default_rng(3)
    plan = []
    live = seeds0
    for r in range(2):
        ins, dele, live = _round_plan(W, H, live, 10_000, 5_000, rng, 1000 + r, 64, 200.0, 3)
        plan.append((ins, dele))
    _say(f"seeds generated in {time.perf_counter() - t0:.1f} s")

    # the session, keeping its fill record: initial build (recorded), then two
    # rounds of updates, each finalised by replaying from the record
    s = dl.Delaunay(W, H, 2_020_000, P)
    s.set_fill_record(True)
    s.insert_deferred(seeds0)
    s.finalise(as_arrays=True)
    full = s.fill_record_stats()[2]
    for ins, dele in plan:
        s.insert_deferred(ins)
        s.remove(dele)
        t0 = time.perf_counter()
        verts_s, grid_s = s.finalise(as_arrays=True)
        held, entries, computed = s.fill_record_stats()
        _say(f"session round finalised in {time.perf_counter() - t0:.2f} s: record {held} "
             f"({entries} entries), {computed} tile computations (recorded fill {full})")
        assert held and computed < full
    assert s.seed_count == live.shape[0]
    s.close()

    # the batch path on the live seeds
    t0 = time.perf_counter()
    grid, rows, pad = _triangulate_c_abi(W, H, live, P)
    _say(f"batch triangulate + copies {time.perf_counter() - t0:.1f} s, {len(rows)} triangles")
    np.testing.assert_array_equal(grid_s, grid)
    np.testing.assert_array_equal(verts_s, rows[:, 2:5])
    del grid_s, verts_s
    assert pad.shape == (H + 2 * P, W + 2 * P, 2)

    srt = O.sort_seeds(W, H, live)
    sx = srt[:, 0].astype(np.int64)
    sy = srt[:, 1].astype(np.int64)
    t0 = time.perf_counter()
    _check_fill(np.ascontiguousarray(grid[:, :, 0]), np.ascontiguousarray(grid[:, :, 1]),
                sx, sy, 0, "image fill")
    _say(f"image fill invariants {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    _check_fill(np.ascontiguousarray(pad[:, :, 0]), np.ascontiguousarray(pad[:, :, 1]),
                sx, sy, P, "padded fill")
    _say(f"padded fill invariants {time.perf_counter() - t0:.1f} s")

    # d^2 >= the exact nearest-seed distance, on sampled rows
    spatial = pytest.importorskip("scipy.spatial")
    tree = spatial.cKDTree(np.stack([sx, sy], 1).astype(np.float64))
    ys = np.random.default_rng(7).choice(H, size=48, replace=False)
    xs = np.arange(W)
    q = np.stack([np.tile(xs, ys.size), np.repeat(ys, W)], 1)
    _, nn = tree.query(q.astype(np.float64))
    near = (q[:, 0] - sx[nn]) ** 2 + (q[:, 1] - sy[nn]) ** 2
    got = grid[ys][:, :, 1].reshape(-1).astype(np.int64)
    assert np.all(got >= near)
    _say(f"sampled rows: {np.mean(got == near) * 100:.4f} % of cells at the nearest distance")

    # detection on the GPU's padded canvas and assignment, bit for bit
    t0 = time.perf_counter()
    ref_rows, _ = O.detect(pad[:, :, 0], srt, P)
    _say(f"oracle detection {time.perf_counter() - t0:.1f} s")
    np.testing.assert_array_equal(rows, ref_rows)
    t0 = time.perf_counter()
    tid = O.assign(W, H, srt, ref_rows, "raster")
    _say(f"oracle raster {time.perf_counter() - t0:.1f} s")
    np.testing.assert_array_equal(grid[:, :, 2], tid)
    # give back the default context's 16384^2 workspaces
    _lib.check(_lib.load().dl_context_trim(_lib.context().handle))


# --- fill records: replaying only what the changes reach ----------------------

def test_fill_record_replay_equals_oracle():
    """A session keeping its fill record recomputes, per launch, only the tiles
    whose neighbourhood differs from the recorded trajectory.  Clustered rounds
    at 1536^2 (as above, fewer rounds), then the edge cases: a finalise with no
    change, inserts only, removals only, seeds on the image border and in the
    padding's reach, a change in every tile, and a single seed left.  Every
    result equals the oracle bit for bit, and from the second finalise on the
    fill replays (a record is held; fewer tile computations than the first,
    recorded, fill)."""
    W = H = 1536
    comps, sigma = 16, 20.0
    O.set_threads(0)
    seeds = clustered_seeds(W, H, 20_000, 5, components=comps, sigma=sigma)
    rng = np.random.default_rng(31)
    s = dl.Delaunay(W, H, 60_000, 16)
    s.set_fill_record(True)
    s.insert_deferred(seeds)
    live = seeds

    def check(tag):
        m, grid = s.finalise()
        ref_rows, ref_grid = O.triangulate(W, H, live, 16, return_rows=True)
        held, entries, computed = s.fill_record_stats()
        _say(f"{tag}: {live.shape[0]} seeds, record {held} ({entries} entries), "
             f"{computed} tile computations")
        np.testing.assert_array_equal(grid, ref_grid, err_msg=tag)
        np.testing.assert_array_equal(_rows(m), ref_rows, err_msg=tag)
        assert held, tag
        return computed

    full = check("recorded fill")
    computed = []
    for r in range(4):
        ins, dele, live = _round_plan(W, H, live, 1_000, 500, rng, 200 + r, comps, sigma, 5)
        s.insert_deferred(ins)
        s.remove(dele)
        computed.append(check(f"round {r}"))
    assert max(computed) < full
    # no change at all: a finalise with nothing pending returns the result
    # as it is; force a recompute with an insert and its removal
    extra = clustered_more(W, H, live, 1, 778, components=comps, sigma=sigma, centres_seed=5)
    s.insert_deferred(extra)
    s.remove(extra)
    check("no net change")
    # inserts only / removals only
    ins = clustered_more(W, H, live, 300, 777, components=comps, sigma=sigma, centres_seed=5)
    s.insert_deferred(ins)
    live = np.concatenate([live, ins])
    check("inserts only")
    drop = rng.choice(live.shape[0], size=300, replace=False)
    s.remove(np.ascontiguousarray(live[drop]))
    keep = np.ones(live.shape[0], bool)
    keep[drop] = False
    live = np.ascontiguousarray(live[keep])
    check("removals only")
    # seeds on the border rows and columns (the padding watch's reach)
    have = set(map(tuple, live.tolist()))
    border = [(x, y) for x, y in [(0, 0), (W - 1, 0), (0, H - 1), (W - 1, H - 1), (0, 700),
                                  (W - 1, 31), (500, 0), (900, H - 1), (1, 1), (W - 2, H - 2)]
              if (x, y) not in have]
    border = np.array(border, np.int32)
    s.insert_deferred(border)
    live = np.concatenate([live, border])
    check("border seeds")
    # a change in every tile: uniform inserts over the whole image
    have = set(map(tuple, live.tolist()))
    cand = np.stack([rng.integers(0, W, 6000), rng.integers(0, H, 6000)], 1).astype(np.int32)
    uni = []
    for x, y in cand.tolist():
        if (x, y) not in have:
            have.add((x, y))
            uni.append((x, y))
    uni = np.array(uni, np.int32)
    s.insert_deferred(uni)
    live = np.concatenate([live, uni])
    check("changes everywhere")
    # down to a single seed
    s.remove(np.ascontiguousarray(live[1:]))
    live = np.ascontiguousarray(live[:1])
    check("single seed")
    s.close()


def _record_stats():
    lib = _lib.load()
    rec, rep, ret = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.check(lib.dl_get_record_stats(_lib.context().handle, ctypes.byref(rec), ctypes.byref(rep),
                                       ctypes.byref(ret)))
    return rec.value, rep.value, ret.value


def test_replay_from_the_retained_plane():
    """A replay starts from the label plane as the previous finalise left it
    when nothing has written it since: only tiles near the changed seeds are
    seeded, only tiles the replay touches are written back.  Every round
    equals the oracle; a round after another call on the same context (which
    rewrites the plane) re-initialises instead, and so does a round with the
    reuse switched off; both still equal the oracle."""
    W = H = 1536
    comps, sigma = 16, 20.0
    O.set_threads(0)
    lib = _lib.load()
    seeds = clustered_seeds(W, H, 20_000, 9, components=comps, sigma=sigma)
    rng = np.random.default_rng(41)
    s = dl.Delaunay(W, H, 60_000, 16)
    s.set_fill_record(True)
    s.insert_deferred(seeds)
    live = seeds

    def round_(tag, r, expect_retained):
        nonlocal live
        ins, dele, live = _round_plan(W, H, live, 1_000, 500, rng, 300 + r, comps, sigma, 9)
        s.insert_deferred(ins)
        s.remove(dele)
        before = _record_stats()
        m, grid = s.finalise()
        after = _record_stats()
        ref_rows, ref_grid = O.triangulate(W, H, live, 16, return_rows=True)
        _say(f"{tag}: stats {before} -> {after}")
        np.testing.assert_array_equal(grid, ref_grid, err_msg=tag)
        np.testing.assert_array_equal(_rows(m), ref_rows, err_msg=tag)
        assert after[1] == before[1] + 1, f"{tag}: not replayed"
        assert after[2] == before[2] + (1 if expect_retained else 0), tag

    s.finalise()   # the recorded fill
    round_("first replay", 0, True)
    round_("second replay", 1, True)
    # another image on the same context rewrites the planes
    other = clustered_seeds(700, 500, 3000, 4)
    dl.triangulate(700, 500, other)
    round_("after another call", 2, False)
    round_("retained again", 3, True)
    _lib.check(lib.dl_set_replay_plane_reuse(_lib.context().handle, 0))
    try:
        round_("reuse off", 4, False)
    finally:
        _lib.check(lib.dl_set_replay_plane_reuse(_lib.context().handle, 1))
    round_("reuse back on", 5, True)   # (a full-init replay leaves the plane claimed too)
    s.close()
