"""Parity of the HIP path (libdelauney_hip.so through the package API) with the
reference's golden vectors and the CPU oracle.  Bit-exact throughout: every
quantity is integer."""
import numpy as np
import pytest

import integer_delauney_amd as dl
from integer_delauney_amd import _lib
from oracle import oracle as O
from tests import goldens

pytestmark = pytest.mark.gpu

FULL = goldens.names("full_")
VOR = goldens.names("vor_")
DET = goldens.names("det_")


def _rows(m):
    return np.array([m[i] for i in range(len(m))], dtype=np.int32).reshape(-1, 5)


def _pad_arg(p):
    return -1 if p is None else p


def test_library_loads_and_runs():
    g = dl.Voronoi().compute(5, 5, [(2, 2)])
    assert g.shape == (5, 5, 2) and np.all(g[:, :, 0] == 0)


@pytest.mark.parametrize("name", FULL)
def test_golden_triangulate(name):
    g = goldens.load(name)
    m, tg = dl.triangulate(g["W"], g["H"], g["seeds"], _pad_arg(g["P"]))
    np.testing.assert_array_equal(tg[:, :, :2], g["vgrid"])
    np.testing.assert_array_equal(_rows(m), g["rows"])
    np.testing.assert_array_equal(tg[:, :, 2], g["tid"])


@pytest.mark.parametrize("name", FULL)
def test_golden_two_step(name):
    g = goldens.load(name)
    vg = dl.Voronoi().compute(g["W"], g["H"], g["seeds"])
    np.testing.assert_array_equal(vg, g["vgrid"])
    m, tg = dl.GridTriangulation().compute(vg, g["seeds"], _pad_arg(g["P"]))
    np.testing.assert_array_equal(_rows(m), g["rows"])
    np.testing.assert_array_equal(tg[:, :, 2], g["tid"])
    np.testing.assert_array_equal(tg[:, :, :2], vg)
    verts, tg2 = dl.GridTriangulation().compute(vg, g["seeds"], _pad_arg(g["P"]), as_arrays=True)
    np.testing.assert_array_equal(verts, g["rows"][:, 2:5])
    np.testing.assert_array_equal(tg2, tg)


@pytest.mark.parametrize("name", VOR)
def test_golden_voronoi(name):
    g = goldens.load(name)
    np.testing.assert_array_equal(dl.Voronoi().compute(g["W"], g["H"], g["seeds"]), g["vgrid"])


@pytest.mark.parametrize("name", DET)
def test_golden_detection(name):
    g = goldens.load(name)
    m, _ = dl.triangulate(g["W"], g["H"], g["seeds"], _pad_arg(g["P"]))
    np.testing.assert_array_equal(_rows(m), g["rows"])


def _random_case(rng):
    kind = rng.integers(0, 4)
    W, H = int(rng.integers(2, 200)), int(rng.integers(2, 200))
    if kind == 0:
        n = int(rng.integers(1, max(2, W * H // 30)))
    elif kind == 1:
        n = int(rng.integers(1, max(2, W * H // 3)))
    elif kind == 2:
        n = int(rng.integers(1, 6))
    else:
        sx, sy = int(rng.integers(2, 9)), int(rng.integers(2, 9))
        pts = np.array([(x, y) for x in range(0, W, sx) for y in range(0, H, sy)])
        return W, H, pts.astype(np.int32)
    idx = rng.choice(W * H, size=min(n, W * H), replace=False)
    return W, H, np.stack([idx % W, idx // W], 1).astype(np.int32)


@pytest.mark.parametrize("seed", range(12))
def test_random_against_oracle(seed):
    rng = np.random.default_rng(1000 + seed)
    for _ in range(6):
        W, H, seeds = _random_case(rng)
        P = int(rng.choice([-1, 0, 1, 2, 7, 16]))
        rows, ref = O.triangulate(W, H, seeds, None if P == -1 else P, return_rows=True)
        m, got = dl.triangulate(W, H, seeds, P)
        np.testing.assert_array_equal(got, ref, err_msg=f"W={W} H={H} n={len(seeds)} P={P}")
        np.testing.assert_array_equal(_rows(m), rows)


@pytest.mark.parametrize("cfg", [(16, 4, 4), (16, 4, 8), (8, 8, 4), (8, 8, 5), (8, 8, 6),
                                 (8, 8, 7), (8, 8, 8), (16, 8, 4), (16, 8, 7), (16, 8, 8),
                                 (16, 8, 12),
                                 (16, 8, 16), (8, 16, 8)])
def test_every_tile_configuration_is_exact(cfg):
    ctx = _lib.context()
    lib = _lib.load()
    _lib.check(lib.dl_set_bfs_config(ctx.handle, *cfg))
    try:
        rng = np.random.default_rng(7)
        for W, H, n in [(300, 200, 600), (97, 411, 300), (513, 129, 50), (1000, 700, 20000),
                        (260, 3, 20), (5, 333, 30)]:
            idx = rng.choice(W * H, size=n, replace=False)
            seeds = np.stack([idx % W, idx // W], 1).astype(np.int32)
            ref = O.voronoi(W, H, seeds)
            np.testing.assert_array_equal(dl.Voronoi().compute(W, H, seeds), ref)
            rows, ref3 = O.triangulate(W, H, seeds, None, return_rows=True)
            m, got = dl.triangulate(W, H, seeds)
            np.testing.assert_array_equal(got, ref3)
            np.testing.assert_array_equal(_rows(m), rows)
    finally:
        _lib.check(lib.dl_set_bfs_config(ctx.handle, 0, 0, 0))


def test_simple_fill_kernel_agrees_with_tiled():
    import ctypes
    ctx = _lib.context()
    lib = _lib.load()
    rng = np.random.default_rng(5)
    for W, H, n in [(64, 64, 40), (300, 200, 3000), (50, 400, 10)]:
        idx = rng.choice(W * H, size=n, replace=False)
        s = np.ascontiguousarray(np.stack([idx % W, idx // W], 1).astype(np.int32))
        out = np.empty((H, W, 2), dtype=np.int32)
        it = ctypes.c_int64(0)
        _lib.check(lib.dl_voronoi_simple(ctx.handle, W, H, _lib.i32p(s), n, _lib.i32p(out),
                                         ctypes.byref(it)))
        ref, ref_it = O.voronoi_sorted(W, H, O.sort_seeds(W, H, s), return_iters=True)
        np.testing.assert_array_equal(out, ref)
        assert it.value == ref_it
        np.testing.assert_array_equal(dl.Voronoi().compute(W, H, s), ref)


def test_forced_second_fill_gives_the_same_grid():
    ctx = _lib.context()
    lib = _lib.load()
    rng = np.random.default_rng(9)
    W, H = 257, 190
    idx = rng.choice(W * H, size=900, replace=False)
    s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    m0, g0 = dl.triangulate(W, H, s)
    _lib.check(lib.dl_set_force_unpadded(ctx.handle, 1))
    try:
        m1, g1 = dl.triangulate(W, H, s)
    finally:
        _lib.check(lib.dl_set_force_unpadded(ctx.handle, 0))
    np.testing.assert_array_equal(g0, g1)
    assert m0 == m1


def test_compute_debug_padded_canvas():
    rng = np.random.default_rng(2)
    W, H = 120, 90
    idx = rng.choice(W * H, size=150, replace=False)
    s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    vg = dl.Voronoi().compute(W, H, s)
    m, tg, padded = dl.GridTriangulation().compute_debug(vg, s, 12)
    ref = O.voronoi(W + 24, H + 24, s + 12)
    np.testing.assert_array_equal(padded, ref)
    m2, tg2, t = dl.GridTriangulation().compute_timed(vg, s, 12)
    assert m2 == m and np.array_equal(tg2, tg)
    assert set(t) == {"detect_ms", "dedup_ms", "assign_ms"} and all(v >= 0 for v in t.values())


@pytest.mark.parametrize("bad", [[], [(2, 2), (2, 2)], [(10, 1)], [(1, 10)], [(-1, 0)],
                                 [(1, 2, 3)]])
def test_value_errors(bad):
    with pytest.raises(ValueError):
        dl.Voronoi().compute(5, 5, bad)
    with pytest.raises(ValueError):
        dl.triangulate(5, 5, bad)
    # the pipeline reads seed validation back late (with the flood fill's first
    # wait): the context must be fully usable after the error
    good = [(0, 0), (4, 1), (2, 4), (1, 2)]
    rows, ref = O.triangulate(5, 5, good, None, return_rows=True)
    m, got = dl.triangulate(5, 5, good)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(_rows(m), rows)


@pytest.mark.parametrize("W,H", [(0, 10), (10, 0), (-4, 10)])
def test_non_positive_dimensions(W, H):
    with pytest.raises(ValueError):
        dl.triangulate(W, H, [(0, 0), (1, 1), (2, 0)])


def test_oversized_canvas_rejected():
    # A canvas the library cannot hold is a RuntimeError, as the reference's
    # failing allocation of one is (tests/test_cuda_error_checking.py:36-46).
    with pytest.raises(RuntimeError, match="exceeds"):
        dl.triangulate(32000, 10, [(0, 0), (5, 5), (9, 1)])
    with pytest.raises(RuntimeError):
        dl.Delaunay(200_000, 200_000, 1_000, 0)


def test_device_allocation_failure_is_a_runtime_error_and_recoverable():
    """A real hipMalloc failure (a seed list of 2^40 entries, 8 TiB) raises
    RuntimeError naming hipMalloc, leaks nothing that breaks later work, and
    the process and a new session work afterwards
    (tests/test_cuda_error_checking.py:28-58)."""
    for _ in range(2):
        with pytest.raises(RuntimeError, match="hipMalloc") as ei:
            dl.Delaunay(64, 64, 1 << 40, 0)
        assert "dl_api.hip" in str(ei.value)
    inc = dl.Delaunay(64, 64, 256, 0)
    tri_map, tgrid = inc.insert([(10, 10), (50, 10), (30, 50)])
    assert len(tri_map) > 0 and tgrid.shape == (64, 64, 3)
    rows, ref = O.triangulate(64, 64, [(10, 10), (50, 10), (30, 50)], 0, return_rows=True)
    np.testing.assert_array_equal(tgrid, ref)
    m, g = dl.triangulate(64, 64, [(10, 10), (50, 10), (30, 50)], 0)
    np.testing.assert_array_equal(g, ref)


@pytest.mark.parametrize("P", [-5, -2, -1000])
def test_negative_padding_clamps_to_zero(P):
    """border_padding below -1 clamps to 0 (triangulation.py:75-78; the
    reference test tests/test_border_padding.py:91-96), -1 selects the
    default; a session takes the default for any negative value
    (delaunay.cu:72)."""
    rng = np.random.default_rng(5000 - P)
    for W, H, n in [(96, 64, 20), (200, 150, 300), (31, 257, 60)]:
        idx = rng.choice(W * H, size=n, replace=False)
        seeds = np.stack([idx % W, idx // W], 1).astype(np.int32)
        rows0, ref0 = O.triangulate(W, H, seeds, 0, return_rows=True)
        m, got = dl.triangulate(W, H, seeds, P)
        np.testing.assert_array_equal(got, ref0)
        np.testing.assert_array_equal(_rows(m), rows0)
        vg = dl.Voronoi().compute(W, H, seeds)
        a, tga = dl.GridTriangulation().compute(vg, seeds, border_padding=P)
        b, tgb = dl.GridTriangulation().compute(vg, seeds, border_padding=0)
        assert a == b and np.array_equal(tga, tgb) and len(a) == len(rows0)
    s = dl.Delaunay(96, 64, 100, border_padding=P)
    assert s.border_padding == dl.DEFAULT_BORDER_PADDING


def test_timed_dictionaries():
    """compute_timed and the session's *_timed methods carry the reference's
    keys (py_grid_triangulation.hpp:39-46, py_delaunay.hpp:268-277) with
    stage times measured on the device, and leave the results unchanged."""
    rng = np.random.default_rng(60)
    W, H = 640, 480
    idx = rng.choice(W * H, size=4000, replace=False)
    seeds = np.stack([idx % W, idx // W], 1).astype(np.int32)
    rows, ref = O.triangulate(W, H, seeds, None, return_rows=True)
    vg = dl.Voronoi().compute(W, H, seeds)
    m, tg, t = dl.GridTriangulation().compute_timed(vg, seeds)
    assert set(t) >= {"detect_ms", "dedup_ms", "assign_ms"}
    assert all(t[k] > 0 for k in ("detect_ms", "dedup_ms", "assign_ms"))
    np.testing.assert_array_equal(tg, ref)
    s = dl.Delaunay(W, H, 10_000)
    t1 = s.insert_deferred_timed(seeds[:2000])
    assert set(t1) >= {"bfs_ms", "bfs_iters", "detect_ms", "dedup_ms", "assign_ms"}
    assert t1["update_ms"] > 0
    m1, g1, t2 = s.insert_timed(seeds[2000:])
    keys = ("bfs_ms", "detect_ms", "dedup_ms", "assign_ms")
    assert all(t2[k] > 0 for k in keys), t2
    assert t2["bfs_iters"] > 0 and t2["fill_launches"] > 0
    assert sum(t2[k] for k in keys) <= t2["total_ms"] + 1e-3
    np.testing.assert_array_equal(g1, ref)
    np.testing.assert_array_equal(_rows(m1), rows)
    s.remove(seeds[:10])
    m2, g2, t3 = s.finalise_timed()
    assert all(t3[k] > 0 for k in keys)
    rows2, ref2 = O.triangulate(W, H, seeds[10:], None, return_rows=True)
    np.testing.assert_array_equal(g2, ref2)
    _, _, t4 = s.finalise_timed()   # nothing pending: nothing ran
    assert t4["bfs_iters"] == 0 and t4["bfs_ms"] == 0.0


def test_reusing_the_context_across_sizes():
    rng = np.random.default_rng(4)
    for W, H, n in [(400, 300, 2000), (20, 20, 5), (640, 480, 5000), (33, 700, 100)]:
        idx = rng.choice(W * H, size=n, replace=False)
        s = np.stack([idx % W, idx // W], 1).astype(np.int32)
        rows, ref = O.triangulate(W, H, s, None, return_rows=True)
        m, got = dl.triangulate(W, H, s)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(_rows(m), rows)


def test_run_twice_identical():
    rng = np.random.default_rng(8)
    W, H = 700, 500
    idx = rng.choice(W * H, size=8000, replace=False)
    s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    a = dl.triangulate(W, H, s, as_arrays=True)
    b = dl.triangulate(W, H, s, as_arrays=True)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def _fill_watch(W, H, seeds, margin):
    import ctypes
    ctx = _lib.context()
    lib = _lib.load()
    s = np.ascontiguousarray(seeds, dtype=np.int32)
    out = np.empty((H, W, 2), dtype=np.int32)
    esc = ctypes.c_int(0)
    _lib.check(lib.dl_debug_fill_watch(ctx.handle, W, H, _lib.i32p(s), s.shape[0], margin,
                                       _lib.i32p(out), ctypes.byref(esc)))
    return out, bool(esc.value)


@pytest.mark.parametrize("seed", range(4))
def test_fill_watch_never_misses_a_crossing(seed):
    """The watch that lets the pipeline skip the second (unpadded) fill: on
    boxes with seeds on both sides, every update that the oracle sees carried
    into the box from outside must raise the GPU flag."""
    rng = np.random.default_rng(300 + seed)
    seen_cross = seen_clear = 0
    for _ in range(10):
        W, H = int(rng.integers(40, 700)), int(rng.integers(40, 500))
        margin = int(rng.choice([1, 2, 5, 13, min(W, H) // 3]))
        n = int(rng.integers(2, max(3, W * H // int(rng.choice([20, 300, 5000])))))
        idx = rng.choice(W * H, size=n, replace=False)
        s = np.stack([idx % W, idx // W], 1).astype(np.int32)
        cross = O.voronoi_crossings(W, H, O.sort_seeds(W, H, s), margin)
        got, esc = _fill_watch(W, H, s, margin)
        np.testing.assert_array_equal(got, O.voronoi(W, H, s))
        assert esc or cross == 0, f"W={W} H={H} n={n} margin={margin} crossings={cross}"
        seen_cross += cross > 0
        seen_clear += not esc
    assert seen_cross > 0


@pytest.mark.parametrize("cfg", [(16, 4, 8), (8, 8, 4), (16, 8, 16), (8, 8, 7), (16, 8, 8)])
def test_fill_watch_on_padded_canvases(cfg):
    """On the pipeline's own padded canvas (seeds only inside the image) the
    oracle sees no crossing in these cases; the GPU flag must stay clear for
    most of them, or the pipeline would run its second fill needlessly."""
    ctx = _lib.context()
    lib = _lib.load()
    _lib.check(lib.dl_set_bfs_config(ctx.handle, *cfg))
    try:
        rng = np.random.default_rng(11)
        flagged = 0
        cases = 0
        for W, H, n, P in [(300, 200, 600, 16), (97, 411, 300, 3), (513, 129, 50, 1),
                           (1000, 700, 20000, 16), (800, 800, 12, 16), (260, 30, 20, 7)]:
            idx = rng.choice(W * H, size=n, replace=False)
            s = np.stack([idx % W, idx // W], 1).astype(np.int32) + P
            Wc, Hc = W + 2 * P, H + 2 * P
            cross = O.voronoi_crossings(Wc, Hc, O.sort_seeds(Wc, Hc, s), P)
            got, esc = _fill_watch(Wc, Hc, s, P)
            np.testing.assert_array_equal(got, O.voronoi(Wc, Hc, s))
            assert esc or cross == 0
            flagged += esc
            cases += 1
        assert flagged <= cases // 2, f"{flagged} of {cases} padded fills flagged"
    finally:
        _lib.check(lib.dl_set_bfs_config(ctx.handle, 0, 0, 0))


@pytest.mark.parametrize("mode", [1, 2])
def test_speculative_sizes_overflowing_fall_back_exactly(mode):
    """Detection and assignment are sized from the seed count without a host
    round trip; when the sizes are too small the pipeline reruns with exact
    ones.  Forcing that path must not change any result.  Mode 1 makes every
    speculative size too small; mode 2 only the registration buckets', so the
    candidate list is complete and only the bucket overflow sends the rerun
    through the hash table."""
    ctx = _lib.context()
    lib = _lib.load()
    rng = np.random.default_rng(44)
    cases = []
    for W, H, n in [(300, 200, 900), (64, 64, 3), (512, 40, 2000), (17, 300, 40)]:
        idx = rng.choice(W * H, size=n, replace=False)
        cases.append((W, H, np.stack([idx % W, idx // W], 1).astype(np.int32)))
    base = [dl.triangulate(W, H, s) for W, H, s in cases]
    batch = dl.triangulate_batch(cases)
    _lib.check(lib.dl_set_tiny_capacities(ctx.handle, mode))
    try:
        tiny = [dl.triangulate(W, H, s) for W, H, s in cases]
        tiny_batch = dl.triangulate_batch(cases)
    finally:
        _lib.check(lib.dl_set_tiny_capacities(ctx.handle, 0))
    for (m0, g0), (m1, g1), (m2, g2), (m3, g3) in zip(base, tiny, batch, tiny_batch):
        assert m0 == m1 == m2 == m3
        np.testing.assert_array_equal(g0, g1)
        np.testing.assert_array_equal(g0, g2)
        np.testing.assert_array_equal(g0, g3)
    for (W, H, s), (m, g) in zip(cases, base):
        rows, ref = O.triangulate(W, H, s, None, return_rows=True)
        np.testing.assert_array_equal(g, ref)
        np.testing.assert_array_equal(_rows(m), rows)


def test_predicted_fill_and_its_shortfall_are_exact():
    """A padded fill of a canvas class seen before is issued as the launch
    count the last one needed (plus a margin) without waiting on the host.
    When a harder input of the same class needs more, the fill continues after
    the status read and detection and assignment run again.  Every call,
    predicted, continued or fresh, equals the oracle."""
    from integer_delauney_amd.workloads import clustered_seeds, uniform_seeds
    W = H = 512
    n = 2000
    easy = uniform_seeds(W, H, n, 11)
    hard = clustered_seeds(W, H, n, 12)
    refs = {}
    for name, seeds in (("easy", easy), ("easy", easy), ("hard", hard), ("hard", hard),
                        ("easy", easy), ("hard", hard)):
        if name not in refs:
            refs[name] = O.triangulate(W, H, seeds, None, return_rows=True)
        rows, ref = refs[name]
        m, got = dl.triangulate(W, H, seeds)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(_rows(m), rows)


@pytest.mark.parametrize("case", range(12))
def test_active_tile_list_fill_equals_oracle(case):
    """Large fills run their later launches over a device-built list of the
    active tiles.  Forced on every fill here (threshold 0), on small cases
    against the oracle: uniform, clustered and dense seeds, ragged sizes, every
    padding rule, batches of several canvases."""
    from integer_delauney_amd import _lib
    from integer_delauney_amd.workloads import clustered_seeds
    lib = _lib.load()
    ctx = _lib.context()
    rng = np.random.default_rng(900 + case)
    W, H = int(rng.integers(40, 700)), int(rng.integers(3, 500))
    n = int(rng.integers(1, max(2, W * H // 40)))
    if case % 3 == 1 and W * H > 2000:
        s = clustered_seeds(W, H, min(n, W * H // 4), 900 + case)
    else:
        idx = rng.choice(W * H, size=min(n, W * H), replace=False)
        s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    P = [None, 0, 3, 16][case % 4]
    _lib.check(lib.dl_set_fill_list_threshold(ctx.handle, 0))
    try:
        rows, ref = O.triangulate(W, H, s, P, return_rows=True)
        m, got = dl.triangulate(W, H, s, -1 if P is None else P)
        np.testing.assert_array_equal(got, ref, err_msg=f"W={W} H={H} n={len(s)} P={P}")
        np.testing.assert_array_equal(_rows(m), rows)
        if case % 4 == 0:
            imgs = [(W, H, s), (H + 5, W + 2, np.ascontiguousarray(s[:, ::-1]))]
            res = dl.triangulate_batch(imgs)
            for (w, h, ss), (mb, gb) in zip(imgs, res):
                r2, ref2 = O.triangulate(w, h, ss, None, return_rows=True)
                np.testing.assert_array_equal(gb, ref2)
                np.testing.assert_array_equal(_rows(mb), r2)
    finally:
        _lib.check(lib.dl_set_fill_list_threshold(ctx.handle, -1))


def _aligned_used():
    import ctypes
    v = ctypes.c_int(-1)
    _lib.check(_lib.load().dl_get_fill_aligned_watch(_lib.context().handle, ctypes.byref(v)))
    return v.value


@pytest.mark.parametrize("seed", range(4))
def test_aligned_watch_equals_the_per_row_watch(seed):
    """Canvases whose watched box starts and ends on the tile kernel's
    wave-row boundaries (margin and height multiples of 8) take the aligned
    padding watch: one row per y pass, no row condition in the x passes.  On
    boxes with seeds on both sides its flag must equal the per-row watch's
    (dl_set_fill_aligned_watch(0)) on the same input, catch every crossing
    the oracle counts, and the labels must equal the oracle's.  A box that
    is not aligned keeps the per-row form."""
    lib = _lib.load()
    ctx = _lib.context()
    rng = np.random.default_rng(900 + seed)
    seen_cross = seen_clear = 0
    for k in range(10):
        margin = int(rng.choice([8, 16, 24, 40]))
        W = int(rng.integers(2 * margin + 8, 700))
        H = 8 * int(rng.integers((2 * margin) // 8 + 2, 64))
        n = int(rng.integers(2, max(3, W * H // int(rng.choice([20, 300, 5000])))))
        idx = rng.choice(W * H, size=n, replace=False)
        s = np.stack([idx % W, idx // W], 1).astype(np.int32)
        cross = O.voronoi_crossings(W, H, O.sort_seeds(W, H, s), margin)
        ref = O.voronoi(W, H, s)
        got, esc = _fill_watch(W, H, s, margin)
        assert _aligned_used() == 1, f"W={W} H={H} margin={margin}"
        _lib.check(lib.dl_set_fill_aligned_watch(ctx.handle, 0))
        try:
            got_r, esc_r = _fill_watch(W, H, s, margin)
            assert _aligned_used() == 0
        finally:
            _lib.check(lib.dl_set_fill_aligned_watch(ctx.handle, 1))
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got_r, ref)
        assert esc == esc_r, f"W={W} H={H} n={n} margin={margin}: {esc} vs {esc_r}"
        assert esc or cross == 0, f"W={W} H={H} n={n} margin={margin} crossings={cross}"
        seen_cross += cross > 0
        seen_clear += not esc
    assert seen_cross > 0
    # not aligned: the height leaves the box's end off a wave boundary
    W, H, margin = 200, 8 * 20 + 3, 16
    idx = rng.choice(W * H, size=300, replace=False)
    s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    got, _ = _fill_watch(W, H, s, margin)
    assert _aligned_used() == 0
    np.testing.assert_array_equal(got, O.voronoi(W, H, s))


def test_pipeline_takes_the_aligned_watch_exactly():
    """The pipeline's padded canvas (P = 16, image height a multiple of 8, as
    at 4096 x 4096) takes the aligned watch; goldens and oracle cases agree
    with the per-row watch bit for bit."""
    lib = _lib.load()
    ctx = _lib.context()
    rng = np.random.default_rng(77)
    W, H = 520, 384
    idx = rng.choice(W * H, size=3000, replace=False)
    s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    rows, ref = O.triangulate(W, H, s, 16, return_rows=True)
    m, got = dl.triangulate(W, H, s, 16)
    assert _aligned_used() == 1
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(_rows(m), rows)
    _lib.check(lib.dl_set_fill_aligned_watch(ctx.handle, 0))
    try:
        m2, got2 = dl.triangulate(W, H, s, 16)
        assert _aligned_used() == 0
        for name in FULL:
            g = goldens.load(name)
            mg, tg = dl.triangulate(g["W"], g["H"], g["seeds"], _pad_arg(g["P"]))
            np.testing.assert_array_equal(tg[:, :, :2], g["vgrid"])
            np.testing.assert_array_equal(_rows(mg), g["rows"])
    finally:
        _lib.check(lib.dl_set_fill_aligned_watch(ctx.handle, 1))
    np.testing.assert_array_equal(got2, got)
    np.testing.assert_array_equal(_rows(m2), _rows(m))


@pytest.mark.parametrize("case", range(10))
def test_wide_key_encoding_equals_oracle(case):
    """Canvases with sides beyond about 23,170 use the compare-and-select key
    update; smaller ones the min-based one.  Forced onto the former here
    (dl_set_fill_min_encoding(0)) on small cases against the oracle, with every
    padding rule, the active-tile list on odd cases, and the goldens on the
    first case; then the same input with both updates, bit for bit."""
    from integer_delauney_amd import _lib
    from integer_delauney_amd.workloads import clustered_seeds
    lib = _lib.load()
    ctx = _lib.context()
    rng = np.random.default_rng(1300 + case)
    W, H = int(rng.integers(30, 600)), int(rng.integers(3, 450))
    n = int(rng.integers(1, max(2, W * H // 30)))
    if case % 3 == 1 and W * H > 2000:
        s = clustered_seeds(W, H, min(n, W * H // 4), 1300 + case)
    else:
        idx = rng.choice(W * H, size=min(n, W * H), replace=False)
        s = np.stack([idx % W, idx // W], 1).astype(np.int32)
    P = [None, 0, 5, 16, -3][case % 5]
    _lib.check(lib.dl_set_fill_min_encoding(ctx.handle, 0))
    if case % 2:
        _lib.check(lib.dl_set_fill_list_threshold(ctx.handle, 0))
    try:
        rows, ref = O.triangulate(W, H, s, P, return_rows=True)
        m, got = dl.triangulate(W, H, s, -1 if P is None else P)
        np.testing.assert_array_equal(got, ref, err_msg=f"W={W} H={H} n={len(s)} P={P}")
        np.testing.assert_array_equal(_rows(m), rows)
        if case == 0:
            for name in FULL:
                g = goldens.load(name)
                mg, tg = dl.triangulate(g["W"], g["H"], g["seeds"], _pad_arg(g["P"]))
                np.testing.assert_array_equal(tg[:, :, :2], g["vgrid"])
                np.testing.assert_array_equal(_rows(mg), g["rows"])
                np.testing.assert_array_equal(tg[:, :, 2], g["tid"])
    finally:
        _lib.check(lib.dl_set_fill_min_encoding(ctx.handle, 1))
        _lib.check(lib.dl_set_fill_list_threshold(ctx.handle, -1))
    m2, got2 = dl.triangulate(W, H, s, -1 if P is None else P)
    np.testing.assert_array_equal(got2, got)
    np.testing.assert_array_equal(_rows(m2), _rows(m))


def test_dense_cluster_bins_fall_back_exactly():
    """Uniform seeds plus one dense cluster: the cluster's tiles hold far more
    triangles than the one-pass binning's fixed per-tile slices (sized from
    the image's average density), so the first call reruns the assignment
    with exact bins, and later calls of the class bin exactly from the start.
    Every call equals the oracle."""
    rng = np.random.default_rng(4242)
    W, H = 640, 480
    idx = rng.choice(W * H, size=400, replace=False)
    pts = {(int(i % W), int(i // W)) for i in idx}
    cx, cy = 300, 200
    for x in range(cx, cx + 120):
        for y in range(cy, cy + 90):
            if rng.random() < 0.5:
                pts.add((x, y))
    seeds = np.array(sorted(pts), dtype=np.int32)
    rows, ref = O.triangulate(W, H, seeds, None, return_rows=True)
    for _ in range(3):
        m, got = dl.triangulate(W, H, seeds)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(_rows(m), rows)
