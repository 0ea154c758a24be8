"""Session queries on the device: point location, the lifted in-circumsphere
predicate, edges, edge scores, midpoint selection, per-seed values and the
device-resident result of finalise_device.

Modelled on the reference's tests/test_scalar_field.py and
tests/test_finalise_device.py, with the bar raised to bit-exact agreement with
the CPU restatement (oracle/queries.py, pinned in test_oracle_queries.py) on
the same seed sets; the reference's own checks (host formulas, the finalised
grid, containment, staleness) are kept as well.
"""
import ctypes
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import integer_delauney_amd as dl
from integer_delauney_amd import _lib
from oracle import oracle as O
from oracle import queries as Q

pytestmark = pytest.mark.gpu

W, H, P = 240, 180, 16
MAX_SEEDS = 4000
MIN_LEN = 2.0


def make_mesh(n=400, seed=0, w=W, h=H, p=P, max_seeds=MAX_SEEDS):
    """test_scalar_field.py:39-51: distinct points in a random order, values
    sampled from a random image."""
    rng = np.random.default_rng(seed)
    cand = np.array(sorted({(int(x), int(y)) for x, y in
                            rng.integers([0, 0], [w, h], size=(n * 4, 2))}), dtype=np.int32)
    pts = cand[np.sort(rng.choice(len(cand), size=min(n, len(cand)), replace=False))]
    pts = pts[rng.permutation(len(pts))]
    img = rng.random((h, w)).astype(np.float32)
    vals = img[pts[:, 1], pts[:, 0]]
    mesh = dl.Delaunay(w, h, max_seeds, p)
    mesh.insert_deferred(pts, vals)
    return mesh, pts, vals


def oracle_state(pts, w=W, h=H, p=P):
    rows, grid = O.triangulate(w, h, pts, p, return_rows=True)
    return rows, grid, O.sort_seeds(w, h, pts)


def queries(rng, n, w=W, h=H, margin=0):
    return np.stack([rng.integers(-margin, w + margin, n), rng.integers(-margin, h + margin, n)],
                    axis=1).astype(np.int32)


# ------------------------------------------------------------------ edges
class TestEdges:

    @pytest.mark.parametrize("n", [20, 120, 400])
    def test_edges_and_scores_equal_the_oracle(self, n):
        mesh, pts, vals = make_mesh(n, seed=n)
        rows, _, _ = oracle_state(pts)
        e = mesh.get_edges()
        np.testing.assert_array_equal(e, Q.edges(rows, pts))
        got = mesh.edge_scores(MIN_LEN)
        np.testing.assert_array_equal(got, Q.edge_scores(pts, vals, e, MIN_LEN))
        # the reference test's host formula (test_scalar_field.py:54-62)
        d = pts[e[:, 1]].astype(np.float64) - pts[e[:, 0]].astype(np.float64)
        length = np.hypot(d[:, 0], d[:, 1])
        want = np.abs(vals[e[:, 1]].astype(np.float64) - vals[e[:, 0]].astype(np.float64)) * length
        want[length < MIN_LEN] = 0.0
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)

    def test_large_mesh_equals_the_oracle(self):
        """120 k edge keys (about 120 tiles of the device radix sort, 2 x 17
        bits in 5 passes) and the selection's two sorts over 64-bit keys, at
        1024 x 768 with 20,000 seeds."""
        w, h = 1024, 768
        mesh, pts, vals = make_mesh(20_000, seed=21, w=w, h=h, max_seeds=20_000)
        rows, _, _ = oracle_state(pts, w, h)
        e = mesh.get_edges()
        np.testing.assert_array_equal(e, Q.edges(rows, pts))
        for take in (5_000, 100_000):
            got = mesh.select_midpoints(MIN_LEN, take)
            np.testing.assert_array_equal(
                got, Q.select_midpoints(w, h, P, pts, vals, e, MIN_LEN, take))

    def test_edges_are_the_triangles_edges(self):
        mesh, pts, _ = make_mesh(300, seed=3)
        tris = mesh.get_triangles()
        want = set()
        for a, b, c in tris.tolist():
            for u, v in ((a, b), (b, c), (c, a)):
                want.add((min(u, v), max(u, v)))
        assert [tuple(x) for x in mesh.get_edges().tolist()] == sorted(want)

    def test_short_edges_score_zero(self):
        mesh, pts, _ = make_mesh(400, seed=5)
        e = mesh.get_edges()
        d = pts[e[:, 1]].astype(np.float64) - pts[e[:, 0]].astype(np.float64)
        length = np.hypot(d[:, 0], d[:, 1])
        s = mesh.edge_scores(8.0)
        assert (s[length < 8.0] == 0).all() and (s[length >= 8.0] >= 0).all()

    def test_order_matches_get_edges(self):
        mesh, _, _ = make_mesh(200, seed=7)
        e1 = mesh.get_edges()
        s = mesh.edge_scores(MIN_LEN)
        np.testing.assert_array_equal(e1, mesh.get_edges())
        assert len(s) == len(e1)

    def test_without_values_it_refuses(self):
        mesh = dl.Delaunay(W, H, MAX_SEEDS, P)
        mesh.insert_deferred(np.array([[10, 10], [60, 12], [30, 70]], dtype=np.int32))
        assert len(mesh.get_edges()) == 3
        with pytest.raises(RuntimeError):
            mesh.edge_scores(MIN_LEN)
        with pytest.raises(RuntimeError):
            mesh.select_midpoints(MIN_LEN, 3)
        assert mesh.get_values().shape == (0,)

    def test_values_must_match_the_batch(self):
        mesh = dl.Delaunay(W, H, MAX_SEEDS, P)
        with pytest.raises(ValueError):
            mesh.insert_deferred(np.array([[10, 10], [60, 12], [30, 70]], dtype=np.int32),
                                 np.array([1.0, 2.0], dtype=np.float32))
        assert mesh.seed_count == 0

    def test_no_triangles_no_edges(self):
        mesh = dl.Delaunay(W, H, MAX_SEEDS, P)
        mesh.insert_deferred(np.array([[10, 10], [60, 12]], dtype=np.int32), np.ones(2))
        assert mesh.get_edges().shape == (0, 2)
        assert mesh.edge_scores().shape == (0,)
        assert mesh.select_midpoints(0.0, 10).shape == (0, 2)

    def test_edges_follow_removals(self):
        mesh, pts, vals = make_mesh(400, seed=9)
        rng = np.random.default_rng(10)
        drop = pts[rng.choice(len(pts), 60, replace=False)]
        mesh.remove(drop)
        keep = ~np.isin(pts[:, 0] * 1000 + pts[:, 1], drop[:, 0] * 1000 + drop[:, 1])
        live, lv = pts[keep], vals[keep]
        np.testing.assert_array_equal(mesh.get_seeds(), live)
        np.testing.assert_array_equal(mesh.get_values(), lv)
        rows, _, _ = oracle_state(live)
        e = mesh.get_edges()
        np.testing.assert_array_equal(e, Q.edges(rows, live))
        np.testing.assert_array_equal(mesh.edge_scores(1.0), Q.edge_scores(live, lv, e, 1.0))


# ------------------------------------------------------ select_midpoints
class TestSelectMidpoints:

    @pytest.mark.parametrize("take", [1, 10, 75, 300, 100000])
    @pytest.mark.parametrize("p", [16, 5])
    def test_equals_the_oracle(self, take, p):
        mesh, pts, vals = make_mesh(400, seed=11, p=p)
        e = mesh.get_edges()
        got = mesh.select_midpoints(MIN_LEN, take)
        np.testing.assert_array_equal(got, Q.select_midpoints(W, H, p, pts, vals, e, MIN_LEN, take))

    def test_threshold(self):
        mesh, pts, vals = make_mesh(400, seed=16)
        e = mesh.get_edges()
        scores = mesh.edge_scores(MIN_LEN)
        cut = float(np.median(scores[scores > 0]))
        got = mesh.select_midpoints(MIN_LEN, 10_000, cut)
        np.testing.assert_array_equal(
            got, Q.select_midpoints(W, H, P, pts, vals, e, MIN_LEN, 10_000, cut))
        assert len(got) <= int((scores > cut).sum())
        assert len(mesh.select_midpoints(MIN_LEN, 10_000, float(scores.max()))) == 0

    def test_never_a_seed_unique_in_bounds_deterministic(self):
        mesh, pts, _ = make_mesh(400, seed=17)
        mid = mesh.select_midpoints(MIN_LEN, 300)
        assert not ({tuple(p) for p in mid.tolist()} & {tuple(p) for p in pts.tolist()})
        assert len(np.unique(mid, axis=0)) == len(mid)
        assert (mid >= 0).all() and (mid[:, 0] < W).all() and (mid[:, 1] < H).all()
        for _ in range(3):
            np.testing.assert_array_equal(mesh.select_midpoints(MIN_LEN, 300), mid)

    def test_inserting_them_back_is_accepted(self):
        mesh, pts, vals = make_mesh(400, seed=29)
        mid = mesh.select_midpoints(MIN_LEN, 150)
        before = mesh.seed_count
        new = np.arange(len(mid), dtype=np.float32)
        mesh.insert_deferred(mid, new)
        assert mesh.seed_count == before + len(mid)
        np.testing.assert_array_equal(mesh.get_seeds(), np.concatenate([pts, mid]))
        np.testing.assert_array_equal(mesh.get_values(), np.concatenate([vals, new]))
        # and a second round sees the new seeds
        allp, allv = np.concatenate([pts, mid]), np.concatenate([vals, new])
        rows, _, _ = oracle_state(allp)
        e = mesh.get_edges()
        np.testing.assert_array_equal(e, Q.edges(rows, allp))
        np.testing.assert_array_equal(mesh.select_midpoints(MIN_LEN, 200),
                                      Q.select_midpoints(W, H, P, allp, allv, e, MIN_LEN, 200))


# ------------------------------------------------------------------ locate
class TestLocate:

    @pytest.mark.parametrize("n,nq", [(50, 200), (400, 2000), (400, 9000)])
    def test_matches_the_finalised_grid(self, n, nq):
        mesh, pts, _ = make_mesh(n, seed=n + nq)
        q = queries(np.random.default_rng(n + nq + 1), nq)
        got = mesh.locate(q)
        _, grid = mesh.finalise(as_arrays=True)
        np.testing.assert_array_equal(got, grid[q[:, 1], q[:, 0], 2])

    def test_valid_before_any_finalise(self):
        mesh, pts, _ = make_mesh(400, seed=71)
        q = queries(np.random.default_rng(72), 3000)
        before = mesh.locate(q)
        _, grid = mesh.finalise(as_arrays=True)
        np.testing.assert_array_equal(before, grid[q[:, 1], q[:, 0], 2])

    def test_ids_index_get_triangles(self):
        mesh, pts, _ = make_mesh(400, seed=73)
        q = queries(np.random.default_rng(74), 2000)
        tids = mesh.locate(q)
        tris = mesh.get_triangles()
        hit = tids >= 0
        assert hit.any() and tids[hit].max() < len(tris)
        seeds = mesh.get_seeds().astype(np.int64)
        for i in np.flatnonzero(hit)[:300]:
            a, b, c = (seeds[j] for j in tris[tids[i]])
            p = q[i].astype(np.int64)
            z = [int((v[0] - u[0]) * (p[1] - u[1]) - (v[1] - u[1]) * (p[0] - u[0]))
                 for u, v in ((a, b), (b, c), (c, a))]
            assert not (min(z) < 0 < max(z))

    def test_ring_and_far_points_equal_the_oracle(self):
        mesh, pts, _ = make_mesh(300, seed=75)
        q = queries(np.random.default_rng(76), 4000, margin=40)
        rows, grid, sxy = oracle_state(pts)
        np.testing.assert_array_equal(mesh.locate(q),
                                      Q.locate(W, H, P, sxy, rows, grid[:, :, 2], q))

    def test_corners_and_empty(self):
        mesh, pts, _ = make_mesh(60, seed=77)
        far = np.array([[0, 0], [W - 1, 0], [0, H - 1], [W - 1, H - 1]], dtype=np.int32)
        _, grid = mesh.finalise(as_arrays=True)
        np.testing.assert_array_equal(mesh.locate(far), grid[far[:, 1], far[:, 0], 2])
        assert len(mesh.locate(np.zeros((0, 2), dtype=np.int32))) == 0
        with pytest.raises(ValueError):
            mesh.locate(np.zeros((3, 3), dtype=np.int32))

    def test_large_canvas_equals_the_oracle(self):
        w, h = 1024, 768
        mesh, pts, _ = make_mesh(20000, seed=78, w=w, h=h, max_seeds=40000)
        q = queries(np.random.default_rng(79), 200000, w, h, margin=20)
        rows, grid, sxy = oracle_state(pts, w, h)
        np.testing.assert_array_equal(mesh.locate(q), Q.locate(w, h, P, sxy, rows, grid[:, :, 2], q))


# --------------------------------------------------------- in_circumsphere
class TestInCircumsphere:

    @pytest.mark.parametrize("t_val", [0.0, 1.0, 4.0, 12.0, 40.0, 2.75])
    @pytest.mark.parametrize("p", [16, 7])
    def test_equals_the_oracle(self, t_val, p):
        mesh, pts, _ = make_mesh(400, seed=int(t_val) + 3, p=p)
        q = queries(np.random.default_rng(int(t_val) + 5), 3000)
        t = np.full(len(q), t_val)
        got, tids = mesh.in_circumsphere(q, t)
        rows, grid, sxy = oracle_state(pts, p=p)
        want_tid = Q.locate(W, H, p, sxy, rows, grid[:, :, 2], q)
        np.testing.assert_array_equal(tids, want_tid)
        np.testing.assert_array_equal(got, Q.in_circumsphere(p, sxy, rows, tids, q, t))

    def test_mixed_t_large(self):
        w, h = 1024, 768
        mesh, pts, _ = make_mesh(20000, seed=80, w=w, h=h, max_seeds=40000)
        rng = np.random.default_rng(81)
        q = queries(rng, 100000, w, h)
        t = rng.random(len(q)) * 8.0
        got, tids = mesh.in_circumsphere(q, t)
        rows, grid, sxy = oracle_state(pts, w, h)
        np.testing.assert_array_equal(tids, Q.locate(w, h, P, sxy, rows, grid[:, :, 2], q))
        np.testing.assert_array_equal(got, Q.in_circumsphere(P, sxy, rows, tids, q, t))

    def test_t_shrinks_the_admissible_radius(self):
        mesh, _, _ = make_mesh(400, seed=31)
        q = queries(np.random.default_rng(32), 3000)
        counts = [int(mesh.in_circumsphere(q, np.full(len(q), tv))[0].sum())
                  for tv in (0.0, 2.0, 5.0, 10.0, 25.0)]
        assert counts == sorted(counts, reverse=True) and counts[0] > counts[-1]

    def test_far_enough_in_time_admits_nothing(self):
        mesh, _, _ = make_mesh(400, seed=33)
        q = queries(np.random.default_rng(34), 500)
        m, _ = mesh.in_circumsphere(q, np.full(len(q), 1e6))
        assert not m.any()

    def test_reports_the_containing_triangle(self):
        mesh, _, _ = make_mesh(400, seed=35)
        q = queries(np.random.default_rng(36), 2000)
        m, tids = mesh.in_circumsphere(q, np.zeros(len(q)))
        np.testing.assert_array_equal(tids, mesh.locate(q))
        assert not m[tids < 0].any() and m.dtype == bool

    def test_length_mismatch_raises(self):
        mesh, _, _ = make_mesh(60, seed=39)
        with pytest.raises(ValueError):
            mesh.in_circumsphere(np.array([[10, 10], [20, 20]], dtype=np.int32), np.zeros(1))


# ---------------------------------------------------------- finalise_device
def _read(view, dtype):
    cai = view.__cuda_array_interface__
    out = np.empty(cai["shape"][0], dtype=dtype)
    _lib.check(_lib.load().dl_device_to_host(out.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_void_p(cai["data"][0]), out.nbytes))
    return out


class TestFinaliseDevice:

    @pytest.mark.parametrize("n,w,h", [(10, 64, 48), (40, 63, 47), (150, 65, 49), (2000, 640, 480),
                                       (300, 201, 3)])
    def test_matches_finalise(self, n, w, h):
        mesh, pts, _ = make_mesh(n, seed=n, w=w, h=h)
        verts, tids, sids, mask = mesh.finalise_device()
        ref_verts, grid = mesh.finalise(as_arrays=True)
        # finalise with nothing pending leaves the generation (and views) alone
        t_ref, s_ref, m_ref = Q.split_grid(grid)
        np.testing.assert_array_equal(verts, ref_verts)
        np.testing.assert_array_equal(_read(tids, np.int32), t_ref)
        np.testing.assert_array_equal(_read(sids, np.int32), s_ref)
        np.testing.assert_array_equal(_read(mask, np.uint8), m_ref)
        np.testing.assert_array_equal(tids.to_host(), t_ref)
        np.testing.assert_array_equal(sids.to_host(), s_ref)
        np.testing.assert_array_equal(mask.to_host(), m_ref)

    def test_interface_shape(self):
        mesh, _, _ = make_mesh(30, seed=0, w=64, h=48)
        _, *views = mesh.finalise_device()
        for v, ts in zip(views, ("<i4", "<i4", "|u1")):
            cai = v.__cuda_array_interface__
            assert cai["shape"] == (64 * 48,) and cai["typestr"] == ts
            assert cai["data"][1] is True and cai["version"] == 3 and cai["strides"] is None

    def test_dlpack_into_torch(self):
        """torch.from_dlpack over the views.  Runs in a child process that
        initialises torch first: torch ships its own HIP runtime, and the
        library then binds to that one (two runtimes in one process do not
        share the device)."""
        pytest.importorskip("torch")
        script = textwrap.dedent("""
            import numpy as np, torch
            torch.cuda.init()
            import integer_delauney_amd as dl
            rng = np.random.default_rng(2)
            pts = np.unique(rng.integers([0, 0], [200, 150], size=(1200, 2)), axis=0)[:300]
            mesh = dl.Delaunay(200, 150, 1000, 16)
            mesh.insert_deferred(pts.astype(np.int32), rng.random(len(pts)))
            verts, tids, sids, mask = mesh.finalise_device()
            _, grid = mesh.finalise(as_arrays=True)
            g = grid.reshape(-1, 3)
            t = torch.from_dlpack(tids)
            assert t.device.type == "cuda" and t.dtype == torch.int32 and t.shape == (30000,)
            assert np.array_equal(t.cpu().numpy(), np.where(g[:, 2] < 0, 0, g[:, 2]))
            assert np.array_equal(torch.from_dlpack(sids).cpu().numpy(), g[:, 0])
            m = torch.from_dlpack(mask)
            assert m.dtype == torch.uint8
            assert np.array_equal(m.cpu().numpy(), (g[:, 2] < 0).astype(np.uint8))
            assert int((t + 1).sum()) == int(np.where(g[:, 2] < 0, 0, g[:, 2]).sum()) + 30000
            del t, m
            tids.__dlpack__()          # a capsule nobody consumes is released
            print("dlpack ok")
        """)
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        r = subprocess.run([sys.executable, "-c", script], cwd=root, capture_output=True, text=True,
                           timeout=240, env=dict(os.environ, PYTHONPATH=root))
        assert r.returncode == 0 and "dlpack ok" in r.stdout, r.stdout + r.stderr

    @pytest.mark.parametrize("mutate", ["insert", "remove", "finalise_device"])
    def test_stale_views_raise(self, mutate):
        mesh, pts, _ = make_mesh(20, seed=2, w=64, h=48)
        _, *views = mesh.finalise_device()
        if mutate == "insert":
            free = [(x, y) for x in range(64) for y in range(48)
                    if (x, y) not in {tuple(p) for p in pts.tolist()}][:3]
            mesh.insert_deferred(np.array(free, dtype=np.int32), np.zeros(3))
        elif mutate == "remove":
            mesh.remove(pts[:2])
        else:
            _, *fresh = mesh.finalise_device()
            fresh[0].__cuda_array_interface__
        for v in views:
            with pytest.raises(RuntimeError):
                v.__cuda_array_interface__
            with pytest.raises(RuntimeError):
                v.to_host()
            with pytest.raises(RuntimeError):
                v.__dlpack__()


# ------------------------------------------------------------ atomic updates
def test_failed_update_leaves_the_session_unchanged():
    mesh, pts, vals = make_mesh(200, seed=41)
    m0, g0 = mesh.finalise(as_arrays=True)
    lib = _lib.load()
    rm = np.ascontiguousarray(pts[:5])
    bad = np.ascontiguousarray(np.concatenate([pts[10:11], pts[10:11]]))   # repeats
    vv = np.zeros(2, dtype=np.float32)
    with pytest.raises(ValueError):
        _lib.check(lib.dl_session_update(_lib.context().handle, mesh._h, _lib.i32p(rm), 5,
                                         _lib.i32p(bad), _lib.f32p(vv), 2))
    np.testing.assert_array_equal(mesh.device_seeds(), pts)
    np.testing.assert_array_equal(mesh.get_values(), vals)
    mesh._mutated()
    m1, g1 = mesh.finalise(as_arrays=True)
    np.testing.assert_array_equal(m1, m0)
    np.testing.assert_array_equal(g1, g0)
    # a good update: removal then insertion, values appended
    new = np.array([[1, 1], [2, 1]], dtype=np.int32)
    new = new[~np.isin(new[:, 0] * 1000 + new[:, 1], pts[:, 0] * 1000 + pts[:, 1])]
    nv = np.full(len(new), 7.0, dtype=np.float32)
    _lib.check(lib.dl_session_update(_lib.context().handle, mesh._h, _lib.i32p(rm), 5,
                                     _lib.i32p(np.ascontiguousarray(new)), _lib.f32p(nv), len(new)))
    np.testing.assert_array_equal(mesh.device_seeds(), np.concatenate([pts[5:], new]))
