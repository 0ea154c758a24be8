"""The incremental `Delaunay` session (insert / insert_deferred / finalise /
remove): after any sequence of updates its state equals the batch pipeline on
the current seed set, bit for bit, and its id conventions follow the
reference's (bindings.cpp:86-227; API_REFERENCE.md:315-340)."""
import numpy as np
import pytest

import integer_delauney_amd as dl
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rows(m):
    return np.array([m[i] for i in range(len(m))], dtype=np.int32).reshape(-1, 5)


def _distinct(rng, W, H, n, exclude=()):
    taken = {tuple(p) for p in exclude}
    out = []
    while len(out) < n:
        p = (int(rng.integers(0, W)), int(rng.integers(0, H)))
        if p not in taken:
            taken.add(p)
            out.append(p)
    return np.array(out, dtype=np.int32)


def test_insert_remove_sequence_equals_batch():
    rng = np.random.default_rng(31)
    W, H = 300, 220
    s = dl.Delaunay(W, H, 5000)
    live = np.empty((0, 2), np.int32)
    for step in range(6):
        add = _distinct(rng, W, H, int(rng.integers(50, 400)), live)
        s.insert_deferred(add)
        live = np.concatenate([live, add])
        if step % 2 and live.shape[0] > 100:
            drop = live[rng.choice(live.shape[0], size=40, replace=False)]
            s.remove(drop)
            keep = ~np.isin(live[:, 0] * 100000 + live[:, 1], drop[:, 0] * 100000 + drop[:, 1])
            live = live[keep]
        assert s.has_pending and s.seed_count == live.shape[0]
        m, grid = s.finalise()
        ref_rows, ref_grid = O.triangulate(W, H, live, None, return_rows=True)
        np.testing.assert_array_equal(grid, ref_grid)
        np.testing.assert_array_equal(_rows(m), ref_rows)
        m2, grid2 = dl.triangulate(W, H, live)
        assert m2 == m and np.array_equal(grid2, grid)
        assert not s.has_pending


def test_id_conventions():
    rng = np.random.default_rng(32)
    W, H = 120, 90
    s = dl.Delaunay(W, H, 1000, border_padding=8)
    a = _distinct(rng, W, H, 200)
    s.insert(a)
    b = _distinct(rng, W, H, 150, a)
    m, grid = s.insert(b, as_arrays=False)
    xy = np.concatenate([a, b])
    order = np.lexsort((xy[:, 1], xy[:, 0]))
    rank = np.empty(xy.shape[0], np.int32)
    rank[order] = np.arange(xy.shape[0])
    np.testing.assert_array_equal(s.sorted_rank, rank)
    np.testing.assert_array_equal(s.get_seeds(), xy)
    rows = _rows(m)
    inv = np.empty_like(rank)
    inv[rank] = np.arange(rank.shape[0])
    np.testing.assert_array_equal(s.get_triangles(), inv[rows[:, 2:5]])
    np.testing.assert_array_equal(s.get_voronoi_grid(), grid[:, :, :2])
    verts, _ = s.finalise(as_arrays=True)
    np.testing.assert_array_equal(verts, rows[:, 2:5])
    assert s.border_padding == 8 and s.width == W and s.height == H


@pytest.mark.parametrize("case", ["capacity", "duplicate_live", "duplicate_batch", "bounds",
                                  "remove_missing"])
def test_session_errors(case):
    s = dl.Delaunay(50, 40, 5)
    s.insert([(1, 1), (10, 10), (20, 5)])
    with pytest.raises(ValueError):
        if case == "capacity":
            s.insert([(2, 2), (3, 3), (4, 4)])
        elif case == "duplicate_live":
            s.insert([(10, 10)])
        elif case == "duplicate_batch":
            s.insert([(7, 7), (7, 7)])
        elif case == "bounds":
            s.insert([(50, 1)])
        else:
            s.remove([(30, 30)])
    # the failed call left the session unchanged
    assert s.seed_count == 3
    m, g = s.finalise()
    m2, g2 = dl.triangulate(50, 40, [(1, 1), (10, 10), (20, 5)])
    assert m == m2 and np.array_equal(g, g2)


def test_failed_calls_roll_back_on_the_device():
    s = dl.Delaunay(40, 30, 100)
    s.insert([(1, 1), (10, 10), (20, 5)])
    # a batch mixing a fresh position with a live one: the fresh bit set by the
    # failing call must be cleared again
    with pytest.raises(ValueError):
        s.insert([(7, 7), (10, 10)])
    s.insert([(7, 7)])                       # succeeds: (7, 7) was rolled back
    # a removal mixing a live position with an absent one: the live bit must
    # be restored
    with pytest.raises(ValueError):
        s.remove([(1, 1), (3, 3)])
    with pytest.raises(ValueError):
        s.remove([(20, 5), (20, 5)])
    s.remove([(1, 1)])                       # still live, so this succeeds
    np.testing.assert_array_equal(s.device_seeds(), s.get_seeds())
    np.testing.assert_array_equal(s.get_seeds(), np.array([(10, 10), (20, 5), (7, 7)], np.int32))
    m, g = s.finalise()
    m2, g2 = dl.triangulate(40, 30, s.get_seeds())
    assert m == m2 and np.array_equal(g, g2)


def test_remove_everything_then_refill():
    rng = np.random.default_rng(33)
    W, H = 64, 48
    s = dl.Delaunay(W, H, 500)
    a = _distinct(rng, W, H, 200)
    s.insert(a)
    s.remove(a)
    assert s.seed_count == 0
    m, g = s.finalise()
    assert len(m) == 0 and np.all(g == -1)
    b = _distinct(rng, W, H, 150)
    m, g = s.insert(b)
    m2, g2 = dl.triangulate(W, H, b)
    assert m == m2 and np.array_equal(g, g2)
    np.testing.assert_array_equal(s.device_seeds(), b)
