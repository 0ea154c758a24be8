"""Pins the CPU restatement of the session queries (oracle/queries.py) before
it is trusted as the checker of the device queries.

The reference has no host implementation of these queries; its tests
(tests/test_scalar_field.py) check them against host formulas and the
finalised grid.  Here each function is checked against something it was not
written from:

* locate: a brute-force raster of the padded canvas (the C oracle's
  O(H*W*T) assignment loop on seeds shifted by P), at every in-image and ring
  point;
* in_circumsphere: exact rational arithmetic (fractions) away from the
  boundary, and closed forms (right triangle, collinear triple);
* edges: a Python set built triangle by triangle, and a closed case;
* edge_scores: closed forms and the reference test's own host formula
  (test_scalar_field.py:55-62);
* select_midpoints: the reference test's host selection restated with sets
  and Python sorting (test_scalar_field.py:78-99), closed cases for seed
  collisions, midpoint collisions and the odd-padding tie;
* planted mistakes that these checks catch.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from oracle import queries as Q


def _mesh(W, H, n, seed, P=16):
    rng = np.random.default_rng(seed)
    cand = np.unique(rng.integers([0, 0], [W, H], size=(n * 4, 2)), axis=0)
    pts = cand[np.sort(rng.choice(len(cand), size=min(n, len(cand)), replace=False))]
    pts = pts[rng.permutation(len(pts))].astype(np.int32)   # insertion order
    rows, grid = O.triangulate(W, H, pts, P, return_rows=True)
    sorted_xy = O.sort_seeds(W, H, pts)
    vals = rng.random(len(pts)).astype(np.float32)
    return pts, vals, rows, grid, sorted_xy


# ---------------------------------------------------------------- locate
def _synthetic(W, H, n_tri, seed):
    """Arbitrary triangles over random image points plus collinear triples (a
    horizontal run and a slope-1/2 run): those contain their whole line, which
    runs into the padding ring, where no other triangle reaches (a triangle
    lies inside the image with its vertices).  The in-image grid is the
    brute-force assignment."""
    rng = np.random.default_rng(seed)
    line = [(x, 7) for x in range(2, W, 9)] + [(3 + 2 * k, 1 + k) for k in range(3)]
    rand = rng.integers([0, 0], [W, H], size=(4 * n_tri, 2)).tolist()
    pts = np.array(sorted(set(map(tuple, rand)) | set(line)), dtype=np.int32)
    at = {p: i for i, p in enumerate(map(tuple, pts.tolist()))}
    ids = [rng.choice(len(pts), 3, replace=False) for _ in range(n_tri)]
    ids += [[at[line[k]], at[line[k + 1]], at[line[k + 2]]] for k in range(0, len(line) - 5, 2)]
    ids.append([at[p] for p in line[-3:]])
    rows = np.zeros((len(ids), 5), dtype=np.int32)
    rows[:, 2:5] = np.array(ids)
    return pts, rows, O.assign(W, H, pts, rows, method="brute")


@pytest.mark.parametrize("seed,P", [(1, 16), (2, 5), (3, 0)])
def test_locate_equals_padded_brute_force(seed, P):
    W, H = 60, 44
    pts, _, rows, grid, sorted_xy = _mesh(W, H, 40, seed, P)
    cases = [(sorted_xy, rows, grid[:, :, 2])] + [_synthetic(W, H, 30, seed)]
    for sx, rw, gt in cases:
        brute = O.assign(W + 2 * P, H + 2 * P, sx + P, rw, method="brute")   # padded raster
        ys, xs = np.mgrid[-P - 2:H + P + 2, -P - 2:W + P + 2]
        q = np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.int32)
        got = Q.locate(W, H, P, sx, rw, gt, q)
        inside = (q[:, 0] >= -P) & (q[:, 0] < W + P) & (q[:, 1] >= -P) & (q[:, 1] < H + P)
        want = np.full(len(q), -1, dtype=np.int32)
        want[inside] = brute[q[inside, 1] + P, q[inside, 0] + P]
        np.testing.assert_array_equal(got, want)
        assert (got[inside] >= 0).any()
    if P > 0:
        ring = inside & ~((q[:, 0] >= 0) & (q[:, 0] < W) & (q[:, 1] >= 0) & (q[:, 1] < H))
        assert (got[ring] >= 0).any()      # the collinear triples reach the ring


def test_locate_planted_mistake_is_caught():
    """Min tid instead of max tid on overlaps must disagree in the ring."""
    W, H, P = 60, 44, 8
    sx, rows, gt = _synthetic(W, H, 30, 4)
    tri = Q.triangle_xy(sx, rows)
    ring = [(x, y) for x in range(-P, W + P) for y in range(-P, H + P)
            if not (0 <= x < W and 0 <= y < H)]
    bad = []
    for x, y in ring:
        hit = np.flatnonzero(Q._contains(x, y, tri))
        bad.append(int(hit[0]) if hit.size else -1)
    good = Q.locate(W, H, P, sx, rows, gt, np.array(ring))
    assert (np.array(bad) != good).any()


# ------------------------------------------------------- in_circumsphere
def _exact_inside(tri, p, t):
    """Exact d^2 + t^2 - R^2 (Fraction) for a non-degenerate triangle."""
    (ax, ay), (bx, by), (cx, cy) = [(Fraction(int(u)), Fraction(int(v))) for u, v in tri]
    d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay)
          + (cx * cx + cy * cy) * (ay - by)) / d
    uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx)
          + (cx * cx + cy * cy) * (bx - ax)) / d
    r2 = (ax - ux) ** 2 + (ay - uy) ** 2
    px, py = Fraction(int(p[0])), Fraction(int(p[1]))
    return (px - ux) ** 2 + (py - uy) ** 2 + Fraction(t) ** 2 - r2, r2


@pytest.mark.parametrize("P", [0, 16, 7])
def test_in_circumsphere_matches_exact_arithmetic(P):
    rng = np.random.default_rng(5 + P)
    n, checked = 400, 0
    sorted_xy = rng.integers(0, 300, size=(3 * n, 2)).astype(np.int32)
    rows = np.zeros((n, 5), dtype=np.int32)
    rows[:, 2:5] = np.arange(3 * n).reshape(n, 3)
    q = rng.integers(-20, 320, size=(n, 2)).astype(np.int32)
    t = rng.choice([0.0, 0.5, 3.0, 17.25], size=n)
    got = Q.in_circumsphere(P, sorted_xy, rows, np.arange(n), q, t)
    for i in range(n):
        tri = sorted_xy[3 * i:3 * i + 3]
        area2 = int((tri[1, 0] - tri[0, 0]) * (tri[2, 1] - tri[0, 1])
                    - (tri[1, 1] - tri[0, 1]) * (tri[2, 0] - tri[0, 0]))
        if area2 == 0:
            continue
        diff, r2 = _exact_inside(tri, q[i], t[i])
        if abs(diff) <= r2 * Fraction(1, 10 ** 9):
            continue   # on the boundary: float rounding decides
        assert got[i] == (diff < 0), (tri, q[i], t[i])
        checked += 1
    assert checked > 350


def test_in_circumsphere_closed_forms():
    # right triangle: centre at the hypotenuse midpoint (2, 1.5), R = 2.5
    sorted_xy = np.array([[0, 0], [4, 0], [0, 3], [0, 5], [2, 5], [4, 5]], dtype=np.int32)
    rows = np.array([[0, 0, 0, 1, 2], [0, 0, 3, 4, 5]], dtype=np.int32)
    q = np.array([[2, 1], [2, 1], [2, 1], [4, 0], [3, 5], [3, 5], [1, 5]], dtype=np.int32)
    tid = np.array([0, 0, 0, 0, 1, 1, 1])
    t = np.array([0.0, 2.4, 2.5, 0.0, 0.0, 1.8, 1.0])
    got = Q.in_circumsphere(0, sorted_xy, rows, tid, q, t)
    # sqrt(.25) < 2.5; sqrt(.25 + 5.76) = 2.452 < 2.5; sqrt(6.5) > 2.5; a vertex is on it;
    # collinear (0,5),(2,5),(4,5): centroid (2,5), farthest vertex R = 2:
    # |(3,5)| = 1 < 2; sqrt(1 + 3.24) > 2; sqrt(1 + 1) < 2
    np.testing.assert_array_equal(got, [True, True, False, False, True, False, True])
    assert not Q.in_circumsphere(0, sorted_xy, rows, np.array([-1]), q[:1], t[:1]).any()


def test_in_circumsphere_planted_mistake_is_caught():
    """A sign error in the circumcentre numerator must change answers."""
    rng = np.random.default_rng(9)
    n = 200
    sorted_xy = rng.integers(0, 200, size=(3 * n, 2)).astype(np.int32)
    rows = np.zeros((n, 5), dtype=np.int32)
    rows[:, 2:5] = np.arange(3 * n).reshape(n, 3)
    q = rng.integers(0, 200, size=(n, 2)).astype(np.int32)
    t = np.zeros(n)
    good = Q.in_circumsphere(0, sorted_xy, rows, np.arange(n), q, t)
    flipped = sorted_xy.copy()
    flipped[:, 1] = -flipped[:, 1]          # mirror: same geometry, other orientation
    qf = q.copy()
    qf[:, 1] = -qf[:, 1]
    mirrored = Q.in_circumsphere(0, flipped, rows, np.arange(n), qf, t)
    np.testing.assert_array_equal(good, mirrored)   # invariant under reflection
    exact = []
    for i in range(n):
        tri = sorted_xy[3 * i:3 * i + 3]
        area2 = int((tri[1, 0] - tri[0, 0]) * (tri[2, 1] - tri[0, 1])
                    - (tri[1, 1] - tri[0, 1]) * (tri[2, 0] - tri[0, 0]))
        d, _ = _exact_inside(tri, q[i], 0) if area2 else (1, 0)
        exact.append(d < 0)
    # a centre computed with one numerator term negated disagrees somewhere
    v = sorted_xy.astype(np.float64)
    ax, ay, bx, by, cx, cy = v[0::3, 0], v[0::3, 1], v[1::3, 0], v[1::3, 1], v[2::3, 0], v[2::3, 1]
    nz = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
    ok = nz != 0
    ox_bad = ax + ((cx - ax) ** 2 + (cy - ay) ** 2) * (nz * (by - ay)) / (2 * nz * nz)
    ox_good = ax + (((cx - ax) ** 2 + (cy - ay) ** 2) * (-(nz * (by - ay)))
                    + ((bx - ax) ** 2 + (by - ay) ** 2) * ((cy - ay) * nz)) / (2 * nz * nz)
    assert (np.abs(ox_bad - ox_good)[ok] > 1e-6).any()
    assert (good[ok] == np.array(exact)[ok]).mean() > 0.99


# ------------------------------------------------------------------ edges
def test_edges_match_python_set():
    W, H = 120, 90
    pts, _, rows, _, sorted_xy = _mesh(W, H, 120, 11)
    ins = {tuple(sorted_xy[k]): k for k in range(len(sorted_xy))}
    order = [ins[tuple(p)] for p in pts]            # insertion index -> sorted id
    sorted_to_ins = {s: i for i, s in enumerate(order)}
    want = set()
    for r in rows:
        v = [sorted_to_ins[int(s)] for s in r[2:5]]
        for a, b in ((v[0], v[1]), (v[1], v[2]), (v[2], v[0])):
            want.add((min(a, b), max(a, b)))
    got = Q.edges(rows, pts)
    assert [tuple(e) for e in got.tolist()] == sorted(want)


def test_edges_closed_case():
    # two triangles sharing the edge (1, 2): five distinct edges
    xy = np.array([[0, 0], [10, 0], [0, 10], [10, 10]], dtype=np.int32)
    rows = np.array([[0, 0, 0, 1, 2], [0, 0, 3, 2, 1]], dtype=np.int32)   # sorted ids
    ins = Q.insertion_of_sorted(xy)
    np.testing.assert_array_equal(ins, [0, 2, 1, 3])   # sorted by (x, y)
    e = Q.edges(rows, xy)
    # sorted ids 0,1,2,3 are insertion 0,2,1,3
    assert e.tolist() == [[0, 1], [0, 2], [1, 2], [1, 3], [2, 3]]


# ------------------------------------------------------------ edge scores
def test_edge_scores_closed_forms():
    xy = np.array([[0, 0], [3, 4], [3, 5], [100, 4]], dtype=np.int32)
    vals = np.array([1.0, 3.0, 3.0, 0.5], dtype=np.float32)
    e = np.array([[0, 1], [1, 2], [1, 3], [0, 3]], dtype=np.int32)
    s = Q.edge_scores(xy, vals, e, 2.0)
    # |3-1| * 5 = 10; length 1 < 2 -> 0; 2.5 * 97; 0.5 * sqrt(10016)
    np.testing.assert_array_equal(s, np.float32([10.0, 0.0, 242.5, np.float32(0.5 * np.sqrt(10016.0))]))
    assert Q.edge_scores(xy, vals, e, 0.0)[1] == 0.0   # equal values: no contrast


def test_edge_scores_match_reference_host_formula():
    W, H = 240, 180
    pts, vals, rows, _, _ = _mesh(W, H, 400, 12)
    e = Q.edges(rows, pts)
    got = Q.edge_scores(pts, vals, e, 2.0)
    d = pts[e[:, 1]].astype(np.float64) - pts[e[:, 0]].astype(np.float64)
    length = np.hypot(d[:, 0], d[:, 1])
    want = np.abs(vals[e[:, 1]].astype(np.float64) - vals[e[:, 0]].astype(np.float64)) * length
    want[length < 2.0] = 0.0
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


# ------------------------------------------------------- select_midpoints
def _host_selection(W, H, pts, e, scores, take, threshold=0.0):
    """test_scalar_field.py:78-99, restated with Python sorting and sets."""
    cand = sorted((-float(scores[i]), i) for i in range(len(e))
                  if scores[i] > np.float32(threshold) and scores[i] > 0)[:take]
    taken = {tuple(p) for p in pts.tolist()}
    seen, out = set(), []
    for _, i in cand:
        a, b = pts[e[i, 0]], pts[e[i, 1]]
        t = (int(np.rint((a[0] + b[0]) / 2.0)), int(np.rint((a[1] + b[1]) / 2.0)))
        t = (min(max(t[0], 0), W - 1), min(max(t[1], 0), H - 1))
        if t in taken or t in seen:
            continue
        seen.add(t)
        out.append(t)
    return out


@pytest.mark.parametrize("take", [1, 10, 75, 300, 5000])
def test_select_matches_reference_host_selection(take):
    W, H = 240, 180
    pts, vals, rows, _, _ = _mesh(W, H, 400, 13)
    e = Q.edges(rows, pts)
    s = Q.edge_scores(pts, vals, e, 2.0)
    got = Q.select_midpoints(W, H, 16, pts, vals, e, 2.0, take)
    want = _host_selection(W, H, pts, e, s, take)
    assert {tuple(p) for p in got.tolist()} == set(want)
    assert len(got) == len(want)
    pix = got[:, 1].astype(np.int64) * W + got[:, 0]
    assert (np.diff(pix) > 0).all()           # raster order, unique


def test_select_threshold_and_closed_cases():
    W, H = 40, 40
    # edges: (0,1) midpoint (5,0) score 10*|2-0|=20; (0,2) midpoint (0,5) lands on seed 3;
    # (1,2) midpoint (5,5) score lower; (4,5) midpoint (5,5) too: collision keeps the better
    xy = np.array([[0, 0], [10, 0], [0, 10], [0, 5], [2, 5], [8, 5]], dtype=np.int32)
    vals = np.array([0.0, 2.0, 1.0, 0.0, 0.0, 5.0], dtype=np.float32)
    e = np.array([[0, 1], [0, 2], [1, 2], [4, 5]], dtype=np.int32)
    s = Q.edge_scores(xy, vals, e, 0.0)
    assert s[0] == 20.0 and s[1] == 10.0 and s[3] == 30.0
    got = Q.select_midpoints(W, H, 16, xy, vals, e, 0.0, 10)
    assert got.tolist() == [[5, 0], [5, 5]]            # (0,5) is a seed; (5,5) once
    assert Q.select_midpoints(W, H, 16, xy, vals, e, 0.0, 10, threshold=25.0).tolist() == [[5, 5]]
    assert Q.select_midpoints(W, H, 16, xy, vals, e, 0.0, 10, threshold=30.0).tolist() == []
    assert Q.select_midpoints(W, H, 16, xy, vals, e, 0.0, 1).tolist() == [[5, 5]]


def test_select_odd_padding_tie_rounds_on_the_padded_canvas():
    """0.5 * (0 + 5) = 2.5 rounds to 2 in image coordinates but the reference
    rounds (1 + 6) / 2 = 3.5 to 4 on a canvas padded by 1, i.e. 3."""
    xy = np.array([[0, 0], [5, 0]], dtype=np.int32)
    vals = np.array([0.0, 1.0], dtype=np.float32)
    e = np.array([[0, 1]], dtype=np.int32)
    assert Q.select_midpoints(20, 20, 1, xy, vals, e, 0.0, 1).tolist() == [[3, 0]]
    assert Q.select_midpoints(20, 20, 2, xy, vals, e, 0.0, 1).tolist() == [[2, 0]]
    assert Q.select_midpoints(20, 20, 0, xy, vals, e, 0.0, 1).tolist() == [[2, 0]]


def test_split_grid():
    g = np.array([[[0, 0, 3], [1, 4, -1]], [[2, 9, 0], [1, 1, 7]]], dtype=np.int32)
    t, sid, m = Q.split_grid(g)
    assert t.tolist() == [3, 0, 0, 7] and sid.tolist() == [0, 1, 2, 1] and m.tolist() == [0, 1, 0, 0]
