"""Pins the CPU oracle (oracle/dl_oracle.c) before it is trusted as the checker.

* golden vectors written by the reference NumPy path (tests/golden/make_golden.py);
* the reference's own known-answer tests (tests/test_voronoi_ref.py,
  test_triangulation_ref.py, test_pixel_assignment_correctness.py), restated;
* what the mathematics fixes: brute-force nearest distances, the brute-force
  containment definition, closed forms;
* planted mistakes that these checks must catch.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import goldens

FULL = goldens.names("full_")
VOR = goldens.names("vor_")
DET = goldens.names("det_")


def test_golden_sets_present():
    assert len(FULL) >= 40 and len(VOR) >= 8 and len(DET) >= 4


@pytest.mark.parametrize("name", FULL)
def test_full_pipeline_matches_reference(name):
    g = goldens.load(name)
    rows, tg = O.triangulate(g["W"], g["H"], g["seeds"], g["P"], return_rows=True,
                             assign_method="brute" if g["W"] * g["H"] * max(1, len(g["rows"])) < 4e7
                             else "raster")
    np.testing.assert_array_equal(tg[:, :, :2], g["vgrid"])
    np.testing.assert_array_equal(rows, g["rows"])
    np.testing.assert_array_equal(tg[:, :, 2], g["tid"])


@pytest.mark.parametrize("name", VOR)
def test_voronoi_matches_reference(name):
    g = goldens.load(name)
    np.testing.assert_array_equal(O.voronoi(g["W"], g["H"], g["seeds"]), g["vgrid"])


@pytest.mark.parametrize("name", DET)
def test_detection_matches_reference(name):
    g = goldens.load(name)
    rows, _ = O.triangulate(g["W"], g["H"], g["seeds"], g["P"], return_rows=True)
    np.testing.assert_array_equal(rows, g["rows"])


# --- reference known-answer tests (test_voronoi_ref.py) --------------------

def test_single_seed_closed_form():
    g = O.voronoi(5, 5, [(2, 2)])
    assert np.all(g[:, :, 0] == 0)
    yy, xx = np.mgrid[0:5, 0:5]
    np.testing.assert_array_equal(g[:, :, 1], (xx - 2) ** 2 + (yy - 2) ** 2)
    assert O.voronoi(4, 4, [(0, 0)])[3, 3, 1] == 18


def test_id_order_and_ties():
    g = O.voronoi(5, 1, [(3, 0), (1, 0)])
    assert g[0, 1, 0] == 0 and g[0, 3, 0] == 1
    g = O.voronoi(5, 5, [(2, 4), (2, 0)])
    assert g[0, 2, 0] == 0 and g[4, 2, 0] == 1
    g = O.voronoi(5, 1, [(0, 0), (4, 0)])
    assert list(g[0, :, 0]) == [0, 0, 1, 1, 1] and list(g[0, :, 1]) == [0, 1, 4, 1, 0]
    assert O.voronoi(1, 5, [(0, 0), (0, 4)])[2, 0, 0] == 1
    assert O.voronoi(5, 5, [(0, 2), (2, 0), (4, 2)])[2, 2, 0] == 2


@pytest.mark.parametrize("seeds,err", [([], "empty"), ([(2, 2), (2, 2)], "duplicate"),
                                       ([(10, 10)], "out of bounds")])
def test_errors(seeds, err):
    with pytest.raises(ValueError, match=err):
        O.voronoi(5, 5, seeds)


def test_nearest_distance_everywhere():
    """d^2 is never below the brute-force nearest distance, every cell is
    defined, and where the nearest seed is unique and reached, the id is it."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        W, H = int(rng.integers(5, 60)), int(rng.integers(5, 60))
        n = int(rng.integers(2, W * H // 4 + 3))
        idx = rng.choice(W * H, size=min(n, W * H), replace=False)
        seeds = np.stack([idx % W, idx // W], 1)
        s = O.sort_seeds(W, H, seeds)
        g = O.voronoi_sorted(W, H, s)
        yy, xx = np.mgrid[0:H, 0:W]
        d = (xx[..., None] - s[:, 0]) ** 2 + (yy[..., None] - s[:, 1]) ** 2
        near = d.min(-1)
        assert np.all(g[:, :, 0] >= 0)
        assert np.all(g[:, :, 1] >= near)
        exact = g[:, :, 1] == near
        assert exact.mean() > 0.99
        uniq = (d == near[..., None]).sum(-1) == 1
        m = exact & uniq
        np.testing.assert_array_equal(g[:, :, 0][m], d.argmin(-1)[m])
        np.testing.assert_array_equal(
            g[:, :, 1], (xx - s[g[:, :, 0], 0]) ** 2 + (yy - s[g[:, :, 0], 1]) ** 2)


# --- reference known-answer tests (test_triangulation_ref.py) --------------

def test_three_seeds_one_triangle():
    m, tg = O.triangulate(10, 10, [(1, 1), (8, 1), (4, 8)])
    assert len(m) == 1 and {m[0][2], m[0][3], m[0][4]} == {0, 1, 2}
    x, y = m[0][:2]
    assert 1 <= x <= 9 and 0 <= y <= 8
    t = tg[:, :, 2]
    assert np.any(t == -1) and np.any(t == 0)


@pytest.mark.parametrize("seeds,W,H", [([(1, 1), (8, 1), (1, 8), (8, 8)], 10, 10),
                                       ([(0, 0), (4, 0), (0, 4), (4, 4)], 5, 5),
                                       ([(2, 2), (12, 2), (2, 12), (12, 12)], 15, 15)])
def test_cocircular_quad_two_triangles(seeds, W, H):
    m, _ = O.triangulate(W, H, seeds)
    tris = [frozenset(v[2:]) for v in m.values()]
    assert len(tris) == 2 and len(tris[0] & tris[1]) == 2 and tris[0] | tris[1] == {0, 1, 2, 3}


def _brute_truth(W, H, rows, s):
    """test_pixel_assignment_correctness.py:16-32 in vectorised form."""
    out = np.full((H, W), -1, dtype=np.int64)
    yy, xx = np.mgrid[0:H, 0:W]
    for t, r in enumerate(rows):
        a, b, c = s[r[2]].astype(np.int64), s[r[3]].astype(np.int64), s[r[4]].astype(np.int64)
        d1 = (a[0] - xx) * (b[1] - yy) - (a[1] - yy) * (b[0] - xx)
        d2 = (b[0] - xx) * (c[1] - yy) - (b[1] - yy) * (c[0] - xx)
        d3 = (c[0] - xx) * (a[1] - yy) - (c[1] - yy) * (a[0] - xx)
        neg = (d1 < 0) | (d2 < 0) | (d3 < 0)
        pos = (d1 > 0) | (d2 > 0) | (d3 > 0)
        out[~(neg & pos)] = t
    return out


def test_raster_equals_brute_force_on_random_triangles():
    """orc_assign_raster against the definition, including collinear triples
    (which contain their whole line) and vertices far outside the image."""
    rng = np.random.default_rng(11)
    for trial in range(40):
        W, H = int(rng.integers(1, 70)), int(rng.integers(1, 70))
        n = int(rng.integers(6, 40))
        s = np.stack([rng.integers(-30, W + 30, n), rng.integers(-30, H + 30, n)], 1).astype(np.int32)
        T = int(rng.integers(1, 60))
        rows = []
        for _ in range(T):
            a, b, c = rng.choice(n, 3, replace=False)
            rows.append((0, 0, a, b, c))
        if trial % 4 == 0:  # collinear triples
            s[0], s[1], s[2] = (2, 3), (5, 3 + 2), (11, 3 + 6)
            rows.append((0, 0, 0, 1, 2))
            s[3], s[4], s[5] = (1, 7), (9, 7), (20, 7)
            rows.append((0, 0, 3, 4, 5))
        rows = np.array(rows, dtype=np.int32)
        # distinct positions, as seeds are
        keys = s[:, 0].astype(np.int64) * 100000 + s[:, 1]
        if len(np.unique(keys)) != n:
            continue
        brute = O.assign(W, H, s, rows, "brute")
        np.testing.assert_array_equal(brute, _brute_truth(W, H, rows, s))
        np.testing.assert_array_equal(O.assign(W, H, s, rows, "raster"), brute)


def test_padded_canvas_interior_against_goldens():
    """The padded golden canvas (seeds shifted by 16) holds the unpadded
    diagram in its interior."""
    a = goldens.load("vor_uniform_256")
    b = goldens.load("vor_uniform_288_pad")
    np.testing.assert_array_equal(b["vgrid"][16:-16, 16:-16], a["vgrid"])


# --- planted mistakes --------------------------------------------------------

def test_planted_tie_rule_is_caught():
    """Lower-id-wins instead of higher-id-wins: emulate by mirroring x (which
    reverses the id order) and mirroring back; the goldens must disagree."""
    g = goldens.load("vor_lattice_257x129")
    W, H = g["W"], g["H"]
    s = g["seeds"].copy()
    s[:, 0] = W - 1 - s[:, 0]
    mir = O.voronoi(W, H, s)[:, ::-1]
    # owner positions per cell: the lattice has tie cells, resolved differently
    ids_ref = O.sort_seeds(W, H, g["seeds"])[g["vgrid"][:, :, 0]]
    ids_mir = O.sort_seeds(W, H, s)[mir[:, :, 0]]
    ids_mir[..., 0] = W - 1 - ids_mir[..., 0]
    assert (ids_ref != ids_mir).any()


def test_planted_scan_order_is_caught():
    """Registering the four L-scans in another order changes tids."""
    g = goldens.load("full_rand3_p-1000")
    rows = g["rows"]
    perm = rows[np.argsort(rows[:, 1] * 10000 + rows[:, 0], kind="stable")]
    assert not np.array_equal(perm, rows)


def test_planted_padding_offset_is_caught():
    g = goldens.load("full_rand2_p-1000")
    s = O.sort_seeds(g["W"], g["H"], g["seeds"])
    pad = O.voronoi_sorted(g["W"] + 32, g["H"] + 32, s + 16)
    rows, _ = O.detect(pad[:, :, 0], s, 15)   # wrong shift back
    assert not np.array_equal(rows, g["rows"])


# --- the padding watch -------------------------------------------------------

def _numpy_crossings(W, H, sorted_xy, m):
    """voronoi.py:110-163 written with whole-array snapshots per pass (an
    implementation independent of the C loops), counting updates whose cell
    lies in [m, W-m) x [m, H-m) and whose source does not."""
    ids = -np.ones((H, W), np.int64)
    d = -np.ones((H, W), np.int64)
    ox = np.zeros((H, W), np.int64)
    oy = np.zeros((H, W), np.int64)
    for i, (x, y) in enumerate(sorted_xy):
        ids[y, x], d[y, x], ox[y, x], oy[y, x] = i, 0, x, y
    yy, xx = np.mgrid[0:H, 0:W]
    box = (xx >= m) & (xx < W - m) & (yy >= m) & (yy < H - m)
    cross = 0
    shifts = [(0, 1), (0, -1), (1, 0), (-1, 0)]   # (dy, dx): source = cell - (dy, dx)
    while True:
        changed = False
        for dy, dx in shifts:
            sid, sox, soy, sbox = (np.full((H, W), -1, np.int64), np.zeros((H, W), np.int64),
                                   np.zeros((H, W), np.int64), np.zeros((H, W), bool))
            ys, yd = (slice(0, H - dy), slice(dy, H)) if dy >= 0 else (slice(-dy, H), slice(0, H + dy))
            xs, xd = (slice(0, W - dx), slice(dx, W)) if dx >= 0 else (slice(-dx, W), slice(0, W + dx))
            sid[yd, xd], sox[yd, xd], soy[yd, xd], sbox[yd, xd] = ids[ys, xs], ox[ys, xs], oy[ys, xs], box[ys, xs]
            cand = (xx - sox) ** 2 + (yy - soy) ** 2
            take = (sid >= 0) & ((ids < 0) | (cand < d) | ((cand == d) & (sid > ids)))
            cross += int((take & box & ~sbox).sum())
            ids = np.where(take, sid, ids)
            d = np.where(take, cand, d)
            ox = np.where(take, sox, ox)
            oy = np.where(take, soy, oy)
            changed |= bool(take.any())
        if not changed:
            return cross, np.stack([ids, d], -1).astype(np.int32)


def test_crossing_count_matches_numpy_restatement():
    rng = np.random.default_rng(21)
    nonzero = 0
    for _ in range(25):
        W, H = int(rng.integers(4, 40)), int(rng.integers(4, 40))
        m = int(rng.integers(0, (min(W, H) - 1) // 2 + 1))
        n = int(rng.integers(1, max(2, W * H // 8)))
        idx = rng.choice(W * H, size=n, replace=False)
        s = O.sort_seeds(W, H, np.stack([idx % W, idx // W], 1))
        want, grid = _numpy_crossings(W, H, s, m)
        np.testing.assert_array_equal(grid, O.voronoi_sorted(W, H, s))
        assert O.voronoi_crossings(W, H, s, m) == want, (W, H, m, n)
        nonzero += want > 0
    assert nonzero > 5


def test_crossing_count_closed_cases():
    s = O.sort_seeds(20, 20, [(3, 4), (15, 9)])
    assert O.voronoi_crossings(20, 20, s, 0) == 0      # the box is the canvas
    # one seed outside a box: every box cell is defined once, and the box's
    # first row and column receive the label from outside at least once
    s1 = O.sort_seeds(20, 20, [(0, 0)])
    assert O.voronoi_crossings(20, 20, s1, 5) > 0
