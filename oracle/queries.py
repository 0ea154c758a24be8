"""CPU restatement of the session queries -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of the device queries of
``integer_delauney_amd.Delaunay``; the product package never imports it.

The reference implements these queries on the device only (src/cuda); they
are restated here in NumPy from its kernels, one function each:

* :func:`locate`          -- delaunay_query.cu:113-133 (``locate_points_kernel``),
  with containment as the batch path defines it (triangulation.py:193-253:
  closed triangles, max tid on overlaps)
* :func:`in_circumsphere` -- delaunay_query.cu:33-100 (``in_circumsphere_kernel``)
* :func:`edges`           -- delaunay_scalar_field.cu:35-60, 185-214, 296-312
* :func:`edge_scores`     -- delaunay_scalar_field.cu:95-113
* :func:`select_midpoints` -- delaunay_scalar_field.cu:124-171, 230-294
* :func:`split_grid`      -- delaunay_output.cu:119-151 (``crop_pixel_arrays_kernel``)

Triangles are the batch rows ``(x, y, a, b, c)`` in sorted seed ids (what
``oracle.triangulate`` returns); ``sorted_xy`` are the seeds in sorted id
order, ``xy`` / ``values`` the live seeds in insertion order.  Where the
reference holds seeds on the padded canvas (image + P) and the rounding
depends on it (the circumcentre, midpoint ties), the same shift is applied.
"""
from __future__ import annotations

import numpy as np


def _cross(px, py, ax, ay, bx, by):
    return (ax - px) * (by - py) - (ay - py) * (bx - px)


def _contains(px, py, tri_xy):
    """(T,) bool: closed containment of point (px, py) in each triangle
    (triangulation.py:226-253).  tri_xy: (T, 3, 2) int64."""
    a, b, c = tri_xy[:, 0], tri_xy[:, 1], tri_xy[:, 2]
    d1 = _cross(px, py, a[:, 0], a[:, 1], b[:, 0], b[:, 1])
    d2 = _cross(px, py, b[:, 0], b[:, 1], c[:, 0], c[:, 1])
    d3 = _cross(px, py, c[:, 0], c[:, 1], a[:, 0], a[:, 1])
    neg = (d1 < 0) | (d2 < 0) | (d3 < 0)
    pos = (d1 > 0) | (d2 > 0) | (d3 > 0)
    return ~(neg & pos)


def triangle_xy(sorted_xy: np.ndarray, rows: np.ndarray) -> np.ndarray:
    """(T, 3, 2) int64 vertex coordinates of the rows, in row vertex order."""
    s = np.asarray(sorted_xy, dtype=np.int64)
    r = np.asarray(rows, dtype=np.int64).reshape(-1, 5)
    return s[r[:, 2:5]]


def locate(width: int, height: int, padding: int, sorted_xy, rows, grid_tid, points) -> np.ndarray:
    """tid of the closed triangle containing each point, max tid on overlaps,
    -1 for none; points outside the padded canvas get -1
    (delaunay_query.cu:128-131).  Inside the image this is the grid's tid
    channel ``grid_tid`` ((H, W) int32).  Outside it only collinear triples
    can contain a point (a non-degenerate closed triangle lies inside the
    bounding box of its vertices, which are image points), so ring points are
    tested against those."""
    q = np.asarray(points, dtype=np.int64).reshape(-1, 2)
    out = np.full(q.shape[0], -1, dtype=np.int32)
    rows = np.asarray(rows).reshape(-1, 5)
    if rows.shape[0] == 0:
        return out
    x, y = q[:, 0], q[:, 1]
    img = (x >= 0) & (x < width) & (y >= 0) & (y < height)
    out[img] = np.asarray(grid_tid)[y[img], x[img]]
    ring = ~img & (x >= -padding) & (x < width + padding) & (y >= -padding) & (y < height + padding)
    tri = triangle_xy(sorted_xy, rows)
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    S = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    lines = np.flatnonzero(S == 0)
    for i in np.flatnonzero(ring):
        hit = lines[_contains(x[i], y[i], tri[lines])]
        out[i] = int(hit[-1]) if hit.size else -1
    return out


def in_circumsphere(padding: int, sorted_xy, rows, tids, points, t) -> np.ndarray:
    """mask[i]: sqrt(dx^2 + dy^2 + t^2) < sqrt(R^2) about the circumcentre of
    triangle tids[i] (False where tids[i] < 0), in float64 with the operation
    order of delaunay_query.cu:58-100 on padded-canvas coordinates."""
    tids = np.asarray(tids)
    q = np.asarray(points, dtype=np.int64).reshape(-1, 2)
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    mask = np.zeros(q.shape[0], dtype=bool)
    ok = tids >= 0
    if not ok.any():
        return mask
    v = triangle_xy(sorted_xy, rows)[tids[ok]].astype(np.float64) + float(padding)
    ax, ay = v[:, 0, 0], v[:, 0, 1]
    bx, by = v[:, 1, 0], v[:, 1, 1]
    cx, cy = v[:, 2, 0], v[:, 2, 1]
    abx, aby = bx - ax, by - ay
    acx, acy = cx - ax, cy - ay
    nz = abx * acy - aby * acx
    collinear = nz * nz < 1e-12
    ab2 = abx * abx + aby * aby
    ac2 = acx * acx + acy * acy
    den = 2.0 * (nz * nz)
    num_x = ac2 * (-(nz * aby)) + ab2 * (acy * nz)
    num_y = ac2 * (nz * abx) + ab2 * (-(acx * nz))
    with np.errstate(divide="ignore", invalid="ignore"):
        ox = np.where(collinear, (ax + bx + cx) / 3.0, ax + num_x / den)
        oy = np.where(collinear, (ay + by + cy) / 3.0, ay + num_y / den)
    ra = (ax - ox) * (ax - ox) + (ay - oy) * (ay - oy)
    rb = (bx - ox) * (bx - ox) + (by - oy) * (by - oy)
    rc = (cx - ox) * (cx - ox) + (cy - oy) * (cy - oy)
    r2 = np.fmax(np.fmax(ra, rb), rc)
    dx = (q[ok, 0] + padding).astype(np.float64) - ox
    dy = (q[ok, 1] + padding).astype(np.float64) - oy
    tt = t[ok]
    mask[ok] = np.sqrt(dx * dx + dy * dy + tt * tt) < np.sqrt(r2)
    return mask


def insertion_of_sorted(xy: np.ndarray) -> np.ndarray:
    """sorted id -> insertion index (ids sort by x, then y)."""
    xy = np.asarray(xy).reshape(-1, 2)
    return np.lexsort((xy[:, 1], xy[:, 0])).astype(np.int64)


def edges(rows, xy) -> np.ndarray:
    """(E, 2) int32 distinct undirected edges (a < b) in insertion-order ids,
    ascending (a, b): three per triangle, deduplicated."""
    rows = np.asarray(rows).reshape(-1, 5)
    if rows.shape[0] == 0:
        return np.empty((0, 2), dtype=np.int32)
    ins = insertion_of_sorted(xy)
    v = ins[rows[:, 2:5].astype(np.int64)]
    p = np.concatenate([v[:, 0], v[:, 1], v[:, 2]])
    r = np.concatenate([v[:, 1], v[:, 2], v[:, 0]])
    pair = np.stack([np.minimum(p, r), np.maximum(p, r)], axis=1)
    return np.unique(pair, axis=0).astype(np.int32)


def edge_scores(xy, values, edge_list, min_length: float) -> np.ndarray:
    """(E,) float32: float(|v_a - v_b| * sqrt(len2)) in float64, 0 where
    len2 < min_length^2 (delaunay_scalar_field.cu:95-113)."""
    xy = np.asarray(xy, dtype=np.int64).reshape(-1, 2)
    values = np.asarray(values, dtype=np.float32)
    e = np.asarray(edge_list, dtype=np.int64).reshape(-1, 2)
    a, b = e[:, 0], e[:, 1]
    dx = xy[a, 0].astype(np.float64) - xy[b, 0].astype(np.float64)
    dy = xy[a, 1].astype(np.float64) - xy[b, 1].astype(np.float64)
    len2 = dx * dx + dy * dy
    dv = np.abs(values[a].astype(np.float64) - values[b].astype(np.float64))
    s = (dv * np.sqrt(len2)).astype(np.float32)
    s[len2 < float(min_length) * float(min_length)] = 0.0
    return s


def select_midpoints(width: int, height: int, padding: int, xy, values, edge_list,
                     min_length: float, count: int, threshold: float = 0.0) -> np.ndarray:
    """(M, 2) int32: midpoints of the `count` best edges in (score descending,
    edge index ascending) order among scores > threshold and > 0, rounded half
    to even on the padded canvas and clamped to the image; positions holding a
    seed dropped; colliding midpoints keep the better-ranked edge; output in
    raster (y, x) order (delaunay_scalar_field.cu:124-171, 230-294)."""
    xy = np.asarray(xy, dtype=np.int64).reshape(-1, 2)
    e = np.asarray(edge_list, dtype=np.int64).reshape(-1, 2)
    if e.shape[0] == 0 or count <= 0:
        return np.empty((0, 2), dtype=np.int32)
    s = edge_scores(xy, values, e, min_length)
    eligible = np.flatnonzero((s > np.float32(threshold)) & (s > 0))
    if eligible.size == 0:
        return np.empty((0, 2), dtype=np.int32)
    order = eligible[np.lexsort((eligible, -s[eligible]))][:min(int(count), eligible.size)]
    a, b = e[order, 0], e[order, 1]
    P = int(padding)
    mx = np.rint(0.5 * ((xy[a, 0] + P).astype(np.float64) + (xy[b, 0] + P))).astype(np.int64)
    my = np.rint(0.5 * ((xy[a, 1] + P).astype(np.float64) + (xy[b, 1] + P))).astype(np.int64)
    mx = np.clip(mx, P, width - 1 + P) - P
    my = np.clip(my, P, height - 1 + P) - P
    live = set(map(tuple, xy.tolist()))
    best = {}
    for r, (x, y) in enumerate(zip(mx.tolist(), my.tolist())):
        if (x, y) in live:
            continue
        pix = y * width + x
        if pix not in best:
            best[pix] = r
    pix = np.array(sorted(best), dtype=np.int64)
    if pix.size == 0:
        return np.empty((0, 2), dtype=np.int32)
    return np.stack([pix % width, pix // width], axis=1).astype(np.int32)


def split_grid(grid: np.ndarray):
    """(H, W, 3) grid -> (tids with 0 outside, seed ids, outside mask), each
    flattened to H*W (delaunay_output.cu:119-151)."""
    g = np.asarray(grid).reshape(-1, 3)
    outside = (g[:, 2] == -1)
    return (np.where(outside, 0, g[:, 2]).astype(np.int32), g[:, 0].astype(np.int32),
            outside.astype(np.uint8))
