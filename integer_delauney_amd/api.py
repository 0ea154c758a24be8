"""The batch `delauney` API on the gfx950 backend.

Same names, arguments, return layouts and errors as the reference module
`_delauney_cuda` (src/cuda/bindings.cpp:17-74, bindings/triangulate.hpp,
py_voronoi.hpp, py_grid_triangulation.hpp), computing the results of the
NumPy reference path (src/delauney/reference/) bit for bit:

* ``Voronoi().compute(width, height, seeds)`` -> (H, W, 2) int32
* ``GridTriangulation().compute(voronoi_grid, seed_positions, border_padding=-1,
  as_arrays=False)`` -> (triangle_map | (T, 3) array, (H, W, 3) int32)
  plus ``compute_timed`` and ``compute_debug``
* ``triangulate(width, height, seeds, border_padding=-1, as_arrays=False)``

``border_padding``: -1 (or None, the reference's spelling) selects
DEFAULT_BORDER_PADDING; other negative values clamp to 0
(src/delauney/reference/triangulation.py:75-78).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

DEFAULT_BORDER_PADDING = 16
AUTO_BORDER_PADDING = -1


def max_boundary_spacing(border_padding: int = DEFAULT_BORDER_PADDING) -> int:
    """Largest boundary seed spacing s with s^2/8 <= border_padding
    (src/delauney/__init__.py:18-24)."""
    if border_padding <= 0:
        return 0
    return int((8 * border_padding) ** 0.5)


def as_seed_array(seeds) -> np.ndarray:
    """(N, 2) int32 C-contiguous array from a sequence of (x, y) pairs or an
    array (src/cuda/bindings/py_delaunay.hpp:279-304 accepts both)."""
    if seeds is None:
        raise ValueError("seeds must not be empty")
    if isinstance(seeds, np.ndarray):
        arr = seeds
    else:
        seeds = list(seeds)
        if not seeds:
            raise ValueError("seeds must not be empty")
        arr = np.asarray(seeds)
    if arr.size == 0:
        raise ValueError("seeds must not be empty")
    if arr.ndim != 2 or arr.shape[1] != 2:
        raise ValueError("each seed must be an (x, y) pair")
    if arr.dtype != np.int32:
        if not np.issubdtype(arr.dtype, np.integer):
            if np.issubdtype(arr.dtype, np.floating) and np.all(np.floor(arr) == arr):
                pass
            else:
                raise ValueError("seed coordinates must be integers")
        lo, hi = np.iinfo(np.int32).min, np.iinfo(np.int32).max
        if arr.size and (arr.min() < lo or arr.max() > hi):
            raise ValueError("seed coordinate out of bounds")
        arr = arr.astype(np.int32)
    return np.ascontiguousarray(arr)


def resolve_padding(border_padding) -> int:
    if border_padding is None:
        return AUTO_BORDER_PADDING
    return int(border_padding)


def rows_to_map(rows: np.ndarray) -> dict:
    """{tid: (x, y, id_a, id_b, id_c)} with Python ints
    (src/cuda/bindings/binding_helpers.hpp:61-87), built in C by the
    _rowsmap extension (built with libdelauney_hip.so)."""
    try:
        from . import _rowsmap
    except ImportError as e:
        raise ImportError("integer_delauney_amd/_rowsmap is not built; run "
                          "`python -c 'import __graft_entry__ as g; g.build()'`") from e
    return _rowsmap.rows_to_map(np.ascontiguousarray(rows, dtype=np.int32).reshape(-1, 5))


def _pack(rows: np.ndarray, as_arrays: bool):
    if as_arrays:
        return np.ascontiguousarray(rows[:, 2:5])
    return rows_to_map(rows)


def _fetch_rows(ctx, n_tri: int) -> np.ndarray:
    rows = np.empty((int(n_tri), 5), dtype=np.int32)
    if n_tri:
        _lib.check(_lib.load().dl_copy_triangles(ctx.handle, _lib.i32p(rows)))
    return rows


def _timings(ctx) -> np.ndarray:
    out = np.zeros(9, dtype=np.float64)
    _lib.check(_lib.load().dl_get_timings(ctx.handle,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


class Voronoi:
    """Stateless L2 Voronoi diagram (src/cuda/bindings/py_voronoi.hpp)."""

    def compute(self, width: int, height: int, seeds) -> np.ndarray:
        s = as_seed_array(seeds)
        out = np.empty((int(height), int(width), 2), dtype=np.int32) if width > 0 and height > 0 \
            else np.empty((0, 0, 2), dtype=np.int32)
        ctx = _lib.context()
        _lib.check(_lib.load().dl_voronoi(ctx.handle, int(width), int(height), _lib.i32p(s),
                                          s.shape[0], _lib.i32p(out)))
        return out


class GridTriangulation:
    """Triangles and pixel assignment from a Voronoi grid
    (src/cuda/bindings/py_grid_triangulation.hpp)."""

    def _run(self, voronoi_grid, seed_positions, border_padding, timing=False):
        vg = np.asarray(voronoi_grid)
        if vg.ndim != 3 or vg.shape[2] != 2:
            raise ValueError("voronoi_grid must have shape (H, W, 2)")
        vg = np.ascontiguousarray(vg, dtype=np.int32)
        H, W = int(vg.shape[0]), int(vg.shape[1])
        s = as_seed_array(seed_positions)
        ctx = _lib.context()
        lib = _lib.load()
        tid = np.empty((H, W), dtype=np.int32)
        n = ctypes.c_int64(0)
        _lib.check(lib.dl_set_timing(ctx.handle, 1 if timing else 0))
        _lib.check(lib.dl_grid_triangulation(ctx.handle, _lib.i32p(vg), W, H, _lib.i32p(s),
                                             s.shape[0], resolve_padding(border_padding),
                                             _lib.i32p(tid), ctypes.byref(n)))
        rows = _fetch_rows(ctx, n.value)
        tgrid = np.empty((H, W, 3), dtype=np.int32)
        tgrid[:, :, 0:2] = vg
        tgrid[:, :, 2] = tid
        return ctx, rows, tgrid

    def compute(self, voronoi_grid, seed_positions, border_padding=AUTO_BORDER_PADDING,
                as_arrays: bool = False):
        _, rows, tgrid = self._run(voronoi_grid, seed_positions, border_padding)
        return _pack(rows, as_arrays), tgrid

    def compute_timed(self, voronoi_grid, seed_positions, border_padding=AUTO_BORDER_PADDING):
        ctx, rows, tgrid = self._run(voronoi_grid, seed_positions, border_padding, timing=True)
        t = _timings(ctx)
        _lib.check(_lib.load().dl_set_timing(ctx.handle, 0))
        timings = {"detect_ms": float(t[2]), "dedup_ms": float(t[3]), "assign_ms": float(t[4])}
        return rows_to_map(rows), tgrid, timings

    def compute_debug(self, voronoi_grid, seed_positions, border_padding=AUTO_BORDER_PADDING):
        ctx, rows, tgrid = self._run(voronoi_grid, seed_positions, border_padding)
        lib = _lib.load()
        w, h = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(lib.dl_copy_padded_voronoi(ctx.handle, None, ctypes.byref(w), ctypes.byref(h)))
        if w.value == tgrid.shape[1] and h.value == tgrid.shape[0]:
            # padding 0: the detection canvas is the caller's grid
            padded = np.ascontiguousarray(np.asarray(voronoi_grid, dtype=np.int32))
        else:
            padded = np.empty((h.value, w.value, 2), dtype=np.int32)
            _lib.check(lib.dl_copy_padded_voronoi(ctx.handle, _lib.i32p(padded),
                                                  ctypes.byref(w), ctypes.byref(h)))
        return rows_to_map(rows), tgrid, padded


def triangulate(width: int, height: int, seeds, border_padding=AUTO_BORDER_PADDING,
                as_arrays: bool = False):
    """Voronoi + triangles + pixel assignment in one call
    (src/cuda/bindings/triangulate.hpp:15-43); equal to
    ``GridTriangulation().compute(Voronoi().compute(width, height, seeds), seeds,
    border_padding, as_arrays)``."""
    if int(width) <= 0 or int(height) <= 0:
        raise ValueError("width and height must be positive")
    s = as_seed_array(seeds)
    ctx = _lib.context()
    out = np.empty((int(height), int(width), 3), dtype=np.int32)
    n = ctypes.c_int64(0)
    _lib.check(_lib.load().dl_triangulate(ctx.handle, int(width), int(height), _lib.i32p(s),
                                          s.shape[0], resolve_padding(border_padding),
                                          _lib.i32p(out), ctypes.byref(n)))
    rows = _fetch_rows(ctx, n.value)
    return _pack(rows, as_arrays), out


def triangulate_batch(images, border_padding=AUTO_BORDER_PADDING, as_arrays: bool = False):
    """Triangulate many independent images in one pipeline run.

    ``images``: a sequence of ``(width, height, seeds)``.  ``border_padding``:
    one value for every image or a sequence with one per image.  Returns a list
    with, per image, what ``triangulate(width, height, seeds, border_padding,
    as_arrays)`` returns for that image alone -- bit for bit.  Every stage runs
    once for the whole batch (no reference counterpart: the reference
    triangulates one image per call, src/cuda/bindings/triangulate.hpp).
    """
    images = list(images)
    if not images:
        return []
    k = len(images)
    widths = np.empty(k, dtype=np.int32)
    heights = np.empty(k, dtype=np.int32)
    counts = np.empty(k, dtype=np.int64)
    parts = []
    for i, (w, h, seeds) in enumerate(images):
        if int(w) <= 0 or int(h) <= 0:
            raise ValueError(f"image {i}: width and height must be positive")
        try:
            s = as_seed_array(seeds)
        except ValueError as e:
            raise ValueError(f"image {i}: {e}") from None
        widths[i], heights[i], counts[i] = int(w), int(h), s.shape[0]
        parts.append(s)
    if np.ndim(border_padding) == 0:
        pads = np.full(k, resolve_padding(border_padding), dtype=np.int32)
    else:
        pads = np.array([resolve_padding(p) for p in border_padding], dtype=np.int32)
        if pads.shape[0] != k:
            raise ValueError("border_padding needs one value per image")
    xy = np.ascontiguousarray(np.concatenate(parts))
    px = widths.astype(np.int64) * heights.astype(np.int64)
    out = np.empty(int(3 * px.sum()), dtype=np.int32)
    ntri = np.zeros(k, dtype=np.int64)
    ctx = _lib.context()
    lib = _lib.load()
    _lib.check(lib.dl_triangulate_batch(
        ctx.handle, k, _lib.i32p(widths), _lib.i32p(heights),
        counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _lib.i32p(xy), _lib.i32p(pads),
        _lib.i32p(out), ntri.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    rows = _fetch_rows(ctx, int(ntri.sum()))
    res = []
    r0 = p0 = 0
    for i in range(k):
        t = int(ntri[i])
        grid = out[p0:p0 + 3 * int(px[i])].reshape(int(heights[i]), int(widths[i]), 3)
        res.append((_pack(rows[r0:r0 + t], as_arrays), grid))
        r0 += t
        p0 += 3 * int(px[i])
    return res
