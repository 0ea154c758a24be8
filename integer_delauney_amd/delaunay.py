"""Incremental `Delaunay` session (insert / remove / finalise) on the gfx950 backend.

Mirrors the reference class `_delauney_cuda.Delaunay` (src/cuda/bindings.cpp:86-227,
bindings/py_delaunay.hpp, delaunay.cuh:23-378) with one change of contract:
the reference's incremental path breaks Voronoi ties and cocircular quads by
*insertion* order and so drifts from the batch result
(tests/test_incremental_deferred.py:333-355).  Here the state after any
sequence of inserts and removals is, bit for bit, the batch pipeline
(`triangulate`) on the current seed set; ``remove`` is new (the reference has
no removal).  The live seeds, the last grid and the last triangle rows stay on
the device between calls (``dl_session_*`` in include/delauney_hip.h); the
host keeps copies of the live seed list and the per-seed values for queries.

Seed numbering follows the reference:
* ``insert`` / ``finalise`` report sorted (x asc, y asc) ids;
* ``get_triangles`` and ``get_seeds`` use insertion order (of the live seeds),
  and ``sorted_rank`` maps insertion order to sorted ids
  (API_REFERENCE.md:315-340).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .api import DEFAULT_BORDER_PADDING, _pack, as_seed_array


def _keys(xy: np.ndarray) -> np.ndarray:
    return (xy[:, 0].astype(np.int64) << 32) | (xy[:, 1].astype(np.int64) & 0xFFFFFFFF)


class Delaunay:
    def __init__(self, width: int, height: int, max_seeds: int, border_padding: int = -1):
        if width <= 0 or height <= 0 or max_seeds <= 0:
            raise ValueError("dimensions and max_seeds must be positive")
        self._W = int(width)
        self._H = int(height)
        self._max = int(max_seeds)
        # delaunay.cu:72: any negative value selects the default
        self._P = int(border_padding) if border_padding is not None and border_padding >= 0 \
            else DEFAULT_BORDER_PADDING
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.dl_session_create(_lib.context().handle, self._W, self._H, self._max,
                                               self._P, ctypes.byref(h)))
        self._h = h
        self._xy = np.empty((0, 2), dtype=np.int32)
        self._values = np.empty(0, dtype=np.float32)
        self._pending = True
        self._rows = None
        self._grid = None
        self._rank = None
        self._last = {"bfs_ms": 0.0, "bfs_iters": 0, "detect_ms": 0.0, "dedup_ms": 0.0,
                      "assign_ms": 0.0}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dl_session_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    # -- properties (bindings.cpp:220-227) --------------------------------
    @property
    def width(self) -> int:
        return self._W

    @property
    def height(self) -> int:
        return self._H

    @property
    def border_padding(self) -> int:
        return self._P

    @property
    def seed_count(self) -> int:
        return int(self._xy.shape[0])

    @property
    def has_pending(self) -> bool:
        return bool(self._pending) and self.seed_count > 0

    # -- mutation ----------------------------------------------------------
    def insert_deferred(self, seeds, values=None) -> None:
        """Append a batch (delaunay.cu:274-321); the raster is deferred.
        The batch is validated on the device (bounds, already live, repeated
        within the batch, capacity); on error nothing changes."""
        if seeds is None or (not isinstance(seeds, np.ndarray) and len(seeds) == 0) or \
                (isinstance(seeds, np.ndarray) and seeds.size == 0):
            return
        xy = as_seed_array(seeds)
        if values is not None:
            vals = np.asarray(values, dtype=np.float32).reshape(-1)
            if vals.shape[0] != xy.shape[0]:
                raise ValueError("values must have one entry per seed in the batch")
        else:
            vals = np.zeros(xy.shape[0], dtype=np.float32)
        _lib.check(self._lib.dl_session_insert(_lib.context().handle, self._h, _lib.i32p(xy),
                                               xy.shape[0]))
        self._xy = np.concatenate([self._xy, xy])
        self._values = np.concatenate([self._values, vals])
        self._pending = True

    def remove(self, seeds) -> None:
        """Remove live seeds by position.  ValueError (nothing removed) for a
        position that is not live or repeats.  No reference counterpart."""
        xy = as_seed_array(seeds)
        _lib.check(self._lib.dl_session_remove(_lib.context().handle, self._h, _lib.i32p(xy),
                                               xy.shape[0]))
        drop = np.isin(_keys(self._xy), _keys(xy))
        self._xy = self._xy[~drop]
        self._values = self._values[~drop]
        self._pending = True

    def _recompute(self):
        t0 = time.perf_counter()
        n = ctypes.c_int64(0)
        _lib.check(self._lib.dl_session_finalise(_lib.context().handle, self._h, ctypes.byref(n)))
        grid = np.empty((self._H, self._W, 3), dtype=np.int32)
        rows = np.empty((int(n.value), 5), dtype=np.int32)
        _lib.check(self._lib.dl_session_copy_result(self._h, _lib.i32p(grid),
                                                    _lib.i32p(rows) if n.value else None))
        self._last["bfs_ms"] = (time.perf_counter() - t0) * 1e3
        self._rows = rows
        self._grid = grid
        order = np.lexsort((self._xy[:, 1], self._xy[:, 0]))
        rank = np.empty(self.seed_count, dtype=np.int32)
        rank[order] = np.arange(self.seed_count, dtype=np.int32)
        self._rank = rank
        self._pending = False

    def finalise(self, as_arrays: bool = False):
        """Materialise (triangle_map, (H, W, 3) grid) in sorted ids (delaunay.cu:323-345)."""
        if self._pending or self._rows is None:
            self._recompute()
        return _pack(self._rows, as_arrays), self._grid.copy()

    def insert(self, seeds, as_arrays: bool = False):
        self.insert_deferred(seeds)
        return self.finalise(as_arrays)

    def _timings(self):
        return dict(self._last)

    def insert_timed(self, seeds, as_arrays: bool = False):
        m, g = self.insert(seeds, as_arrays)
        return m, g, self._timings()

    def insert_deferred_timed(self, seeds):
        self.insert_deferred(seeds)
        return {"bfs_ms": 0.0, "bfs_iters": 0, "detect_ms": 0.0, "dedup_ms": 0.0, "assign_ms": 0.0}

    def finalise_timed(self, as_arrays: bool = False):
        m, g = self.finalise(as_arrays)
        return m, g, self._timings()

    # -- queries -----------------------------------------------------------
    def _ensure(self):
        if self._pending or self._rows is None:
            self._recompute()

    @property
    def sorted_rank(self) -> np.ndarray:
        """insertion-order id -> sorted id (delaunay_output.cu:39-57)."""
        self._ensure()
        return self._rank.copy()

    def get_triangles(self) -> np.ndarray:
        """(T, 3) vertex ids in insertion order (delaunay_output.cu:195-207)."""
        self._ensure()
        inv = np.empty_like(self._rank)
        inv[self._rank] = np.arange(self._rank.shape[0], dtype=np.int32)
        return inv[self._rows[:, 2:5]].astype(np.int32) if self._rows.shape[0] else \
            np.empty((0, 3), dtype=np.int32)

    def get_voronoi_grid(self) -> np.ndarray:
        """(H, W, 2): sorted seed id and squared distance (delaunay_output.cu:169-189)."""
        self._ensure()
        return np.ascontiguousarray(self._grid[:, :, 0:2])

    def get_seeds(self) -> np.ndarray:
        return self._xy.copy()

    def get_values(self) -> np.ndarray:
        return self._values.copy()

    def device_seeds(self) -> np.ndarray:
        """The live seed list as held on the device (insertion order)."""
        out = np.empty((self.seed_count, 2), dtype=np.int32)
        if self.seed_count:
            _lib.check(self._lib.dl_session_copy_seeds(self._h, _lib.i32p(out)))
        return out
