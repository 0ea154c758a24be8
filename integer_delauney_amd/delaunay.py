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
from .device_views import DeviceArrayView


def _stage_timings(ctx, wall_ms: float) -> dict:
    """The *_timed dictionary of the context's last timed call (all zero
    without one)."""
    d = {"bfs_ms": 0.0, "bfs_iters": 0, "detect_ms": 0.0, "dedup_ms": 0.0, "assign_ms": 0.0,
         "seeds_ms": 0.0, "output_ms": 0.0, "total_ms": wall_ms, "fill_launches": 0}
    if ctx is None:
        return d
    lib = _lib.load()
    t = np.zeros(9, dtype=np.float64)
    _lib.check(lib.dl_get_timings(ctx.handle, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    issued, conv, visits = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    k = ctypes.c_int(0)
    _lib.check(lib.dl_get_fill_stats(ctx.handle, ctypes.byref(issued), ctypes.byref(conv),
                                     ctypes.byref(visits), ctypes.byref(k), None))
    d.update(bfs_ms=float(t[0] + t[1]), bfs_iters=int(conv.value) * int(k.value),
             detect_ms=float(t[2]), dedup_ms=float(t[3]), assign_ms=float(t[4]),
             seeds_ms=float(t[0]), output_ms=float(t[5]), fill_launches=int(issued.value))
    return d


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
        ctx = _lib.context()
        self._device = ctx.device
        h = ctypes.c_void_p()
        _lib.check(self._lib.dl_session_create(ctx.handle, self._W, self._H, self._max,
                                               self._P, ctypes.byref(h)))
        self._h = h
        self._xy = np.empty((0, 2), dtype=np.int32)
        self._pending = True
        self._rows = None      # rows of the current result (host copy)
        self._grid = None      # its (H, W, 3) grid, copied on first use
        self._rank = None
        # Bumped by every call that changes what finalise_device's arrays show
        # (delaunay.cu:287-289, 337-339, 349-351).
        self._generation = 0

    def close(self):
        if getattr(self, "_h", None):
            self._generation += 1
            self._lib.dl_session_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def _handle(self):
        if not self._h:
            raise RuntimeError("the session is closed")
        return self._h

    # -- fill records (no reference counterpart) ---------------------------
    def set_fill_record(self, enabled: bool = True):
        """Keep each finalise's flood-fill trajectory on the device, so that
        the next finalise recomputes only the tiles the inserted and removed
        seeds reach (the result is the same, bit for bit; see
        dl_session_set_fill_record).  Costs device memory; False frees it."""
        _lib.check(self._lib.dl_session_set_fill_record(self._handle(), 1 if enabled else 0))

    def fill_record_stats(self):
        """(record held, its entries, tile computations of the last finalise's
        flood fill)."""
        v = ctypes.c_int(0)
        e = ctypes.c_int64(0)
        t = ctypes.c_int64(0)
        _lib.check(self._lib.dl_session_fill_record_stats(self._handle(), ctypes.byref(v),
                                                          ctypes.byref(e), ctypes.byref(t)))
        return bool(v.value), int(e.value), int(t.value)

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
        ``values``: one float per seed of the batch, the scalar field that
        edge_scores / select_midpoints read (ValueError on a length mismatch).
        The batch is validated on the device (bounds, already live, repeated
        within the batch, capacity); on error nothing changes."""
        if seeds is None or (not isinstance(seeds, np.ndarray) and len(seeds) == 0) or \
                (isinstance(seeds, np.ndarray) and seeds.size == 0):
            if values is not None and np.asarray(values).size != 0:
                raise ValueError("values must have one entry per seed in the batch")
            return
        xy = as_seed_array(seeds)
        vp = None
        if values is not None:
            vals = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(-1))
            if vals.shape[0] != xy.shape[0]:
                raise ValueError("values must have one entry per seed in the batch")
            vp = _lib.f32p(vals)
        _lib.check(self._lib.dl_session_insert(_lib.context().handle, self._handle(), _lib.i32p(xy),
                                               vp, xy.shape[0]))
        self._xy = np.concatenate([self._xy, xy])
        self._mutated()

    def remove(self, seeds) -> None:
        """Remove live seeds by position.  ValueError (nothing removed) for a
        position that is not live or repeats.  No reference counterpart."""
        xy = as_seed_array(seeds)
        _lib.check(self._lib.dl_session_remove(_lib.context().handle, self._handle(), _lib.i32p(xy),
                                               xy.shape[0]))
        drop = np.isin(_keys(self._xy), _keys(xy))
        self._xy = self._xy[~drop]
        self._mutated()

    def _mutated(self):
        self._pending = True
        self._generation += 1

    def _recompute(self, timed: bool = False):
        """Runs the pipeline on the live seeds; keeps the rows, leaves the grid
        on the device until asked for.  With ``timed`` the stages are timed
        with HIP events and the timings dictionary is returned."""
        ctx = _lib.context()
        n = ctypes.c_int64(0)
        t0 = time.perf_counter()
        if timed:
            _lib.check(self._lib.dl_set_timing(ctx.handle, 1))
        try:
            _lib.check(self._lib.dl_session_finalise(ctx.handle, self._handle(), ctypes.byref(n)))
        finally:
            if timed:
                _lib.check(self._lib.dl_set_timing(ctx.handle, 0))
        wall_ms = (time.perf_counter() - t0) * 1e3
        timings = _stage_timings(ctx, wall_ms) if timed else None
        rows = np.empty((int(n.value), 5), dtype=np.int32)
        if n.value:
            _lib.check(self._lib.dl_session_copy_result(self._handle(), None, _lib.i32p(rows)))
        self._rows = rows
        self._grid = None
        order = np.lexsort((self._xy[:, 1], self._xy[:, 0]))
        rank = np.empty(self.seed_count, dtype=np.int32)
        rank[order] = np.arange(self.seed_count, dtype=np.int32)
        self._rank = rank
        self._pending = False
        return timings

    def _ensure(self):
        if self._pending or self._rows is None:
            self._recompute()

    def _host_grid(self) -> np.ndarray:
        self._ensure()
        if self._grid is None:
            grid = np.empty((self._H, self._W, 3), dtype=np.int32)
            _lib.check(self._lib.dl_session_copy_result(self._handle(), _lib.i32p(grid), None))
            self._grid = grid
        return self._grid

    def _finalise(self, as_arrays: bool, timed: bool):
        timings = None
        if self._pending or self._rows is None:
            timings = self._recompute(timed)
            self._generation += 1
        if timed and timings is None:   # nothing pending: no stage ran
            timings = _stage_timings(None, 0.0)
        return _pack(self._rows, as_arrays), self._host_grid().copy(), timings

    def finalise(self, as_arrays: bool = False):
        """Materialise (triangle_map, (H, W, 3) grid) in sorted ids (delaunay.cu:323-345)."""
        m, g, _ = self._finalise(as_arrays, False)
        return m, g

    def finalise_device(self):
        """As finalise(as_arrays=True), with the per-pixel result left on the
        device (delaunay.cu:347-360, py_delaunay.hpp:65-88): returns
        (triangle_verts (T, 3), pixel_tids, pixel_seed_ids, outside_hull_mask),
        the last three as :class:`DeviceArrayView` of H*W elements (tid 0 where
        the mask is 1).  The views are valid until the next insert,
        insert_deferred, remove, finalise or finalise_device."""
        self._generation += 1
        self._ensure()
        n = ctypes.c_int64(0)
        p_t, p_s, p_m = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(self._lib.dl_session_finalise_device(
            _lib.context().handle, self._handle(), ctypes.byref(n), ctypes.byref(p_t),
            ctypes.byref(p_s), ctypes.byref(p_m)))
        count = self._W * self._H
        g = self._generation
        views = tuple(DeviceArrayView(self, p.value, count, ts, g, self._device)
                      for p, ts in ((p_t, "<i4"), (p_s, "<i4"), (p_m, "|u1")))
        return (_pack(self._rows, True),) + views

    def insert(self, seeds, as_arrays: bool = False):
        self.insert_deferred(seeds)
        return self.finalise(as_arrays)

    # The timings dictionaries of py_delaunay.hpp:268-277 (delaunay.cuh:8-14):
    # bfs_ms (seed ordering + flood fill), bfs_iters, detect_ms, dedup_ms
    # (registration and tid ranking) and assign_ms, from HIP events on the
    # context's stream, plus seeds_ms, output_ms, total_ms and fill_launches.
    def insert_timed(self, seeds, as_arrays: bool = False):
        self.insert_deferred(seeds)
        return self._finalise(as_arrays, True)

    def insert_deferred_timed(self, seeds):
        """insert_deferred with its timings.  Here a deferred insert only
        updates the live seed set on the device (validation, presence bits,
        list); the flood fill, detection and assignment all run at finalise,
        so those keys are 0 and ``update_ms`` holds the update's own time."""
        t0 = time.perf_counter()
        self.insert_deferred(seeds)
        t = _stage_timings(None, 0.0)
        t["update_ms"] = t["total_ms"] = (time.perf_counter() - t0) * 1e3
        return t

    def finalise_timed(self, as_arrays: bool = False):
        return self._finalise(as_arrays, True)

    # -- topology and seeds --------------------------------------------------
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
        return np.ascontiguousarray(self._host_grid()[:, :, 0:2])

    def get_seeds(self) -> np.ndarray:
        """(N, 2) live seeds in insertion order (delaunay_output.cu:209-219)."""
        return self._xy.copy()

    def get_values(self) -> np.ndarray:
        """(N,) float32 scalar field in insertion order; empty when no insert
        carried values (delaunay_output.cu:221-224).  Seeds inserted without
        values hold 0."""
        flag = ctypes.c_int(0)
        _lib.check(self._lib.dl_session_has_values(self._handle(), ctypes.byref(flag)))
        if not flag.value:
            return np.empty(0, dtype=np.float32)
        out = np.empty(self.seed_count, dtype=np.float32)
        if out.size:
            _lib.check(self._lib.dl_session_copy_values(self._handle(), _lib.f32p(out)))
        return out

    def device_seeds(self) -> np.ndarray:
        """The live seed list as held on the device (insertion order)."""
        n = ctypes.c_int64(0)
        _lib.check(self._lib.dl_session_seed_count(self._handle(), ctypes.byref(n)))
        out = np.empty((int(n.value), 2), dtype=np.int32)
        if out.shape[0]:
            _lib.check(self._lib.dl_session_copy_seeds(self._handle(), _lib.i32p(out)))
        return out

    # -- point queries (delaunay_query.cu) -----------------------------------
    @staticmethod
    def _points(points) -> np.ndarray:
        a = np.asarray(points)
        if a.ndim != 2 or a.shape[1] != 2:
            if a.size == 0:
                return np.empty((0, 2), dtype=np.int32)
            raise ValueError("points must have shape (N, 2)")
        return np.ascontiguousarray(a, dtype=np.int32)

    def locate(self, points) -> np.ndarray:
        """Triangle containing each (x, y) point (image coordinates) as an (N,)
        int32 array, -1 where none does; ids index get_triangles() and
        finalise()'s map (bindings.cpp:196-205)."""
        q = self._points(points)
        out = np.empty(q.shape[0], dtype=np.int32)
        if q.shape[0]:
            _lib.check(self._lib.dl_session_locate(_lib.context().handle, self._handle(),
                                                   _lib.i32p(q), q.shape[0], _lib.i32p(out)))
        return out

    def in_circumsphere(self, points, t):
        """(mask, triangle_ids): mask[i] iff dx^2 + dy^2 + t[i]^2 < R^2 about the
        circumcircle of the triangle containing point i (bindings.cpp:184-195)."""
        q = self._points(points)
        tt = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(-1))
        if tt.shape[0] != q.shape[0]:
            raise ValueError("t must have one entry per point")
        n = q.shape[0]
        mask = np.zeros(n, dtype=np.uint8)
        tids = np.full(n, -1, dtype=np.int32)
        if n:
            _lib.check(self._lib.dl_session_in_circumsphere(
                _lib.context().handle, self._handle(), _lib.i32p(q), _lib.f64p(tt), n,
                _lib.u8p(mask), _lib.i32p(tids)))
        return mask.astype(bool), tids

    # -- edges and the scalar field (delaunay_scalar_field.cu) ----------------
    def _n_edges(self) -> int:
        n = ctypes.c_int64(0)
        _lib.check(self._lib.dl_session_edges(_lib.context().handle, self._handle(),
                                              ctypes.byref(n)))
        return int(n.value)

    def get_edges(self) -> np.ndarray:
        """(E, 2) distinct undirected edges (a < b, insertion-order ids),
        ascending (bindings.cpp:210-219)."""
        e = self._n_edges()
        out = np.empty((e, 2), dtype=np.int32)
        if e:
            _lib.check(self._lib.dl_session_copy_edges(self._handle(), _lib.i32p(out)))
        return out

    def edge_scores(self, min_length: float = 0.0) -> np.ndarray:
        """(E,) float32 |v_a - v_b| * |ab| in get_edges() order, 0 for edges
        shorter than min_length (bindings.cpp:166-172)."""
        e = self._n_edges()
        out = np.empty(e, dtype=np.float32)
        if e:
            _lib.check(self._lib.dl_session_edge_scores(_lib.context().handle, self._handle(),
                                                        float(min_length), _lib.f32p(out)))
        return out

    def select_midpoints(self, min_length: float, count: int, threshold: float = 0.0) -> np.ndarray:
        """(M, 2) int32 midpoints of the best `count` edges (score descending,
        edge index ascending), without positions that hold a seed and one per
        pixel, in raster order (bindings.cpp:173-183)."""
        e = self._n_edges()
        count = int(count)
        if e == 0 or count <= 0:
            return np.empty((0, 2), dtype=np.int32)
        cap = min(count, e)
        buf = np.empty((cap, 2), dtype=np.int32)
        n = ctypes.c_int64(0)
        _lib.check(self._lib.dl_session_select_midpoints(
            _lib.context().handle, self._handle(), float(min_length), count, float(threshold),
            _lib.i32p(buf), cap, ctypes.byref(n)))
        return buf[:int(n.value)].copy()
