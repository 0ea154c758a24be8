"""Serving loop: continuous batching of triangulation requests and session
updates on one GPU.

Requests arrive at any time (``submit`` returns a future).  Each of ``slots``
workers owns a library context (its own HIP stream and workspaces) and loops:
take the waiting requests that fit the HBM budget, run them as one pipeline
pass (``dl_run_batch``: every stage once for the whole batch), complete their
futures, repeat.  Requests that arrive while a batch is on the GPU are admitted
by the next free worker as soon as memory allows, so batches of different
workers overlap on the device (the library releases the GIL).

* ``TriangulateRequest(width, height, seeds, border_padding)`` -- result
  ``(rows, grid)``: (T, 5) int32 rows ``(x, y, a, b, c)`` in tid order and the
  (H, W, 3) int32 grid, exactly ``triangulate(..., as_arrays)``'s data.  Grids
  are returned in page-locked host memory from a pool (a grid's block goes
  back to the pool when the array is gone).
* ``UpdateRequest(session_id, insert, remove)`` -- applies the removals, then
  the insertions, to a resident session (its seeds and last result stay on the
  device) and returns the session's new ``(rows, grid)``.  Requests on one
  session run in arrival order and never in two batches at once.

Failures stay with their request:

* A triangulation request is validated on the host when it is submitted
  (sizes, seed bounds, duplicates -- the checks ``triangulate`` makes); an
  invalid one fails its own future with the ``ValueError`` (or
  ``RuntimeError`` for an oversized canvas) that ``triangulate`` raises and
  never joins a batch.
* An invalid update fails its own future and leaves the session unchanged.
* If a batch fails on the device anyway, its images run again one at a time,
  so only the failing one fails.  An update that was applied is never
  reported as failed for a reason that is not its own: if its session's
  recompute fails, the future gets :class:`UpdateAppliedError`, which says that
  the seeds changed.

Admission by HBM budget: a request's device cost is :func:`footprint`, an
upper bound (tested against measured workspace growth) on what a worker's
context reserves for it.  A worker keeps at most ``idle_bytes_per_slot`` of
workspace between batches (it trims its context after a larger batch), so
the server's device memory stays below resident session bytes (measured) +
``slots * idle_bytes_per_slot`` + the costs of the requests in flight, and a
worker admits waiting requests in arrival order (those of a busy session are
skipped) only while that total stays within ``hbm_budget``, its batch within
``max_batch_bytes``.  A request that could never be admitted -- its cost
exceeds ``hbm_budget`` minus the resident sessions and the slots' idle
reserve -- fails with :class:`AdmissionError` instead of waiting or running
out of device memory; so does a session the budget cannot hold.  Each result
is the same as the request's result computed alone: batching never changes a
result (tests/test_gpu_serving.py).

No reference counterpart: the reference library has no serving layer.
"""
from __future__ import annotations

import ctypes
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .api import AUTO_BORDER_PADDING, DEFAULT_BORDER_PADDING, as_seed_array

MAX_CANVAS = 32000   # DL_MAX_CANVAS (include/delauney_hip.h)


class AdmissionError(RuntimeError):
    """A request (or session) that the server's HBM budget can never hold."""


class UpdateAppliedError(RuntimeError):
    """An update was applied to its session (the live seeds changed) but the
    session's recompute failed; the session recomputes at its next update."""


def footprint(width: int, height: int, n_seeds: int, padding: int = DEFAULT_BORDER_PADDING) -> int:
    """Upper bound on the device bytes a context's workspaces grow by for one
    image inside a batch: per padded canvas cell 8 B of label planes, 1 B of
    registration bitmap and scan, 1/8 B of detection counts; per image pixel
    12 B of staged output and 8 B of label planes for the unpadded fill (run
    only after a padding crossing), plus the seed index; per seed about
    720 B of candidates, registration buckets, hash table slots and triplets
    (the exact rerun's), triangle rows and bin entries (two triangles per seed); and a constant for the speculative
    capacities' fixed terms and allocation rounding."""
    wc, hc = width + 2 * padding, height + 2 * padding
    return int(10 * wc * hc + 21 * width * height + 780 * n_seeds + (4 << 20))


def _resolve_padding(p) -> int:
    if p is None or p == -1:
        return DEFAULT_BORDER_PADDING
    return max(0, int(p))


def validate_request(req) -> None:
    """The checks triangulate makes (api.py, dl_api.hip check_dims and seed
    validation), on the host: ValueError for bad sizes or seeds, RuntimeError
    for a canvas beyond the library's limit."""
    if int(req.width) <= 0 or int(req.height) <= 0:
        raise ValueError("width and height must be positive")
    P = _resolve_padding(req.border_padding)
    if req.width + 2 * P > MAX_CANVAS or req.height + 2 * P > MAX_CANVAS:
        raise RuntimeError(f"canvas (image plus twice the border padding) exceeds {MAX_CANVAS} "
                           "cells per side")
    xy = req.seeds
    if xy[:, 0].min() < 0 or xy[:, 0].max() >= req.width:
        raise ValueError("seed x coordinate out of bounds")
    if xy[:, 1].min() < 0 or xy[:, 1].max() >= req.height:
        raise ValueError("seed y coordinate out of bounds")
    keys = xy[:, 0].astype(np.int64) * int(req.height) + xy[:, 1]
    if np.unique(keys).shape[0] != keys.shape[0]:
        raise ValueError("duplicate seed positions are not allowed")


@dataclass
class TriangulateRequest:
    width: int
    height: int
    seeds: np.ndarray
    border_padding: int = AUTO_BORDER_PADDING


@dataclass
class UpdateRequest:
    session_id: int
    insert: np.ndarray = field(default_factory=lambda: np.empty((0, 2), np.int32))
    remove: np.ndarray = field(default_factory=lambda: np.empty((0, 2), np.int32))


@dataclass
class BatchRecord:
    """One pipeline pass of a worker: its requests, admitted cost, device time
    of the pipeline stages and of the copy of results to the host (HIP events
    on the worker's stream), and host wall time."""
    requests: int
    cost: int
    gpu_ms: float
    copy_ms: float
    wall_ms: float


class _Session:
    def __init__(self, handle, width, height, padding, max_seeds):
        self.handle = handle
        self.width = width
        self.height = height
        self.padding = padding
        self.max_seeds = max_seeds
        self.n_live = 0
        self.busy = False
        self.resident = 0


def _session_bytes(handle) -> int:
    b = ctypes.c_size_t(0)
    _lib.check(_lib.load().dl_session_device_bytes(handle, ctypes.byref(b)))
    return int(b.value)


def session_estimate(width: int, height: int, max_seeds: int) -> int:
    """Device bytes a session holds once finalised: presence bitmap, seed
    lists and values with their spares and undo copies (40 B per seed of
    capacity), the (H, W, 3) result and per-seed rows and labels."""
    return int(12 * width * height + width * height // 8 + 40 * max_seeds + 48 * max_seeds)


class Server:
    def __init__(self, hbm_budget: int = 64 << 30, slots: int = 3, max_batch: int = 256,
                 max_batch_bytes: int = 2 << 30, idle_bytes_per_slot: int = 4 << 30,
                 device: int = 0):
        self._lib = _lib.load()
        self._budget = int(hbm_budget)
        self._max_batch = int(max_batch)
        # Batches stay small enough that one slot's results travel back to
        # the host while another slot's batch computes.
        self._max_batch_bytes = int(max_batch_bytes)
        self._idle_cap = int(idle_bytes_per_slot)
        self._reserve = int(slots) * self._idle_cap
        self._ctxs = [_lib.Context(device) for _ in range(slots)]
        for c in self._ctxs:
            # Mixed request sizes: a steady latency matters more than the
            # first-use tuning of tile shapes or the launch-count prediction
            # (both would occasionally spend a request's time on extra work).
            _lib.check(self._lib.dl_set_autotune(c.handle, 0))
            _lib.check(self._lib.dl_set_fill_prediction(c.handle, 0))
            # stage events: the batch records split device time from copy-back
            _lib.check(self._lib.dl_set_timing(c.handle, 1))
        # results land in page-locked memory: the copy back runs at link speed
        self._pool = _lib.PinnedPool()
        self._cond = threading.Condition()
        self._queue: deque = deque()
        self._sessions: dict[int, _Session] = {}
        self._next_sid = 0
        self._inflight = 0
        self._resident = 0
        self._stop = False
        self.batches: list[BatchRecord] = []
        self.rejected = 0
        self.peak_inflight = 0   # largest admitted in-flight cost
        self._workers = [threading.Thread(target=self._work, args=(i,), daemon=True)
                         for i in range(slots)]
        for w in self._workers:
            w.start()

    # -- budget --------------------------------------------------------------
    def _limit(self) -> int:
        """Bytes requests in flight may take (called with the lock held)."""
        return self._budget - self._resident - self._reserve

    @property
    def resident_bytes(self) -> int:
        return self._resident

    def set_budget(self, hbm_budget: int) -> None:
        """Change the HBM budget; waiting requests it can no longer admit
        fail with AdmissionError when a worker next takes a batch."""
        with self._cond:
            self._budget = int(hbm_budget)
            self._cond.notify_all()

    # -- sessions ------------------------------------------------------------
    def create_session(self, width: int, height: int, max_seeds: int,
                       border_padding: int = AUTO_BORDER_PADDING, seeds=None) -> int:
        """A resident session (device memory held until close); its seeds, if
        given, are inserted and finalised at once.  AdmissionError when the
        budget cannot hold it next to the resident sessions and the slots'
        reserve."""
        pad = DEFAULT_BORDER_PADDING if border_padding is None or border_padding < 0 \
            else int(border_padding)
        est = session_estimate(int(width), int(height), int(max_seeds))
        with self._cond:
            if self._resident + est + self._reserve > self._budget:
                raise AdmissionError(f"a {width}x{height} session of {max_seeds} seeds needs about "
                                     f"{est} device bytes; the HBM budget holds "
                                     f"{self._budget - self._resident - self._reserve} more")
        h = ctypes.c_void_p()
        ctx = self._ctxs[0]
        _lib.check(self._lib.dl_session_create(ctx.handle, int(width), int(height), int(max_seeds),
                                               pad, ctypes.byref(h)))
        s = _Session(h, int(width), int(height), pad, int(max_seeds))
        try:
            if seeds is not None:
                xy = as_seed_array(seeds)
                _lib.check(self._lib.dl_session_insert(ctx.handle, h, _lib.i32p(xy), None,
                                                       xy.shape[0]))
                s.n_live = xy.shape[0]
                n = ctypes.c_int64()
                _lib.check(self._lib.dl_session_finalise(ctx.handle, h, ctypes.byref(n)))
        except Exception:
            self._lib.dl_session_destroy(h)
            raise
        s.resident = max(_session_bytes(h), est)
        with self._cond:
            sid = self._next_sid
            self._next_sid += 1
            self._sessions[sid] = s
            self._resident += s.resident
        return sid

    # -- requests --------------------------------------------------------------
    def submit(self, req, keep: bool = True) -> Future:
        """Queue a request.  keep=False completes the future with None once the
        result has been produced and copied back, releasing it at once (load
        generators that only time requests).  An invalid or inadmissible
        request gets a failed future at once."""
        fut: Future = Future()
        fut.t_submit = time.perf_counter()
        fut.keep = keep
        try:
            if isinstance(req, TriangulateRequest):
                req.seeds = as_seed_array(req.seeds)
                validate_request(req)
                cost = footprint(req.width, req.height, req.seeds.shape[0],
                                 _resolve_padding(req.border_padding))
            elif isinstance(req, UpdateRequest):
                if req.session_id not in self._sessions:
                    raise ValueError("unknown session")
                req.insert = np.ascontiguousarray(
                    np.asarray(req.insert, dtype=np.int32).reshape(-1, 2))
                req.remove = np.ascontiguousarray(
                    np.asarray(req.remove, dtype=np.int32).reshape(-1, 2))
                s = self._sessions[req.session_id]
                cost = footprint(s.width, s.height, min(s.max_seeds, s.n_live + req.insert.shape[0]),
                                 s.padding)
            else:
                raise TypeError("unknown request type")
        except (ValueError, RuntimeError) as e:
            fut.set_exception(e)
            return fut
        with self._cond:
            if self._stop:
                raise RuntimeError("server is closed")
            if cost > self._limit():
                self.rejected += 1
                fut.set_exception(self._inadmissible(cost))
                return fut
            self._queue.append((req, fut, cost))
            self._cond.notify()
        return fut

    def _inadmissible(self, cost: int) -> AdmissionError:
        return AdmissionError(f"request needs {cost} device bytes; the HBM budget of "
                              f"{self._budget} bytes leaves {self._limit()} after "
                              f"{self._resident} resident session bytes and "
                              f"{self._reserve} reserved for idle workers")

    def _take(self):
        """Admit waiting requests for one batch (called with the lock held)."""
        batch, total = [], 0
        blocked = set()
        limit = self._limit()
        rest = deque()
        while self._queue:
            item = self._queue.popleft()
            req, fut, cost = item
            if cost > limit:   # the budget shrank (new sessions) below what it needs
                self.rejected += 1
                fut.set_exception(self._inadmissible(cost))
                continue
            if isinstance(req, UpdateRequest):
                s = self._sessions[req.session_id]
                if s.busy or req.session_id in blocked:
                    blocked.add(req.session_id)
                    rest.append(item)
                    continue
            over_budget = self._inflight + total + cost > limit
            over_batch = bool(batch) and total + cost > self._max_batch_bytes
            if len(batch) >= self._max_batch or over_batch or over_budget:
                rest.append(item)
                rest.extend(self._queue)
                self._queue.clear()
                break
            if isinstance(req, UpdateRequest):
                self._sessions[req.session_id].busy = True
                blocked.add(req.session_id)
            batch.append(item)
            total += cost
        self._queue.extendleft(reversed(rest))
        self._inflight += total
        self.peak_inflight = max(self.peak_inflight, self._inflight)
        return batch, total

    def _work(self, slot: int):
        ctx = self._ctxs[slot]
        while True:
            with self._cond:
                while not self._stop:
                    batch, total = self._take()
                    if batch:
                        break
                    self._cond.wait()
                if self._stop:
                    return
            rec = None
            try:
                try:
                    rec = self._run(ctx, batch, total)
                except Exception as e:   # never leave a future pending
                    for _, fut, _ in batch:
                        if not fut.done():
                            fut.set_exception(e)
                held = ctypes.c_size_t(0)
                _lib.check(self._lib.dl_context_device_bytes(ctx.handle, ctypes.byref(held)))
                if held.value > self._idle_cap:
                    _lib.check(self._lib.dl_context_trim(ctx.handle))
            finally:
                with self._cond:
                    self._inflight -= total
                    for req, _, _ in batch:
                        if isinstance(req, UpdateRequest):
                            s = self._sessions[req.session_id]
                            s.busy = False
                            now = max(_session_bytes(s.handle), s.resident)
                            self._resident += now - s.resident
                            s.resident = now
                    if rec is not None:
                        self.batches.append(rec)
                    self._cond.notify_all()

    def _images(self, items):
        """DLImage records for (req, fut, session, grid) items."""
        images = []
        for req, _, s, grid in items:
            im = _lib.DLImage()
            if s is not None:
                im.session = s.handle
            else:
                im.width, im.height = req.width, req.height
                im.border_padding = -1 if req.border_padding is None else int(req.border_padding)
                im.host = 1
                im.n_seeds = req.seeds.shape[0]
                im.seeds_xy = req.seeds.ctypes.data
                im.out = grid.ctypes.data
            images.append(im)
        return (_lib.DLImage * len(images))(*images)

    def _run_pass(self, ctx, items):
        """One dl_run_batch over items; returns the image records and the
        batch's rows (fresh images' slices are theirs)."""
        arr = self._images(items)
        _lib.check(self._lib.dl_run_batch(ctx.handle, len(items), ctypes.cast(arr, ctypes.c_void_p),
                                          None))
        total_rows = max(int(arr[i].tri_offset + arr[i].n_tri) for i in range(len(items)))
        rows_all = np.empty((total_rows, 5), dtype=np.int32)
        if total_rows:
            _lib.check(self._lib.dl_copy_triangles(ctx.handle, _lib.i32p(rows_all)))
        return arr, rows_all

    def _deliver(self, items, arr, rows_all):
        """Results of a pass to the futures; returns the host time spent
        copying session results back (ms)."""
        lib = self._lib
        t_copy = 0.0
        for i, (req, fut, s, grid) in enumerate(items):
            try:
                if s is None:
                    a, t = int(arr[i].tri_offset), int(arr[i].n_tri)
                    res = (rows_all[a:a + t].copy(), grid)
                else:
                    t0 = time.perf_counter()
                    g = self._pool.array((s.height, s.width, 3), np.int32)
                    r = np.empty((int(arr[i].n_tri), 5), dtype=np.int32)
                    _lib.check(lib.dl_session_copy_result(s.handle, _lib.i32p(g),
                                                          _lib.i32p(r) if r.shape[0] else None))
                    t_copy += (time.perf_counter() - t0) * 1e3
                    res = (r, g)
                fut.t_done = time.perf_counter()
                fut.set_result(res if fut.keep else None)
            except Exception as e:
                fut.set_exception(e)
        return t_copy

    def _run(self, ctx, batch, total) -> BatchRecord | None:
        lib = self._lib
        t_start = time.perf_counter()
        items = []
        for req, fut, _ in batch:
            if isinstance(req, UpdateRequest):
                s = self._sessions[req.session_id]
                try:
                    # removal then insertion, all or nothing
                    _lib.check(lib.dl_session_update(ctx.handle, s.handle, _lib.i32p(req.remove),
                                                     req.remove.shape[0], _lib.i32p(req.insert),
                                                     None, req.insert.shape[0]))
                except Exception as e:   # invalid update: this request fails alone
                    fut.set_exception(e)
                    continue
                s.n_live += req.insert.shape[0] - req.remove.shape[0]
                items.append((req, fut, s, None))
            else:
                grid = self._pool.array((req.height, req.width, 3), np.int32)
                items.append((req, fut, None, grid))
        if not items:
            return None
        gpu_ms = copy_ms = 0.0
        try:
            arr, rows_all = self._run_pass(ctx, items)
            gpu_ms, copy_ms = self._stage_ms(ctx)
            copy_ms += self._deliver(items, arr, rows_all)
        except Exception:
            # Isolate the failure: every image again on its own.
            for it in items:
                req, fut, s, _ = it
                try:
                    arr, rows_all = self._run_pass(ctx, [it])
                except Exception as e:
                    if s is not None:
                        e = UpdateAppliedError(f"the update was applied to session "
                                               f"{req.session_id}, but its recompute failed: {e}")
                    fut.set_exception(e)
                    continue
                g, c = self._stage_ms(ctx)
                gpu_ms += g
                copy_ms += c + self._deliver([it], arr, rows_all)
        return BatchRecord(len(batch), total, gpu_ms, copy_ms,
                           (time.perf_counter() - t_start) * 1e3)

    def _stage_ms(self, ctx):
        """(pipeline stages, copy back) device ms of the context's last pass."""
        t = np.zeros(9, dtype=np.float64)
        _lib.check(self._lib.dl_get_timings(ctx.handle,
                                            t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return float(t[:5].sum()), float(t[5])

    def close(self):
        with self._cond:
            self._stop = True
            self._cond.notify_all()
        for w in self._workers:
            w.join()
        for s in self._sessions.values():
            self._lib.dl_session_destroy(s.handle)
        self._sessions.clear()
        for c in self._ctxs:
            c.close()
        self._pool.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
