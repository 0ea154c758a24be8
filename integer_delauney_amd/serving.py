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
  session run in arrival order and never in two batches at once.  An invalid
  update fails its own future and leaves the session unchanged.

Admission: a request's device footprint is estimated from its size
(:func:`footprint`).  A worker admits waiting requests in arrival order (those
of a busy session are skipped) while the bytes in flight over all workers stay
within ``hbm_budget - resident session bytes`` and its batch within
``max_batch_bytes``; a request that does not fit waits.  One request runs alone even if it exceeds the budget, so nothing
starves.  Each result is the same as the request's result computed alone:
batching never changes a result (tests/test_gpu_serving.py).

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


def footprint(width: int, height: int, n_seeds: int, padding: int = DEFAULT_BORDER_PADDING) -> int:
    """Device bytes one image needs inside a batch: canvas planes and
    detection counters per padded cell, seed maps and the staged output per
    pixel, and seed/triangle records."""
    wc, hc = width + 2 * padding, height + 2 * padding
    return int(14 * wc * hc + 20 * width * height + 400 * n_seeds)


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


class _Session:
    def __init__(self, handle, width, height, padding, max_seeds):
        self.handle = handle
        self.width = width
        self.height = height
        self.padding = padding
        self.max_seeds = max_seeds
        self.n_live = 0
        self.busy = False
        self.resident = 12 * width * height + width * height // 8 + 16 * max_seeds


class Server:
    def __init__(self, hbm_budget: int = 64 << 30, slots: int = 3, max_batch: int = 256,
                 max_batch_bytes: int = 2 << 30, device: int = 0):
        self._lib = _lib.load()
        self._budget = int(hbm_budget)
        self._max_batch = int(max_batch)
        # Batches stay small enough that one slot's results travel back to
        # the host while another slot's batch computes.
        self._max_batch_bytes = int(max_batch_bytes)
        self._ctxs = [_lib.Context(device) for _ in range(slots)]
        for c in self._ctxs:
            # Mixed request sizes: a steady latency matters more than the
            # first-use tuning of tile shapes or the launch-count prediction
            # (both would occasionally spend a request's time on extra work).
            _lib.check(self._lib.dl_set_autotune(c.handle, 0))
            _lib.check(self._lib.dl_set_fill_prediction(c.handle, 0))
        # results land in page-locked memory: the copy back runs at link speed
        self._pool = _lib.PinnedPool()
        self._cond = threading.Condition()
        self._queue: deque = deque()
        self._sessions: dict[int, _Session] = {}
        self._next_sid = 0
        self._inflight = 0
        self._resident = 0
        self._stop = False
        self.batches = []   # (n_requests, bytes) per batch run, for reporting
        self._workers = [threading.Thread(target=self._work, args=(i,), daemon=True)
                         for i in range(slots)]
        for w in self._workers:
            w.start()

    # -- sessions ------------------------------------------------------------
    def create_session(self, width: int, height: int, max_seeds: int,
                       border_padding: int = AUTO_BORDER_PADDING, seeds=None) -> int:
        """A resident session (device memory held until close); its seeds, if
        given, are inserted and finalised at once."""
        pad = DEFAULT_BORDER_PADDING if border_padding is None or border_padding < 0 \
            else int(border_padding)
        h = ctypes.c_void_p()
        ctx = self._ctxs[0]
        _lib.check(self._lib.dl_session_create(ctx.handle, int(width), int(height), int(max_seeds),
                                               pad, ctypes.byref(h)))
        s = _Session(h, int(width), int(height), pad, int(max_seeds))
        with self._cond:
            sid = self._next_sid
            self._next_sid += 1
            self._sessions[sid] = s
            self._resident += s.resident
        if seeds is not None:
            xy = as_seed_array(seeds)
            _lib.check(self._lib.dl_session_insert(ctx.handle, h, _lib.i32p(xy), None, xy.shape[0]))
            s.n_live = xy.shape[0]
            n = ctypes.c_int64()
            _lib.check(self._lib.dl_session_finalise(ctx.handle, h, ctypes.byref(n)))
        return sid

    # -- requests --------------------------------------------------------------
    def submit(self, req, keep: bool = True) -> Future:
        """Queue a request.  keep=False completes the future with None once the
        result has been produced and copied back, releasing it at once (load
        generators that only time requests)."""
        fut: Future = Future()
        fut.t_submit = time.perf_counter()
        fut.keep = keep
        if isinstance(req, TriangulateRequest):
            req.seeds = as_seed_array(req.seeds)
            pad = DEFAULT_BORDER_PADDING if req.border_padding is None or req.border_padding == -1 \
                else max(0, int(req.border_padding))
            cost = footprint(req.width, req.height, req.seeds.shape[0], pad)
        elif isinstance(req, UpdateRequest):
            if req.session_id not in self._sessions:
                raise ValueError("unknown session")
            req.insert = np.ascontiguousarray(np.asarray(req.insert, dtype=np.int32).reshape(-1, 2))
            req.remove = np.ascontiguousarray(np.asarray(req.remove, dtype=np.int32).reshape(-1, 2))
            s = self._sessions[req.session_id]
            cost = footprint(s.width, s.height, s.max_seeds, s.padding)
        else:
            raise TypeError("unknown request type")
        with self._cond:
            if self._stop:
                raise RuntimeError("server is closed")
            self._queue.append((req, fut, cost))
            self._cond.notify()
        return fut

    def _take(self):
        """Admit waiting requests for one batch (called with the lock held)."""
        batch, total = [], 0
        blocked = set()
        limit = self._budget - self._resident
        rest = deque()
        while self._queue:
            item = self._queue.popleft()
            req, fut, cost = item
            if isinstance(req, UpdateRequest):
                s = self._sessions[req.session_id]
                if s.busy or req.session_id in blocked:
                    blocked.add(req.session_id)
                    rest.append(item)
                    continue
            over_budget = self._inflight + total + cost > limit
            over_batch = bool(batch) and total + cost > self._max_batch_bytes
            alone = not batch and self._inflight == 0   # runs even above the budget
            if len(batch) >= self._max_batch or over_batch or (over_budget and not alone):
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
            try:
                self._run(ctx, batch)
            finally:
                with self._cond:
                    self._inflight -= total
                    for req, _, _ in batch:
                        if isinstance(req, UpdateRequest):
                            self._sessions[req.session_id].busy = False
                    self.batches.append((len(batch), total))
                    self._cond.notify_all()

    def _run(self, ctx, batch):
        lib = self._lib
        images, slots, outs = [], [], []
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
                im = _lib.DLImage()
                im.session = s.handle
                images.append(im)
                slots.append((req, fut, s))
                outs.append(None)
            else:
                im = _lib.DLImage()
                im.width, im.height = req.width, req.height
                im.border_padding = -1 if req.border_padding is None else int(req.border_padding)
                im.host = 1
                im.n_seeds = req.seeds.shape[0]
                im.seeds_xy = req.seeds.ctypes.data
                grid = self._pool.array((req.height, req.width, 3), np.int32)
                im.out = grid.ctypes.data
                images.append(im)
                slots.append((req, fut, None))
                outs.append(grid)
        if not images:
            return
        arr = (_lib.DLImage * len(images))(*images)
        try:
            _lib.check(lib.dl_run_batch(ctx.handle, len(images), ctypes.cast(arr, ctypes.c_void_p),
                                        None))
            # rows of the whole batch, image after image (fresh images' slices
            # are returned; sessions keep their own copy on the device)
            total_rows = max(int(arr[i].tri_offset + arr[i].n_tri) for i in range(len(images)))
            rows_all = np.empty((total_rows, 5), dtype=np.int32)
            if total_rows:
                _lib.check(lib.dl_copy_triangles(ctx.handle, _lib.i32p(rows_all)))
        except Exception as e:
            for _, fut, _ in slots:
                if not fut.done():
                    fut.set_exception(e)
            return
        now = time.perf_counter
        for i, ((req, fut, s), grid) in enumerate(zip(slots, outs)):
            try:
                if s is None:
                    a, t = int(arr[i].tri_offset), int(arr[i].n_tri)
                    res = (rows_all[a:a + t].copy(), grid)
                else:
                    g = self._pool.array((s.height, s.width, 3), np.int32)
                    r = np.empty((int(arr[i].n_tri), 5), dtype=np.int32)
                    _lib.check(lib.dl_session_copy_result(s.handle, _lib.i32p(g),
                                                          _lib.i32p(r) if r.shape[0] else None))
                    res = (r, g)
                fut.t_done = now()
                fut.set_result(res if fut.keep else None)
            except Exception as e:
                fut.set_exception(e)

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
