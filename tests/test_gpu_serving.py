"""The serving loop: requests submitted concurrently (fresh triangulations and
updates of resident sessions) each get exactly the result they get alone and
the oracle's result; invalid requests fail alone; admission by HBM budget
rejects what it can never hold and never exceeds what it can."""
import ctypes
import threading
import time

import numpy as np
import pytest

import integer_delauney_amd as dl
from integer_delauney_amd import _lib
from integer_delauney_amd.serving import (AdmissionError, Server, TriangulateRequest,
                                          UpdateRequest, footprint)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rows(m):
    return np.array([m[i] for i in range(len(m))], dtype=np.int32).reshape(-1, 5)


def _seeds(rng, W, H, n, exclude=None):
    taken = set() if exclude is None else {tuple(p) for p in exclude.tolist()}
    out = []
    while len(out) < n:
        p = (int(rng.integers(0, W)), int(rng.integers(0, H)))
        if p not in taken:
            taken.add(p)
            out.append(p)
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def _stream(rng, srv, n_req, invalid=False):
    """Requests plus, for each, how to compute its result alone.  With
    invalid, some fresh triangulations are invalid (duplicate or
    out-of-range seeds)."""
    sessions = {}
    for k in range(3):
        W, H = int(rng.integers(60, 300)), int(rng.integers(60, 300))
        s0 = _seeds(rng, W, H, int(rng.integers(20, 400)))
        sid = srv.create_session(W, H, 5000, seeds=s0)
        alone = dl.Delaunay(W, H, 5000)
        alone.insert(s0)
        sessions[sid] = {"live": s0, "alone": alone, "W": W, "H": H}
    reqs = []
    for i in range(n_req):
        if rng.random() < 0.3:
            sid = int(rng.choice(list(sessions)))
            st = sessions[sid]
            live = st["live"]
            k_del = int(rng.integers(0, max(1, live.shape[0] // 10)))
            drop = live[rng.choice(live.shape[0], size=k_del, replace=False)] if k_del else \
                np.empty((0, 2), np.int32)
            add = _seeds(rng, st["W"], st["H"], int(rng.integers(1, 30)), live)
            if i == n_req // 2:   # one invalid update: removes a position that is not live
                bad = _seeds(rng, st["W"], st["H"], 1, np.concatenate([live, add]))
                reqs.append((UpdateRequest(sid, insert=add, remove=np.concatenate([drop, bad])),
                             ("bad", sid)))
                continue
            keep = ~np.isin(live[:, 0] * 100000 + live[:, 1], drop[:, 0] * 100000 + drop[:, 1])
            st["live"] = np.concatenate([live[keep], add])
            reqs.append((UpdateRequest(sid, insert=add, remove=drop),
                         ("upd", sid, drop, add, st["live"].copy())))
        else:
            W, H = int(rng.integers(2, 400)), int(rng.integers(2, 300))
            n = int(rng.integers(1, max(2, W * H // int(rng.choice([10, 100, 1000])))))
            P = int(rng.choice([-1, 0, 4, 16]))
            s = _seeds(rng, W, H, min(n, W * H))
            if invalid and rng.random() < 0.25:
                if s.shape[0] > 1 and rng.random() < 0.5:
                    s[-1] = s[0]                       # a duplicate
                else:
                    s[rng.integers(0, s.shape[0])] = (W, 0)   # x out of range
                reqs.append((TriangulateRequest(W, H, s, P), ("bad", None)))
                continue
            reqs.append((TriangulateRequest(W, H, s, P), ("tri", W, H, s, P)))
    return reqs, sessions


def _check(reqs, futs, sessions, oracle_every=0):
    k = 0
    for (req, how), fut in zip(reqs, futs):
        if how[0] == "bad":
            with pytest.raises(ValueError):
                fut.result(timeout=120)
            continue
        rows, grid = fut.result(timeout=120)
        if how[0] == "tri":
            _, W, H, s, P = how
            m, g = dl.triangulate(W, H, s, P)
            np.testing.assert_array_equal(grid, g)
            np.testing.assert_array_equal(rows, _rows(m))
            if oracle_every and k % oracle_every == 0:
                ref_rows, ref = O.triangulate(W, H, s, None if P == -1 else P, return_rows=True)
                np.testing.assert_array_equal(grid, ref)
                np.testing.assert_array_equal(rows, ref_rows)
        else:
            _, sid, drop, add, live = how
            alone = sessions[sid]["alone"]
            if drop.shape[0]:
                alone.remove(drop)
            m, g = alone.insert(add)
            np.testing.assert_array_equal(grid, g)
            np.testing.assert_array_equal(rows, _rows(m))
            if oracle_every and k % oracle_every == 0:
                ref_rows, ref = O.triangulate(sessions[sid]["W"], sessions[sid]["H"], live, None,
                                              return_rows=True)
                np.testing.assert_array_equal(grid, ref)
                np.testing.assert_array_equal(rows, ref_rows)
        k += 1


def _feed(srv, reqs, rng):
    futs = [None] * len(reqs)

    def feed():
        for i, (req, _) in enumerate(reqs):
            futs[i] = srv.submit(req)
            if rng.random() < 0.3:
                time.sleep(0.002)

    t = threading.Thread(target=feed)
    t.start()
    t.join()
    for f in futs:
        f.exception(timeout=120)
    return futs


def test_results_are_the_same_alone_and_in_the_mix():
    rng = np.random.default_rng(900)
    with Server(slots=2) as srv:
        reqs, sessions = _stream(rng, srv, 60)
        futs = _feed(srv, reqs, rng)
        _check(reqs, futs, sessions, oracle_every=3)
        assert max(b.requests for b in srv.batches) > 1   # requests were batched together
        assert all(b.gpu_ms > 0 and b.copy_ms >= 0 for b in srv.batches)


def test_invalid_requests_fail_alone():
    """Out-of-range and duplicate seeds are caught when submitted: the
    request's own future fails with ValueError and the valid requests and
    session updates batched around it are exact."""
    rng = np.random.default_rng(903)
    with Server(slots=2) as srv:
        reqs, sessions = _stream(rng, srv, 50, invalid=True)
        assert sum(1 for _, h in reqs if h[0] == "bad") >= 3
        futs = _feed(srv, reqs, rng)
        _check(reqs, futs, sessions, oracle_every=4)
    with Server(slots=1) as srv:
        f = srv.submit(TriangulateRequest(40_000, 10, np.array([[0, 0], [1, 1]], np.int32)))
        with pytest.raises(RuntimeError):
            f.result(timeout=60)


def test_tight_budget_never_exceeded():
    rng = np.random.default_rng(901)
    with Server(slots=2, idle_bytes_per_slot=0) as srv:
        reqs, sessions = _stream(rng, srv, 12)
        costs = [footprint(r.width, r.height, len(r.seeds), 16) for r, h in reqs
                 if h[0] == "tri"]
        limit = max(costs + [footprint(300, 300, 5000, 16)])
        srv.set_budget(srv.resident_bytes + limit)
        futs = [srv.submit(r) for r, _ in reqs]
        _check(reqs, futs, sessions)
        assert srv.peak_inflight <= limit
        assert srv.rejected == 0


def test_requests_over_the_budget_are_rejected():
    rng = np.random.default_rng(902)
    with Server(slots=2, idle_bytes_per_slot=1 << 20) as srv:
        srv.set_budget(srv.resident_bytes + (2 << 20) + footprint(256, 256, 600, 16))
        small = TriangulateRequest(200, 150, _seeds(rng, 200, 150, 300))
        big = TriangulateRequest(2000, 2000, _seeds(rng, 2000, 2000, 40_000))
        fs, fb = srv.submit(small), srv.submit(big)
        with pytest.raises(AdmissionError, match="budget"):
            fb.result(timeout=60)
        rows, grid = fs.result(timeout=60)
        ref_rows, ref = O.triangulate(200, 150, small.seeds, None, return_rows=True)
        np.testing.assert_array_equal(grid, ref)
        np.testing.assert_array_equal(rows, ref_rows)
        assert srv.rejected == 1
        with pytest.raises(AdmissionError):
            srv.create_session(4000, 4000, 100_000)


@pytest.mark.parametrize("W,H,density", [(512, 512, 0.01), (1024, 768, 0.02), (2048, 2048, 0.01),
                                         (4096, 2048, 0.005), (8192, 8192, 0.01)])
def test_footprint_bounds_measured_device_bytes(W, H, density):
    """footprint() is an upper bound on what one image makes a fresh context
    reserve (the sum of its workspaces) and on the device memory that
    hipMemGetInfo sees taken (allocation rounding included, 64 MiB slack)."""
    lib = _lib.load()
    ctx = _lib.Context(0)
    try:
        rng = np.random.default_rng(W + H)
        n = int(density * W * H)
        idx = rng.choice(W * H, size=n, replace=False)
        seeds = np.ascontiguousarray(np.stack([idx % W, idx // W], 1).astype(np.int32))
        out = np.empty((H, W, 3), np.int32)
        im = _lib.DLImage()
        im.width, im.height, im.border_padding, im.host = W, H, 16, 1
        im.n_seeds, im.seeds_xy, im.out = n, seeds.ctypes.data, out.ctypes.data
        arr = (_lib.DLImage * 1)(im)
        free0, tot = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _lib.check(lib.dl_device_memory(0, ctypes.byref(free0), ctypes.byref(tot)))
        _lib.check(lib.dl_run_batch(ctx.handle, 1, ctypes.cast(arr, ctypes.c_void_p), None))
        free1 = ctypes.c_size_t(0)
        _lib.check(lib.dl_device_memory(0, ctypes.byref(free1), ctypes.byref(tot)))
        held = ctypes.c_size_t(0)
        _lib.check(lib.dl_context_device_bytes(ctx.handle, ctypes.byref(held)))
        fp = footprint(W, H, n, 16)
        taken = free0.value - free1.value
        print(f"{W}x{H} n={n}: reserved {held.value / 2**20:.1f} MiB, hipMemGetInfo "
              f"{taken / 2**20:.1f} MiB, footprint {fp / 2**20:.1f} MiB", flush=True)
        assert held.value <= fp
        assert taken <= fp + (64 << 20)
        assert held.value >= fp // 4   # and not wildly above it
        _lib.check(lib.dl_context_trim(ctx.handle))
        _lib.check(lib.dl_context_device_bytes(ctx.handle, ctypes.byref(held)))
        assert held.value == 0
        # the trimmed context still works
        _lib.check(lib.dl_run_batch(ctx.handle, 1, ctypes.cast(arr, ctypes.c_void_p), None))
        m, g = dl.triangulate(W, H, seeds, 16)
        np.testing.assert_array_equal(out, g)
    finally:
        ctx.close()
