"""The serving loop: requests submitted concurrently (fresh triangulations and
updates of resident sessions) each get exactly the result they get alone."""
import threading
import time

import numpy as np
import pytest

import integer_delauney_amd as dl
from integer_delauney_amd.serving import Server, TriangulateRequest, UpdateRequest

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


def _stream(rng, srv, n_req):
    """Requests plus, for each, how to compute its result alone."""
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
            reqs.append((UpdateRequest(sid, insert=add, remove=drop), ("upd", sid, drop, add)))
        else:
            W, H = int(rng.integers(2, 400)), int(rng.integers(2, 300))
            n = int(rng.integers(1, max(2, W * H // int(rng.choice([10, 100, 1000])))))
            P = int(rng.choice([-1, 0, 4, 16]))
            s = _seeds(rng, W, H, min(n, W * H))
            reqs.append((TriangulateRequest(W, H, s, P), ("tri", W, H, s, P)))
    return reqs, sessions


def _check(reqs, futs, sessions):
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
        else:
            _, sid, drop, add = how
            alone = sessions[sid]["alone"]
            if drop.shape[0]:
                alone.remove(drop)
            m, g = alone.insert(add)
            np.testing.assert_array_equal(grid, g)
            np.testing.assert_array_equal(rows, _rows(m))


def test_results_are_the_same_alone_and_in_the_mix():
    rng = np.random.default_rng(900)
    with Server(slots=2) as srv:
        reqs, sessions = _stream(rng, srv, 60)
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
        _check(reqs, futs, sessions)
        assert max(n for n, _ in srv.batches) > 1   # requests were batched together


def test_tight_budget_runs_requests_one_at_a_time():
    rng = np.random.default_rng(901)
    with Server(slots=2, hbm_budget=1) as srv:
        reqs, sessions = _stream(rng, srv, 12)
        futs = [srv.submit(r) for r, _ in reqs]
        _check(reqs, futs, sessions)
        assert all(n == 1 for n, _ in srv.batches)
