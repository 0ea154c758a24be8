"""Synthetic seed sets for the benchmark workloads.

Every generator is deterministic in its ``seed`` argument and returns an
(N, 2) int32 array of distinct (x, y) integer positions in draw order (callers
may pass it straight to ``triangulate``; the id order is set by the library).

* :func:`uniform_seeds`   -- N distinct points uniform on [0,W)x[0,H): the
  single-image workloads (256x256/1,000 and 4096x4096/200,000).
* :func:`clustered_seeds` -- a Gaussian mixture, rounded and deduplicated, for
  the 16384x16384 / 2,000,000-seed incremental workload.
* :func:`batch_shapes`    -- uneven image sizes and densities for the batched
  workload (sides uniform in [128, 2048], density 0.5-2% of pixels).
"""
from __future__ import annotations

import numpy as np


def _distinct_in_order(chunks, n):
    """Keep the first n distinct rows over an iterator of (k,2) int arrays."""
    seen = set()
    out = []
    for c in chunks:
        for x, y in c.tolist():
            key = (x, y)
            if key in seen:
                continue
            seen.add(key)
            out.append(key)
            if len(out) == n:
                return np.asarray(out, dtype=np.int32).reshape(-1, 2)
    return np.asarray(out, dtype=np.int32).reshape(-1, 2)


def uniform_seeds(width: int, height: int, n: int, seed: int) -> np.ndarray:
    """First n distinct draws of integers(0, [width, height]) from default_rng(seed)."""
    if n > width * height:
        raise ValueError("more seeds than pixels")
    rng = np.random.default_rng(seed)

    def chunks():
        while True:
            k = max(1024, n - 0)
            yield rng.integers(0, [width, height], size=(k, 2))

    if n > 200_000:
        return _uniform_fast(width, height, n, rng)
    return _distinct_in_order(chunks(), n)


def _uniform_fast(width, height, n, rng):
    """Vectorised first-n-distinct for large n (same draw-order semantics)."""
    kept = np.empty((0, 2), dtype=np.int64)
    keys_seen = np.empty(0, dtype=np.int64)
    while kept.shape[0] < n:
        k = max(n - kept.shape[0], 1 << 16) * 2
        c = rng.integers(0, [width, height], size=(k, 2)).astype(np.int64)
        keys = c[:, 0] * height + c[:, 1]
        _, first = np.unique(keys, return_index=True)
        first.sort()
        c, keys = c[first], keys[first]
        fresh = ~np.isin(keys, keys_seen)
        c, keys = c[fresh], keys[fresh]
        kept = np.concatenate([kept, c])
        keys_seen = np.concatenate([keys_seen, keys])
    return kept[:n].astype(np.int32)


def clustered_seeds(width: int, height: int, n: int, seed: int, components: int = 64,
                    sigma: float = 200.0) -> np.ndarray:
    """n distinct points from an equal-weight Gaussian mixture.

    Centres are uniform on the canvas; samples are rounded with np.rint,
    out-of-range samples are rejected, duplicates dropped (first draw kept),
    and the set is topped up until it holds n points.
    """
    rng = np.random.default_rng(seed)
    centres = rng.uniform([0, 0], [width, height], size=(components, 2))
    kept = np.empty((0, 2), dtype=np.int64)
    keys_seen = np.empty(0, dtype=np.int64)
    while kept.shape[0] < n:
        k = max(n - kept.shape[0], 1 << 14) * 2
        comp = rng.integers(0, components, size=k)
        pts = np.rint(centres[comp] + rng.normal(0.0, sigma, size=(k, 2))).astype(np.int64)
        ok = (pts[:, 0] >= 0) & (pts[:, 0] < width) & (pts[:, 1] >= 0) & (pts[:, 1] < height)
        pts = pts[ok]
        keys = pts[:, 0] * height + pts[:, 1]
        _, first = np.unique(keys, return_index=True)
        first.sort()
        pts, keys = pts[first], keys[first]
        fresh = ~np.isin(keys, keys_seen)
        pts, keys = pts[fresh], keys[fresh]
        kept = np.concatenate([kept, pts])
        keys_seen = np.concatenate([keys_seen, keys])
    return kept[:n].astype(np.int32)


def clustered_more(width: int, height: int, existing: np.ndarray, n: int, seed: int,
                   components: int = 64, sigma: float = 200.0,
                   centres_seed: int | None = None) -> np.ndarray:
    """n new distinct mixture samples not present in ``existing``."""
    rng_c = np.random.default_rng(centres_seed if centres_seed is not None else seed)
    centres = rng_c.uniform([0, 0], [width, height], size=(components, 2))
    rng = np.random.default_rng(seed)
    have = set(map(tuple, np.asarray(existing).tolist()))
    out = []
    while len(out) < n:
        k = max(n - len(out), 1024) * 2
        comp = rng.integers(0, components, size=k)
        pts = np.rint(centres[comp] + rng.normal(0.0, sigma, size=(k, 2))).astype(np.int64)
        for x, y in pts.tolist():
            if 0 <= x < width and 0 <= y < height and (x, y) not in have:
                have.add((x, y))
                out.append((x, y))
                if len(out) == n:
                    break
    return np.asarray(out, dtype=np.int32).reshape(-1, 2)


def batch_shapes(count: int, seed: int, side_lo: int = 128, side_hi: int = 2048,
                 dens_lo: float = 0.005, dens_hi: float = 0.02):
    """(width, height, n_seeds, rng_seed) per image of the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w = int(rng.integers(side_lo, side_hi + 1))
        h = int(rng.integers(side_lo, side_hi + 1))
        dens = float(rng.uniform(dens_lo, dens_hi))
        n = max(1, int(round(dens * w * h)))
        out.append((w, h, n, int(seed) * 100_003 + i))
    return out


def batch_images(count: int, seed: int, **kw):
    """(width, height, seeds) per image of the batched workload: shapes from
    :func:`batch_shapes`, seeds the first n distinct uniform draws of each
    image's own generator (vectorised)."""
    out = []
    for w, h, n, sd in batch_shapes(count, seed, **kw):
        out.append((w, h, _uniform_fast(w, h, n, np.random.default_rng(sd))))
    return out
