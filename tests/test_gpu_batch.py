"""Batched triangulation: every image of a batch gets, bit for bit, what the
single-image call (and the oracle) gives it alone."""
import numpy as np
import pytest

import integer_delauney_amd as dl
from integer_delauney_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rows(m):
    return np.array([m[i] for i in range(len(m))], dtype=np.int32).reshape(-1, 5)


def _seeds(rng, W, H, n):
    idx = rng.choice(W * H, size=min(n, W * H), replace=False)
    return np.stack([idx % W, idx // W], 1).astype(np.int32)


def _mixed_batch(rng, k):
    imgs = []
    for _ in range(k):
        kind = rng.integers(0, 5)
        if kind == 0:      # degenerate strips
            W, H = (1, int(rng.integers(2, 60))) if rng.integers(0, 2) else (int(rng.integers(2, 60)), 1)
            n = int(rng.integers(1, 5))
        elif kind == 1:    # tiny
            W, H = int(rng.integers(2, 12)), int(rng.integers(2, 12))
            n = int(rng.integers(1, W * H + 1))
        else:
            W, H = int(rng.integers(20, 260)), int(rng.integers(20, 260))
            n = int(rng.integers(1, max(2, W * H // int(rng.choice([5, 40, 400])))))
        imgs.append((W, H, _seeds(rng, W, H, n)))
    return imgs


@pytest.mark.parametrize("seed", range(4))
def test_batch_equals_single_and_oracle(seed):
    rng = np.random.default_rng(500 + seed)
    imgs = _mixed_batch(rng, 14)
    pads = [int(rng.choice([-1, 0, 1, 3, 16])) for _ in imgs]
    res = dl.triangulate_batch(imgs, pads)
    assert len(res) == len(imgs)
    for (W, H, s), P, (m, grid) in zip(imgs, pads, res):
        m1, g1 = dl.triangulate(W, H, s, P)
        np.testing.assert_array_equal(grid, g1, err_msg=f"W={W} H={H} n={len(s)} P={P}")
        assert m == m1
        rows, ref = O.triangulate(W, H, s, None if P == -1 else P, return_rows=True)
        np.testing.assert_array_equal(grid, ref)
        np.testing.assert_array_equal(_rows(m), rows)


def test_batch_of_uneven_larger_images():
    from integer_delauney_amd.workloads import batch_shapes, uniform_seeds
    shapes = batch_shapes(10, 3, side_lo=128, side_hi=700)
    imgs = [(w, h, uniform_seeds(w, h, n, sd)) for (w, h, n, sd) in shapes]
    res = dl.triangulate_batch(imgs, as_arrays=True)
    for (W, H, s), (verts, grid) in zip(imgs, res):
        v1, g1 = dl.triangulate(W, H, s, as_arrays=True)
        np.testing.assert_array_equal(grid, g1)
        np.testing.assert_array_equal(verts, v1)


def test_batch_with_many_detection_blocks():
    """A batch with more than 8,192 detection thread blocks (each covers 256 x
    32 2x2 blocks) scans its per-block candidate counts with the multi-block
    scan: every image equals its single-image result, and a sample the
    oracle's."""
    from integer_delauney_amd.workloads import uniform_seeds
    rng = np.random.default_rng(41)
    imgs = []
    for i in range(44):
        W, H = int(rng.integers(900, 1300)), int(rng.integers(1000, 1400))
        imgs.append((W, H, uniform_seeds(W, H, int(W * H * rng.uniform(0.003, 0.02)), 900 + i)))
    det_blocks = sum(-(-(W + 31) // 256) * -(-(H + 31) // 32) for W, H, _ in imgs)
    assert det_blocks > 8192
    res = dl.triangulate_batch(imgs, as_arrays=True)
    for i, ((W, H, s), (verts, grid)) in enumerate(zip(imgs, res)):
        v1, g1 = dl.triangulate(W, H, s, as_arrays=True)
        np.testing.assert_array_equal(grid, g1, err_msg=f"image {i}")
        np.testing.assert_array_equal(verts, v1, err_msg=f"image {i}")
        if i % 11 == 0:
            rows, ref = O.triangulate(W, H, s, 16, return_rows=True)
            np.testing.assert_array_equal(grid, ref, err_msg=f"image {i}")
            np.testing.assert_array_equal(verts, rows[:, 2:5], err_msg=f"image {i}")


def test_batch_with_forced_second_fill():
    ctx = _lib.context()
    lib = _lib.load()
    rng = np.random.default_rng(77)
    imgs = _mixed_batch(rng, 8)
    base = dl.triangulate_batch(imgs)
    _lib.check(lib.dl_set_force_unpadded(ctx.handle, 1))
    try:
        forced = dl.triangulate_batch(imgs)
    finally:
        _lib.check(lib.dl_set_force_unpadded(ctx.handle, 0))
    for (m0, g0), (m1, g1) in zip(base, forced):
        assert m0 == m1
        np.testing.assert_array_equal(g0, g1)


def test_batch_repeated_image_and_order():
    rng = np.random.default_rng(78)
    a = (90, 70, _seeds(rng, 90, 70, 120))
    b = (33, 150, _seeds(rng, 33, 150, 60))
    r = dl.triangulate_batch([a, b, a, b, a])
    for i in (2, 4):
        assert r[i][0] == r[0][0] and np.array_equal(r[i][1], r[0][1])
    assert r[3][0] == r[1][0] and np.array_equal(r[3][1], r[1][1])


@pytest.mark.parametrize("bad", ["bounds", "duplicate", "empty", "size"])
def test_batch_errors_name_the_image(bad):
    good = (20, 20, [(1, 1), (5, 9), (13, 2)])
    if bad == "bounds":
        img = (20, 20, [(1, 1), (25, 3), (4, 4)])
    elif bad == "duplicate":
        img = (20, 20, [(1, 1), (1, 1), (4, 4)])
    elif bad == "empty":
        img = (20, 20, [])
    else:
        img = (0, 20, [(0, 0)])
    with pytest.raises(ValueError, match="image 1"):
        dl.triangulate_batch([good, img, good])
    # the context still works afterwards
    m, g = dl.triangulate(*good)
    m2, g2 = dl.triangulate_batch([good])[0]
    assert m == m2 and np.array_equal(g, g2)
