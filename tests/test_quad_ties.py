"""Four-region blocks whose corner seeds lie in exactly the same direction from
their centroid: the one sub-case of the cocircular quad rule where the
reference's arctan2 ordering (triangulation.py:120-124, a stable sorted() on
np.arctan2 angles) and the exact comparator could part.

* libm's atan2 (which NumPy's arctan2 is when it does not dispatch its
  AVX-512 kernels) returns identical angles for positively proportional
  integer vectors, so such corners tie and the stable sort keeps their
  TL, TR, BL, BR order -- what the oracle (libm atan2, stable insertion sort)
  and the GPU's exact comparator (a tie, stable insertion sort) do.
* NumPy's AVX-512 arctan2 differs in the last ulp for a few tenths of a
  percent of such pairs, so on AVX-512 hosts the reference's own order there
  is decided by rounding noise of its SIMD kernel.
* For four exactly cocircular seeds two corners on one ray from the centroid
  would put the centroid outside their circle; the discrete diagram obeys this
  only approximately, and a seeded search over small dense inputs finds none.
"""
import math
import os
import subprocess
import sys

import numpy as np

from oracle import oracle as O


def _proportional_pairs(rng, n):
    x = rng.integers(-2**14, 2**14, n)
    y = rng.integers(-2**14, 2**14, n)
    g = np.gcd(x, y)
    ok = g > 0
    x, y = x[ok] // g[ok], y[ok] // g[ok]   # primitive directions
    a = rng.integers(1, 64, x.size)
    b = rng.integers(1, 64, x.size)
    # 4 * (corner - centroid) stays below 2^17 on canvases below 32768
    m = (np.abs(x) * np.maximum(a, b) < 2**17) & (np.abs(y) * np.maximum(a, b) < 2**17)
    return x[m], y[m], a[m], b[m]


def test_libm_atan2_ties_positively_proportional_vectors():
    x, y, a, b = _proportional_pairs(np.random.default_rng(11), 20_000)
    for k in range(x.size):
        assert math.atan2(float(a[k] * y[k]), float(a[k] * x[k])) == \
            math.atan2(float(b[k] * y[k]), float(b[k] * x[k]))


def test_numpy_arctan2_without_avx512_is_libm():
    """In a fresh interpreter with NumPy's AVX-512 dispatch switched off (the
    reference's arctan2 on a host without AVX-512): equal angles for
    proportional vectors, and the same values as libm."""
    code = (
        "import math, numpy as np, sys\n"
        "sys.path.insert(0, %r)\n"
        "from tests.test_quad_ties import _proportional_pairs\n"
        "x, y, a, b = _proportional_pairs(np.random.default_rng(12), 400_000)\n"
        "t1 = np.arctan2((a * y).astype(float), (a * x).astype(float))\n"
        "t2 = np.arctan2((b * y).astype(float), (b * x).astype(float))\n"
        "assert int(np.count_nonzero(t1 != t2)) == 0, 'unequal'\n"
        "for k in range(0, x.size, max(1, x.size // 3000)):\n"
        "    s = np.arctan2(np.float64(a[k] * y[k]), np.float64(a[k] * x[k]))\n"
        "    assert s == math.atan2(float(a[k] * y[k]), float(a[k] * x[k]))\n"
        "print(x.size)\n" % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    env = dict(os.environ)
    env["NPY_DISABLE_CPU_FEATURES"] = ("AVX512F AVX512CD AVX512_SKX AVX512_CLX AVX512_CNL "
                                       "AVX512_ICL AVX512_SPR")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert int(r.stdout.strip().splitlines()[-1]) > 100_000


def _equal_direction_blocks(W, H, seeds, P):
    srt = O.sort_seeds(W, H, seeds).astype(np.int64)
    if P > 0:
        ids = O.voronoi_sorted(W + 2 * P, H + 2 * P, (srt + P).astype(np.int32))[:, :, 0]
    else:
        ids = O.voronoi_sorted(W, H, srt.astype(np.int32))[:, :, 0]
    tl, tr, bl, br = ids[:-1, :-1], ids[:-1, 1:], ids[1:, :-1], ids[1:, 1:]
    quad = (tl != tr) & (tl != bl) & (tl != br) & (tr != bl) & (tr != br) & (bl != br)
    n = 0
    for yy, xx in zip(*np.nonzero(quad)):
        p = srt[[tl[yy, xx], tr[yy, xx], bl[yy, xx], br[yy, xx]]]
        v = 4 * p - p.sum(0)
        for i in range(4):
            for j in range(i + 1, 4):
                if v[i, 0] * v[j, 1] == v[i, 1] * v[j, 0] and v[i, 0] * v[j, 0] + v[i, 1] * v[j, 1] > 0:
                    n += 1
    return n, int(quad.sum())


def test_equal_direction_blocks_do_not_arise_in_random_inputs():
    rng = np.random.default_rng(13)
    quads = 0
    for _ in range(1500):
        W, H = int(rng.integers(5, 30)), int(rng.integers(5, 30))
        n = int(rng.integers(4, max(5, W * H // 5)))
        idx = rng.choice(W * H, size=n, replace=False)
        seeds = np.stack([idx % W, idx // W], 1).astype(np.int32)
        P = int(rng.choice([0, 1, 3, 16]))
        hits, q = _equal_direction_blocks(W, H, seeds, P)
        assert hits == 0, (W, H, P, seeds.tolist())
        quads += q
    assert quads > 1000   # four-region blocks did occur
