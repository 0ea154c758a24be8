"""dl_triangulate_device_noalloc, the entry point that allocates nothing, against
dl_triangulate_device (the entry point it is a variant of) and the oracle: two
calls in a row, the second with a smaller input, then a third larger one, all
through the C ABI on device buffers; plus its size limit.

Device buffers for the inputs and outputs come from the HIP runtime the
library itself uses (libamdhip64 through ctypes): the caller allocates them,
the entry point does not."""
import ctypes

import numpy as np
import pytest

from integer_delauney_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    _lib.load()   # its libamdhip64 is the one loaded below
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return hip


class _Dev:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(int(nbytes), 4)) == 0
        self.nbytes = int(nbytes)

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D) == 0

    def get(self, shape, dtype=np.int32):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H) == 0
        return out

    def fill(self, byte):
        assert self.hip.hipMemset(self.ptr, byte, self.nbytes) == 0

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


def _seeds(W, H, n, seed):
    rng = np.random.default_rng(seed)
    idx = rng.choice(W * H, size=n, replace=False)
    return np.ascontiguousarray(np.stack([idx % W, idx // W], 1).astype(np.int32))


def test_noalloc_matches_the_allocating_entry_point():
    lib = _lib.load()
    ctx = _lib.context()
    hip = _hip()
    cases = [(1000, 800, 12_000, 16), (300, 200, 700, 5), (1024, 1024, 50_000, -1)]
    bufs = []
    try:
        for W, H, n, P in cases:
            seeds = _seeds(W, H, n, W + n)
            d_xy = _Dev(hip, seeds.nbytes)
            d_xy.put(seeds)
            d_ref = _Dev(hip, 12 * W * H)
            d_out = _Dev(hip, 12 * W * H)
            d_out.fill(0x5A)   # stale contents must not matter
            cap = 4 * n + 64
            d_rows = _Dev(hip, 20 * cap)
            bufs += [d_xy, d_ref, d_out, d_rows]
            nt_ref, nt = ctypes.c_int64(0), ctypes.c_int64(0)
            _lib.check(lib.dl_triangulate_device(ctx.handle, W, H, d_xy.ptr, n, P, d_ref.ptr,
                                                 None, ctypes.byref(nt_ref)))
            ref_grid = d_ref.get((H, W, 3))
            ref_rows = np.empty((nt_ref.value, 5), np.int32)
            _lib.check(lib.dl_copy_triangles(ctx.handle, _lib.i32p(ref_rows)))
            _lib.check(lib.dl_triangulate_device_noalloc(ctx.handle, W, H, d_xy.ptr, n, P,
                                                         d_out.ptr, d_rows.ptr, cap, None,
                                                         ctypes.byref(nt)))
            assert nt.value == nt_ref.value
            np.testing.assert_array_equal(d_out.get((H, W, 3)), ref_grid)
            np.testing.assert_array_equal(d_rows.get((nt.value, 5)), ref_rows)
            o_rows, o_grid = O.triangulate(W, H, seeds, None if P == -1 else P, return_rows=True)
            np.testing.assert_array_equal(ref_grid, o_grid)
            np.testing.assert_array_equal(ref_rows, o_rows)
            print(f"{W}x{H} n={n} P={P}: {nt.value} triangles, equal", flush=True)
    finally:
        for b in bufs:
            b.free()


def test_noalloc_limits():
    lib = _lib.load()
    ctx = _lib.context()
    hip = _hip()
    W, H, n = 3000, 3000, 1000
    seeds = _seeds(W, H, n, 1)
    d_xy = _Dev(hip, seeds.nbytes)
    d_out = _Dev(hip, 12 * W * H)
    try:
        d_xy.put(seeds)
        nt = ctypes.c_int64(-1)
        rc = lib.dl_triangulate_device_noalloc(ctx.handle, W, H, d_xy.ptr, n, 16, d_out.ptr,
                                               None, 0, None, ctypes.byref(nt))
        assert rc == _lib.DL_EVALUE
        assert "scratch" in lib.dl_last_error().decode()
        assert nt.value == -1   # rejected before anything ran
        # a rows buffer that is too small: the count comes back with the error
        W2, H2, n2 = 200, 150, 400
        s2 = _seeds(W2, H2, n2, 2)
        d_xy.put(s2)
        d_rows = _Dev(hip, 20 * 4)
        try:
            rc = lib.dl_triangulate_device_noalloc(ctx.handle, W2, H2, d_xy.ptr, n2, 16,
                                                   d_out.ptr, d_rows.ptr, 4, None,
                                                   ctypes.byref(nt))
            assert rc == _lib.DL_EVALUE and nt.value > 4
            o_rows, o_grid = O.triangulate(W2, H2, s2, 16, return_rows=True)
            assert nt.value == len(o_rows)
            np.testing.assert_array_equal(d_out.get((H2, W2, 3)), o_grid)
        finally:
            d_rows.free()
    finally:
        d_xy.free()
        d_out.free()
