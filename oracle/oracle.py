"""Python face of the CPU oracle (dl_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product package
(``integer_delauney_amd``) never imports it.

The functions mirror ``delauney.reference`` (src/delauney/reference/voronoi.py
and triangulation.py) one for one:

* :func:`voronoi`            -- ``reference.Voronoi().compute`` (voronoi.py:34-119)
* :func:`grid_triangulation` -- ``reference.GridTriangulation().compute``
  (triangulation.py:26-219)
* :func:`triangulate`        -- the two in sequence, which is what the batch
  ``triangulate`` entry point is defined to equal.

``assign="brute"`` runs the reference's O(H*W*T) assignment loop
(triangulation.py:193-204) in C; ``assign="raster"`` runs the per-triangle
scanline evaluation of the same function, which the tests pin against
``brute`` so that the oracle can answer at sizes where brute force cannot.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdl_oracle.so")
_lib = None

DEFAULT_BORDER_PADDING = 16

_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile dl_oracle.c into oracle/libdl_oracle.so (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_sort_seeds.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64, _I32P]
    lib.orc_sort_seeds.restype = ctypes.c_int
    lib.orc_voronoi.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64, _I32P, _I64P]
    lib.orc_voronoi.restype = ctypes.c_int
    lib.orc_detect.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, ctypes.c_int,
                               ctypes.POINTER(_I32P), _I64P, _I64P]
    lib.orc_detect.restype = ctypes.c_int
    lib.orc_voronoi_watch.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64,
                                      ctypes.c_int, _I32P, _I64P]
    lib.orc_voronoi_watch.restype = ctypes.c_int
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_free.restype = None
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads.restype = ctypes.c_int
    for name in ("orc_assign_brute", "orc_assign_raster"):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, ctypes.c_int64, _I32P]
        fn.restype = None
    _lib = lib
    return lib


def set_threads(n: int) -> int:
    """Threads of the oracle's parallel parts (flood-fill passes, raster row
    bands; results do not depend on it).  n <= 0 restores the OpenMP default
    (OMP_NUM_THREADS, else all cores).  Returns the count now in use."""
    return int(_load().orc_set_threads(int(n)))


def _p(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


def _as_seed_array(seeds) -> np.ndarray:
    arr = np.asarray(seeds if not isinstance(seeds, np.ndarray) else seeds)
    if arr.size == 0:
        raise ValueError("seeds must not be empty")
    arr = np.asarray(arr, dtype=np.int32)
    if arr.ndim != 2 or arr.shape[1] != 2:
        raise ValueError("each seed must be an (x, y) pair")
    return np.ascontiguousarray(arr)


def sort_seeds(width: int, height: int, seeds) -> np.ndarray:
    """voronoi.py:90-108: validate and return seeds in id order, (N, 2) int32."""
    if seeds is None or (not isinstance(seeds, np.ndarray) and len(seeds) == 0):
        raise ValueError("seeds must not be empty")
    arr = _as_seed_array(seeds)
    out = np.empty_like(arr)
    rc = _load().orc_sort_seeds(int(width), int(height), _p(arr), arr.shape[0], _p(out))
    if rc == 1:
        raise ValueError("seeds must not be empty")
    if rc == 2:
        raise ValueError("seed x coordinate out of bounds")
    if rc == 3:
        raise ValueError("seed y coordinate out of bounds")
    if rc == 4:
        raise ValueError("duplicate seed positions are not allowed")
    return out


def voronoi_sorted(width: int, height: int, sorted_xy: np.ndarray, return_iters=False):
    """voronoi.py:110-163 on seeds already in id order."""
    sorted_xy = np.ascontiguousarray(sorted_xy, dtype=np.int32)
    out = np.empty((height, width, 2), dtype=np.int32)
    iters = ctypes.c_int64(0)
    rc = _load().orc_voronoi(int(width), int(height), _p(sorted_xy), sorted_xy.shape[0],
                             _p(out), ctypes.byref(iters))
    if rc != 0:
        raise MemoryError("oracle voronoi allocation failed")
    return (out, int(iters.value)) if return_iters else out


def voronoi_crossings(width: int, height: int, sorted_xy: np.ndarray, box_margin: int) -> int:
    """Number of flood-fill updates that carry a value into the box
    [m, W-m) x [m, H-m) from a cell outside it (0 means the box's cells evolve
    exactly as if the canvas ended at the box)."""
    sorted_xy = np.ascontiguousarray(sorted_xy, dtype=np.int32)
    out = np.empty((height, width, 2), dtype=np.int32)
    cross = ctypes.c_int64(0)
    rc = _load().orc_voronoi_watch(int(width), int(height), _p(sorted_xy), sorted_xy.shape[0],
                                   int(box_margin), _p(out), ctypes.byref(cross))
    if rc != 0:
        raise MemoryError("oracle voronoi allocation failed")
    return int(cross.value)


def voronoi(width: int, height: int, seeds) -> np.ndarray:
    """reference.Voronoi().compute(width, height, seeds)."""
    return voronoi_sorted(width, height, sort_seeds(width, height, seeds))


def detect(id_grid: np.ndarray, sorted_xy: np.ndarray, padding: int):
    """triangulation.py:109-183 on an (Hp, Wp) id grid.

    Returns ((T, 5) int32 rows (x, y, a, b, c) in tid order, number of quads).
    """
    g = np.ascontiguousarray(id_grid, dtype=np.int32)
    s = np.ascontiguousarray(sorted_xy, dtype=np.int32)
    hp, wp = g.shape
    ptr = _I32P()
    cnt = ctypes.c_int64(0)
    quads = ctypes.c_int64(0)
    rc = _load().orc_detect(wp, hp, _p(g), _p(s), int(padding), ctypes.byref(ptr),
                            ctypes.byref(cnt), ctypes.byref(quads))
    if rc != 0:
        raise MemoryError("oracle detect allocation failed")
    n = int(cnt.value)
    rows = np.ctypeslib.as_array(ptr, shape=(max(n, 1) * 5,))[: n * 5].copy().reshape(n, 5)
    _load().orc_free(ptr)
    return rows, int(quads.value)


def assign(width: int, height: int, sorted_xy: np.ndarray, rows: np.ndarray,
           method: str = "raster") -> np.ndarray:
    """triangulation.py:193-204: per-pixel max tid over containing triangles."""
    s = np.ascontiguousarray(sorted_xy, dtype=np.int32)
    r = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1, 5)
    out = np.empty((height, width), dtype=np.int32)
    fn = _load().orc_assign_brute if method == "brute" else _load().orc_assign_raster
    fn(int(width), int(height), _p(s), _p(r), r.shape[0], _p(out))
    return out


def resolve_padding(border_padding) -> int:
    """triangulation.py:75-78 (None -> default, negative clamps to 0), with the
    CUDA binding's -1 sentinel also meaning the default
    (src/cuda/bindings/binding_helpers.hpp:18-20,49-53)."""
    if border_padding is None or border_padding == -1:
        return DEFAULT_BORDER_PADDING
    return max(0, int(border_padding))


def grid_triangulation(voronoi_grid: np.ndarray, seed_positions, border_padding=None,
                       as_arrays: bool = False, assign_method: str = "raster",
                       return_rows: bool = False):
    """reference.GridTriangulation().compute (triangulation.py:26-219)."""
    vg = np.asarray(voronoi_grid, dtype=np.int32)
    H, W = vg.shape[:2]
    sorted_xy = sort_seeds(W, H, seed_positions)
    P = resolve_padding(border_padding)
    if P > 0:
        pad = voronoi_sorted(W + 2 * P, H + 2 * P, sorted_xy + P)
        n_grid = pad[:, :, 0]
    else:
        n_grid = vg[:, :, 0]
    rows, _ = detect(n_grid, sorted_xy, P)
    tid = assign(W, H, sorted_xy, rows, assign_method)
    tgrid = np.empty((H, W, 3), dtype=np.int32)
    tgrid[:, :, 0] = vg[:, :, 0]
    tgrid[:, :, 1] = vg[:, :, 1]
    tgrid[:, :, 2] = tid
    if return_rows:
        return rows, tgrid
    if as_arrays:
        return np.ascontiguousarray(rows[:, 2:5]), tgrid
    return {i: tuple(int(v) for v in rows[i]) for i in range(rows.shape[0])}, tgrid


def triangulate(width: int, height: int, seeds, border_padding=None,
                as_arrays: bool = False, assign_method: str = "raster",
                return_rows: bool = False):
    """Voronoi().compute followed by GridTriangulation().compute."""
    vg = voronoi(width, height, seeds)
    return grid_triangulation(vg, seeds, border_padding, as_arrays, assign_method,
                              return_rows)
