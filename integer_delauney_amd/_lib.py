"""ctypes binding of libdelauney_hip.so (declarations: include/delauney_hip.h).

This is the only place that touches the C ABI.  It loads the in-tree library
built by ``__graft_entry__.build()`` (or ``make -C integer_delauney_amd/csrc``),
maps status codes to the exceptions the reference raises (ValueError for bad
input, RuntimeError for device failures -- src/cuda/cuda_check.cuh:12-44,
binding_helpers.hpp:23-46) and keeps one library context per thread.

There is no fallback: if the library or a GPU is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdelauney_hip.so")

DL_OK, DL_EVALUE, DL_ERUNTIME = 0, 1, 2

_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_U8P = ctypes.POINTER(ctypes.c_uint8)


class DLImage(ctypes.Structure):
    """struct dl_image (include/delauney_hip.h)."""
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("border_padding", ctypes.c_int), ("host", ctypes.c_int), ("n_seeds", ctypes.c_int64),
                ("seeds_xy", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("session", ctypes.c_void_p), ("n_tri", ctypes.c_int64),
                ("tri_offset", ctypes.c_int64)]

_lib = None
_lib_lock = threading.Lock()
_tls = threading.local()

# name -> (restype, argtypes)
_SIGNATURES = {
    "dl_version": (ctypes.c_char_p, []),
    "dl_last_error": (ctypes.c_char_p, []),
    "dl_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_VP)]),
    "dl_context_destroy": (None, [_VP]),
    "dl_voronoi": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64, _I32P]),
    "dl_grid_triangulation": (ctypes.c_int, [_VP, _I32P, ctypes.c_int, ctypes.c_int, _I32P,
                                             ctypes.c_int64, ctypes.c_int, _I32P, _I64P]),
    "dl_triangulate": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64,
                                      ctypes.c_int, _I32P, _I64P]),
    "dl_copy_triangles": (ctypes.c_int, [_VP, _I32P]),
    "dl_copy_padded_voronoi": (ctypes.c_int, [_VP, _I32P, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)]),
    "dl_get_timings": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double)]),
    "dl_set_timing": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_get_fill_stats": (ctypes.c_int, [_VP, _I64P, _I64P, _I64P, ctypes.POINTER(ctypes.c_int),
                                         _I64P]),
    "dl_set_bfs_config": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "dl_get_bfs_config": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "dl_triangulate_device": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP,
                                             ctypes.c_int64, ctypes.c_int, _VP, _VP, _I64P]),
    "dl_triangulate_device_noalloc": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP,
                                                     ctypes.c_int64, ctypes.c_int, _VP, _VP,
                                                     ctypes.c_int64, _VP, _I64P]),
    "dl_device_triangles": (_VP, [_VP]),
    "dl_voronoi_simple": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64,
                                         _I32P, _I64P]),
    "dl_set_force_unpadded": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_set_replay_plane_reuse": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_get_record_stats": (ctypes.c_int, [_VP, _I64P, _I64P, _I64P]),
    "dl_set_kernel_profiling": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_get_kernel_stats": (ctypes.c_int, [_VP, _I64P, ctypes.POINTER(ctypes.c_double)]),
    "dl_reset_kernel_stats": (ctypes.c_int, [_VP]),
    "dl_triangulate_batch": (ctypes.c_int, [_VP, ctypes.c_int, _I32P, _I32P, _I64P, _I32P, _I32P,
                                            _I32P, _I64P]),
    "dl_triangulate_batch_device": (ctypes.c_int, [_VP, ctypes.c_int, _I32P, _I32P, _I64P, _VP,
                                                   _I32P, _VP, _VP, _I64P]),
    "dl_host_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_VP)]),
    "dl_host_free": (None, [_VP]),
    "dl_run_batch": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP]),
    "dl_context_device_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
    "dl_context_trim": (ctypes.c_int, [_VP]),
    "dl_session_device_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
    "dl_device_memory": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "dl_session_create": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                         ctypes.c_int, ctypes.POINTER(_VP)]),
    "dl_session_destroy": (None, [_VP]),
    "dl_session_insert": (ctypes.c_int, [_VP, _VP, _I32P, _F32P, ctypes.c_int64]),
    "dl_session_remove": (ctypes.c_int, [_VP, _VP, _I32P, ctypes.c_int64]),
    "dl_session_update": (ctypes.c_int, [_VP, _VP, _I32P, ctypes.c_int64, _I32P, _F32P,
                                         ctypes.c_int64]),
    "dl_session_seed_count": (ctypes.c_int, [_VP, _I64P]),
    "dl_session_finalise": (ctypes.c_int, [_VP, _VP, _I64P]),
    "dl_session_set_fill_record": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_session_fill_record_stats": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), _I64P, _I64P]),
    "dl_session_copy_result": (ctypes.c_int, [_VP, _I32P, _I32P]),
    "dl_session_copy_seeds": (ctypes.c_int, [_VP, _I32P]),
    "dl_session_device_grid": (_VP, [_VP]),
    "dl_set_autotune": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_set_fill_prediction": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_set_graphs": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_get_graph_stats": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int64)]),
    "dl_autotune_reset": (ctypes.c_int, []),
    "dl_device_to_host": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
    "dl_session_has_values": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    "dl_session_copy_values": (ctypes.c_int, [_VP, _F32P]),
    "dl_session_locate": (ctypes.c_int, [_VP, _VP, _I32P, ctypes.c_int64, _I32P]),
    "dl_session_in_circumsphere": (ctypes.c_int, [_VP, _VP, _I32P, _F64P, ctypes.c_int64, _U8P,
                                                  _I32P]),
    "dl_session_edges": (ctypes.c_int, [_VP, _VP, _I64P]),
    "dl_session_copy_edges": (ctypes.c_int, [_VP, _I32P]),
    "dl_session_edge_scores": (ctypes.c_int, [_VP, _VP, ctypes.c_double, _F32P]),
    "dl_session_select_midpoints": (ctypes.c_int, [_VP, _VP, ctypes.c_double, ctypes.c_int64,
                                                   ctypes.c_float, _I32P, ctypes.c_int64, _I64P]),
    "dl_session_finalise_device": (ctypes.c_int, [_VP, _VP, _I64P, ctypes.POINTER(_VP),
                                                  ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "dl_set_fill_list_threshold": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_set_tiny_capacities": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_set_fill_min_encoding": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_set_fill_aligned_watch": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dl_get_fill_aligned_watch": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    "dl_debug_fill_watch": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64,
                                           ctypes.c_int, _I32P, ctypes.POINTER(ctypes.c_int)]),
}


def exported_symbols():
    """Names the C ABI declares (include/delauney_hip.h)."""
    return sorted(_SIGNATURES)


def load():
    """Load libdelauney_hip.so (no GPU needed to load it)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `make -C integer_delauney_amd/csrc`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc == DL_OK:
        return
    msg = load().dl_last_error().decode("utf-8", "replace")
    if rc == DL_EVALUE:
        raise ValueError(msg)
    raise RuntimeError(msg)


class Context:
    """One library context (HIP stream + device workspaces)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        self.device = int(device)
        h = _VP()
        check(self._lib.dl_context_create(int(device), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._lib.dl_context_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def context() -> Context:
    """The calling thread's default context on device 0."""
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(int(os.environ.get("DELAUNEY_HIP_DEVICE", "0")))
        _tls.ctx = ctx
    return ctx


def i32p(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


def f32p(a: np.ndarray):
    return a.ctypes.data_as(_F32P)


def f64p(a: np.ndarray):
    return a.ctypes.data_as(_F64P)


def u8p(a: np.ndarray):
    return a.ctypes.data_as(_U8P)


class _PinnedBlock:
    __slots__ = ("pool", "ptr", "size", "__weakref__")

    def __init__(self, pool, ptr, size):
        self.pool, self.ptr, self.size = pool, ptr, size

    def __del__(self):
        try:
            self.pool._give_back(self)
        except Exception:   # pragma: no cover - interpreter shutdown
            pass


class PinnedPool:
    """Page-locked host buffers by power-of-two size class, reused: an array
    from :meth:`array` returns its block to the pool when the array (and every
    view of it) is gone.  ``limit`` bytes of idle blocks are kept."""

    def __init__(self, limit: int = 8 << 30):
        self._free = {}
        self._idle = 0
        self._limit = int(limit)
        self._lock = threading.Lock()

    def _give_back(self, blk):
        with self._lock:
            if self._idle + blk.size <= self._limit:
                self._free.setdefault(blk.size, []).append(blk.ptr)
                self._idle += blk.size
                return
        load().dl_host_free(blk.ptr)

    def array(self, shape, dtype=np.int32) -> np.ndarray:
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        size = 1 << max(12, (max(nbytes, 1) - 1).bit_length())
        with self._lock:
            lst = self._free.get(size)
            ptr = lst.pop() if lst else None
            if ptr is not None:
                self._idle -= size
        if ptr is None:
            p = _VP()
            check(load().dl_host_alloc(size, ctypes.byref(p)))
            ptr = p.value
        blk = _PinnedBlock(self, ptr, size)
        buf = (ctypes.c_char * nbytes).from_address(ptr) if nbytes else (ctypes.c_char * 1)()
        buf._dl_block = blk   # the block lives as long as the buffer does
        return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def clear(self):
        with self._lock:
            for size, lst in self._free.items():
                for ptr in lst:
                    load().dl_host_free(ptr)
            self._free.clear()
            self._idle = 0
