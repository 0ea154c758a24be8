"""Read-only views onto a session's device arrays (``Delaunay.finalise_device``).

Mirrors the reference's ``PyDeviceArrayView``
(src/cuda/bindings/py_device_array_view.hpp:1-83): a view exposes one of the
session's per-pixel device arrays through ``__cuda_array_interface__``
(version 3, read-only, shape ``(H*W,)``) and ``to_host()``, keeps the owning
``Delaunay`` object alive, and raises ``RuntimeError`` once the session has
been mutated after the view was made (a generation counter snapshot).

On ROCm the same memory is also offered through DLPack (``__dlpack__`` /
``__dlpack_device__``, device type kDLROCM), which is how PyTorch-ROCm and
other ROCm libraries import device arrays (``torch.from_dlpack(view)``).  A
consumer's tensor keeps the session alive; it aliases the session's buffers,
which the next ``finalise_device`` rewrites.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_KDL_ROCM = 10
_DTYPES = {"<i4": (np.int32, 0, 32), "|u1": (np.uint8, 1, 8)}   # numpy, DLPack code, bits


class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", _DELETER)]

# Exported tensors still referenced by a consumer: manager_ctx -> what they keep alive.
_live: dict = {}
_NAME = b"dltensor"
# A private handle on the interpreter's C API, so that the argument types set
# here do not affect other users of ctypes.pythonapi.
_py = ctypes.PyDLL(None)
_CAPSULE_DTOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
_py.PyCapsule_New.restype = ctypes.py_object
_py.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _CAPSULE_DTOR]
_py.PyCapsule_IsValid.restype = ctypes.c_int
_py.PyCapsule_IsValid.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
_py.PyCapsule_GetPointer.restype = ctypes.c_void_p
_py.PyCapsule_GetPointer.argtypes = [ctypes.c_void_p, ctypes.c_char_p]


@_DELETER
def _release(managed):
    _live.pop(int(managed.contents.manager_ctx or 0), None)


@_CAPSULE_DTOR
def _capsule_gone(capsule):
    # Never consumed (a consumer renames the capsule): release it here.
    if _py.PyCapsule_IsValid(capsule, _NAME):
        ptr = _py.PyCapsule_GetPointer(capsule, _NAME)
        _release(ctypes.cast(ptr, ctypes.POINTER(_DLManagedTensor)))


class DeviceArrayView:
    """One of a session's per-pixel device arrays (H*W elements)."""

    def __init__(self, owner, ptr: int, count: int, typestr: str, generation: int, device: int):
        self._owner = owner          # keeps the session (and its device buffers) alive
        self._ptr = int(ptr)
        self._count = int(count)
        self._typestr = typestr
        self._generation = generation
        self._device = int(device)

    def _check_fresh(self):
        if self._owner._generation != self._generation:
            raise RuntimeError("device view is stale: the mesh was mutated after "
                               "finalise_device() produced this view")

    @property
    def __cuda_array_interface__(self) -> dict:
        self._check_fresh()
        return {"shape": (self._count,), "typestr": self._typestr, "data": (self._ptr, True),
                "strides": None, "version": 3}

    def to_host(self) -> np.ndarray:
        """A host copy of the view's content."""
        self._check_fresh()
        out = np.empty(self._count, dtype=_DTYPES[self._typestr][0])
        if self._count:
            _lib.check(_lib.load().dl_device_to_host(out.ctypes.data_as(ctypes.c_void_p),
                                                      ctypes.c_void_p(self._ptr), out.nbytes))
        return out

    def __dlpack_device__(self):
        return (_KDL_ROCM, self._device)

    def __dlpack__(self, stream=None):
        """A DLPack capsule over the view.  The data is complete when the view
        is made (finalise_device waits for it), so no stream ordering is needed."""
        self._check_fresh()
        _, code, bits = _DTYPES[self._typestr]
        shape = (ctypes.c_int64 * 1)(self._count)
        m = _DLManagedTensor()
        m.dl_tensor.data = self._ptr
        m.dl_tensor.device = _DLDevice(_KDL_ROCM, self._device)
        m.dl_tensor.ndim = 1
        m.dl_tensor.dtype = _DLDataType(code, bits, 1)
        m.dl_tensor.shape = shape
        m.dl_tensor.strides = None
        m.dl_tensor.byte_offset = 0
        key = id(m)
        m.manager_ctx = key
        m.deleter = _release
        _live[key] = (m, shape, self)
        return _py.PyCapsule_New(ctypes.addressof(m), _NAME, _capsule_gone)

    def __len__(self):
        return self._count

    def __repr__(self):
        return f"DeviceArrayView(count={self._count}, typestr={self._typestr!r})"
