"""CPU-side checks of the C ABI: the library loads without a GPU, exports
every entry point include/delauney_hip.h declares, and the ctypes layer binds
exactly those."""
import ctypes
import os
import re

import numpy as np
import pytest

from integer_delauney_amd import _lib
from integer_delauney_amd.api import as_seed_array, max_boundary_spacing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "delauney_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_the_header():
    assert _lib.exported_symbols() == _declared()


def test_version_and_error_string_without_gpu():
    lib = _lib.load()
    assert b"gfx950" in lib.dl_version()
    assert isinstance(lib.dl_last_error(), bytes)


def test_seed_parsing():
    a = as_seed_array([(1, 2), (3, 4)])
    assert a.dtype == np.int32 and a.shape == (2, 2) and a.flags.c_contiguous
    assert as_seed_array(np.array([[5, 6]], dtype=np.int64)).dtype == np.int32
    for bad in ([], None, [(1, 2, 3)], np.zeros((0, 2)), [(1.5, 2)], [(2**40, 1)]):
        with pytest.raises(ValueError):
            as_seed_array(bad)


def test_max_boundary_spacing_inverts_the_bound():
    for p in (1, 2, 8, 16, 32, 128):
        s = max_boundary_spacing(p)
        assert s * s / 8 <= p < (s + 1) ** 2 / 8
    assert max_boundary_spacing(0) == 0 and max_boundary_spacing(-1) == 0
    assert max_boundary_spacing(16) >= 8
