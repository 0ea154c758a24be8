"""The triangle-map builder (a CPython extension): the same dictionary as the
plain Python construction {tid: tuple(row)}, Python ints throughout."""
import numpy as np
import pytest

from integer_delauney_amd.api import rows_to_map


def _plain(rows):
    return {i: tuple(int(v) for v in r) for i, r in enumerate(rows.tolist())}


@pytest.mark.parametrize("n,lo,hi", [(0, 0, 1), (1, -16, 20), (1000, -16, 600), (5000, 0, 1 << 30),
                                     (300, -(1 << 31), (1 << 31) - 1)])
def test_rows_to_map_equals_plain_construction(n, lo, hi):
    rows = np.random.default_rng(n).integers(lo, hi, size=(n, 5), dtype=np.int64).astype(np.int32)
    m = rows_to_map(rows)
    assert m == _plain(rows)
    assert list(m) == list(range(n))
    assert all(type(v) is int for t in m.values() for v in t)


def test_rows_to_map_accepts_views_and_rejects_other_layouts():
    rows = np.arange(50, dtype=np.int32).reshape(10, 5)
    assert rows_to_map(rows[::2]) == _plain(rows[::2])       # copied to contiguous
    assert rows_to_map(rows.astype(np.int64)) == _plain(rows)  # converted to int32
    from integer_delauney_amd import _rowsmap
    with pytest.raises(ValueError):
        _rowsmap.rows_to_map(np.zeros((3, 4), np.int32))
    with pytest.raises(ValueError):
        _rowsmap.rows_to_map(np.zeros((3, 5), np.float32))
