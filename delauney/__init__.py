"""The `delauney` package on the MI355X (gfx950) backend.

Same names as the reference façade (src/delauney/__init__.py:1-32):
``Voronoi``, ``GridTriangulation``, ``Delaunay`` and ``triangulate`` are
re-exported lazily from the backend module ``delauney._delauney_cuda`` on
first attribute access, and ``DEFAULT_BORDER_PADDING`` / ``max_boundary_spacing``
are defined here.  The backend module keeps the reference's module name so
that code importing ``from delauney import _delauney_cuda`` (as the
reference's tests do) finds it; it is the HIP library behind the same
Python surface (``integer_delauney_amd``), not a CUDA module.

The reference's NumPy path ``delauney.reference`` is not part of this
package.
"""
from integer_delauney_amd.api import DEFAULT_BORDER_PADDING, max_boundary_spacing

_CUDA_EXPORTS = {"Voronoi", "GridTriangulation", "Delaunay", "triangulate"}

__all__ = sorted(_CUDA_EXPORTS) + ["DEFAULT_BORDER_PADDING", "max_boundary_spacing"]


def __getattr__(name):
    if name in _CUDA_EXPORTS:
        from delauney import _delauney_cuda as _cu
        globals().update({k: getattr(_cu, k) for k in _CUDA_EXPORTS})
        return globals()[name]
    raise AttributeError(f"module 'delauney' has no attribute {name!r}")
