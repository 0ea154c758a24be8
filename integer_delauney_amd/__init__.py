"""MI355X (gfx950) backend of the `delauney` integer Voronoi / Delaunay pipeline.

Drop-in for the reference package `delauney` (src/delauney/__init__.py) and
its extension `_delauney_cuda`: the same names, signatures and array layouts,
computing the results of the reference's NumPy path (src/delauney/reference)
bit for bit on hand-written HIP kernels (libdelauney_hip.so, C ABI in
include/delauney_hip.h).

Importing the package needs no GPU; the first computation loads the library
and creates a HIP context, and raises if either is unavailable.
"""
from .api import (AUTO_BORDER_PADDING, DEFAULT_BORDER_PADDING, GridTriangulation, Voronoi,
                  max_boundary_spacing, triangulate, triangulate_batch)
from .delaunay import Delaunay
from .device_views import DeviceArrayView

__all__ = ["Voronoi", "GridTriangulation", "Delaunay", "DeviceArrayView", "triangulate",
           "triangulate_batch", "DEFAULT_BORDER_PADDING", "max_boundary_spacing"]
__version__ = "0.2.0"
