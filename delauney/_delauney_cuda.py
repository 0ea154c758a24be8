"""Backend module of the `delauney` package: the classes and functions of the
reference's extension module ``_delauney_cuda`` (src/cuda/bindings.cpp:17-227),
served by the HIP library through ``integer_delauney_amd``.

Importing it needs no GPU; the first computation loads libdelauney_hip.so and
raises RuntimeError if the library or a GPU is missing.
"""
from integer_delauney_amd import (Delaunay, DeviceArrayView, GridTriangulation, Voronoi,
                                  triangulate)

__all__ = ["Voronoi", "GridTriangulation", "Delaunay", "DeviceArrayView", "triangulate"]
