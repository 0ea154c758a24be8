"""The Python surface against the reference's (no GPU needed).

The table below restates the reference extension's names, argument names and
defaults (src/cuda/bindings.cpp:17-227) and façade (src/delauney/__init__.py:
1-32); the tests check the `delauney` package of this repository against it,
so that code written for the reference imports and calls it unchanged.
"""
import inspect

import pytest

import delauney
from delauney import _delauney_cuda as cu

AUTO = -1   # AUTO_BORDER_PADDING, binding_helpers.hpp:18-20

# (qualified name, [(argument, default or REQUIRED)])  -- bindings.cpp line
REQUIRED = object()
METHODS = [
    ("triangulate", [("width", REQUIRED), ("height", REQUIRED), ("seeds", REQUIRED),
                     ("border_padding", AUTO), ("as_arrays", False)]),                 # :29-41
    ("Voronoi.compute", [("width", REQUIRED), ("height", REQUIRED), ("seeds", REQUIRED)]),  # :21-27
    ("GridTriangulation.compute", [("voronoi_grid", REQUIRED), ("seed_positions", REQUIRED),
                                   ("border_padding", AUTO), ("as_arrays", False)]),   # :43-62
    ("GridTriangulation.compute_timed", [("voronoi_grid", REQUIRED),
                                         ("seed_positions", REQUIRED),
                                         ("border_padding", AUTO)]),                   # :63-68
    ("GridTriangulation.compute_debug", [("voronoi_grid", REQUIRED),
                                         ("seed_positions", REQUIRED),
                                         ("border_padding", AUTO)]),                   # :69-74
    ("Delaunay.__init__", [("width", REQUIRED), ("height", REQUIRED), ("max_seeds", REQUIRED),
                           ("border_padding", AUTO)]),                                 # :86-102
    ("Delaunay.insert", [("seeds", REQUIRED), ("as_arrays", False)]),                  # :103-112
    ("Delaunay.insert_timed", [("seeds", REQUIRED), ("as_arrays", False)]),            # :113-116
    ("Delaunay.insert_deferred", [("seeds", REQUIRED), ("values", None)]),             # :117-128
    ("Delaunay.insert_deferred_timed", [("seeds", REQUIRED)]),                         # :129-132
    ("Delaunay.finalise", [("as_arrays", False)]),                                     # :133-140
    ("Delaunay.finalise_timed", [("as_arrays", False)]),                               # :141-143
    ("Delaunay.finalise_device", []),                                                  # :144-152
    ("Delaunay.get_triangles", []),                                                    # :153-159
    ("Delaunay.edge_scores", [("min_length", 0.0)]),                                   # :160-166
    ("Delaunay.select_midpoints", [("min_length", REQUIRED), ("count", REQUIRED),
                                   ("threshold", 0.0)]),                               # :167-178
    ("Delaunay.in_circumsphere", [("points", REQUIRED), ("t", REQUIRED)]),             # :179-190
    ("Delaunay.locate", [("points", REQUIRED)]),                                       # :191-200
    ("Delaunay.get_seeds", []),                                                        # :201-202
    ("Delaunay.get_values", []),                                                       # :203-204
    ("Delaunay.get_edges", []),                                                        # :205-214
    ("Delaunay.get_voronoi_grid", []),                                                 # :218-219
    ("DeviceArrayView.to_host", []),                                                   # :76-85
]
PROPERTIES = [("Delaunay", p) for p in ("sorted_rank", "seed_count", "width", "height",
                                        "border_padding", "has_pending")] + \
             [("DeviceArrayView", "__cuda_array_interface__")]


def _resolve(qual):
    obj = cu
    for part in qual.split("."):
        obj = getattr(obj, part)
    return obj


def test_facade_names():
    assert set(delauney.__all__) >= {"Voronoi", "GridTriangulation", "Delaunay", "triangulate",
                                     "DEFAULT_BORDER_PADDING", "max_boundary_spacing"}
    for name in ("Voronoi", "GridTriangulation", "Delaunay", "triangulate"):
        assert getattr(delauney, name) is getattr(cu, name)
    for name in ("Voronoi", "GridTriangulation", "Delaunay", "DeviceArrayView", "triangulate"):
        assert hasattr(cu, name)
    with pytest.raises(AttributeError):
        delauney.not_a_name  # noqa: B018


def test_padding_constant_and_bound():
    # src/delauney/__init__.py:16-24; BORDER_PADDING_BOUND.md: P >= s^2 / 8
    assert delauney.DEFAULT_BORDER_PADDING == 16
    assert delauney.max_boundary_spacing() == 11
    assert delauney.max_boundary_spacing(0) == 0 and delauney.max_boundary_spacing(-3) == 0
    for p in range(1, 200):
        s = delauney.max_boundary_spacing(p)
        assert s * s <= 8 * p < (s + 1) * (s + 1)


@pytest.mark.parametrize("qual,args", METHODS, ids=[m[0] for m in METHODS])
def test_signature(qual, args):
    sig = inspect.signature(_resolve(qual))
    params = [p for p in sig.parameters.values() if p.name != "self"]
    assert [p.name for p in params] == [a for a, _ in args], qual
    for p, (_, default) in zip(params, args):
        if default is REQUIRED:
            assert p.default is inspect.Parameter.empty, (qual, p.name)
        else:
            assert p.default == default, (qual, p.name, p.default)


@pytest.mark.parametrize("owner,prop", PROPERTIES)
def test_read_only_properties(owner, prop):
    attr = inspect.getattr_static(getattr(cu, owner), prop)
    assert isinstance(attr, property) and attr.fset is None
