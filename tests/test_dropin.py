"""The drop-in surface: the reference's module paths (gquery.shapes.mesh / polyline / primitive,
gquery.core.fwd, gquery.gquery_ext) and its calls from examples/example_2d.py and example_3d.py,
with components-first (DrJit-style) inputs, including a one-triangle mesh whose (3, 3) vertex
array is ambiguous without the type."""
import math

import numpy as np
import pytest
import torch

from gquery.core.fwd import Array2, Array2i, Array3, Array3i, Float
from gquery.gquery_ext import SNCH3D
from gquery_amd.scene import as_points
from oracle import oracle as O

# examples/example_3d.py:6-21 and example_2d.py:6-21
TET_V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1.0, 0.0], [0.5, 0.5, 1.0]])
TET_F = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]])
PENT_V = np.array([[0.0, 0.0], [1.0, 0.0], [1.5, 1.0], [0.5, 1.5], [-0.5, 1.0]])
PENT_E = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 0]])
# one triangle: its vertex and index arrays are both (3, 3)
TRI_V = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.5], [0.0, 3.0, 1.0]], np.float32)
TRI_F = np.array([[0, 1, 2]])


def test_module_paths():
    import gquery.shapes.mesh as m
    import gquery.shapes.polyline as p
    import gquery.shapes.primitive as r
    import gquery_amd.shapes.mesh as am
    import gquery_amd.shapes.polyline as ap
    import gquery_amd.shapes.primitive as ar
    assert m.Mesh is am.Mesh and p.Polyline is ap.Polyline
    for name in ("ClosestPointRecord", "ClosestPointRecord3D", "Intersection", "Intersection3D",
                 "ClosestSilhouettePointRecord", "ClosestSilhouettePointRecord3D"):
        assert getattr(r, name) is getattr(ar, name)


def test_components_first_types():
    a = Array3(TET_V.T)
    assert len(a) == 4 and a.value.shape == (3, 4) and a.value.dtype == np.float32
    np.testing.assert_array_equal(a.rows(), TET_V.astype(np.float32))
    b = Array3(TET_V[:, 0], TET_V[:, 1], TET_V[:, 2])
    np.testing.assert_array_equal(b.rows(), a.rows())
    assert Array3i(TET_F.T).value.dtype == np.int32
    assert Float(float("inf")) == float("inf") and isinstance(Float(1), float)
    t = Array2(torch.tensor(PENT_V.T, dtype=torch.float64))
    assert isinstance(t.value, torch.Tensor) and t.rows().shape == (5, 2)
    with pytest.raises(ValueError):
        Array3(np.zeros((4, 2)))


def test_square_inputs_are_unambiguous_with_the_type():
    cpu = torch.device("cpu")
    # tagged components first: transposed
    got = as_points(Array3(TRI_V.T), 3, cpu).numpy()
    np.testing.assert_array_equal(got, TRI_V)
    # untagged square input is rows (documented), untagged (3, N != 3) is components first
    np.testing.assert_array_equal(as_points(TRI_V, 3, cpu).numpy(), TRI_V)
    np.testing.assert_array_equal(as_points(TET_V.T, 3, cpu).numpy(), TET_V.astype(np.float32))
    with pytest.raises(ValueError):
        as_points(Array2(PENT_V.T), 3, cpu)


def test_one_triangle_mesh_builds_from_components_first_arrays():
    s = SNCH3D(Array3(TRI_V.T), Array3i(TRI_F.T))
    prim = s.primitive_data().reshape(-1, 10)
    np.testing.assert_array_equal(prim[0, :9], TRI_V.reshape(-1))   # a, b, c as given, not transposed
    # the same through plain rows
    s2 = SNCH3D(TRI_V, TRI_F)
    np.testing.assert_array_equal(s2.primitive_data(), s.primitive_data())


@pytest.mark.gpu
def test_example_3d(gpu):
    from gquery.shapes.mesh import Mesh
    mesh = Mesh(Array3(TET_V.T), Array3i(TET_F.T))
    query_point = Array3(np.array([[0.6, 0.6, 0.6]]).T)
    closest = mesh.closest_point_bvh(query_point)
    assert closest.valid[0].item() and closest.prim_id[0].item() == 2
    np.testing.assert_allclose(closest.d[0].item(), 0.2 / math.sqrt(21), rtol=2e-6)
    np.testing.assert_allclose(closest.p[0].cpu().numpy(), np.array([59, 61, 62]) / 105, atol=1e-6)
    direction = Array3(np.array([[1.0, 1.0, 1.0]]).T)
    r_max = Float(float("inf"))
    intersection = mesh.intersect_bvh(query_point, direction, r_max)
    assert not intersection.valid[0].item() and intersection.d[0].item() == 0.0
    query_point = Array3(np.array([[1.5, 0.5, 0.5]]).T)
    silhouette = mesh.closest_silhouette_snch(query_point)
    ref = O.silhouette(3, mesh.snch.node_data(), mesh.snch.silhouette_data(), np.array([[1.5, 0.5, 0.5]], np.float32))
    assert bool(silhouette.valid[0].item()) == bool(ref["valid"][0])
    np.testing.assert_allclose(silhouette.d[0].item(), ref["d"][0], rtol=1e-6)


@pytest.mark.gpu
def test_example_2d(gpu):
    from gquery.shapes.polyline import Polyline
    polyline = Polyline(Array2(PENT_V.T), Array2i(PENT_E.T))
    query_point = Array2(np.array([[0.8, 0.8]]).T)
    closest = polyline.closest_point_bvh(query_point)
    assert closest.prim_id[0].item() == 2
    np.testing.assert_allclose(closest.p[0].cpu().numpy(), [1.02, 1.24], atol=1e-6)
    np.testing.assert_allclose(closest.t[0].item(), 0.48, atol=1e-6)
    direction = Array2(np.array([[1.0, 1.0]]).T)
    intersection = polyline.intersect_bvh(query_point, direction, Float(float("inf")))
    assert intersection.valid[0].item() and intersection.prim_id[0].item() == 2
    np.testing.assert_allclose(intersection.d[0].item(), 0.7 - 1 / 3, rtol=1e-6)
    silhouette = polyline.closest_silhouette_snch(Array2(np.array([[2.0, 0.5]]).T))
    assert silhouette.valid[0].item() and silhouette.prim_id[0].item() == 2
    np.testing.assert_allclose(silhouette.d[0].item(), math.sqrt(0.5), rtol=1e-6)


@pytest.mark.gpu
def test_one_triangle_mesh_queries(gpu):
    from gquery.shapes.mesh import Mesh
    mesh = Mesh(Array3(TRI_V.T), Array3i(TRI_F.T))
    x = np.array([[0.5, 0.5, 3.0], [-1.0, -1.0, 0.0]], np.float32)
    r = mesh.closest_point_bvh(Array3(x.T))
    b = O.bf_closest_point(3, TRI_V, TRI_F, x)
    np.testing.assert_allclose(r.d.cpu().numpy(), b["d"], rtol=1e-6)
    np.testing.assert_allclose(r.p.cpu().numpy(), b["p"], atol=1e-6)
