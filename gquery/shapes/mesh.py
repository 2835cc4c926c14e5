"""gquery.shapes.mesh (gquery/shapes/mesh.py:11-139): Mesh with closest_point_bvh, intersect_bvh,
closest_silhouette_snch, the brute-force *_baseline methods and configure_normal."""
from gquery_amd.shapes.mesh import Mesh  # noqa: F401
