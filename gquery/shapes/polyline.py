"""gquery.shapes.polyline (gquery/shapes/polyline.py:11-227): Polyline with closest_point_bvh,
intersect_bvh, closest_silhouette_snch, the brute-force *_baseline methods and
configure_silhouette."""
from gquery_amd.shapes.polyline import Polyline  # noqa: F401
