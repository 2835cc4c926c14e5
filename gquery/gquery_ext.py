"""gquery.gquery_ext (src/gquery.cpp:180-250): the CPU builders BVH / BVH3D / SNCH / SNCH3D with
node_data / primitive_data / silhouette_data / raw_silhouette_data, and
build_flat_silhouette_edges."""
from gquery_amd.ext import BVH, BVH3D, SNCH, SNCH3D, build_flat_silhouette_edges  # noqa: F401
