"""Mesh (gquery/shapes/mesh.py:11-139): a triangle mesh with BVH and SNCH queries.

Mesh(vertices, indices) builds the SAH BVH and the SNCH (mesh.py:23-24; the reference builds the
same BVH twice, here one SNCH build provides both) and keeps them on the current HIP device.
"""
from __future__ import annotations

from .. import ext
from ..scene import INF, Scene


class Mesh:
    def __init__(self, vertices, indices, device=None, max_prims_in_node: int = 10, leaf_prims: int = 0,
                 packet_budget: int = 0, traversal: int = 0, wave_front: int = 0):
        self.snch = ext.SNCH3D(vertices, indices, max_prims_in_node)
        self.bvh = self.snch
        self.vertices = self.snch._vertices
        self.indices = self.snch._indices
        self.scene = Scene(self.snch, device, leaf_prims, packet_budget, traversal, wave_front)

    def intersect_bvh(self, p, v, r_max=INF, method: str = "engine"):
        """BVH3D.intersect (bvh3d.py:176-239) -> Intersection3D."""
        return self.scene.intersect(p, v, r_max, method)

    def closest_point_bvh(self, x, method: str = "engine"):
        """BVH3D.closest_point (bvh3d.py:108-174) -> ClosestPointRecord3D."""
        return self.scene.closest_point(x, method)

    def closest_silhouette_snch(self, p, r_max=INF, method: str = "engine"):
        """SNCH3D.closest_silhouette (snch3d.py:143-203) -> ClosestSilhouettePointRecord3D."""
        return self.scene.closest_silhouette(p, r_max, method)
