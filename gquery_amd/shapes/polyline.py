"""Polyline (gquery/shapes/polyline.py:11-227): a 2D line-segment set with BVH and SNCH queries.

Polyline(vertices, indices) builds the SAH BVH and the SNCH on the CPU (polyline.py:20-22) and
keeps both on the current HIP device.  The query methods return the reference's records as torch
tensors in input order.
"""
from __future__ import annotations

import numpy as np

from .. import ext
from ..scene import INF, Scene


class Polyline:
    def __init__(self, vertices, indices, device=None, max_prims_in_node: int = 10, leaf_prims: int = 0,
                 packet_budget: int = 0, traversal: int = 0, wave_front: int = 0):
        # One SNCH build serves both queries: its BVH part is the same tree BVH() builds.
        self.snch = ext.SNCH(vertices, indices, max_prims_in_node)
        self.bvh = self.snch
        self.vertices = self.snch._vertices
        self.indices = self.snch._indices
        self.scene = Scene(self.snch, device, leaf_prims, packet_budget, traversal, wave_front)

    def intersect_bvh(self, p, v, r_max=INF, method: str = "engine"):
        """BVH.intersect (bvh.py:200-276) -> Intersection."""
        return self.scene.intersect(p, v, r_max, method)

    def closest_point_bvh(self, p, method: str = "engine"):
        """BVH.closest_point (bvh.py:127-198) -> ClosestPointRecord."""
        return self.scene.closest_point(p, method)

    def closest_silhouette_snch(self, x, r_max=INF, method: str = "engine"):
        """SNCH.closest_silhouette (snch.py:106-164) -> ClosestSilhouettePointRecord."""
        return self.scene.closest_silhouette(x, r_max, method)
