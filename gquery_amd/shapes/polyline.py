"""Polyline (gquery/shapes/polyline.py:11-227): a 2D line-segment set with BVH and SNCH queries.

Polyline(vertices, indices) builds the SAH BVH and the SNCH on the CPU (polyline.py:20-22) and
keeps both on the current HIP device.  The query methods return the reference's records as torch
tensors in input order.
"""
from __future__ import annotations

import numpy as np

from .. import ext
from ..scene import INF, BruteForce, Scene


class Polyline:
    def __init__(self, vertices, indices, device=None, max_prims_in_node: int = 10, leaf_prims: int = 0,
                 packet_budget: int = 0, traversal: int = 0, wave_front: int = 0, sil_pairs: int = 0,
                 structure=None, ray_leaf_prims: int = 0):
        # structure: an ext.SNCH already built or loaded from a scene file (then vertices and
        # indices may be None)
        self.snch = structure if structure is not None else ext.SNCH(vertices, indices, max_prims_in_node)
        if not isinstance(self.snch, ext.SNCH):
            raise ValueError("structure must be an ext.SNCH")
        self.bvh = self.snch
        self.vertices = self.snch._vertices
        self.indices = self.snch._indices
        self.scene = Scene(self.snch, device, leaf_prims, packet_budget, traversal, wave_front, sil_pairs,
                           ray_leaf_prims)
        self._device = self.scene.device
        self._bf = None
        self.configure_silhouette()

    def configure_silhouette(self):
        """Polyline.configure_silhouette (polyline.py:44-122): the silhouette vertices (vertices with
        both neighbours, nearly collinear ones dropped) as (k, 10) float32 records a, b, c, prev,
        vertex, next, index -- the reference's SilhouetteVertices layout -- built on the
        polyline's device (a tensor there)."""
        import torch
        self.silhouettes = ext.configure_silhouette(torch.from_numpy(self.vertices).to(self._device),
                                                    torch.from_numpy(self.indices).to(self._device))
        return self.silhouettes

    def _brute_force(self) -> BruteForce:
        if self._bf is None:
            self._bf = BruteForce(2, self.vertices, self.indices, self._device)
        return self._bf

    def intersect(self, p, v, n=None, on_boundary=False, r_max=INF):
        """Polyline.intersect (polyline.py:125-131): origins flagged on_boundary move by -1e-5 n,
        then the brute-force ray query."""
        return self._brute_force().intersect(p, v, r_max, n, on_boundary)

    def intersect_baseline(self, p, v, r_max=INF):
        """Polyline.intersect_baseline (polyline.py:138-170): every segment for every ray."""
        return self._brute_force().intersect(p, v, r_max)

    def closest_point_baseline(self, p):
        """Polyline.closest_point_baseline (polyline.py:176-209): every segment for every point."""
        return self._brute_force().closest_point(p)

    def closest_silhouette_baseline(self, x, r_max=INF):
        """Polyline.closest_silhouette_baseline (polyline.py:212-223): every silhouette vertex."""
        return self._brute_force().closest_silhouette(x, r_max)

    def intersect_bvh(self, p, v, r_max=INF, method: str = "engine"):
        """BVH.intersect (bvh.py:200-276) -> Intersection."""
        return self.scene.intersect(p, v, r_max, method)

    def closest_point_bvh(self, p, method: str = "engine"):
        """BVH.closest_point (bvh.py:127-198) -> ClosestPointRecord."""
        return self.scene.closest_point(p, method)

    def closest_silhouette_snch(self, x, r_max=INF, method: str = "engine"):
        """SNCH.closest_silhouette (snch.py:106-164) -> ClosestSilhouettePointRecord."""
        return self.scene.closest_silhouette(x, r_max, method)

    def sample_boundary(self, x, R, sampler):
        """The 2D form of BVH3D.sample_boundary (bvh3d.py:246-346) over the polyline's BVH, with
        LineSegment.sample_point (line_segment.py:27-32) -> (hit, BoundarySamplingRecord).
        `sampler`: gquery_amd.sampler.PCG32 with one lane per query, advanced in place."""
        return self.scene.sample_boundary(x, R, sampler)

    def set_options(self, leaf_prims: int = 0, packet_budget: int = 0, traversal: int = 0, wave_front: int = 0,
                    sil_pairs: int = 0, ray_leaf_prims: int = 0):
        """Rebuilds the device scene with other engine options (the autotuner's choice)."""
        self.scene = Scene(self.snch, self._device, leaf_prims, packet_budget, traversal, wave_front, sil_pairs,
                           ray_leaf_prims)

    def save(self, path) -> None:
        """Writes the polyline's trees and geometry to a scene file (Polyline.load reads it back
        without running the builder)."""
        self.snch.save(path)

    @classmethod
    def load(cls, path, device=None, **options) -> "Polyline":
        """A Polyline from a scene file written by save(); options as for the constructor."""
        return cls(None, None, device=device, structure=ext.load(path), **options)
