"""Mesh (gquery/shapes/mesh.py:11-139): a triangle mesh with BVH and SNCH queries.

Mesh(vertices, indices) builds the SAH BVH and the SNCH (mesh.py:23-24; the reference builds the
same BVH twice, here one SNCH build provides both) and keeps them on the current HIP device.
"""
from __future__ import annotations

from .. import ext
from ..scene import INF, BruteForce, Scene


class Mesh:
    def __init__(self, vertices, indices, device=None, max_prims_in_node: int = 10, leaf_prims: int = 0,
                 packet_budget: int = 0, traversal: int = 0, wave_front: int = 0, sil_pairs: int = 0,
                 structure=None, ray_leaf_prims: int = 0):
        # structure: an ext.SNCH3D already built or loaded from a scene file (then vertices and
        # indices may be None)
        self.snch = structure if structure is not None else ext.SNCH3D(vertices, indices, max_prims_in_node)
        if not isinstance(self.snch, ext.SNCH3D):
            raise ValueError("structure must be an ext.SNCH3D")
        self.bvh = self.snch
        self.vertices = self.snch._vertices
        self.indices = self.snch._indices
        self.scene = Scene(self.snch, device, leaf_prims, packet_budget, traversal, wave_front, sil_pairs,
                           ray_leaf_prims)
        self._device = self.scene.device
        self._bf = None
        self._normals = None
        self._silhouettes = None

    def _brute_force(self) -> BruteForce:
        if self._bf is None:
            self._bf = BruteForce(3, self.vertices, self.indices, self._device)
        return self._bf

    def configure_normal(self):
        """Mesh.configure_normal (mesh.py:35-50): face normals normalize(cross(b-a, c-a)) and vertex
        normals (normalised sums of the incident face normals), computed on the device."""
        self._normals = self._brute_force().normals()
        return self._normals

    @property
    def face_normals(self):
        return (self._normals or self.configure_normal())[0]

    @property
    def normals(self):
        return (self._normals or self.configure_normal())[1]

    @property
    def silhouettes(self):
        """The mesh's silhouette edges (mesh.py:26-29): build_flat_silhouette_edges records, built on
        the mesh's device (a flat float32 tensor there; 19 floats per edge)."""
        if self._silhouettes is None:
            import torch
            self._silhouettes = ext.build_flat_silhouette_edges(torch.from_numpy(self.vertices).to(self._device),
                                                                torch.from_numpy(self.indices).to(self._device))
        return self._silhouettes

    def intersect(self, p, v, n=None, on_boundary=False, r_max=INF):
        """Mesh.intersect (mesh.py:53-59): origins flagged on_boundary move by -1e-5 n, then the
        brute-force ray query.  (The reference passes n and on_boundary on to intersect_baseline,
        which does not take them; this runs what the method evidently means.)"""
        return self._brute_force().intersect(p, v, r_max, n, on_boundary)

    def intersect_baseline(self, p, v, r_max=INF):
        """Mesh.intersect_baseline (mesh.py:61-85): every triangle for every ray."""
        return self._brute_force().intersect(p, v, r_max)

    def closest_point_baseline(self, x):
        """Mesh.closest_point_baseline (mesh.py:92-118): every triangle for every point."""
        return self._brute_force().closest_point(x)

    def closest_silhouette_baseline(self, p, r_max=INF):
        """Mesh.closest_silhouette_baseline (mesh.py:124-136): every silhouette edge."""
        return self._brute_force().closest_silhouette(p, r_max)

    def intersect_bvh(self, p, v, r_max=INF, method: str = "engine"):
        """BVH3D.intersect (bvh3d.py:176-239) -> Intersection3D."""
        return self.scene.intersect(p, v, r_max, method)

    def closest_point_bvh(self, x, method: str = "engine"):
        """BVH3D.closest_point (bvh3d.py:108-174) -> ClosestPointRecord3D."""
        return self.scene.closest_point(x, method)

    def closest_silhouette_snch(self, p, r_max=INF, method: str = "engine"):
        """SNCH3D.closest_silhouette (snch3d.py:143-203) -> ClosestSilhouettePointRecord3D."""
        return self.scene.closest_silhouette(p, r_max, method)

    def sample_boundary(self, x, R, sampler):
        """BVH3D.sample_boundary (bvh3d.py:246-346) over the mesh's BVH -> (hit,
        BoundarySamplingRecord3D).  `sampler`: gquery_amd.sampler.PCG32 with one lane per query,
        advanced in place."""
        return self.scene.sample_boundary(x, R, sampler)

    def set_options(self, leaf_prims: int = 0, packet_budget: int = 0, traversal: int = 0, wave_front: int = 0,
                    sil_pairs: int = 0, ray_leaf_prims: int = 0):
        """Rebuilds the device scene with other engine options (the autotuner's choice)."""
        self.scene = Scene(self.snch, self._device, leaf_prims, packet_budget, traversal, wave_front, sil_pairs,
                           ray_leaf_prims)

    def save(self, path) -> None:
        """Writes the mesh's trees and geometry to a scene file (Mesh.load reads it back
        without running the builder)."""
        self.snch.save(path)

    @classmethod
    def load(cls, path, device=None, **options) -> "Mesh":
        """A Mesh from a scene file written by save(); options as for the constructor."""
        return cls(None, None, device=device, structure=ext.load(path), **options)
