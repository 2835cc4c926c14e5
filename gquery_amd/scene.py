"""Device-resident scene (gq_scene) and the three query calls on torch tensors.

Inputs may be torch tensors on any device, numpy arrays or sequences; they are moved to the
scene's HIP device as contiguous float32 (N, dim).  Outputs are freshly allocated device tensors.
All calls are ordered on torch's current stream of the scene's device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .fwd import SoA
from .shapes import primitive as rec

INF = float("inf")


def _device(device=None) -> torch.device:
    if device is None:
        if not torch.cuda.is_available():
            raise RuntimeError("gquery_amd needs a HIP device (torch.cuda.is_available() is False)")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def as_points(x, dim: int, device: torch.device) -> torch.Tensor:
    """(N, dim) float32 points on `device`.  Input: fwd.Array2 / Array3 (components first, as the
    reference's DrJit arrays), or numpy / torch / sequences in (N, dim) rows.  Untagged input in
    (dim, N) layout is accepted only where that is unambiguous (N != dim); a square (dim, dim)
    array is read as rows -- pass components-first data of that shape as Array2 / Array3."""
    if isinstance(x, SoA):
        if x.dim != dim:
            raise ValueError(f"expected {dim}-component points, got {type(x).__name__}")
        x = x.rows()
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(np.asarray(x, dtype=np.float32))
    if x.ndim == 1 and x.shape[0] == dim:
        x = x[None, :]
    if x.ndim != 2:
        raise ValueError(f"expected (N, {dim}) points, got shape {tuple(x.shape)}")
    if x.shape[1] != dim and x.shape[0] == dim:
        x = x.T
    if x.shape[1] != dim:
        raise ValueError(f"expected (N, {dim}) points, got shape {tuple(x.shape)}")
    return x.to(device=device, dtype=torch.float32).contiguous()


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class _DeviceQueries:
    """Record allocation, r_max and stream handling shared by the device objects."""
    dim: int
    device: torch.device

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _alloc(self, n: int, fields):
        D = self.dim
        f32 = dict(device=self.device, dtype=torch.float32)
        out = {}
        for name in fields:
            if name in ("valid", "on_boundary"):
                out[name] = torch.empty(n, device=self.device, dtype=torch.bool)
            elif name in ("p", "n"):
                out[name] = torch.empty((n, D), **f32)
            elif name == "uv":
                out[name] = torch.empty((n, 2), **f32) if D == 3 else torch.empty(n, **f32)
            elif name == "d":
                out[name] = torch.empty(n, **f32)
            elif name == "prim_id":
                out[name] = torch.empty(n, device=self.device, dtype=torch.int32)
        r = _lib.Records()
        for name, t in out.items():
            setattr(r, name, t.data_ptr())
        return out, r

    def _r_max(self, r_max, n):
        if isinstance(r_max, (int, float)):
            return None, float(r_max)
        t = torch.as_tensor(r_max, dtype=torch.float32)
        if t.numel() == 1:
            return None, float(t.reshape(()).item())
        t = t.to(self.device).reshape(-1).contiguous()
        if t.numel() != n:
            raise ValueError("r_max must be a scalar or have one entry per query")
        return t, INF

    def _cp_record(self, out):
        if self.dim == 3:
            return rec.ClosestPointRecord3D(out["valid"], out["p"], out["n"], out["uv"], out["d"], out["prim_id"])
        return rec.ClosestPointRecord(out["valid"], out["p"], out["n"], out["uv"], out["d"], out["prim_id"])

    def _ray_record(self, out):
        cls = rec.Intersection3D if self.dim == 3 else rec.Intersection
        return cls(out["valid"], out["p"], out["n"], out["uv"], out["d"], out["prim_id"], out["on_boundary"])

    def _sil_record(self, out):
        cls = rec.ClosestSilhouettePointRecord3D if self.dim == 3 else rec.ClosestSilhouettePointRecord
        return cls(out["valid"], out["p"], out["d"], out["prim_id"])


class BruteForce(_DeviceQueries):
    """A polyline or mesh on the device for the reference's brute-force methods (gq_shape): every
    query tests every primitive / silhouette (Polyline/Mesh.*_baseline), plus the mesh normals of
    Mesh.configure_normal.  Silhouettes: Polyline.configure_silhouette's vertices (2D) or the
    mesh's flat silhouette edges (3D)."""

    def __init__(self, dim: int, vertices, indices, device=None):
        self.device = _device(device)
        self.dim = int(dim)
        V = np.ascontiguousarray(vertices, np.float32).reshape(-1, self.dim)
        F = np.ascontiguousarray(indices, np.int32).reshape(-1, self.dim)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().gq_shape_create(self.dim, V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0],
                                                  C.byref(h)), "gq_shape_create")
        self._h = h
        self._destroy = _lib.lib().gq_shape_destroy
        self._sync = torch.cuda.synchronize

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._sync(self.device)
            except Exception:
                pass
            self._destroy(h)
            self._h = None

    def info(self) -> dict:
        a = (C.c_int64 * 4)()
        _lib.check(_lib.lib().gq_shape_info(self._h, a))
        return dict(zip(["dim", "n_vertices", "n_prims", "n_silhouettes"], list(a)))

    def normals(self):
        """(face normals (F, 3), vertex normals (V, 3)) as device tensors (Mesh.configure_normal)."""
        i = self.info()
        with torch.cuda.device(self.device):
            fn = torch.empty((i["n_prims"], 3), device=self.device, dtype=torch.float32)
            vn = torch.empty((i["n_vertices"], 3), device=self.device, dtype=torch.float32)
            _lib.check(_lib.lib().gq_shape_normals(self._h, _ptr(fn), _ptr(vn), self._stream()), "normals")
        return fn, vn

    def closest_point(self, x):
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            n = x.shape[0]
            out, r = self._alloc(n, ["valid", "p", "n", "uv", "d", "prim_id"])
            _lib.check(_lib.lib().gq_closest_point_baseline(self._h, _ptr(x), n, C.byref(r), self._stream()),
                       "closest_point_baseline")
        return self._cp_record(out)

    def intersect(self, x, v, r_max=INF, normal=None, on_boundary=None):
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            v = as_points(v, self.dim, self.device)
            if x.shape != v.shape:
                raise ValueError("origins and directions must have the same shape")
            n = x.shape[0]
            nb = ob = None
            if on_boundary is not None:
                ob = torch.as_tensor(on_boundary).to(self.device, torch.bool).reshape(-1).expand(n).contiguous()
                nb = as_points(normal if normal is not None else np.zeros(self.dim, np.float32), self.dim, self.device)
                nb = nb.expand(n, self.dim).contiguous()
            rt, rs = self._r_max(r_max, n)
            out, r = self._alloc(n, ["valid", "p", "n", "uv", "d", "prim_id", "on_boundary"])
            _lib.check(_lib.lib().gq_intersect_baseline(self._h, _ptr(x), _ptr(v), _ptr(nb), _ptr(ob), _ptr(rt), rs,
                                                        n, C.byref(r), self._stream()), "intersect_baseline")
        return self._ray_record(out)

    def closest_silhouette(self, x, r_max=INF):
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            n = x.shape[0]
            rt, rs = self._r_max(r_max, n)
            out, r = self._alloc(n, ["valid", "p", "d", "prim_id"])
            _lib.check(_lib.lib().gq_closest_silhouette_baseline(self._h, _ptr(x), _ptr(rt), rs, n, C.byref(r),
                                                                 self._stream()), "closest_silhouette_baseline")
        return self._sil_record(out)


class Scene(_DeviceQueries):
    """One polyline or mesh on the device: reference-layout BVH/SNCH and the engine layouts."""

    def __init__(self, structure, device=None, leaf_prims: int = 0, packet_budget: int = 0, traversal: int = 0,
                 wave_front: int = 0, sil_pairs: int = 0, ray_leaf_prims: int = 0):
        self.device = _device(device)
        self.dim = structure.info()["dim"]
        self.has_snch = bool(structure.info()["is_snch"])
        self._structure = structure
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            opt = _lib.SceneOptions(leaf_prims=int(leaf_prims), packet_budget=int(packet_budget),
                                    traversal=int(traversal), wave_front=int(wave_front),
                                    sil_pairs=int(sil_pairs), ray_leaf_prims=int(ray_leaf_prims))
            _lib.check(_lib.lib().gq_scene_create_ex(structure.handle, C.byref(opt), C.byref(h)), "gq_scene_create")
        self._h = h
        # held by the object: module globals may already be gone when it is collected at exit
        self._destroy = _lib.lib().gq_scene_destroy
        self._sync = torch.cuda.synchronize

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._sync(self.device)
            except Exception:
                pass
            self._destroy(h)
            self._h = None

    def info(self) -> dict:
        a = (C.c_int64 * 8)()
        _lib.check(_lib.lib().gq_scene_info(self._h, a))
        keys = ["dim", "has_snch", "device_bytes", "n_nodes", "n_wide_nodes", "n_prims", "n_silhouettes"]
        return dict(zip(keys, list(a)[:7]))

    # -- queries -------------------------------------------------------------------------------
    def closest_point(self, x, method: str = "engine"):
        m = _lib.METHODS[method]
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            n = x.shape[0]
            out, r = self._alloc(n, ["valid", "p", "n", "uv", "d", "prim_id"])
            _lib.check(_lib.lib().gq_closest_point_bvh(self._h, m, _ptr(x), n, C.byref(r), self._stream()),
                       "closest_point_bvh")
        return self._cp_record(out)

    def intersect(self, x, v, r_max=INF, method: str = "engine"):
        m = _lib.METHODS[method]
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            v = as_points(v, self.dim, self.device)
            if x.shape != v.shape:
                raise ValueError("origins and directions must have the same shape")
            n = x.shape[0]
            rt, rs = self._r_max(r_max, n)
            out, r = self._alloc(n, ["valid", "p", "n", "uv", "d", "prim_id", "on_boundary"])
            _lib.check(_lib.lib().gq_intersect_bvh(self._h, m, _ptr(x), _ptr(v), _ptr(rt), rs, n, C.byref(r),
                                                   self._stream()), "intersect_bvh")
        return self._ray_record(out)

    def closest_silhouette(self, x, r_max=INF, method: str = "engine"):
        m = _lib.METHODS[method]
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            n = x.shape[0]
            rt, rs = self._r_max(r_max, n)
            out, r = self._alloc(n, ["valid", "p", "d", "prim_id"])
            _lib.check(_lib.lib().gq_closest_silhouette_snch(self._h, m, _ptr(x), _ptr(rt), rs, n, C.byref(r),
                                                             self._stream()), "closest_silhouette_snch")
        return self._sil_record(out)

    def sample_boundary(self, x, R, sampler):
        """BVH3D.sample_boundary (bvh3d.py:246-346) on the scene's reference BVH: per query one
        BVH-guided walk to a leaf, an area-weighted primitive among those meeting the ball (x, R)
        and a uniform point on it.  `sampler`: a gquery_amd.sampler.PCG32 with one lane per query
        on the scene's device, advanced in place.  Returns (hit, BoundarySamplingRecord3D) in 3D,
        (hit, BoundarySamplingRecord) in 2D; a query without a sample is all zero."""
        with torch.cuda.device(self.device):
            x = as_points(x, self.dim, self.device)
            n = x.shape[0]
            if len(sampler) != n:
                raise ValueError(f"sampler has {len(sampler)} lanes for {n} queries")
            if sampler.state.device != self.device or sampler.inc.device != self.device:
                raise ValueError("sampler must live on the scene's device")
            rt, rs = self._r_max(R, n)
            out, _ = self._alloc(n, ["p", "n", "uv", "prim_id"])
            out["hit"] = torch.empty(n, device=self.device, dtype=torch.bool)
            out["pdf"] = torch.empty(n, device=self.device, dtype=torch.float32)
            out["pmf"] = torch.empty(n, device=self.device, dtype=torch.float32)
            r = _lib.SampleRecords()
            for name in ("hit", "p", "n", "uv", "pdf", "pmf", "prim_id"):
                setattr(r, name, out[name].data_ptr())
            _lib.check(_lib.lib().gq_sample_boundary(self._h, _ptr(x), _ptr(rt), rs, _ptr(sampler.state),
                                                     _ptr(sampler.inc), n, C.byref(r), self._stream()),
                       "sample_boundary")
        cls = rec.BoundarySamplingRecord3D if self.dim == 3 else rec.BoundarySamplingRecord
        return out["hit"], cls(out["p"], out["n"], out["uv"], out["pdf"], out["prim_id"], out["pmf"])
