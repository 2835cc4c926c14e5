"""Host-side builders with the surface of gQuery's nanobind module gquery_ext
(src/gquery.cpp:180-250): BVH / BVH3D / SNCH / SNCH3D constructors taking vertices and indices,
and node_data() / primitive_data() / silhouette_data() / raw_silhouette_data() returning the flat
float32 arrays of the reference wire format; build_flat_silhouette_edges(vertices, faces).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .fwd import SoA


def _as_rows(a, dim: int, dtype) -> np.ndarray:
    """(N, dim) C-contiguous array from fwd.Array2 / Array3 / Array2i / Array3i (components first,
    as the reference's DrJit arrays) or numpy / torch / sequence rows.  Untagged (dim, N) input is
    transposed only where that is unambiguous (N != dim): a square array is read as rows."""
    if isinstance(a, SoA):
        if a.dim != dim:
            raise ValueError(f"expected {dim}-component arrays, got {type(a).__name__}")
        a = a.rows()
    try:
        import torch
        if isinstance(a, torch.Tensor):
            a = a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    a = np.asarray(a)
    if a.ndim != 2:
        raise ValueError(f"expected a 2-D array of {dim}-vectors, got shape {a.shape}")
    if a.shape[1] != dim and a.shape[0] == dim:
        a = a.T
    if a.shape[1] != dim:
        raise ValueError(f"expected shape (N, {dim}), got {a.shape}")
    return np.ascontiguousarray(a, dtype=dtype)


class _Structure:
    _dim = 0
    _snch = False

    def __init__(self, vertices, indices, max_prims_in_node: int = 10):
        v = _as_rows(vertices, self._dim, np.float32)
        f = _as_rows(indices, self._dim, np.int32)
        self._vertices, self._indices = v, f
        h = C.c_void_p()
        fn = _lib.lib().gq_build_snch if self._snch else _lib.lib().gq_build_bvh
        _lib.check(fn(self._dim, v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0],
                      int(max_prims_in_node), C.byref(h)), type(self).__name__)
        self._h = h
        # held by the object: module globals may already be gone when it is collected at exit
        self._free = _lib.lib().gq_structure_free

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        a = (C.c_int64 * 8)()
        _lib.check(_lib.lib().gq_structure_info(self._h, a))
        keys = ["dim", "is_snch", "n_nodes", "n_prims", "n_silhouettes", "n_raw_silhouettes", "max_depth"]
        return dict(zip(keys, list(a)[:7]))

    def _fetch(self, fn_name: str) -> np.ndarray:
        fn = getattr(_lib.lib(), fn_name)
        n = C.c_int64(0)
        _lib.check(fn(self._h, None, C.byref(n)), fn_name)
        out = np.empty(n.value, dtype=np.float32)
        if n.value:
            _lib.check(fn(self._h, out.ctypes.data, C.byref(n)), fn_name)
        return out

    def node_data(self) -> np.ndarray:
        return self._fetch("gq_structure_snch_node_data" if self._snch else "gq_structure_node_data")

    def bvh_node_data(self) -> np.ndarray:
        return self._fetch("gq_structure_node_data")

    def primitive_data(self) -> np.ndarray:
        return self._fetch("gq_structure_primitive_data")

    def save(self, path) -> None:
        """Writes the structure (flat arrays, counts, input geometry) to a scene file."""
        _lib.check(_lib.lib().gq_structure_save(self._h, str(path).encode()), "gq_structure_save")

    def engine_layout(self, which: int) -> np.ndarray:
        """Raw bytes of one engine-layout array built on the host (see gq_structure_engine_layout)."""
        L = _lib.lib()
        n = C.c_int64(0)
        _lib.check(L.gq_structure_engine_layout(self._h, which, None, C.byref(n)), "engine_layout")
        out = np.empty(n.value, dtype=np.uint8)
        if n.value:
            _lib.check(L.gq_structure_engine_layout(self._h, which, out.ctypes.data, C.byref(n)), "engine_layout")
        return out


class _SnchMixin:
    def silhouette_data(self) -> np.ndarray:
        return self._fetch("gq_structure_silhouette_data")

    def raw_silhouette_data(self) -> np.ndarray:
        return self._fetch("gq_structure_raw_silhouette_data")


class BVH(_Structure):
    """gquery_ext.BVH (2D line segments)."""
    _dim = 2


class BVH3D(_Structure):
    """gquery_ext.BVH3D (triangles)."""
    _dim = 3


class SNCH(_SnchMixin, _Structure):
    """gquery_ext.SNCH (2D): BVH + normal cones + silhouette vertices."""
    _dim = 2
    _snch = True


class SNCH3D(_SnchMixin, _Structure):
    """gquery_ext.SNCH3D: BVH + normal cones + silhouette edges."""
    _dim = 3
    _snch = True


def load(path) -> _Structure:
    """Reads a scene file written by save(): a BVH, BVH3D, SNCH or SNCH3D without running the
    builder (raises ValueError on a file that is not a complete, uncorrupted scene file)."""
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.gq_structure_load(str(path).encode(), C.byref(h)), "gq_structure_load")
    a = (C.c_int64 * 8)()
    rc = L.gq_structure_info(h, a)
    if rc != _lib.GQ_OK:
        L.gq_structure_free(h)
        _lib.check(rc, "gq_structure_info")
    dim, snch = int(a[0]), bool(a[1])
    cls = {(2, False): BVH, (3, False): BVH3D, (2, True): SNCH, (3, True): SNCH3D}[(dim, snch)]
    obj = cls.__new__(cls)
    obj._h = h
    obj._free = L.gq_structure_free
    nf, ni = C.c_int64(0), C.c_int64(0)
    _lib.check(L.gq_structure_geometry(h, None, C.byref(nf), None, C.byref(ni)), "gq_structure_geometry")
    v = np.empty(nf.value, np.float32)
    f = np.empty(ni.value, np.int32)
    _lib.check(L.gq_structure_geometry(h, v.ctypes.data, C.byref(nf), f.ctypes.data, C.byref(ni)),
               "gq_structure_geometry")
    obj._vertices, obj._indices = v.reshape(-1, dim), f.reshape(-1, dim)
    return obj


def _device_setup(fn, vertices, index, dim, stride, what):
    """Runs a silhouette setup on the device for CUDA tensors (rows, as _as_rows reads them);
    returns the records as a (k, stride) float32 tensor on that device."""
    import torch
    from .fwd import SoA
    from .scene import as_points
    dev = (vertices.value if isinstance(vertices, SoA) else vertices).device
    v = as_points(vertices, dim, dev).contiguous()
    if isinstance(index, SoA):
        index = index.rows()
    i = index if isinstance(index, torch.Tensor) else torch.as_tensor(np.asarray(index))
    i = i.to(device=dev, dtype=torch.int32)
    if i.ndim == 2 and i.shape[0] == dim and i.shape[1] != dim:
        i = i.T
    i = i.reshape(-1, dim).contiguous()
    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    n = C.c_int64(0)
    _lib.check(getattr(L, fn)(v.data_ptr(), v.shape[0], i.data_ptr(), i.shape[0], None, C.byref(n), stream), what)
    out = torch.empty(max(n.value, 1), dtype=torch.float32, device=dev)
    _lib.check(getattr(L, fn)(v.data_ptr(), v.shape[0], i.data_ptr(), i.shape[0], out.data_ptr(), C.byref(n), stream),
               what)
    return out[:n.value].reshape(-1, stride)


def _on_device(a):
    try:
        import torch
        from .fwd import SoA
        if isinstance(a, SoA):
            a = a.value
        return isinstance(a, torch.Tensor) and a.is_cuda
    except ImportError:  # pragma: no cover
        return False


def build_flat_silhouette_edges(vertices, faces):
    """gquery_ext.build_flat_silhouette_edges: all mesh edges as 19-float silhouette records (a flat
    float32 array).  CUDA tensor vertices run the setup on their device and return a flat device
    tensor, bit-identical to the host result."""
    if _on_device(vertices):
        return _device_setup("gq_build_flat_silhouette_edges_device", vertices, faces, 3, 19,
                             "build_flat_silhouette_edges").reshape(-1)
    v = _as_rows(vertices, 3, np.float32)
    f = _as_rows(faces, 3, np.int32)
    n = C.c_int64(0)
    L = _lib.lib()
    _lib.check(L.gq_build_flat_silhouette_edges(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0], None, C.byref(n)))
    out = np.empty(n.value, dtype=np.float32)
    _lib.check(L.gq_build_flat_silhouette_edges(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0],
                                                out.ctypes.data, C.byref(n)))
    return out


def configure_silhouette(vertices, edges):
    """Polyline.configure_silhouette (gquery/shapes/polyline.py:44-122): the polyline's silhouette
    vertices as (k, 10) float32 records (a, b, c, prev, vertex, next, index).  CUDA tensor vertices
    run it on their device (a tensor there, bit-identical to the host result)."""
    if _on_device(vertices):
        return _device_setup("gq_configure_silhouette_device", vertices, edges, 2, 10, "configure_silhouette")
    v = _as_rows(vertices, 2, np.float32)
    e = _as_rows(edges, 2, np.int32)
    n = C.c_int64(0)
    L = _lib.lib()
    _lib.check(L.gq_configure_silhouette(v.ctypes.data, v.shape[0], e.ctypes.data, e.shape[0], None, C.byref(n)))
    out = np.empty(n.value, dtype=np.float32)
    _lib.check(L.gq_configure_silhouette(v.ctypes.data, v.shape[0], e.ctypes.data, e.shape[0], out.ctypes.data,
                                         C.byref(n)))
    return out.reshape(-1, 10)
