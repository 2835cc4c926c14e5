"""Batched queries: many scenes and query batches in one enqueue (gq_batch_enqueue).

    b = Batch([(mesh, "cp", x0), (polyline, "ray", x1, v1), (mesh2, "sil", x2)])
    b.run()                   # eager enqueue on torch's current stream
    b.capture(); b.replay()   # recorded once into a HIP graph (torch.cuda.CUDAGraph), replayed
    b.results[i]              # the record of job i (same dataclasses as the single calls)

A batch owns its device workspace, its pinned staging buffer and its output records; the query
tensors are referenced, not copied, so a replay reads whatever they hold at that time.  Results are
bit-identical to the single-batch calls (Mesh.closest_point_bvh etc.) on the same inputs.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib
from .scene import INF, as_points
from .shapes import primitive as rec

_FIELDS = {"cp": ["valid", "p", "n", "uv", "d", "prim_id"],
           "ray": ["valid", "p", "n", "uv", "d", "prim_id", "on_boundary"],
           "sil": ["valid", "p", "d", "prim_id"]}


def _record(kind, dim, out):
    if kind == "cp":
        cls = rec.ClosestPointRecord3D if dim == 3 else rec.ClosestPointRecord
        return cls(out["valid"], out["p"], out["n"], out["uv"], out["d"], out["prim_id"])
    if kind == "ray":
        cls = rec.Intersection3D if dim == 3 else rec.Intersection
        return cls(out["valid"], out["p"], out["n"], out["uv"], out["d"], out["prim_id"], out["on_boundary"])
    cls = rec.ClosestSilhouettePointRecord3D if dim == 3 else rec.ClosestSilhouettePointRecord
    return cls(out["valid"], out["p"], out["d"], out["prim_id"])


class Batch:
    """A fixed set of jobs (shape or Scene, kind, x[, v][, r_max]) with options per job.

    kind is "cp", "ray" or "sil"; method "engine" (quantized nodes) or "engine_fp32";
    packet_budget / traversal override the scene's settings for that job (0 = scene's); order=1
    traverses the job in the caller's order (no Morton sort); shared=1 says other launches run
    beside this one, so a small launch group keeps the full-launch packet budgets."""

    def __init__(self, jobs, method: str = "engine", device=None, workspace=None, staging=None):
        """workspace / staging: optional reusable buffers (a uint8 device tensor and a pinned uint8
        host tensor) used instead of fresh ones when they are large enough; the caller must not
        hand them to another batch until this one's enqueue has completed on its stream."""
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self._keep = []
        self.results = []
        arr = (_lib.Job * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            j = dict(zip(("shape", "kind", "x", "v", "r_max"), j)) if not isinstance(j, dict) else dict(j)
            shape, kind = j["shape"], j["kind"]
            scene = getattr(shape, "scene", shape)
            if scene.device != self.device:
                raise ValueError(f"job {i}: scene on {scene.device}, batch on {self.device}")
            dim = scene.dim
            x = as_points(j["x"], dim, self.device)
            n = x.shape[0]
            v = as_points(j["v"], dim, self.device) if kind == "ray" else None
            r_max = j.get("r_max", INF)
            r_max = INF if r_max is None else r_max
            rt, rs = scene._r_max(r_max, n)
            out, r = scene._alloc(n, _FIELDS[kind])
            self._keep += [x, v, rt]
            self.results.append(_record(kind, dim, out))
            jb = arr[i]
            jb.scene = scene._h
            jb.kind = _lib.QUERIES[kind]
            jb.method = _lib.METHODS[j.get("method", method)]
            jb.x = x.data_ptr()
            jb.v = v.data_ptr() if v is not None else None
            jb.r_max = rt.data_ptr() if rt is not None else None
            jb.r_max_scalar = rs
            jb.packet_budget = int(j.get("packet_budget", 0))
            jb.n = n
            jb.out = r
            jb.traversal = int(j.get("traversal", 0))
            jb.order = int(j.get("order", 0))
            jb.shared = int(j.get("shared", 0))
        self._jobs = arr
        self._n = len(jobs)
        dev, host = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().gq_batch_sizes(self._n, arr, C.byref(dev), C.byref(host)), "gq_batch_sizes")
        self.device_bytes, self.host_bytes = dev.value, host.value
        if workspace is not None and workspace.numel() >= self.device_bytes and workspace.device == self.device:
            self._ws = workspace
        else:
            self._ws = torch.empty(max(self.device_bytes, 1), dtype=torch.uint8, device=self.device)
        if staging is not None and staging.numel() >= self.host_bytes and staging.is_pinned():
            self._stage = staging
        else:
            self._stage = torch.empty(max(self.host_bytes, 1), dtype=torch.uint8, pin_memory=True)
        self._graph = None

    def run(self):
        """Enqueue every job on torch's current stream (asynchronous)."""
        with torch.cuda.device(self.device):
            st = torch.cuda.current_stream(self.device).cuda_stream
            _lib.check(_lib.lib().gq_batch_enqueue(self._n, self._jobs, C.c_void_p(self._ws.data_ptr()),
                                                   self.device_bytes, C.c_void_p(self._stage.data_ptr()),
                                                   self.host_bytes, C.c_void_p(st)), "gq_batch_enqueue")
        return self.results

    def capture(self):
        """Record the enqueue into a graph (torch.cuda.CUDAGraph on a side stream)."""
        with torch.cuda.device(self.device):
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream(self.device)
            with torch.cuda.stream(s):
                self.run()   # warm the allocator outside capture
            torch.cuda.synchronize(self.device)
            with torch.cuda.graph(g, stream=s):
                self.run()
            self._graph = g
        return self

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph.replay()
        return self.results

    @staticmethod
    def sizes(jobs, method: str = "engine"):
        """(device workspace bytes, pinned staging bytes) gq_batch_sizes reports for `jobs`, without
        allocating records or buffers.  jobs: (shape or Scene, kind, x[, v]) with x, v on the device."""
        arr = (_lib.Job * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            shape, kind, x = j[0], j[1], j[2]
            v = j[3] if len(j) > 3 else None
            scene = getattr(shape, "scene", shape)
            jb = arr[i]
            jb.scene = scene._h
            jb.kind = _lib.QUERIES[kind]
            jb.method = _lib.METHODS[method]
            jb.x = x.data_ptr()
            jb.v = v.data_ptr() if v is not None else None
            jb.r_max_scalar = INF
            jb.n = int(x.shape[0])
        dev, host = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().gq_batch_sizes(len(jobs), arr, C.byref(dev), C.byref(host)), "gq_batch_sizes")
        return dev.value, host.value

    def __len__(self):
        return self._n

    @property
    def n_queries(self) -> int:
        return sum(int(self._jobs[i].n) for i in range(self._n))
