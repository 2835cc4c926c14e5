"""GPU: the default packet budgets and silhouette traversals, which follow the scene size and the
queries per primitive, give the same records as every forced choice: a middle-density 3D
silhouette job (lane per query by default), a dense one (the long budget), a polyline above the
packet threshold (packet silhouettes by default) and closest point on each budget tier."""
import numpy as np
import pytest
import torch

from gquery_amd import synth
from gquery_amd.batch import Batch
from gquery_amd.shapes.mesh import Mesh
from gquery_amd.shapes.polyline import Polyline

pytestmark = pytest.mark.gpu

FIELDS = ("valid", "prim_id", "d", "p", "n", "uv", "t", "on_boundary")


def _np(r):
    torch.cuda.synchronize()
    return {f: getattr(r, f).cpu().numpy() for f in FIELDS if hasattr(r, f)}


def _same(a, b, tag):
    for f in a:
        assert np.array_equal(a[f].view(np.uint8), b[f].view(np.uint8)), (tag, f)


def _run(shape, kind, x, **over):
    b = Batch([dict(shape=shape, kind=kind, x=x, **over)])
    return _np(b.run()[0])


@pytest.mark.parametrize("case", [
    ("mesh", 4, "sil", 1 << 18),      # 5,120 triangles, 51 queries per triangle: lane per query
    ("mesh", 3, "sil", 1 << 18),      # 1,280 triangles, 205 per triangle: dense, long budget
    ("mesh", 5, "cp", 1 << 18),       # 20,480 triangles: the small-tree closest-point budget
    ("mesh", 6, "cp", 1 << 20),       # 81,920 triangles: the mid-tree budget
    ("mesh", 3, "cp", 8000),          # a tiny launch
    ("star", 400_000, "sil", 1 << 17),  # above the 2D packet-silhouette threshold
    ("star", 300_000, "cp", 1 << 19),   # the large-tree 2D closest-point budget
])
def test_defaults_equal_forced_choices(gpu, case):
    shape_kind, size, kind, n = case
    if shape_kind == "mesh":
        V, F = synth.displaced_icosphere(size, seed=21)
        shape, dim = Mesh(V, F, device="cuda:0"), 3
    else:
        V, F = synth.star_polyline(size, seed=22)
        shape, dim = Polyline(V, F, device="cuda:0"), 2
    x = torch.from_numpy(synth.uniform_points(n, dim, seed=23)).cuda()
    print(f"[defaults] {case}", flush=True)
    ref = _run(shape, kind, x)
    forced = [dict(packet_budget=128, traversal=1), dict(packet_budget=3000, traversal=1)]
    if kind == "sil":
        forced.append(dict(traversal=2))
    for over in forced:
        _same(ref, _run(shape, kind, x, **over), (case, over))
    # and the single call (its own sort) agrees with the one-job batch
    single = shape.closest_silhouette_snch(x) if kind == "sil" else shape.closest_point_bvh(x)
    _same(ref, _np(single), (case, "single"))
