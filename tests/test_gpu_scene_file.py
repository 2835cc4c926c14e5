"""GPU: a Mesh / Polyline loaded from a scene file answers every query exactly as the original."""
import numpy as np
import pytest
import torch

from gquery_amd import synth
from gquery_amd.shapes.mesh import Mesh
from gquery_amd.shapes.polyline import Polyline

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_loaded_shape_answers_the_same(gpu, tmp_path, dim):
    cls = Polyline if dim == 2 else Mesh
    V, F = synth.star_polyline(3000, seed=6) if dim == 2 else synth.displaced_icosphere(4, seed=6)
    a = cls(V, F, device="cuda:0")
    a.save(tmp_path / "s.gqs")
    b = cls.load(tmp_path / "s.gqs", device="cuda:0")
    x = torch.from_numpy(synth.uniform_points(20000, dim, seed=61)).cuda()
    v = torch.from_numpy(synth.uniform_directions(20000, dim, seed=62)).cuda()
    for q in ("closest_point_bvh", "closest_silhouette_snch", "closest_point_baseline"):
        ra, rb = getattr(a, q)(x), getattr(b, q)(x)
        torch.cuda.synchronize()
        for f in ("valid", "d", "prim_id", "p"):
            assert torch.equal(getattr(ra, f), getattr(rb, f)), (q, f)
    ra, rb = a.intersect_bvh(x, v), b.intersect_bvh(x, v)
    torch.cuda.synchronize()
    assert torch.equal(ra.d, rb.d) and torch.equal(ra.prim_id, rb.prim_id)
    assert torch.equal(b.silhouettes, a.silhouettes)
