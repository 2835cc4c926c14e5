"""GPU: the autotuner picks among exact candidates; its choice returns the default records."""
import numpy as np
import pytest
import torch

from gquery_amd import synth, tune
from gquery_amd.shapes.mesh import Mesh
from gquery_amd.shapes.polyline import Polyline

pytestmark = pytest.mark.gpu

FIELDS = ("valid", "p", "d", "prim_id")


@pytest.mark.parametrize("dim", [2, 3])
def test_tuned_options_give_the_default_records(gpu, dim):
    V, F = synth.star_polyline(4096, seed=8) if dim == 2 else synth.displaced_icosphere(4, seed=9)
    x = torch.from_numpy(synth.uniform_points(20000, dim, seed=31)).cuda()
    v = torch.from_numpy(synth.uniform_directions(20000, dim, seed=32)).cuda()
    res = tune.tune_shape(V, F, queries={k: (x, v) for k in ("cp", "ray", "sil")},
                          leaves=(2, 16), budgets=(1, 1000), traversals=(1, 2), ray_leaves=(4, 6))
    opt = res["options"]
    assert opt["leaf_prims"] in (2, 16) and opt["packet_budget"] in (1, 1000) and opt["traversal"] in (1, 2)
    assert opt["ray_leaf_prims"] in (4, 6)
    # two ray-tree trials, then per closest-point leaf size two budgets plus lane-per-query silhouettes
    assert len(res["trials"]) == 2 + 2 * 3 and set(res["ms"]) == {"cp", "ray", "sil"}
    cls = Polyline if dim == 2 else Mesh
    base, tuned = cls(V, F, device="cuda:0"), cls(V, F, device="cuda:0", **opt)
    for q in ("cp", "ray", "sil"):
        rec = []
        for shape in (base, tuned):
            r = (shape.closest_point_bvh(x) if q == "cp" else shape.intersect_bvh(x, v) if q == "ray"
                 else shape.closest_silhouette_snch(x))
            torch.cuda.synchronize()
            rec.append({f: getattr(r, f).cpu().numpy() for f in FIELDS})
        for f in FIELDS:
            assert np.array_equal(rec[0][f], rec[1][f]), (q, f)


def test_tune_jobs_groups_and_keeps_results(gpu):
    from gquery_amd.batch import Batch
    shapes = [Polyline(*synth.star_polyline(2000, seed=s), device="cuda:0") for s in (1, 2)]
    shapes.append(Mesh(*synth.displaced_icosphere(3, seed=3), device="cuda:0"))
    jobs = []
    for k, (shape, kind) in enumerate(zip(shapes, ("cp", "cp", "sil"))):
        dim = shape.scene.dim
        jobs.append((shape, kind, torch.from_numpy(synth.uniform_points(9000, dim, seed=70 + k)).cuda(), None))
    ref = [j[0].closest_point_bvh(j[2]) if j[1] == "cp" else j[0].closest_silhouette_snch(j[2]) for j in jobs]
    opts = tune.tune_jobs(jobs, leaves=(2, 16), budgets=(500, 1000), traversals=(1, 2), ray_leaves=(4, 6))
    assert opts[0] == opts[1]                       # same group: same dimension, kind and size class
    for (shape, _, _, _), o in zip(jobs, opts):
        shape.set_options(**o)
    got = Batch(jobs).run()
    torch.cuda.synchronize()
    for a, b in zip(ref, got):
        for f in ("valid", "d", "prim_id"):
            assert torch.equal(getattr(a, f), getattr(b, f))
