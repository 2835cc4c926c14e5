"""GPU: the serving loop returns, for every request, exactly what the request gives alone."""
import numpy as np
import pytest
import torch

from gquery_amd import synth
from gquery_amd.serve import GpuExecutor, Request, Scheduler, percentile, run_open_loop
from tests.test_gpu_parity import scene

pytestmark = pytest.mark.gpu


def test_results_alone_equal_results_in_the_mix(gpu):
    names = ["star4096", "ico4", "oico4", "ostar1024"]
    shapes = [scene(n)[2] for n in names]
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(48):
        k = int(rng.integers(0, len(names)))
        dim = 2 if "star" in names[k] else 3
        kind = ["cp", "ray", "sil"][int(rng.integers(0, 3))]
        n = int(np.exp(rng.uniform(np.log(64), np.log(30000))))
        x = torch.from_numpy(synth.uniform_points(n, dim, seed=300 + i)).cuda()
        v = torch.from_numpy(synth.uniform_directions(n, dim, seed=400 + i)).cuda() if kind == "ray" else None
        reqs.append(Request(rid=i, scene=k, kind=kind, x=x, v=v))
    ex = GpuExecutor(shapes, slots=(2, 1))
    sched = Scheduler(ex, memory_budget=256 << 20, slots=(2, 1))
    # a burst (half the requests at once: they must be coalesced), then arrivals spread over 50 ms
    arrivals = np.concatenate([np.zeros(len(reqs) // 2), np.sort(rng.uniform(0.0, 0.05, size=len(reqs) - len(reqs) // 2))])
    run_open_loop(sched, reqs, arrivals)
    torch.cuda.synchronize()
    assert len(sched.completed) == len(reqs)
    for r in reqs:
        shape = shapes[r.scene]
        alone = (shape.closest_point_bvh(r.x) if r.kind == "cp" else
                 shape.intersect_bvh(r.x, r.v) if r.kind == "ray" else shape.closest_silhouette_snch(r.x))
        torch.cuda.synchronize()
        for f in ("valid", "p", "d", "prim_id", "n", "uv", "t", "on_boundary"):
            if hasattr(alone, f):
                a, b = getattr(alone, f).cpu().numpy(), getattr(r.result, f).cpu().numpy()
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (r.rid, r.kind, f)
    lat = [r.latency for r in reqs]
    assert all(l > 0 for l in lat) and percentile(lat, 50) <= percentile(lat, 99)
    assert ex.launches < len(reqs)   # requests were coalesced into fewer launches
    # workspaces and staging buffers were reused across launches, not allocated per launch
    assert ex.buffer_allocations < 2 * ex.launches
    # admission charged what the library reports: records + the request's own batch workspace
    from gquery_amd.batch import Batch
    from gquery_amd.serve import record_bytes
    for r in reqs[:8]:
        dev, _ = Batch.sizes([(shapes[r.scene], r.kind, r.x, r.v)])
        dim = 2 if "star" in names[r.scene] else 3
        assert r.footprint == record_bytes(r.kind, dim) * r.n + dev and dev > 0
