"""GPU: batched enqueue (gq_batch_enqueue) against the single-batch calls, and graph replays."""
import numpy as np
import pytest
import torch

from gquery_amd import synth
from gquery_amd.batch import Batch
from tests.test_gpu_parity import scene

pytestmark = pytest.mark.gpu

FIELDS = ("valid", "p", "n", "uv", "t", "d", "prim_id", "on_boundary")


def _np(r):
    torch.cuda.synchronize()
    return {f: getattr(r, f).cpu().numpy() for f in FIELDS if hasattr(r, f)}


def _single(shape, kind, x, v, method):
    if kind == "cp":
        return shape.closest_point_bvh(x, method=method)
    if kind == "ray":
        return shape.intersect_bvh(x, v, method=method)
    return shape.closest_silhouette_snch(x, method=method)


def _jobs(seed=0):
    rng = np.random.default_rng(seed)
    jobs = []
    for k, (name, kind, n) in enumerate([("star4096", "cp", 20000), ("ico4", "ray", 3000), ("oico4", "sil", 12000),
                                          ("ico4", "cp", 9000), ("star4096", "sil", 5000), ("ostar1024", "ray", 16000),
                                          ("ico3", "sil", 30000), ("star1024", "cp", 100)]):
        V, F, shape = scene(name)
        dim = V.shape[1]
        x = torch.from_numpy(synth.uniform_points(n, dim, seed=100 + k)).cuda()
        v = torch.from_numpy(synth.uniform_directions(n, dim, seed=200 + k)).cuda() if kind == "ray" else None
        jobs.append((shape, kind, x, v))
    return jobs


@pytest.mark.parametrize("method", ["engine", "engine_fp32"])
def test_batch_equals_single_calls(gpu, method):
    jobs = _jobs()
    b = Batch(jobs, method=method)
    got = b.run()
    for (shape, kind, x, v), r in zip(jobs, got):
        ref = _np(_single(shape, kind, x, v, method))
        g = _np(r)
        for f in ref:
            assert np.array_equal(g[f].view(np.uint8), ref[f].view(np.uint8)), (kind, f)


def test_shared_jobs_equal_single_calls(gpu):
    """shared=1 (other launches beside this one) changes only the packet budgets of small launch
    groups, never a record."""
    jobs = _jobs(2)
    got = Batch([dict(shape=s, kind=k, x=x, v=v, shared=1) for s, k, x, v in jobs]).run()
    for (shape, kind, x, v), r in zip(jobs, got):
        ref = _np(_single(shape, kind, x, v, "engine"))
        g = _np(r)
        for f in ref:
            assert np.array_equal(g[f].view(np.uint8), ref[f].view(np.uint8)), (kind, f)


def test_batch_graph_replays(gpu):
    jobs = _jobs(1)
    b = Batch(jobs)
    b.capture()
    for it in range(3):
        # new inputs in place, eager work in between, then a replay must match single calls
        for k, (shape, kind, x, v) in enumerate(jobs):
            dim = x.shape[1]
            x.copy_(torch.from_numpy(synth.uniform_points(x.shape[0], dim, seed=1000 * it + k)))
        torch.cuda.synchronize()
        got = b.replay()
        torch.cuda.synchronize()
        for (shape, kind, x, v), r in zip(jobs, got):
            ref = _np(_single(shape, kind, x, v, "engine"))
            g = _np(r)
            for f in ref:
                assert np.array_equal(g[f].view(np.uint8), ref[f].view(np.uint8)), (it, kind, f)


def test_batch_rejects_bad_jobs(gpu):
    V, F, shape = scene("star1024")
    x = torch.zeros((10, 2), device="cuda:0")
    with pytest.raises(NotImplementedError):
        Batch([(shape, "cp", x)], method="plain")
    with pytest.raises(ValueError):
        Batch([{"shape": shape, "kind": "cp", "x": x, "traversal": 7}])
    with pytest.raises(ValueError):
        Batch([{"shape": shape, "kind": "cp", "x": x, "shared": 2}])


def test_batch_input_order_equals_sorted(gpu):
    """order=1 (the caller's order, no Morton sort) changes the wavefront grouping only: every
    record equals the sorted batch's bit for bit (ties break to the lowest primitive slot)."""
    from gquery_amd.batch import Batch
    rng = np.random.default_rng(9)
    jobs = []
    for k, (name, kind, n) in enumerate([("ico4", "cp", 20000), ("star4096", "cp", 30000), ("oico4", "sil", 12000),
                                          ("star4096", "sil", 9000)]):
        V, F, shape = scene(name)
        x = torch.from_numpy(synth.uniform_points(n, V.shape[1], seed=500 + k)).cuda()
        jobs.append({"shape": shape, "kind": kind, "x": x})
    sorted_ = Batch(jobs).run()
    unsorted = Batch([dict(j, order=1) for j in jobs]).run()
    for j, a, b in zip(jobs, sorted_, unsorted):
        A, B = _np(a), _np(b)
        for f in A:
            assert np.array_equal(A[f].view(np.uint8), B[f].view(np.uint8)), (j["kind"], f)
    with pytest.raises(ValueError):
        Batch([dict(jobs[0], order=2)])
