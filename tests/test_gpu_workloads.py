"""GPU: the batched and served workloads at their full sizes.

Batched: 64 scenes (half 2D star polylines, half 3D displaced icospheres, 1k-2M primitives, 17.1M in
all), one query batch each (4k-4M queries, 40.3M in all, closest point / ray / silhouette), enqueued
as one batch.  Served: 16 resident scenes (10k-2M primitives), 400 requests (256 to 262,144 queries,
closest point 60% / ray 25% / silhouette 15%) through the serving loop.

Every job / request is compared field by field, bit for bit, with the same query run alone (the
single-batch call), on the device; a sample of every job / request is checked against the oracle
(the FP32 restatement of the reference traversal) and the FP64 brute force with the acceptance rule
of tests/parity.py, and closest-point distances against the float32 brute force (exact pruning).
"""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

from gquery_amd import synth
from oracle import oracle as O
from tests import parity

pytestmark = pytest.mark.gpu

FIELDS = ("valid", "p", "n", "uv", "t", "d", "prim_id", "on_boundary")
SAMPLE = 192


def _make(sp):
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    V, F = synth.scene_geometry(sp)
    torch.cuda.set_device(0)
    return V, F, (Mesh if sp["dim"] == 3 else Polyline)(V, F, device="cuda:0")


def _build(specs):
    # the CPU builders release the GIL: build on host threads
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(_make, specs))


def _single(shape, kind, x, v):
    if kind == "cp":
        return shape.closest_point_bvh(x)
    if kind == "ray":
        return shape.intersect_bvh(x, v)
    return shape.closest_silhouette_snch(x)


def _equal_on_device(a, b, tag):
    for f in FIELDS:
        if hasattr(b, f):
            ga, gb = getattr(a, f), getattr(b, f)
            assert ga.shape == gb.shape and torch.equal(ga.view(torch.uint8), gb.view(torch.uint8)), (tag, f)


def _check_sample(V, F, shape, kind, x, v, rec, idx, tag):
    """Records of the queries `idx` of one job against the oracle and the FP64 brute force."""
    dim = V.shape[1]
    s = shape.snch
    xs = x[idx].cpu().numpy()
    got = {f: getattr(rec, f)[torch.from_numpy(idx).cuda()].cpu().numpy() for f in FIELDS if hasattr(rec, f)}
    if kind == "cp":
        ref = O.closest_point(dim, s.bvh_node_data(), s.primitive_data(), xs)
        b = O.bf_closest_point(dim, V, F, xs)
        gap = parity.gap_mask(b["d"], b["second_d"])
        exact = O.bf_closest_point_f32(dim, s.primitive_data(), xs)
        # the leaf distance of the chosen primitive (the 3D record's d; 2D records use
        # LineSegment.closest_point's formula) equals the float32 brute-force minimum
        ld = got["d"] if dim == 3 else O.leaf_distance(dim, s.primitive_data(), got["prim_id"], xs)
        assert np.array_equal(ld.view(np.int32), exact["d"].view(np.int32)), tag
        assert parity.rel_err(got["d"], b["d"]).max() <= parity.TOL_F64, tag
        res = parity.summarize(tag, got, ref, gap)
        assert res["ok"], parity.brief(res)
    elif kind == "ray":
        vs = v[idx].cpu().numpy()
        ref = O.intersect(dim, s.bvh_node_data(), s.primitive_data(), xs, vs)
        b = O.bf_intersect(dim, V, F, xs, vs)
        gap = parity.gap_mask(b["t"], b["second_t"])
        keep = np.ones(len(idx), bool)
        if dim == 3:
            assert np.array_equal(got["valid"].astype(bool), np.isfinite(b["t"])), tag
            keep = parity.ray3d_vs_fp64(got, ref, b, gap)
        sub = lambda r: {k: a[keep] for k, a in r.items() if isinstance(a, np.ndarray) and len(a) == len(idx)}
        res = parity.summarize(tag, sub(got), sub(ref), gap[keep])
        assert res["ok"], parity.brief(res)
    else:
        ref = O.silhouette(dim, s.node_data(), s.silhouette_data(), xs)
        b = O.bf_silhouette(dim, s.raw_silhouette_data(), xs)
        gap = parity.gap_mask(b["d"], b["second_d"]) & (np.abs(ref["d"] - b["d"]) <= 1e-6 * np.maximum(b["d"], 1.0))
        # the engine's FP64 sign predicate may overrule a float32 sign of the reference (nearly
        # degenerate edges / vertices): every such difference must be the FP64 answer
        keep = np.ones(len(idx), bool)
        if dim == 2:
            diff, ok = parity.sil2d_differences_explained(s.silhouette_data(), xs, got, ref)
            assert ok.all(), (tag, diff[~ok])
        else:
            gv, rv = got["valid"].astype(bool), ref["valid"].astype(bool)
            diff = np.flatnonzero((gv != rv) | (gv & rv & (parity.rel_err(got["d"], ref["d"]) > parity.TOL_REF)))
            ok = gv[diff] & (~rv[diff] | (got["d"][diff] < ref["d"][diff])) & (
                parity.rel_err(got["d"][diff], b["d"][diff]) <= parity.TOL_F64)
            assert ok.all(), (tag, diff[~ok])
        assert len(diff) <= max(2, len(idx) // 100), (tag, len(diff))
        keep[diff] = False
        sub = lambda r: {k: a[keep] for k, a in r.items() if isinstance(a, np.ndarray) and len(a) == len(idx)}
        res = parity.summarize(tag, sub(got), sub(ref), gap[keep])
        assert res["ok"], parity.brief(res)


@pytest.fixture(scope="module")
def batched():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    specs = synth.batched_workload()
    scenes = _build(specs)
    jobs = []
    for sp, (V, F, shape) in zip(specs, scenes):
        n, dim = sp["queries"], sp["dim"]
        x = torch.from_numpy(synth.uniform_points(n, dim, seed=sp["query_seed"])).cuda()
        v = (torch.from_numpy(synth.uniform_directions(n, dim, seed=sp["query_seed"] + 1)).cuda()
             if sp["kind"] == "ray" else None)
        jobs.append((shape, sp["kind"], x, v))
    return specs, scenes, jobs


def test_config4_full_batch_equals_single_calls(gpu, batched):
    from gquery_amd.batch import Batch
    specs, scenes, jobs = batched
    assert sum(sp["prims"] for sp in specs) > 15_000_000 and sum(sp["queries"] for sp in specs) > 35_000_000
    print(f"[config4] {len(jobs)} jobs, eager batch", flush=True)
    b = Batch(jobs)
    got = b.run()
    torch.cuda.synchronize()
    for k, ((shape, kind, x, v), r) in enumerate(zip(jobs, got)):
        _equal_on_device(r, _single(shape, kind, x, v), (k, kind))
    torch.cuda.synchronize()


def test_config4_full_graph_replay_vs_oracle(gpu, batched):
    from gquery_amd.batch import Batch
    specs, scenes, jobs = batched
    print(f"[config4] graph capture + 2 replays", flush=True)
    b = Batch(jobs).capture()
    b.replay()
    torch.cuda.synchronize()
    first = [{f: getattr(r, f).clone() for f in FIELDS if hasattr(r, f)} for r in b.results]
    torch.randn(1 << 20, device="cuda:0").sum().item()   # eager work between the replays
    b.replay()
    torch.cuda.synchronize()
    rng = np.random.default_rng(11)
    for k, ((shape, kind, x, v), r, sp, (V, F, _)) in enumerate(zip(jobs, b.results, specs, scenes)):
        for f, a in first[k].items():
            assert torch.equal(getattr(r, f).view(torch.uint8), a.view(torch.uint8)), (k, f)
        idx = np.sort(rng.choice(x.shape[0], min(SAMPLE, x.shape[0]), replace=False))
        _check_sample(V, F, shape, kind, x, v, r, idx, f"config4/job{k}/{kind}/{sp['dim']}D/{sp['prims']}")


def test_config5_served_requests_equal_alone_and_oracle(gpu):
    from gquery_amd.serve import GpuExecutor, Request, Scheduler, run_open_loop
    scene_specs, req_specs = synth.served_workload()
    scenes = _build(scene_specs)
    shapes = [s for _, _, s in scenes]
    reqs = []
    for i, rs in enumerate(req_specs):
        dim = scene_specs[rs["scene"]]["dim"]
        x = torch.from_numpy(synth.uniform_points(rs["n"], dim, seed=rs["query_seed"])).cuda()
        v = (torch.from_numpy(synth.uniform_directions(rs["n"], dim, seed=rs["query_seed"] + 1)).cuda()
             if rs["kind"] == "ray" else None)
        reqs.append(Request(rid=i, scene=rs["scene"], kind=rs["kind"], x=x, v=v))
    assert sum(r.n for r in reqs) > 10_000_000
    print(f"[config5] {len(reqs)} requests over {len(shapes)} scenes", flush=True)
    ex = GpuExecutor(shapes, slots=(2, 1))
    free, _ = torch.cuda.mem_get_info()
    sched = Scheduler(ex, memory_budget=free // 4, slots=(2, 1))
    arrivals = np.sort(np.random.default_rng(6).uniform(0.0, 1.0, size=len(reqs)))
    run_open_loop(sched, reqs, arrivals)
    torch.cuda.synchronize()
    assert len(sched.completed) == len(reqs) and ex.launches < len(reqs)
    rng = np.random.default_rng(12)
    for r in reqs:
        V, F, shape = scenes[r.scene]
        _equal_on_device(r.result, _single(shape, r.kind, r.x, r.v), (r.rid, r.kind))
        idx = np.sort(rng.choice(r.n, min(48, r.n), replace=False))
        _check_sample(V, F, shape, r.kind, r.x, r.v, r.result, idx, f"config5/req{r.rid}/{r.kind}")
