"""Autotuner: picks the engine's per-scene settings by timing candidates on the device.

Knobs (all exact: every candidate returns the same records):
  * leaf_prims     -- primitives per leaf of the closest-point BVH4 (8, 16, 24 or 32);
  * ray_leaf_prims -- primitives per leaf of the rays' BVH4 (2, 3, 4 or 6);
  * packet_budget  -- packet steps a wavefront takes before its queries move to the
                      wave-per-query pass (how much work runs wave-per-query vs 64-lane packets);
  * traversal      -- silhouettes: packet + wave-per-query pass, or lane per query.
The traversal stack depth is not a knob: it is fixed by the tree (the engine sizes the LDS stacks
and frontiers from the scene's depth at launch).

    best = tune_shape(V, F, queries={"cp": x_cp, "sil": x_sil})
    mesh = Mesh(V, F, **best["options"])        # leaf_prims, ray_leaf_prims, packet_budget, traversal

Timing uses a sample of the caller's queries (or uniform points in the scene box enlarged by
25%), median of a few runs per candidate after a warm-up.
"""
from __future__ import annotations

import concurrent.futures as cf
import itertools
import time

import numpy as np
import torch

LEAVES = (8, 16, 24, 32)
RAY_LEAVES = (2, 3, 4, 6)
BUDGETS = (128, 350, 1000, 3000)
TRAVERSALS = (1, 2)


def _time(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def _run(scene, kind, x, v):
    if kind == "cp":
        return scene.closest_point(x)
    if kind == "ray":
        return scene.intersect(x, v)
    return scene.closest_silhouette(x)


def sample_queries(V, n, seed=0):
    """Uniform points in the scene's box enlarged by 25% (and unit directions)."""
    V = np.asarray(V, np.float32)
    lo, hi = V.min(0), V.max(0)
    c, h = 0.5 * (lo + hi), 0.625 * (hi - lo)
    rng = np.random.default_rng(seed)
    x = (c + rng.uniform(-1.0, 1.0, size=(n, V.shape[1])) * h).astype(np.float32)
    d = rng.standard_normal(size=(n, V.shape[1])).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return x, d


def tune_shape(V, F, queries=None, kinds=("cp", "ray", "sil"), n_sample: int = 1 << 18,
               device="cuda:0", leaves=LEAVES, budgets=BUDGETS, traversals=TRAVERSALS, log=None,
               ray_leaves=RAY_LEAVES):
    """Returns {"options": {...}, "ms": {kind: ms}, "trials": [...]} for the polyline (V: n x 2) or
    mesh (V: n x 3) with primitives F.

    queries: {kind: (x, v)} device tensors to time on (a sample of the real workload); otherwise
    uniform points.  Cost: one CPU build of the SNCH, one device scene (engine layout + upload) per
    candidate, and a few timed calls per candidate and kind."""
    from . import ext
    from .scene import Scene
    dim = np.asarray(V).shape[1]
    structure = (ext.SNCH3D if dim == 3 else ext.SNCH)(V, F)
    if queries is None:
        x, v = sample_queries(V, n_sample)
        x, v = torch.from_numpy(x).to(device), torch.from_numpy(v).to(device)
        queries = {k: (x, v) for k in kinds}
    trials = []
    # rays: their own tree's leaf size (the packet budget and traversal do not apply)
    ray_best = None
    if "ray" in kinds:
        for rl in ray_leaves:
            scene = Scene(structure, device, ray_leaf_prims=rl)
            x, v = queries["ray"]
            t_ms = _time(lambda: _run(scene, "ray", x, v))
            trials.append({"ray_leaf_prims": rl, "ms": {"ray": t_ms}})
            if ray_best is None or t_ms < ray_best[1]:
                ray_best = (rl, t_ms)
            del scene
    for leaf in leaves:
        for budget, trav in itertools.product(budgets, traversals):
            if trav == 2 and budget != budgets[0]:
                continue        # lane per query ignores the budget: time it once
            scene = Scene(structure, device, leaf, budget, trav)
            ms = {}
            for k in kinds:
                if k == "ray" or (k != "sil" and trav == 2):
                    continue
                x, v = queries[k]
                ms[k] = _time(lambda: _run(scene, k, x, v))
            trials.append({"leaf_prims": leaf, "packet_budget": budget, "traversal": trav, "ms": ms})
            if log:
                log(f"tune leaf={leaf} budget={budget} traversal={trav}: {ms}")
            del scene
    # closest-point leaf size and budget: minimise the closest-point time; silhouettes pick their
    # traversal (and budget) on the chosen leaf size
    packet = [t for t in trials if t.get("traversal") == 1]
    if "cp" in kinds:
        best = min(packet, key=lambda t: t["ms"]["cp"])
    else:
        best = min(packet, key=lambda t: t["ms"].get("sil", 0.0))
    opts = {"leaf_prims": best["leaf_prims"], "packet_budget": best["packet_budget"], "traversal": 1}
    ms = {"cp": best["ms"]["cp"]} if "cp" in kinds else {}
    if ray_best is not None:
        opts["ray_leaf_prims"] = ray_best[0]
        ms["ray"] = ray_best[1]
    if "sil" in kinds:
        same_leaf = [t for t in trials if t.get("leaf_prims") == best["leaf_prims"] and "sil" in t["ms"]]
        s = min(same_leaf, key=lambda t: t["ms"]["sil"])
        opts["traversal"] = s["traversal"]
        ms["sil"] = s["ms"]["sil"]
        if "cp" not in kinds:
            opts["packet_budget"] = s["packet_budget"]
    return {"options": opts, "ms": ms, "trials": trials}



def tune_jobs(jobs, leaves=LEAVES, budgets=BUDGETS, traversals=TRAVERSALS, method: str = "engine", log=None,
              ray_leaves=RAY_LEAVES):
    """Options for a batch of jobs [(shape, kind, x, v), ...] (Mesh / Polyline shapes), timed as they
    will run: jobs are grouped by (dimension, query kind, primitive count within a factor of 4), and
    for every leaf size each group's jobs run as one gq_batch_enqueue batch with every packet
    budget / traversal candidate as per-job overrides (one device scene per shape and leaf size; ray
    groups take their leaf size from ray_leaves).  Returns one options dict per job (leaf_prims or
    ray_leaf_prims, packet_budget, traversal)."""
    from .batch import Batch
    from .scene import Scene
    groups = {}
    for k, (shape, kind, x, v) in enumerate(jobs):
        n_prims = int(shape.indices.shape[0])
        key = (shape.scene.dim, kind, int(np.log2(max(n_prims, 1)) // 2))
        groups.setdefault(key, []).append(k)
    best = {key: (float("inf"), None) for key in groups}
    for i in range(max(len(leaves), len(ray_leaves))):
        leaf, ray_leaf = leaves[min(i, len(leaves) - 1)], ray_leaves[min(i, len(ray_leaves) - 1)]
        t0 = time.perf_counter()

        def make(shape):
            # host threads: the engine layout build releases the GIL
            torch.cuda.set_device(shape.scene.device)
            return Scene(shape.snch, shape.scene.device, leaf, ray_leaf_prims=ray_leaf)

        with cf.ThreadPoolExecutor(max_workers=8) as ex:
            scenes = dict(enumerate(ex.map(make, [j[0] for j in jobs])))
        if log:
            log(f"tune leaf={leaf} ray_leaf={ray_leaf}: {len(scenes)} scenes in {time.perf_counter() - t0:.0f}s")
        for key, ks in groups.items():
            kind = key[1]
            for budget, trav in itertools.product(budgets, traversals):
                if (kind != "sil" and trav == 2) or ((trav == 2 or kind == "ray") and budget != budgets[0]):
                    continue    # traversal only matters for silhouettes; lane-per-query kernels ignore the budget
                b = Batch([dict(shape=scenes[k], kind=kind, x=jobs[k][2], v=jobs[k][3], packet_budget=budget,
                                traversal=trav) for k in ks], method=method)
                ms = _time(b.run)
                if ms < best[key][0]:
                    lk = {"ray_leaf_prims": ray_leaf} if kind == "ray" else {"leaf_prims": leaf}
                    best[key] = (ms, {**lk, "packet_budget": budget, "traversal": trav})
                del b
            if log:
                log(f"tune leaf={leaf} ray_leaf={ray_leaf} group dim={key[0]} kind={key[1]} "
                    f"prims~4^{key[2]}: best so far {best[key][0]:.3f} ms")
        del scenes
    out = [None] * len(jobs)
    for key, ks in sorted(groups.items()):
        if log:
            log(f"tune group dim={key[0]} kind={key[1]} prims~4^{key[2]} ({len(ks)} jobs): {best[key][1]} "
                f"{best[key][0]:.3f} ms")
        for k in ks:
            out[k] = dict(best[key][1])
    return out
