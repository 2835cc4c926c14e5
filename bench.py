"""Benchmark: Mqueries/s of 3D closest-point queries on one MI355X.

Workload (the headline): displaced icosphere, 7 subdivisions (163,842 vertices, 327,680
triangles), 1,048,576 query points uniform in [-1.5, 1.5]^3, all resident in HBM before timing.
A step is one Mesh.closest_point_bvh call over the whole batch.  The CPU baseline is the FP32
restatement of the reference traversal (oracle/gq_oracle.c) on the host cores, over a bounded
sample of the same queries.  The same run times the plain FP32 kernels on the same batch
(`vs_plain`), every other query path on its configuration workload (`paths`: 3D rays and
silhouettes on the same mesh, 2D closest point / rays / silhouettes on the 65,536-segment star,
closest point unsorted and below the sort threshold), the uneven batched workload (`batched`: 64
scenes, one graph replay per step) and the serving loop (16 resident scenes, Poisson requests at
25/50/90% of the saturated rate: p50/p99 latency under `served`).

The roofline names the dominant kernel's bound: the closest-point packet kernel is limited by VALU
issue, so `roofline` is its VALU instruction rate (instructions per launch from the committed
rocprofv3 SQ counter summary, kernel time measured in this run) against the issue peak, with the
HBM traffic of the traversal kernels (counter summary) against 8 TB/s beside it.

    python bench.py [--steps K] [--warmup W] [--method engine|engine_fp32|plain]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# VALU issue peak of MI355X in wave64 instructions per second.  The published FP32 vector peak,
# 157.3 TFLOP/s, is packed FMA (v_pk_fma_f32: 2 FMAs per lane) at one wave64 instruction per SIMD
# every 4 clocks: 1,024 SIMDs x 2.4e9 / 4 = 614.4e9 instructions per second.  The probe built with
# the library (gquery_amd/valu_probe, source csrc/valu_probe.hip; profiles/valu_issue_rate.md)
# measures the same 4-clock cadence for v_fma_f32 and about 2 clocks for v_sub_f32 / v_max_f32, so
# instructions x 4 clocks over the SIMDs' clocks is an upper bound on the VALU busy fraction (exact
# for an FMA-only stream, twice too high for an add/max-only one).
VALU_PEAK_GIPS = 157.3e12 / (64 * 2 * 2) / 1e9
PATH_STEPS = 10
PLAIN_STEPS = 3
N_QUERIES = 1 << 20
SUBDIV = 7

# Reference float32 record sizes (bytes): node 36 B (BVH3D, 9 floats), triangle 40 B (10 floats),
# query point 12 B, ClosestPointRecord3D 41 B (valid 1 + p 12 + n 12 + uv 8 + d 4 + prim_id 4).
S_NODE, S_PRIM, S_IN, S_OUT = 36, 40, 12, 41


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_query(scene, x_sample):
    """B_q = S_in + S_out + S_node * (1 + child-box tests) + S_prim * (leaf primitive tests): the
    reference-layout bytes the reference traversal reads and writes for one query, with the work
    counts taken from the CPU restatement of that traversal on a sample of the queries.  The figure
    is fixed by the workload, not by the engine's own (smaller, quantized) layout."""
    from oracle import oracle as O
    s = scene.snch
    r = O.closest_point(3, s.bvh_node_data(), s.primitive_data(), x_sample)
    visits = float(r["visits"].mean())
    child = float(r["child_tests"].mean())
    leaf = float(r["leaf_tests"].mean())
    bq = S_IN + S_OUT + S_NODE * (1.0 + child) + S_PRIM * leaf
    return bq, {"visits": visits, "child_tests": child, "leaf_tests": leaf}


def cpu_baseline(scene, x, budget_s=10.0):
    """Oracle FP32 traversal on all host threads over a bounded prefix of the batch."""
    from oracle import oracle as O
    s = scene.snch
    nodes, prims = s.bvh_node_data(), s.primitive_data()
    n = 16384
    t0 = time.perf_counter()
    O.closest_point(3, nodes, prims, x[:n])
    dt = time.perf_counter() - t0
    n = int(min(len(x), max(n, n * budget_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    O.closest_point(3, nodes, prims, x[:n])
    dt = time.perf_counter() - t0
    return {"value": n / dt / 1e6, "unit": "Mqueries/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"first {n} of the {len(x)} queries, FP32 restatement of BVH3D.closest_point, "
                      f"OpenMP over {O.num_threads()} host threads, {dt:.1f} s"}


def _pmc_table(d):
    """{kernel name: {counter: mean value per dispatch, counter + "_max": largest dispatch}} from
    rocprofv3 csv counter output (a kernel dispatched twice per call, such as a wave-per-query pass
    with its hand-off pass, has its main dispatch in the _max entry)."""
    import csv
    acc = {}
    for f in Path(d).rglob("*counter_collection.csv"):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                k = (row["Kernel_Name"], row["Counter_Name"])
                acc.setdefault(k, {}).setdefault(row.get("Dispatch_Id", len(acc)), 0.0)
                acc[k][row.get("Dispatch_Id")] += float(row["Counter_Value"])
    out = {}
    for (kn, cn), per in acc.items():
        out.setdefault(kn, {})[cn] = sum(per.values()) / max(len(per), 1)
        out[kn][cn + "_max"] = max(per.values()) if per else 0.0
    return out


def pmc_summary(args):
    """Per-launch HBM traffic of the traversal kernels, each counter scaled by what it reads for
    the elementwise copy of known size (copy_mb MiB read + copy_mb MiB written)."""
    fetch, write = _pmc_table(args.pmc_summary[0]), _pmc_table(args.pmc_summary[1])
    copy_bytes = float(args.copy_mb << 20)
    ck = [k for k in fetch if "elementwise" in k.lower() or "mul" in k.lower()]
    if not ck:
        raise SystemExit("no elementwise copy kernel in the counter output (run with --copy-mb)")
    ck = max(ck, key=lambda k: fetch[k].get("FETCH_SIZE", 0.0))
    f_scale = copy_bytes / (fetch[ck]["FETCH_SIZE"] * 1024.0)
    w_scale = copy_bytes / (write[ck]["WRITE_SIZE"] * 1024.0)
    trav = [k for k in fetch if k.startswith("void gqa::") or "gqa::" in k]
    rows, tot_f, tot_w = [], 0.0, 0.0
    for k in sorted(trav):
        if not any(t in k for t in ("k_cp_", "k_ray_", "k_sil_", "k_zero")):
            continue
        fb = fetch[k].get("FETCH_SIZE", 0.0) * 1024.0 * f_scale
        wb = write.get(k, {}).get("WRITE_SIZE", 0.0) * 1024.0 * w_scale
        rows.append((k, fb, wb))
        tot_f += fb
        tot_w += wb
    tag = f"traffic_cp3d_s7_1M_{args.method}"
    res = {"hbm_bytes_per_launch": tot_f + tot_w, "fetch_bytes": tot_f, "write_bytes": tot_w,
           "fetch_scale": f_scale, "write_scale": w_scale, "copy_kernel": ck[:120],
           "kernels": {k[:120]: {"fetch_bytes": fb, "write_bytes": wb} for k, fb, wb in rows}}
    if len(args.pmc_summary) > 2:
        # L2 (TCC) hit rate per kernel: hits / (hits + misses) over all L2 requests of a dispatch
        l2 = _pmc_table(args.pmc_summary[2])
        hits = misses = 0.0
        for k, _, _ in rows:
            h, m = l2.get(k, {}).get("TCC_HIT_sum", 0.0), l2.get(k, {}).get("TCC_MISS_sum", 0.0)
            res["kernels"][k[:120]]["l2_hit_rate"] = h / (h + m) if h + m else None
            hits, misses = hits + h, misses + m
        res["l2_hit_rate"] = hits / (hits + misses) if hits + misses else None
    (ROOT / "profiles" / f"{tag}.json").write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))


def _engine_sha():
    import hashlib
    return hashlib.sha256((ROOT / "gquery_amd" / "csrc" / "engine.hip").read_bytes()).hexdigest()[:16]


def sq_summary(args):
    """No GPU: SQ counters of rocprofv3 --pmc runs (csv; each its own pass) of a bench command into
    profiles/counters_<tag>.json: mean per dispatch for every traversal kernel, with the hash of the
    engine source the run was built from (bench.py reports whether it still matches)."""
    merged = {}
    for d in args.sq_summary:
        for k, c in _pmc_table(d).items():
            if "gqa::" not in k or not any(t in k for t in ("k_cp_", "k_ray_", "k_sil_")):
                continue
            merged.setdefault(k[:160], {}).update(c)
    res = {"engine_src_sha": _engine_sha(), "tag": args.sq_tag, "kernels": merged}
    (ROOT / "profiles" / f"counters_{args.sq_tag}.json").write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))


def valu_roofline(method, packet_ms, traversal_ms, kernel_name):
    """VALU-issue roofline of the packet kernel from the committed SQ counter summary (instructions
    per launch) and the kernel time measured in this run, with the HBM traffic of all traversal
    kernels (FETCH_SIZE + WRITE_SIZE summary) against the HBM peak beside it."""
    out = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_GIPS, "unit": "G wave64 VALU instr/s",
           "frac": None, "traffic": None, "kernel": kernel_name, "kernel_ms": packet_ms,
           "traversal_ms": traversal_ms}
    cf = ROOT / "profiles" / f"counters_cp3d_s7_1M_{method}.json"
    if cf.exists():
        c = json.loads(cf.read_text())
        pk = [k for k in c["kernels"] if "k_cp_packet<3" in k]
        if pk and packet_ms:
            valu = c["kernels"][pk[0]].get("SQ_INSTS_VALU")
            out["valu_instr_per_launch"] = valu
            out["achieved"] = valu / (packet_ms * 1e-3) / 1e9
            out["frac"] = out["achieved"] / VALU_PEAK_GIPS
            out["issue_model"] = ("4 clocks per wave64 VALU instruction (the FMA rate); add/max/compare "
                                  "issue in about 2, so frac bounds the VALU busy fraction from above and "
                                  "frac/2 from below")
            out["counters"] = f"profiles/{cf.name}"
            out["counters_match_build"] = c.get("engine_src_sha") == _engine_sha()
    tf = ROOT / "profiles" / f"traffic_cp3d_s7_1M_{method}.json"
    if tf.exists():
        t = json.loads(tf.read_text())
        out["traffic"] = t.get("hbm_bytes_per_launch")
        if out["traffic"] and traversal_ms:
            gbs = out["traffic"] / (traversal_ms * 1e-3) / 1e9
            out["hbm"] = {"achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
                          "l2_hit_rate": t.get("l2_hit_rate"), "summary": f"profiles/{tf.name}"}
    return out


def time_call(fn, steps=PATH_STEPS, warmup=2):
    """(ms per step from events around the steps, traversal ms and first-kernel ms per launch from
    the library's own events on the launch stream)."""
    import torch
    from gquery_amd import _lib
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    _lib.profile_reset()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    tot, first, n, name = _lib.profile_read_split()
    return e0.elapsed_time(e1) / steps, tot / max(n, 1), first / max(n, 1), name


def query_paths(mesh, poly, x3, v3, x2, v2, log):
    """Every query path on its configuration workload: 3D on the s=7 icosphere and 2D on the
    65,536-segment star, 1,048,576 queries each; plus the closest-point query unsorted (caller's
    order, incoherent packets) and below the sort threshold."""
    from gquery_amd.batch import Batch
    from gquery_amd.shapes.mesh import Mesh
    nq = x3.shape[0]
    from gquery_amd.shapes.polyline import Polyline
    mesh_q8 = Mesh(None, None, device="cuda:0", structure=mesh.snch, sil_pairs=1)
    poly_q16 = Polyline(None, None, device="cuda:0", structure=poly.snch, sil_pairs=1)
    cases = {
        "cp3_engine_fp32": lambda: mesh.closest_point_bvh(x3, method="engine_fp32"),
        "ray3": lambda: mesh.intersect_bvh(x3, v3),
        "sil3": lambda: mesh.closest_silhouette_snch(x3),
        "sil3_pairs_q8": lambda: mesh_q8.closest_silhouette_snch(x3),
        "cp2": lambda: poly.closest_point_bvh(x2),
        "ray2": lambda: poly.intersect_bvh(x2, v2),
        "sil2": lambda: poly.closest_silhouette_snch(x2),
        "sil2_engine_fp32": lambda: poly.closest_silhouette_snch(x2, method="engine_fp32"),
        "sil2_pairs_q16": lambda: poly_q16.closest_silhouette_snch(x2),
    }
    b_unsorted = Batch([{"shape": mesh, "kind": "cp", "x": x3, "order": 1}])
    cases["cp3_input_order"] = b_unsorted.run
    small = x3[:8000]
    cases["cp3_8000_queries"] = lambda: mesh.closest_point_bvh(small)
    out = {}
    for name, fn in cases.items():
        ms, trav, first, kern = time_call(fn)
        n = 8000 if name == "cp3_8000_queries" else nq
        out[name] = {"ms_per_step": ms, "Mq_per_s": n / (ms * 1e-3) / 1e6, "traversal_ms": trav,
                     "first_kernel_ms": first, "queries": n}
        log(f"path {name}: {ms:.3f} ms/step, {out[name]['Mq_per_s']:.1f} Mq/s (traversal {trav:.3f} ms, "
            f"first kernel {first:.3f} ms)")
    return out


def accuracy(args):
    """Per query type and method: errors against the FP64 brute force (oracle, double precision over
    every primitive) and the FP32 restatement of the reference traversal, on N queries sampled from
    the headline distribution.  Prints one JSON object per (scene, query, method)."""
    import numpy as np
    import torch
    from gquery_amd import synth
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    from oracle import oracle as O
    n = args.accuracy
    scenes = [("mesh s=%d" % args.subdiv, 3, *synth.displaced_icosphere(args.subdiv)),
              ("polyline %d" % args.segments, 2, *synth.star_polyline(args.segments))]
    for name, dim, V, F in scenes:
        shape = (Mesh if dim == 3 else Polyline)(V, F, device="cuda:0")
        s = shape.snch
        shape_q = (Mesh if dim == 3 else Polyline)(None, None, device="cuda:0", structure=s, sil_pairs=1)
        x = synth.uniform_points(n, dim, seed=101)
        v = synth.uniform_directions(n, dim, seed=102)
        for kind in ("cp", "ray", "sil"):
            t0 = time.perf_counter()
            if kind == "cp":
                ref = O.closest_point(dim, s.bvh_node_data(), s.primitive_data(), x)
                bf = O.bf_closest_point(dim, V, F, x)
                bd, b2 = bf["d"], bf["second_d"]
            elif kind == "ray":
                ref = O.intersect(dim, s.bvh_node_data(), s.primitive_data(), x, v)
                bf = O.bf_intersect(dim, V, F, x, v)
                bd, b2 = bf["t"], bf["second_t"]
            else:
                ref = O.silhouette(dim, s.node_data(), s.silhouette_data(), x)
                bf = O.bf_silhouette(dim, s.raw_silhouette_data(), x)
                bd, b2 = bf["d"], bf["second_d"]
            bvalid = np.isfinite(bd)
            with np.errstate(invalid="ignore"):
                gap = bvalid & (~np.isfinite(b2) | (b2 - bd > 1e-5 * np.maximum(bd, 1e-30)))
            log(f"{name} {kind}: oracle + FP64 brute force {time.perf_counter() - t0:.1f}s")
            methods = ("plain", "engine_fp32", "engine") + (("engine_quantized_pairs",) if kind == "sil" else ())
            for method in methods:
                if kind == "cp":
                    r = shape.closest_point_bvh(x, method=method)
                elif kind == "ray":
                    r = shape.intersect_bvh(x, v, method=method)
                elif method == "engine_quantized_pairs":
                    # the quantized pair records (8-bit boxes in 3D, 16-bit in 2D, int16 cones)
                    r = shape_q.closest_silhouette_snch(x, method="engine")
                else:
                    r = shape.closest_silhouette_snch(x, method=method)
                torch.cuda.synchronize()
                gv = r.valid.cpu().numpy().astype(bool)
                gd = r.d.cpu().numpy().astype(np.float64)
                gid = r.prim_id.cpu().numpy()
                both = gv & bvalid
                same = both & gap & (gid == bf["prim_id"])
                rv = gv & ref["valid"].astype(bool)
                # error relative to max(d, 1): relative beyond the scene scale (radius ~1), absolute
                # below it, where float32 coordinates carry absolute (not relative) error
                rel64 = np.abs(gd - bd) / np.maximum(np.abs(np.nan_to_num(bd)), 1.0)
                rel32 = np.abs(gd - ref["d"]) / np.maximum(np.abs(ref["d"]), 1.0)
                out = {"scene": name, "query": kind, "method": method, "n": n,
                       "valid_vs_fp64": float((gv == bvalid).mean()),
                       "id_vs_fp64_where_gap": float(same.sum() / max((both & gap).sum(), 1)),
                       "id_vs_fp32_restatement_where_gap": float((rv & gap & (gid == ref["prim_id"])).sum()
                                                                 / max((rv & gap).sum(), 1)),
                       "n_gap": int((both & gap).sum()),
                       "max_err_d_vs_fp64": float(rel64[same].max()) if same.any() else 0.0,
                       "mean_err_d_vs_fp64": float(rel64[same].mean()) if same.any() else 0.0,
                       "p99_err_d_vs_fp64": float(np.percentile(rel64[same], 99)) if same.any() else 0.0,
                       "max_err_d_vs_fp32_restatement": float(rel32[rv].max()) if rv.any() else 0.0,
                       "mean_err_d_vs_fp32_restatement": float(rel32[rv].mean()) if rv.any() else 0.0}
                if kind == "sil":
                    # where the reference traversal itself (its cone test) misses the FP64 nearest
                    # silhouette; the engine follows the reference there
                    ok_ref = ref["valid"].astype(bool) & bvalid & gap
                    out["reference_misses_fp64_nearest"] = float((ok_ref & (ref["prim_id"] != bf["prim_id"])).sum()
                                                                 / max(ok_ref.sum(), 1))
                if "p" in bf:
                    # point error: Euclidean distance to the FP64 point, over max(d, 1) like d
                    gp = r.p.cpu().numpy().astype(np.float64).reshape(n, dim)
                    pe = np.linalg.norm(gp - bf["p"], axis=1) / np.maximum(np.abs(np.nan_to_num(bd)), 1.0)
                    out["max_err_p_vs_fp64"] = float(pe[same].max()) if same.any() else 0.0
                    out["mean_err_p_vs_fp64"] = float(pe[same].mean()) if same.any() else 0.0
                print(json.dumps(out), flush=True)


def batched(args):
    """64 scenes (half 2D stars, half 3D icospheres, 1k-2M primitives), one query batch each
    (4k-4M queries, closest point / ray / silhouette), all enqueued as one batch recorded in a
    graph; a step is one graph replay.  The same jobs as sequential single calls are timed too."""
    import numpy as np
    import torch
    from gquery_amd import synth
    from gquery_amd.batch import Batch
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    torch.cuda.set_device(0)
    specs = synth.batched_workload(scale=args.scale)
    import concurrent.futures as cf
    t0 = time.perf_counter()

    def make(sp):
        V, F = synth.scene_geometry(sp)
        torch.cuda.set_device(0)
        return (Mesh if sp["dim"] == 3 else Polyline)(V, F, device="cuda:0")

    # the CPU builders release the GIL: build the scenes on host threads
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        shapes = list(ex.map(make, specs))
    jobs = []
    for sp, shape in zip(specs, shapes):
        n, dim = sp["queries"], sp["dim"]
        x = torch.from_numpy(synth.uniform_points(n, dim, seed=sp["query_seed"])).cuda()
        v = torch.from_numpy(synth.uniform_directions(n, dim, seed=sp["query_seed"] + 1)).cuda() if sp["kind"] == "ray" else None
        jobs.append((shape, sp["kind"], x, v))
    n_prims = sum(sp["prims"] for sp in specs)
    n_q = sum(sp["queries"] for sp in specs)
    log(f"{len(specs)} scenes, {n_prims / 1e6:.2f}M primitives, {n_q / 1e6:.2f}M queries, built in "
        f"{time.perf_counter() - t0:.0f}s")
    tuned = None
    if args.tune:
        from gquery_amd import tune
        t1 = time.perf_counter()
        tuned = tune.tune_jobs(jobs, log=log)

        def apply(a):
            torch.cuda.set_device(0)
            a[0][0].set_options(**a[1])

        with cf.ThreadPoolExecutor(max_workers=8) as ex:
            list(ex.map(apply, zip(jobs, tuned)))
        log(f"autotuned {len(set(map(str, tuned)))} option sets in {time.perf_counter() - t1:.0f}s")
    stream = torch.cuda.current_stream()

    def single_all():
        for shape, kind, x, v in jobs:
            if kind == "cp":
                shape.closest_point_bvh(x, method=args.method)
            elif kind == "ray":
                shape.intersect_bvh(x, v, method=args.method)
            else:
                shape.closest_silhouette_snch(x, method=args.method)

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    method = args.method if args.method != "plain" else "engine"
    b = Batch(jobs, method=method)
    b.capture()
    ms_graph = timed(b.replay, args.steps)
    ms_seq = timed(single_all, max(1, args.steps // 2))
    # the graph's results equal the single calls' (checked on a few jobs)
    b.replay()
    torch.cuda.synchronize()
    for (shape, kind, x, v), r in list(zip(jobs, b.results))[:6]:
        ref = (shape.closest_point_bvh(x, method=method) if kind == "cp" else
               shape.intersect_bvh(x, v, method=method) if kind == "ray" else
               shape.closest_silhouette_snch(x, method=method))
        torch.cuda.synchronize()
        assert torch.equal(ref.d, r.d) and torch.equal(ref.prim_id, r.prim_id), "batch differs from single calls"
    by_kind = {k: sum(sp["queries"] for sp in specs if sp["kind"] == k) for k in ("cp", "ray", "sil")}
    out = {"metric": "Mqueries/s (uneven batch: 64 scenes, 1k-2M primitives, 4k-4M queries each, one graph)",
           "value": n_q / (ms_graph * 1e-3) / 1e6, "unit": "Mqueries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_graph, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic: seeded stars / displaced icospheres, uniform queries",
           "config": {"workload": f"batched: {len(specs)} scenes, {n_prims} primitives, {n_q} queries "
                                  f"(cp {by_kind['cp']}, ray {by_kind['ray']}, sil {by_kind['sil']}), scale {args.scale}",
                      "method": method, "parallelism": "single GPU"},
           "sequential_single_calls_ms": ms_seq, "graph_speedup_vs_sequential": ms_seq / ms_graph,
           "device_workspace_bytes": b.device_bytes,
           "autotuned_options": None if tuned is None else sorted({json.dumps(o, sort_keys=True) for o in tuned})}
    return out


def served(args, loads=(0.25, 0.5, 0.9)):
    """Serving workload: 16 scenes resident, requests (scene, kind 60/25/15 % cp/ray/sil, 256 to
    262,144 queries) admitted by free device memory and coalesced into batches on two stream
    slots.  The saturated rate is measured with every request present at once; then Poisson
    streams at fractions of it give p50 / p99 latency (completion - arrival) and throughput."""
    import concurrent.futures as cf
    import numpy as np
    import torch
    from gquery_amd import synth
    from gquery_amd.serve import GpuExecutor, Request, Scheduler, percentile, run_open_loop
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    torch.cuda.set_device(0)
    scene_specs, req_specs = synth.served_workload(n_requests=args.requests, scale=args.scale)
    t0 = time.perf_counter()

    cache = Path(args.scene_cache) if args.scene_cache else None

    def make(k_sp):
        k, sp = k_sp
        torch.cuda.set_device(0)
        cls = Mesh if sp["dim"] == 3 else Polyline
        f = cache / f"served_{k}_{sp['dim']}d_{sp['prims']}.gqs" if cache else None
        if f is not None and f.exists():
            loaded.append(k)
            return cls.load(f, device="cuda:0")
        V, F = synth.scene_geometry(sp)
        shape = cls(V, F, device="cuda:0")
        if f is not None:
            shape.save(f)
        return shape

    loaded = []
    if cache:
        cache.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        shapes = list(ex.map(make, enumerate(scene_specs)))
    inputs = []
    for q in req_specs:
        dim = scene_specs[q["scene"]]["dim"]
        x = torch.from_numpy(synth.uniform_points(q["n"], dim, seed=q["query_seed"])).cuda()
        v = torch.from_numpy(synth.uniform_directions(q["n"], dim, seed=q["query_seed"] + 1)).cuda() if q["kind"] == "ray" else None
        inputs.append((x, v))
    torch.cuda.synchronize()
    n_q = sum(q["n"] for q in req_specs)
    setup_s = time.perf_counter() - t0
    log(f"served: {len(shapes)} scenes ({sum(sp['prims'] for sp in scene_specs) / 1e6:.2f}M primitives) "
        f"set up in {setup_s:.1f}s ({len(loaded)} loaded from scene files); "
        f"{len(req_specs)} requests, {n_q / 1e6:.2f}M queries per run")
    slots = tuple(int(a) for a in args.slots.split(","))
    executor = GpuExecutor(shapes, slots=slots, method=args.method if args.method != "plain" else "engine")
    budget = int(torch.cuda.mem_get_info()[0] * 0.5)

    def run(arrivals):
        reqs = [Request(rid=i, scene=q["scene"], kind=q["kind"], x=inputs[i][0], v=inputs[i][1])
                for i, q in enumerate(req_specs)]
        launches0 = executor.launches
        sched = Scheduler(executor, memory_budget=budget, slots=slots, max_batch_queries=args.max_batch)
        run_open_loop(sched, reqs, arrivals)
        torch.cuda.synchronize()
        lat = [r.latency * 1e3 for r in reqs]
        fast = [r.latency * 1e3 for r in reqs if r.kind != "sil"]
        sil = [r.latency * 1e3 for r in reqs if r.kind == "sil"]
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        return {"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99), "mean_ms": float(np.mean(lat)),
                "cp_ray_p50_ms": percentile(fast, 50), "cp_ray_p99_ms": percentile(fast, 99),
                "sil_p50_ms": percentile(sil, 50), "sil_p99_ms": percentile(sil, 99),
                "Mq_per_s": n_q / span / 1e6, "launches": executor.launches - launches0}

    run([0.0] * len(req_specs))                       # warm-up (allocator, code objects)
    sat = run([0.0] * len(req_specs))
    sat_qps = sat["Mq_per_s"] * 1e6
    out = {"saturated": sat}
    rng = np.random.default_rng(synth.WORKLOAD_SEED + 2)
    mean_n = n_q / len(req_specs)
    for load in loads:
        lam = load * sat_qps / mean_n                   # requests per second
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(req_specs)))
        out[f"load_{int(load * 100)}"] = run(list(arrivals - arrivals[0]))
        log(f"load {load:.2f}: {out[f'load_{int(load * 100)}']}")
    p99 = float(out["load_90"]["p99_ms"]) if "load_90" in out else None
    for k in out:
        out[k] = {a: float(b) for a, b in out[k].items()}
    return {"metric": "served p99 latency (ms) at 90% of saturated load", "value": p99, "unit": "ms",
            "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: seeded scenes and Poisson request stream",
            "config": {"workload": f"served: {len(shapes)} scenes, {len(req_specs)} requests per load, "
                                   f"{n_q} queries, memory budget {budget} B, stream slots {slots} (cp+ray, sil)",
                       "scale": args.scale, "scene_setup_s": setup_s, "scenes_loaded": len(loaded)},
            "served": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--method", default="engine", choices=["engine", "engine_fp32", "plain"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain FP32 kernels")
    ap.add_argument("--no-served", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--leaf-prims", type=int, default=0, help="engine BVH4 leaf size (0 = default)")
    ap.add_argument("--traversal", type=int, default=0, help="silhouettes: 0 auto, 1 packet, 2 lane per query")
    ap.add_argument("--wave-front", type=int, default=0,
                    help="first wave-per-query pass frontier entries (0 = default)")
    ap.add_argument("--packet-budget", type=int, default=0,
                    help="packet steps per wavefront before deferral (0 = default, -1 = never)")
    ap.add_argument("--query", default="cp", choices=["cp", "ray", "sil", "sample"],
                    help="query type to time (the headline metric is cp; sample = BVH-guided boundary "
                         "sampling in a ball of --radius)")
    ap.add_argument("--radius", type=float, default=0.5, help="--query sample: ball radius")
    ap.add_argument("--sil-pairs", type=int, default=0, choices=[0, 1],
                    help="3D silhouettes under method engine: 0 float32 pair records, 1 8-bit")
    ap.add_argument("--n", type=int, default=N_QUERIES, help="queries per step")
    ap.add_argument("--subdiv", type=int, default=SUBDIV)
    ap.add_argument("--shape", default="mesh", choices=["mesh", "polyline"],
                    help="polyline: 2D star polyline with --segments segments")
    ap.add_argument("--segments", type=int, default=65536)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after timing, run 3 elementwise float copies of this many MiB (the known-size "
                         "kernel that calibrates FETCH_SIZE / WRITE_SIZE counter runs)")
    ap.add_argument("--sq-summary", nargs="+", metavar="DIR",
                    help="no GPU: rocprofv3 --pmc csv runs with SQ counters into profiles/counters_<--sq-tag>.json")
    ap.add_argument("--sq-tag", default="cp3d_s7_1M_engine")
    ap.add_argument("--no-paths", action="store_true", help="skip timing the other query paths")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-workload leg")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="no GPU: turn two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; csv) of "
                         "`bench.py --copy-mb M`, and optionally a third (TCC_HIT_sum TCC_MISS_sum), "
                         "into profiles/traffic_<workload>_<method>.json")
    ap.add_argument("--accuracy", type=int, default=0, metavar="N",
                    help="instead of timing: error of every method and query type against the FP64 "
                         "brute force and the FP32 restatement on N sampled queries")
    ap.add_argument("--batched", action="store_true",
                    help="uneven batched workload: 64 scenes, one batch each, one graph replay per step")
    ap.add_argument("--scale", type=float, default=1.0, help="--batched / --served: size multiplier")
    ap.add_argument("--served", action="store_true",
                    help="serving loop: 16 resident scenes, Poisson request stream at 25/50/90%% of the "
                         "saturated rate, p50/p99 latency")
    ap.add_argument("--requests", type=int, default=400, help="--served: requests per load level")
    ap.add_argument("--slots", default="2,1", help="--served: stream slots of the (cp+ray, silhouette) classes")
    ap.add_argument("--max-batch", type=int, default=1 << 20, help="--served: queries per coalesced launch at most")
    ap.add_argument("--scene-cache", default="",
                    help="--served: directory of scene files (loaded if present, else built and written)")
    ap.add_argument("--tune", action="store_true",
                    help="autotune leaf size / packet budget / silhouette traversal for the scene and "
                         "query kind given by --shape/--subdiv/--segments/--query, then time the best; "
                         "with --batched: tune each group of similar jobs and run the batch with the choices")
    args = ap.parse_args()
    if args.served:
        print(json.dumps(served(args)))
        return
    if args.batched:
        print(json.dumps(batched(args)))
        return
    if args.tune:
        import torch
        from gquery_amd import synth, tune
        torch.cuda.set_device(0)
        dim = 2 if args.shape == "polyline" else 3
        V, F = synth.star_polyline(args.segments) if dim == 2 else synth.displaced_icosphere(args.subdiv)
        x = torch.from_numpy(synth.uniform_points(args.n, dim)).cuda()
        v = torch.from_numpy(synth.uniform_directions(args.n, dim)).cuda()
        res = tune.tune_shape(V, F, queries={args.query: (x, v)}, kinds=(args.query,), log=log)
        print(json.dumps({"shape": args.shape, "query": args.query, "n": args.n, "best": res["options"],
                          "ms": res["ms"], "trials": res["trials"]}))
        return

    if args.accuracy:
        accuracy(args)
        return
    if args.pmc_summary:
        pmc_summary(args)
        return
    if args.sq_summary:
        sq_summary(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; queries shard by batch with no "
                                   "exchange, so multi-GPU numbers are not measured here"}))
        sys.exit(2)

    import numpy as np
    import torch

    from gquery_amd import _lib, synth
    from gquery_amd.shapes.mesh import Mesh

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    dim = 3
    if args.shape == "polyline":
        from gquery_amd.shapes.polyline import Polyline
        dim = 2
        V, F = synth.star_polyline(args.segments)
        mesh = Polyline(V, F, device="cuda:0", leaf_prims=args.leaf_prims, packet_budget=args.packet_budget,
                       traversal=args.traversal, wave_front=args.wave_front)
    else:
        V, F = synth.displaced_icosphere(args.subdiv)
        mesh = Mesh(V, F, device="cuda:0", leaf_prims=args.leaf_prims, packet_budget=args.packet_budget,
                    traversal=args.traversal, wave_front=args.wave_front, sil_pairs=args.sil_pairs)
    log(f"scene built in {time.perf_counter() - t0:.1f}s: {F.shape[0]} primitives")
    nq = args.n
    x_host = synth.uniform_points(nq, dim)
    x = torch.from_numpy(x_host).cuda()
    v = torch.from_numpy(synth.uniform_directions(nq, dim)).cuda()
    torch.cuda.synchronize()

    method = args.method
    stream = torch.cuda.current_stream()
    if args.query == "sample":
        from gquery_amd.sampler import PCG32
        gen = PCG32(nq, device="cuda:0")

    def step():
        if args.query == "sample":
            return mesh.sample_boundary(x, args.radius, gen)
        if args.query == "cp":
            return mesh.closest_point_bvh(x, method=method)
        if args.query == "ray":
            return mesh.intersect_bvh(x, v, method=method)
        return mesh.closest_silhouette_snch(x, method=method)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.profile_reset()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        rec = step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    if args.copy_mb:
        src = torch.ones(args.copy_mb << 18, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    kern_ms, first_ms, kern_launches, kern_name = _lib.profile_read_split()
    ms_per_step = ms_total / args.steps
    value = nq / (ms_per_step * 1e-3) / 1e6
    log(f"{method} {args.shape} {args.query} s={args.subdiv} n={nq} leaf={args.leaf_prims} budget={args.packet_budget}: {ms_per_step:.3f} ms/step ({wall / args.steps * 1e3:.3f} ms wall), {value:.1f} Mq/s; "
        f"traversal kernel {kern_name} {kern_ms / max(kern_launches, 1):.3f} ms/launch")

    if args.query != "cp" or args.subdiv != SUBDIV or nq != N_QUERIES or dim != 3:
        line = {"shape": args.shape, "query": args.query, "method": method, "prims": int(F.shape[0]), "n": nq,
                "ms_per_step": ms_per_step, "Mq_per_s": value, "kernel_ms": kern_ms / max(kern_launches, 1)}
        if args.query == "sample":
            line["radius"] = args.radius
            line["hit_fraction"] = float(rec[0].float().mean())
            if not args.no_cpu:
                # the CPU restatement of the walk on the same queries, all host threads
                from oracle import oracle as O
                st, inc = O.pcg32_seed(nq)
                t1 = time.perf_counter()
                O.sample_boundary(dim, mesh.snch.bvh_node_data(), mesh.snch.primitive_data(), x_host,
                                  args.radius, st, inc)
                cpu_s = time.perf_counter() - t1
                line["cpu_baseline"] = {"value": nq / cpu_s / 1e6, "unit": "Msamples/s",
                                        "cores": O.num_threads(), "kind": "port",
                                        "sample": f"all {nq} queries, {cpu_s:.2f} s"}
        print(json.dumps(line))
        return

    # roofline of the dominant kernel (the packet pass, k_cp_packet): VALU issue, with HBM beside it
    rng = np.random.default_rng(0)
    trav_ms = kern_ms / max(kern_launches, 1)
    packet_ms = first_ms / max(kern_launches, 1)
    roof = valu_roofline(method, packet_ms, trav_ms,
                         f"k_cp_packet<3, {'true' if method == 'engine' else 'false'}> ({kern_name})")
    # informational: the reference-layout bytes the reference traversal would move (B_q per query,
    # work counts from the CPU restatement on 20,000 sampled queries) per second of traversal
    sample = x_host[rng.choice(N_QUERIES, 20000, replace=False)]
    bq, counts = algorithmic_bytes_per_query(mesh, sample)
    roof["reference_layout"] = {"bytes_per_query": bq, "work_per_query": counts,
                                "rate_GBs": bq * N_QUERIES / (trav_ms * 1e-3) / 1e9 if trav_ms else None}
    log(f"roofline: packet kernel {packet_ms:.3f} ms of {trav_ms:.3f} ms traversal; VALU frac "
        f"{roof['frac']}; HBM {roof.get('hbm')}")

    # sanity: a sample of the timed results against the oracle
    from oracle import oracle as O
    idx = rng.choice(N_QUERIES, 2000, replace=False)
    ref = O.closest_point(3, mesh.snch.bvh_node_data(), mesh.snch.primitive_data(), x_host[idx])
    got_d = rec.d.cpu().numpy()[idx]
    max_err = float(np.max(np.abs(got_d - ref["d"]) / np.maximum(ref["d"], 1.0)))
    log(f"sample check vs oracle: max rel err d = {max_err:.2e}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, x_host)
        log(f"cpu baseline: {cpu['value']:.3f} Mq/s ({cpu['sample']})")

    # end to end from host memory: pinned queries in, query, the records back to pinned host memory
    xh = torch.from_numpy(x_host).pin_memory()
    pinned = {}
    with torch.no_grad():
        r0 = mesh.closest_point_bvh(x, method=method)
        for f in ("valid", "p", "n", "uv", "d", "prim_id"):
            t = getattr(r0, f)
            pinned[f] = torch.empty(t.shape, dtype=t.dtype).pin_memory()
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(PLAIN_STEPS):
        xd = xh.to("cuda:0", non_blocking=True)
        r0 = mesh.closest_point_bvh(xd, method=method)
        for f, t in pinned.items():
            t.copy_(getattr(r0, f), non_blocking=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    host_ms = ev0.elapsed_time(ev1) / PLAIN_STEPS
    host_rate = {"ms_per_step": host_ms, "value": nq / (host_ms * 1e-3) / 1e6,
                 "bytes": int(x_host.nbytes + sum(t.numel() * t.element_size() for t in pinned.values()))}
    log(f"from host memory (pinned, PCIe both ways): {host_ms:.2f} ms/step, {host_rate['value']:.1f} Mq/s")
    del xh, pinned, r0

    plain = None
    if not args.no_plain and method != "plain":
        # the plain FP32 kernels (reference layout, lane per query) on the same resident batch
        mesh.closest_point_bvh(x, method="plain")
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(PLAIN_STEPS):
            mesh.closest_point_bvh(x, method="plain")
        ev1.record(stream)
        torch.cuda.synchronize()
        p_ms = ev0.elapsed_time(ev1) / PLAIN_STEPS
        plain = {"ms_per_step": p_ms, "value": nq / (p_ms * 1e-3) / 1e6, "steps": PLAIN_STEPS,
                 "speedup": p_ms / ms_per_step}
        log(f"plain FP32 kernels: {p_ms:.2f} ms/step, {plain['value']:.1f} Mq/s; engine speedup {plain['speedup']:.2f}x")

    paths = None
    if not args.no_paths:
        from gquery_amd.shapes.polyline import Polyline
        V2, F2 = synth.star_polyline(65536)
        poly = Polyline(V2, F2, device="cuda:0")
        x2 = torch.from_numpy(synth.uniform_points(N_QUERIES, 2)).cuda()
        v2 = torch.from_numpy(synth.uniform_directions(N_QUERIES, 2)).cuda()
        paths = query_paths(mesh, poly, x, v, x2, v2, log)
        del poly, x2, v2
    # the serving leg runs before the batched one: a batch of 4M queries or more creates the
    # library's auxiliary streams, which then share the process's four hardware queues with the
    # serving loop's stream slots
    srv = None
    if not args.no_served:
        del rec
        s = served(args)
        srv = {"p99_ms_at_90pct_load": s["value"], "workload": s["config"]["workload"],
               **{k: {a: round(b, 3) for a, b in v.items()} for k, v in s["served"].items()}}

    bat = None
    if not args.no_batched:
        bsteps, bwarm = args.steps, args.warmup
        args.steps, args.warmup = 5, 1
        b = batched(args)
        args.steps, args.warmup = bsteps, bwarm
        bat = {"ms_per_step": b["ms_per_step"], "Mq_per_s": b["value"], "workload": b["config"]["workload"],
               "sequential_single_calls_ms": b["sequential_single_calls_ms"],
               "graph_speedup_vs_sequential": b["graph_speedup_vs_sequential"],
               "device_workspace_bytes": b["device_workspace_bytes"], "steps": 5}
        torch.cuda.empty_cache()

    out = {
        "metric": "Mqueries/s (3D closest-point, 327k tris, 1M q) vs plain; served p99 latency",
        "value": value,
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded displaced icosphere + uniform queries",
        "config": {"workload": "closest_point_bvh, displaced icosphere s=7 (327,680 triangles), "
                               "1,048,576 uniform queries in [-1.5,1.5]^3",
                   "method": method, "global_batch": N_QUERIES, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "vs_plain": plain["speedup"] if plain else None,
        "plain": plain,
        "host_memory": host_rate,
        "paths": paths,
        "batched": bat,
        "served": srv,
        "max_rel_err_vs_oracle": max_err,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
