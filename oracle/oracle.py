"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of oracle/liboracle.so (oracle/gq_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, as the
checker; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    return LIB


class _Out(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ["valid", "p", "n", "uv", "d", "prim_id", "on_boundary", "visits", "child_tests", "leaf_tests"]]


class _SampleOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ["hit", "p", "n", "uv", "pdf", "pmf", "prim_id"]]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        vp, i64, f32 = C.c_void_p, C.c_int64, C.c_float
        h.or_closest_point.argtypes = [C.c_int, vp, vp, vp, i64, _Out]
        h.or_intersect.argtypes = [C.c_int, vp, vp, vp, vp, vp, f32, i64, _Out]
        h.or_silhouette.argtypes = [C.c_int, vp, vp, vp, vp, f32, i64, _Out]
        for f in (h.or_closest_point, h.or_intersect, h.or_silhouette):
            f.restype = i64
        h.or_bf_closest_point_f64.argtypes = [C.c_int, vp, vp, i64, vp, i64, vp, vp, vp, vp]
        h.or_bf_intersect_f64.argtypes = [C.c_int, vp, vp, i64, vp, vp, i64, vp, vp, vp, vp]
        h.or_bf_silhouette_f64.argtypes = [C.c_int, vp, i64, vp, i64, vp, vp, vp, vp]
        h.or_bf_closest_point_f32.argtypes = [C.c_int, vp, i64, vp, i64, vp, vp]
        h.or_cp_triangle_pairs.argtypes = [vp, vp, i64, vp]
        h.or_leaf_distance.argtypes = [C.c_int, vp, vp, vp, i64, vp]
        h.or_num_threads.restype = C.c_int
        h.or_normals.argtypes = [vp, i64, vp, i64, vp, vp]
        h.or_cp_baseline.argtypes = [C.c_int, vp, vp, i64, vp, vp, i64, _Out]
        h.or_ray_baseline.argtypes = [C.c_int, vp, vp, i64, vp, vp, vp, vp, vp, f32, i64, _Out]
        h.or_sil_baseline.argtypes = [C.c_int, vp, i64, vp, vp, f32, i64, _Out]
        u64 = C.c_uint64
        h.or_pcg32_seed.argtypes = [i64, u64, u64, vp, vp]
        h.or_pcg32_stream.argtypes = [vp, vp, i64, i64, vp]
        h.or_sample_boundary.argtypes = [C.c_int, vp, vp, vp, vp, f32, vp, vp, i64, _SampleOut]
        _lib = h
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _records(n, dim, kind):
    r = {
        "valid": np.zeros(n, np.uint8), "p": np.zeros((n, dim), np.float32), "d": np.zeros(n, np.float32),
        "prim_id": np.zeros(n, np.int32),
        "visits": np.zeros(n, np.int32), "child_tests": np.zeros(n, np.int32), "leaf_tests": np.zeros(n, np.int32),
    }
    if kind != "sil":
        # the records' field names: uv (3D barycentrics), t (2D segment parameter)
        r["n"] = np.zeros((n, dim), np.float32)
        r["uv" if dim == 3 else "t"] = np.zeros((n, 2), np.float32) if dim == 3 else np.zeros(n, np.float32)
    if kind == "ray":
        r["on_boundary"] = np.zeros(n, np.uint8)
    o = _Out()
    for k, v in r.items():
        setattr(o, "uv" if k == "t" else k, v.ctypes.data)
    return r, o


def closest_point(dim, node_data, primitive_data, x):
    x = _c(x, np.float32)
    nodes, prims = _c(node_data, np.float32), _c(primitive_data, np.float32)
    r, o = _records(x.shape[0], dim, "cp")
    ov = lib().or_closest_point(dim, nodes.ctypes.data, prims.ctypes.data, x.ctypes.data, x.shape[0], o)
    r["overflows"] = ov
    return r


def intersect(dim, node_data, primitive_data, x, v, r_max=float("inf")):
    x, v = _c(x, np.float32), _c(v, np.float32)
    nodes, prims = _c(node_data, np.float32), _c(primitive_data, np.float32)
    r, o = _records(x.shape[0], dim, "ray")
    rm = None
    if np.ndim(r_max) > 0:
        rm = _c(r_max, np.float32)
    ov = lib().or_intersect(dim, nodes.ctypes.data, prims.ctypes.data, x.ctypes.data, v.ctypes.data,
                            None if rm is None else rm.ctypes.data, float(r_max) if rm is None else 0.0,
                            x.shape[0], o)
    r["overflows"] = ov
    return r


def silhouette(dim, snch_node_data, silhouette_data, x, r_max=float("inf")):
    x = _c(x, np.float32)
    nodes, sils = _c(snch_node_data, np.float32), _c(silhouette_data, np.float32)
    r, o = _records(x.shape[0], dim, "sil")
    rm = None
    if np.ndim(r_max) > 0:
        rm = _c(r_max, np.float32)
    ov = lib().or_silhouette(dim, nodes.ctypes.data, sils.ctypes.data, x.ctypes.data,
                             None if rm is None else rm.ctypes.data, float(r_max) if rm is None else 0.0,
                             x.shape[0], o)
    r["overflows"] = ov
    return r


def bf_closest_point(dim, V, F, x):
    V, F, x = _c(V, np.float32), _c(F, np.int32), _c(x, np.float32)
    n = x.shape[0]
    d, p, i, d2 = np.zeros(n), np.zeros((n, dim)), np.zeros(n, np.int32), np.zeros(n)
    lib().or_bf_closest_point_f64(dim, V.ctypes.data, F.ctypes.data, F.shape[0], x.ctypes.data, n,
                                  d.ctypes.data, p.ctypes.data, i.ctypes.data, d2.ctypes.data)
    return {"d": d, "p": p, "prim_id": i, "second_d": d2}


def bf_intersect(dim, V, F, x, v):
    V, F, x, v = _c(V, np.float32), _c(F, np.int32), _c(x, np.float32), _c(v, np.float32)
    n = x.shape[0]
    t, i, t2, uv = np.zeros(n), np.zeros(n, np.int32), np.zeros(n), np.zeros((n, 2))
    lib().or_bf_intersect_f64(dim, V.ctypes.data, F.ctypes.data, F.shape[0], x.ctypes.data, v.ctypes.data, n,
                              t.ctypes.data, i.ctypes.data, t2.ctypes.data, uv.ctypes.data)
    out = {"t": t, "prim_id": i, "second_t": t2}
    if dim == 3:
        out["uv"] = uv   # barycentrics (1 - v - w, v) of the first hit
    return out


def bf_silhouette(dim, raw_silhouette_data, x):
    raw, x = _c(raw_silhouette_data, np.float32), _c(x, np.float32)
    ss = 10 if dim == 2 else 19
    n = x.shape[0]
    d, p, i, d2 = np.zeros(n), np.zeros((n, dim)), np.zeros(n, np.int32), np.zeros(n)
    lib().or_bf_silhouette_f64(dim, raw.ctypes.data, raw.size // ss, x.ctypes.data, n,
                               d.ctypes.data, p.ctypes.data, i.ctypes.data, d2.ctypes.data)
    return {"d": d, "p": p, "prim_id": i, "second_d": d2}


def bf_closest_point_f32(dim, primitive_data, x):
    """Smallest float32 leaf distance over every primitive (the traversal's own leaf formula) and
    the index of the first primitive reaching it."""
    prims, x = _c(primitive_data, np.float32), _c(x, np.float32)
    n = x.shape[0]
    d, i = np.zeros(n, np.float32), np.zeros(n, np.int32)
    lib().or_bf_closest_point_f32(dim, prims.ctypes.data, prims.size // (5 if dim == 2 else 10),
                                  x.ctypes.data, n, d.ctypes.data, i.ctypes.data)
    return {"d": d, "prim_id": i}


def leaf_distance(dim, primitive_data, prim_id, x):
    """float32 leaf distance (the traversal's own formula) from each query to the primitive whose
    original index is prim_id[i]."""
    ps = 5 if dim == 2 else 10
    prims, x = _c(primitive_data, np.float32), _c(x, np.float32)
    idx = prims.reshape(-1, ps)[:, -1].astype(np.int64)
    row_of = np.empty(idx.size, np.int64)
    row_of[idx] = np.arange(idx.size)
    rows = _c(row_of[np.asarray(prim_id, np.int64)], np.int64)
    d = np.zeros(x.shape[0], np.float32)
    lib().or_leaf_distance(dim, prims.ctypes.data, rows.ctypes.data, x.ctypes.data, x.shape[0], d.ctypes.data)
    return d


def cp_triangle_pairs(p, tri):
    """float32 closest_point_triangle distance for point/triangle pairs: p (n, 3), tri (n, 3, 3)."""
    p, tri = _c(p, np.float32), _c(tri, np.float32)
    d = np.zeros(p.shape[0], np.float32)
    lib().or_cp_triangle_pairs(p.ctypes.data, tri.ctypes.data, p.shape[0], d.ctypes.data)
    return d


def num_threads() -> int:
    return lib().or_num_threads()


# ---- brute-force baselines, normals and silhouette configuration (FP32 restatements) ----------

def normals(V, F):
    """Mesh.configure_normal (mesh.py:35-50): (face normals, vertex normals)."""
    V, F = _c(V, np.float32), _c(F, np.int32)
    fn = np.zeros((F.shape[0], 3), np.float32)
    vn = np.zeros((V.shape[0], 3), np.float32)
    lib().or_normals(V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0], fn.ctypes.data, vn.ctypes.data)
    return fn, vn


def closest_point_baseline(dim, V, F, x, face_normals=None):
    """Polyline/Mesh.closest_point_baseline (polyline.py:176-209, mesh.py:92-118)."""
    V, F, x = _c(V, np.float32), _c(F, np.int32), _c(x, np.float32)
    fn = _c(normals(V, F)[0] if (dim == 3 and face_normals is None) else
            (face_normals if face_normals is not None else np.zeros((1, 3))), np.float32)
    r, o = _records(x.shape[0], dim, "cp")
    lib().or_cp_baseline(dim, V.ctypes.data, F.ctypes.data, F.shape[0], fn.ctypes.data, x.ctypes.data,
                         x.shape[0], o)
    return r


def intersect_baseline(dim, V, F, x, v, r_max=float("inf"), normal=None, on_boundary=None):
    """Polyline/Mesh.intersect_baseline (polyline.py:138-170, mesh.py:61-85); with normal and
    on_boundary, Polyline/Mesh.intersect (polyline.py:125-131, mesh.py:53-59)."""
    V, F, x, v = _c(V, np.float32), _c(F, np.int32), _c(x, np.float32), _c(v, np.float32)
    r, o = _records(x.shape[0], dim, "ray")
    rm = _c(r_max, np.float32) if np.ndim(r_max) > 0 else None
    nb = _c(normal, np.float32) if normal is not None else None
    ob = _c(on_boundary, np.uint8) if on_boundary is not None else None
    lib().or_ray_baseline(dim, V.ctypes.data, F.ctypes.data, F.shape[0], x.ctypes.data, v.ctypes.data,
                          None if nb is None else nb.ctypes.data, None if ob is None else ob.ctypes.data,
                          None if rm is None else rm.ctypes.data, float(r_max) if rm is None else 0.0,
                          x.shape[0], o)
    return r


def silhouette_baseline(dim, sils, x, r_max=float("inf")):
    """Polyline/Mesh.closest_silhouette_baseline (polyline.py:212-223, mesh.py:124-136) over the
    configure_silhouette records (2D, stride 10) or the flat silhouette edges (3D, stride 19)."""
    x = _c(x, np.float32)
    sils = _c(sils, np.float32).reshape(-1)
    ns = sils.size // (10 if dim == 2 else 19)
    r, o = _records(x.shape[0], dim, "sil")
    rm = _c(r_max, np.float32) if np.ndim(r_max) > 0 else None
    lib().or_sil_baseline(dim, sils.ctypes.data, ns, x.ctypes.data, None if rm is None else rm.ctypes.data,
                          float(r_max) if rm is None else 0.0, x.shape[0], o)
    return r


def configure_silhouette(V, E):
    """Polyline.configure_silhouette (polyline.py:44-122) restated with Python containers: records
    of 10 floats (a, b, c, prev, vertex, next, first-appearance index), nearly collinear vertices
    dropped (float32 unit edge directions, |dot| >= 1 - 1e-5)."""
    V = np.asarray(V, np.float32)
    nbr = {}
    for a, b in np.asarray(E, np.int64):
        a, b = int(a), int(b)
        for u in (a, b):
            if u not in nbr:
                nbr[u] = [-1, u, -1]
        nbr[a][2] = b
        nbr[b][0] = a
    def dot(a, b):   # fma(a.y, b.y, a.x * b.x): the product of two float32 is exact in float64
        return np.float32(np.float64(a[1]) * np.float64(b[1]) + np.float64(np.float32(a[0] * b[0])))

    out = []
    one = np.float32(1.0)
    for i, u in enumerate(nbr):
        p, _, q = nbr[u]
        if p == -1 or q == -1:
            continue
        a, b, c = V[p], V[u], V[q]
        e1, e2 = (b - a).astype(np.float32), (c - b).astype(np.float32)
        e1 = e1 * (one / np.sqrt(dot(e1, e1)))
        e2 = e2 * (one / np.sqrt(dot(e2, e2)))
        col = np.abs(dot(e1, e2))
        if not col < np.float32(1.0 - 1e-5):
            continue
        out.append([a[0], a[1], b[0], b[1], c[0], c[1], p, u, q, i])
    return np.asarray(out, np.float32).reshape(-1, 10)


def pcg32_seed(n, initstate=0x853C49E6748FEA9B, initseq=0xDA3E39CB94B95BDB):
    """(state, inc) uint64 arrays of a PCG32 of size n seeded as DrJit does (stream initseq + lane)."""
    state = np.zeros(n, np.uint64)
    inc = np.zeros(n, np.uint64)
    lib().or_pcg32_seed(n, initstate & (2**64 - 1), initseq & (2**64 - 1), state.ctypes.data, inc.ctypes.data)
    return state, inc


def pcg32_stream(state, inc, k):
    """k next_uint32 outputs per lane (n, k); advances `state` in place."""
    state = np.ascontiguousarray(state, np.uint64)
    inc = np.ascontiguousarray(inc, np.uint64)
    out = np.zeros((len(state), k), np.uint32)
    lib().or_pcg32_stream(state.ctypes.data, inc.ctypes.data, len(state), k, out.ctypes.data)
    return out, state


def sample_boundary(dim, node_data, primitive_data, x, R, state, inc):
    """BVH-guided boundary sampling (bvh3d.py:246-346 and its 2D analogue) over the reference
    BVH arrays.  R: scalar or per-query array.  Returns (records, new_state)."""
    x = _c(x, np.float32).reshape(-1, dim)
    n = x.shape[0]
    nodes = _c(node_data, np.float32)
    prims = _c(primitive_data, np.float32)
    st = np.array(state, dtype=np.uint64, copy=True).reshape(-1)
    inc = _c(inc, np.uint64).reshape(-1)
    assert st.shape == (n,) and inc.shape == (n,)
    Ra = None
    Rs = 0.0
    if np.ndim(R) == 0:
        Rs = float(R)
    else:
        Ra = _c(R, np.float32).reshape(-1)
        assert Ra.shape == (n,)
    r = {"hit": np.zeros(n, np.uint8), "p": np.zeros((n, dim), np.float32), "n": np.zeros((n, dim), np.float32),
         "uv": np.zeros((n, 2), np.float32) if dim == 3 else np.zeros(n, np.float32),
         "pdf": np.zeros(n, np.float32), "pmf": np.zeros(n, np.float32), "prim_id": np.zeros(n, np.int32)}
    o = _SampleOut()
    for k, v in r.items():
        setattr(o, k, v.ctypes.data)
    lib().or_sample_boundary(dim, nodes.ctypes.data, prims.ctypes.data, x.ctypes.data,
                             None if Ra is None else Ra.ctypes.data, Rs, st.ctypes.data, inc.ctypes.data, n, o)
    return r, st
