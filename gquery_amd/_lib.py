"""ctypes binding of libgquery_amd.so (declarations: include/gquery_amd.h).

The library is loaded from the package directory only; a missing library is an error, never a
fallback to another implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgquery_amd.so"

GQ_OK = 0
GQ_ERR_INVALID = -1
GQ_ERR_OOM = -2
GQ_ERR_DEVICE = -3
GQ_ERR_UNSUPPORTED = -4

METHOD_PLAIN = 0
METHOD_ENGINE_FP32 = 1
METHOD_ENGINE = 2
METHODS = {"plain": METHOD_PLAIN, "engine_fp32": METHOD_ENGINE_FP32, "engine": METHOD_ENGINE}


class SceneOptions(C.Structure):
    _fields_ = [("leaf_prims", C.c_int32), ("packet_budget", C.c_int32), ("traversal", C.c_int32),
                ("wave_front", C.c_int32), ("sil_pairs", C.c_int32), ("ray_leaf_prims", C.c_int32),
                ("reserved", C.c_int32 * 2)]


class Records(C.Structure):
    _fields_ = [
        ("valid", C.c_void_p),
        ("p", C.c_void_p),
        ("n", C.c_void_p),
        ("uv", C.c_void_p),
        ("d", C.c_void_p),
        ("prim_id", C.c_void_p),
        ("on_boundary", C.c_void_p),
    ]


class SampleRecords(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in ("hit", "p", "n", "uv", "pdf", "pmf", "prim_id")]


QUERY_CLOSEST_POINT = 0
QUERY_INTERSECT = 1
QUERY_SILHOUETTE = 2
QUERIES = {"cp": QUERY_CLOSEST_POINT, "ray": QUERY_INTERSECT, "sil": QUERY_SILHOUETTE}


class Job(C.Structure):
    """gq_job: one query batch of a gq_batch_enqueue call."""
    _fields_ = [
        ("scene", C.c_void_p),
        ("kind", C.c_int32),
        ("method", C.c_int32),
        ("x", C.c_void_p),
        ("v", C.c_void_p),
        ("r_max", C.c_void_p),
        ("r_max_scalar", C.c_float),
        ("packet_budget", C.c_int32),
        ("n", C.c_int64),
        ("out", Records),
        ("traversal", C.c_int32),
        ("order", C.c_int32),
        ("shared", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]


_i64p = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "gq_version": (C.c_int, []),
    "gq_last_error": (C.c_char_p, []),
    "gq_build_bvh": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "gq_build_snch": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "gq_structure_free": (None, [C.c_void_p]),
    "gq_structure_info": (C.c_int, [C.c_void_p, _i64p]),
    "gq_structure_node_data": (C.c_int, [C.c_void_p, C.c_void_p, _i64p]),
    "gq_structure_snch_node_data": (C.c_int, [C.c_void_p, C.c_void_p, _i64p]),
    "gq_structure_primitive_data": (C.c_int, [C.c_void_p, C.c_void_p, _i64p]),
    "gq_structure_silhouette_data": (C.c_int, [C.c_void_p, C.c_void_p, _i64p]),
    "gq_structure_raw_silhouette_data": (C.c_int, [C.c_void_p, C.c_void_p, _i64p]),
    "gq_build_flat_silhouette_edges": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, _i64p]),
    "gq_scene_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "gq_scene_destroy": (C.c_int, [C.c_void_p]),
    "gq_scene_create_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "gq_scene_info": (C.c_int, [C.c_void_p, _i64p]),
    "gq_closest_point_bvh": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_intersect_bvh": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_closest_silhouette_snch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_sample_boundary": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(SampleRecords), C.c_void_p]),
    "gq_structure_engine_layout": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, _i64p]),
    "gq_batch_sizes": (C.c_int, [C.c_int, C.POINTER(Job), _i64p, _i64p]),
    "gq_batch_enqueue": (C.c_int, [C.c_int, C.POINTER(Job), C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]),
    "gq_shape_create": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]),
    "gq_shape_destroy": (C.c_int, [C.c_void_p]),
    "gq_shape_info": (C.c_int, [C.c_void_p, _i64p]),
    "gq_shape_normals": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gq_closest_point_baseline": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_intersect_baseline": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_closest_silhouette_baseline": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_configure_silhouette": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, _i64p]),
    "gq_configure_silhouette_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, _i64p,
                                                 C.c_void_p]),
    "gq_build_flat_silhouette_edges_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                                        _i64p, C.c_void_p]),
    "gq_structure_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "gq_structure_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "gq_structure_geometry": (C.c_int, [C.c_void_p, C.c_void_p, _i64p, C.c_void_p, _i64p]),
    "gq_profile_begin": (C.c_int, []),
    "gq_closest_point_bvh_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.POINTER(Records), C.c_void_p]),
    "gq_profile_end": (C.c_int, [C.POINTER(C.c_double), _i64p, C.c_char_p, C.c_int32]),
    "gq_profile_end_split": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), _i64p, C.c_char_p, C.c_int32]),
}

_lib = None


def lib() -> C.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m gquery_amd.build` "
                "(there is no fallback implementation)")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == GQ_OK:
        return
    msg = lib().gq_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == GQ_ERR_INVALID:
        raise ValueError(text)
    if rc == GQ_ERR_OOM:
        raise MemoryError(text)
    if rc == GQ_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise RuntimeError(text)


def profile_reset() -> None:
    """Start timing every traversal kernel launch (hipEvents on the query's stream)."""
    check(lib().gq_profile_begin(), "gq_profile_begin")


def profile_read():
    """-> (total traversal-kernel ms, launches, kernel name) since profile_reset()."""
    ms, _, n, name = profile_read_split()
    return ms, n, name


def profile_read_split():
    """-> (total traversal ms, first-kernel ms, launches, kernel name) since profile_reset(); the
    first kernel is the packet (or lane) kernel of each launch, the wave passes follow it."""
    ms = C.c_double(0)
    first = C.c_double(0)
    n = C.c_int64(0)
    buf = C.create_string_buffer(128)
    check(lib().gq_profile_end_split(C.byref(ms), C.byref(first), C.byref(n), buf, 128), "gq_profile_end_split")
    return ms.value, first.value, n.value, buf.value.decode()
