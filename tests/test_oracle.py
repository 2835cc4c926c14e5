"""The CPU oracle (oracle/gq_oracle.c) pinned against hand-derived answers on the reference's own
fixtures and against FP64 brute force on synthetic scenes; a planted mistake must be caught."""
import math

import numpy as np
import pytest

from gquery_amd import ext, synth
from oracle import oracle as O
from tests import parity
from tests.test_builder import PENTAGON_E, PENTAGON_V, TET_F, TET_V


# ---- known answers (examples/example_2d.py, examples/example_3d.py) ---------------------------

def test_pentagon_closest_point():
    # Query (0.8, 0.8): edge 2 from (1.5, 1) to (0.5, 1.5); u = (-1, 0.5), w = (-0.7, -0.2),
    # t = (0.7 - 0.1) / 1.25 = 0.48, p = (1.02, 1.24), d = sqrt(0.22^2 + 0.44^2) = sqrt(0.242).
    s = ext.SNCH(PENTAGON_V, PENTAGON_E)
    r = O.closest_point(2, s.bvh_node_data(), s.primitive_data(), np.array([[0.8, 0.8]], np.float32))
    assert r["valid"][0] == 1 and r["prim_id"][0] == 2
    np.testing.assert_allclose(r["p"][0], [1.02, 1.24], atol=1e-6)
    np.testing.assert_allclose(r["d"][0], math.sqrt(0.242), rtol=1e-6)
    np.testing.assert_allclose(r["t"][0], 0.48, atol=1e-6)
    # n = (t.y, -t.x) of normalize(b - a) = normalize(-1, 0.5)
    np.testing.assert_allclose(r["n"][0], [0.5 / math.sqrt(1.25), 1 / math.sqrt(1.25)], atol=1e-6)


def test_pentagon_ray():
    # Ray (0.8, 0.8) + t (1, 1) meets edge 2 at s = 1/3: 0.8 + t = 1.5 - s -> t = 0.7 - 1/3.
    s = ext.SNCH(PENTAGON_V, PENTAGON_E)
    r = O.intersect(2, s.bvh_node_data(), s.primitive_data(), np.array([[0.8, 0.8]], np.float32),
                    np.array([[1.0, 1.0]], np.float32))
    t = 0.7 - 1.0 / 3.0
    assert r["valid"][0] == 1 and r["prim_id"][0] == 2 and r["on_boundary"][0] == 1
    np.testing.assert_allclose(r["d"][0], t, rtol=1e-6)
    np.testing.assert_allclose(r["p"][0], [0.8 + t, 0.8 + t], atol=1e-6)
    assert r["t"][0] == -1.0


def test_pentagon_silhouette():
    # From (2, 0.5) only vertices 1 and 2 are silhouettes (cross products of the two incident
    # edges change sign); vertex 2 = (1.5, 1) is nearer: d = |(0.5, -0.5)| = sqrt(0.5).
    s = ext.SNCH(PENTAGON_V, PENTAGON_E)
    r = O.silhouette(2, s.node_data(), s.silhouette_data(), np.array([[2.0, 0.5]], np.float32))
    assert r["valid"][0] == 1 and r["prim_id"][0] == 2
    np.testing.assert_allclose(r["p"][0], [1.5, 1.0])
    np.testing.assert_allclose(r["d"][0], math.sqrt(0.5), rtol=1e-6)


def test_tetrahedron_closest_point():
    # Query (0.6, 0.6, 0.6) lies 0.05/|n| outside face 2 = (1, 2, 3), n = (1, 0.5, 0.25):
    # d = 0.05 / sqrt(21/16) = 0.2 / sqrt(21); foot q = p - (0.05 / 1.3125) n = (59, 61, 62)/105,
    # barycentric (b, c) weights (2/7, 62/105).
    s = ext.SNCH3D(TET_V, TET_F)
    r = O.closest_point(3, s.bvh_node_data(), s.primitive_data(), np.array([[0.6, 0.6, 0.6]], np.float32))
    assert r["valid"][0] == 1 and r["prim_id"][0] == 2
    np.testing.assert_allclose(r["d"][0], 0.2 / math.sqrt(21), rtol=2e-6)
    np.testing.assert_allclose(r["p"][0], np.array([59, 61, 62]) / 105, atol=1e-6)
    np.testing.assert_allclose(r["uv"][0], [2 / 7, 62 / 105], atol=1e-6)
    np.testing.assert_allclose(r["n"][0], np.array([1, 0.5, 0.25]) / math.sqrt(1.3125), atol=1e-6)


def test_tetrahedron_ray_misses():
    # From (0.6, 0.6, 0.6) along (1, 1, 1) every face plane is behind the origin or the ray leaves
    # face 2's plane outward (n . d = 1.75 > 0 with the origin already outside): no hit, zero record.
    s = ext.SNCH3D(TET_V, TET_F)
    r = O.intersect(3, s.bvh_node_data(), s.primitive_data(), np.array([[0.6, 0.6, 0.6]], np.float32),
                    np.array([[1.0, 1.0, 1.0]], np.float32))
    assert r["valid"][0] == 0 and r["d"][0] == 0 and r["prim_id"][0] == 0 and not r["p"].any()


# Two parallel triangles: a = (0,0,0), b = (2,0,0), c = (0,2,0) and the same 1 below.  The ray from
# (0.5, 0.25, 1) along (0, 0, -1) meets the upper one first.  Moller-Trumbore (math.py:110-128):
# v1 = b - a = (2,0,0), v2 = c - a = (0,2,0), p = d x v2 = (2,0,0), det = v1.p = 4, s = x - a,
# v = s.p / det = 0.25, q = s x v1 = (0, 2, -0.5), w = d.q / det = 0.125, t = v2.q / det = 1.
# Record (triangle.py:48-73): p = a + v1 v + v2 w = (0.5, 0.25, 0), n = normalize(v1 x v2) = (0,0,1),
# uv = (1 - v - w, v) = (0.625, 0.25), d = t = 1, on_boundary.
TWO_TRI_V = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [0, 0, -1], [2, 0, -1], [0, 2, -1]], np.float32)
TWO_TRI_F = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
TWO_TRI_RAY = (np.array([[0.5, 0.25, 1.0]], np.float32), np.array([[0.0, 0.0, -1.0]], np.float32))
TWO_TRI_HIT = {"prim_id": 0, "d": 1.0, "p": [0.5, 0.25, 0.0], "n": [0.0, 0.0, 1.0], "uv": [0.625, 0.25]}


def test_two_triangle_ray_record():
    s = ext.SNCH3D(TWO_TRI_V, TWO_TRI_F)
    r = O.intersect(3, s.bvh_node_data(), s.primitive_data(), *TWO_TRI_RAY)
    assert r["valid"][0] == 1 and r["on_boundary"][0] == 1 and r["prim_id"][0] == TWO_TRI_HIT["prim_id"]
    np.testing.assert_allclose(r["d"][0], TWO_TRI_HIT["d"], rtol=1e-7)
    for f in ("p", "n", "uv"):
        np.testing.assert_allclose(r[f][0], TWO_TRI_HIT[f], atol=1e-7)
    # the ray from below meets the lower triangle first: v = 0.25, w = 0.125 again, t = 1
    r = O.intersect(3, s.bvh_node_data(), s.primitive_data(), np.array([[0.5, 0.25, -2.0]], np.float32),
                    np.array([[0.0, 0.0, 1.0]], np.float32))
    assert r["valid"][0] == 1 and r["prim_id"][0] == 1
    np.testing.assert_allclose(r["p"][0], [0.5, 0.25, -1.0], atol=1e-7)
    np.testing.assert_allclose(r["uv"][0], [0.625, 0.25], atol=1e-7)
    # the FP64 brute force reports the same record fields
    b = O.bf_intersect(3, TWO_TRI_V, TWO_TRI_F, *TWO_TRI_RAY)
    assert b["prim_id"][0] == 0 and b["t"][0] == 1.0 and b["second_t"][0] == 2.0
    np.testing.assert_allclose(b["uv"][0], [0.625, 0.25], atol=1e-15)


def test_ray_barycentrics_vs_fp64():
    """The FP32 restatement's ray barycentrics (1 - v - w, v) against the FP64 brute force's, away
    from grazing angles (|cos| of the hit angle above 0.2) and near ties."""
    V, F = synth.displaced_icosphere(3, seed=4)
    s = ext.SNCH3D(V, F)
    x = synth.uniform_points(6000, 3, seed=13)
    v = synth.uniform_directions(6000, 3, seed=14)
    r = O.intersect(3, s.bvh_node_data(), s.primitive_data(), x, v)
    b = O.bf_intersect(3, V, F, x, v)
    hit = np.isfinite(b["t"]) & parity.gap_mask(b["t"], b["second_t"]) & (r["prim_id"] == b["prim_id"])
    tri = V[F[np.where(hit, b["prim_id"], 0)]].astype(np.float64)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    cos = np.abs((nrm * v).sum(1)) / np.linalg.norm(nrm, axis=1)
    m = hit & (cos > 0.2)
    assert m.sum() > 1000
    assert np.abs(r["uv"][m] - b["uv"][m]).max() <= 1e-5
    # p = a + v (b - a) + w (c - a) of the FP64 barycentrics is the FP64 hit point x + t d
    u = b["uv"][m]
    T = tri[m]
    p = T[:, 0] + u[:, 1:] * (T[:, 1] - T[:, 0]) + (1 - u[:, :1] - u[:, 1:]) * (T[:, 2] - T[:, 0])
    np.testing.assert_allclose(p, x[m] + b["t"][m][:, None] * v[m], atol=1e-9)


def test_tetrahedron_silhouette_matches_brute_force():
    s = ext.SNCH3D(TET_V, TET_F)
    x = np.array([[1.5, 0.5, 0.5]], np.float32)
    r = O.silhouette(3, s.node_data(), s.silhouette_data(), x)
    b = O.bf_silhouette(3, s.raw_silhouette_data(), x)
    assert bool(r["valid"][0]) == bool(np.isfinite(b["d"][0]))
    if r["valid"][0]:
        np.testing.assert_allclose(r["d"][0], b["d"][0], rtol=1e-6)


# ---- FP64 brute force on synthetic scenes ------------------------------------------------------

SCENES = {
    "star1024": lambda: synth.star_polyline(1024),
    "star4096": lambda: synth.star_polyline(4096, seed=3),
    "ico3": lambda: synth.displaced_icosphere(3),
    "ico4": lambda: synth.displaced_icosphere(4, seed=5),
}


def _bf_records(kind, dim, s, V, F, x, v=None):
    if kind == "cp":
        b = O.bf_closest_point(dim, V, F, x)
        rec = {"valid": b["prim_id"] >= 0, "prim_id": b["prim_id"], "d": b["d"], "p": b["p"]}
        return rec, parity.gap_mask(b["d"], b["second_d"])
    if kind == "ray":
        b = O.bf_intersect(dim, V, F, x, v)
        hit = np.isfinite(b["t"])
        rec = {"valid": hit, "prim_id": np.where(hit, b["prim_id"], 0), "d": np.where(hit, b["t"], 0),
               "p": np.where(hit[:, None], x + v * np.where(hit, b["t"], 0)[:, None], 0)}
        return rec, parity.gap_mask(b["t"], b["second_t"])
    b = O.bf_silhouette(dim, s.raw_silhouette_data(), x)
    hit = np.isfinite(b["d"])
    rec = {"valid": hit, "prim_id": np.where(hit, b["prim_id"], 0), "d": np.where(hit, b["d"], 0), "p": b["p"]}
    return rec, parity.gap_mask(b["d"], b["second_d"])


@pytest.mark.parametrize("scene", sorted(SCENES))
@pytest.mark.parametrize("kind", ["cp", "ray", "sil"])
def test_oracle_vs_fp64(scene, kind):
    V, F = SCENES[scene]()
    dim = V.shape[1]
    s = (ext.SNCH if dim == 2 else ext.SNCH3D)(V, F)
    n = 3000 if dim == 2 else 1500
    x = synth.uniform_points(n, dim, seed=11)
    v = synth.uniform_directions(n, dim, seed=12)
    if kind == "cp":
        got = O.closest_point(dim, s.bvh_node_data(), s.primitive_data(), x)
    elif kind == "ray":
        got = O.intersect(dim, s.bvh_node_data(), s.primitive_data(), x, v)
    else:
        got = O.silhouette(dim, s.node_data(), s.silhouette_data(), x)
    assert got["overflows"] == 0
    ref, gap = _bf_records(kind, dim, s, V, F, x, v)
    res = parity.summarize(f"{scene}/{kind}", got, ref, gap, tol=parity.TOL_F64 if kind != "ray" else 1e-5)
    assert res["ok"], res


def test_planted_mistake_is_caught():
    """Shrinking every node box by 1% (a pruning bug) must break agreement with brute force."""
    V, F = synth.displaced_icosphere(3)
    s = ext.SNCH3D(V, F)
    nodes = s.bvh_node_data().reshape(-1, 9).copy()
    c = 0.5 * (nodes[:, :3] + nodes[:, 3:6])
    h = 0.5 * (nodes[:, 3:6] - nodes[:, :3]) * 0.99
    nodes[:, :3], nodes[:, 3:6] = c - h, c + h
    x = synth.uniform_points(1500, 3, seed=11)
    got = O.closest_point(3, nodes.reshape(-1), s.primitive_data(), x)
    ref, gap = _bf_records("cp", 3, s, V, F, x)
    assert not parity.summarize("planted", got, ref, gap, tol=parity.TOL_F64)["ok"]
