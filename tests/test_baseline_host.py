"""CPU: Polyline.configure_silhouette on the host, and the oracle's brute-force baselines and mesh
normals pinned against FP64 (the baseline GPU kernels are checked against these in
test_gpu_baseline.py)."""
import numpy as np
import pytest

from gquery_amd import ext, synth
from oracle import oracle as O
from tests import parity


def _star_edges(n):
    return np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)


def _polylines():
    V, E = synth.star_polyline(300, seed=3)
    open_E = E[:-7]                                   # an open polyline: both ends have one neighbour
    rng = np.random.default_rng(5)
    shuffled = E[rng.permutation(len(E))]             # first-appearance order differs from vertex order
    # collinear vertices: midpoints inserted on every third segment (dropped by the filter)
    Vc, Ec = [], []
    for k, (a, b) in enumerate(E):
        Vc.append(V[a])
        if k % 3 == 0:
            Vc.append(0.5 * (V[a] + V[b]))
    Vc = np.asarray(Vc, np.float32)
    Ec = _star_edges(len(Vc))
    return {"closed": (V, E), "open": (V, open_E), "shuffled": (V, shuffled), "collinear": (Vc, Ec)}


@pytest.mark.parametrize("name", ["closed", "open", "shuffled", "collinear"])
def test_configure_silhouette_matches_restatement(name):
    V, E = _polylines()[name]
    got = ext.configure_silhouette(V, E)
    want = O.configure_silhouette(V, E)
    assert got.shape == want.shape and got.shape[1] == 10
    np.testing.assert_array_equal(got, want)
    if name == "closed":
        # first-appearance order is vertex order; only nearly straight joints are dropped
        np.testing.assert_array_equal(got[:, 9], got[:, 7])
        assert len(got) > 0.9 * len(V)
    if name == "open":
        ends = {int(E[0, 0]), int(E[-1, 1])}          # one neighbour each: never silhouettes
        assert not ends & set(got[:, 7].astype(int).tolist()) and len(got) < len(E)
    if name == "collinear":
        # the inserted midpoints are exactly collinear with their neighbours: all dropped
        mids = {sum(1 for j in range(k) if j % 3 == 0) + k + 1 for k in range(0, 300, 3)}
        assert not mids & set(got[:, 7].astype(int).tolist())


def test_configure_silhouette_index_is_first_appearance():
    # vertex 5 first appears in the second edge: its record's index is 2, not 5
    V = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [2, 2], [3, 1]], np.float32)
    E = np.array([[0, 1], [1, 5], [5, 2], [2, 0]], np.int32)
    got = ext.configure_silhouette(V, E)
    by_vertex = {int(r[7]): int(r[9]) for r in got}
    assert by_vertex == {0: 0, 1: 1, 5: 2, 2: 3}


def test_normals_fp64():
    V, F = synth.displaced_icosphere(3)
    fn, vn = O.normals(V, F)
    Vd = V.astype(np.float64)
    n = np.cross(Vd[F[:, 1]] - Vd[F[:, 0]], Vd[F[:, 2]] - Vd[F[:, 0]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.abs(fn - n).max() < 1e-6
    acc = np.zeros_like(Vd)
    for k in range(3):
        np.add.at(acc, F[:, k], n)
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    assert np.abs(vn - acc).max() < 1e-6
    # outward on the (star-shaped) displaced sphere
    assert ((vn * V).sum(1) > 0).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_baselines_vs_fp64(dim):
    if dim == 2:
        V, F = synth.star_polyline(512, seed=4)
    else:
        V, F = synth.displaced_icosphere(3, seed=4)
    x = synth.uniform_points(3000, dim, seed=9)
    v = synth.uniform_directions(3000, dim, seed=10)
    cp = O.closest_point_baseline(dim, V, F, x)
    b = O.bf_closest_point(dim, V, F, x)
    assert cp["valid"].all()
    assert parity.rel_err(cp["d"], b["d"]).max() <= parity.TOL_F64
    gap = parity.gap_mask(b["d"], b["second_d"])
    np.testing.assert_array_equal(cp["prim_id"][gap], b["prim_id"][gap])
    ray = O.intersect_baseline(dim, V, F, x, v)
    b = O.bf_intersect(dim, V, F, x, v)
    bv = np.isfinite(b["t"])
    np.testing.assert_array_equal(ray["valid"].astype(bool), bv)
    assert parity.rel_err(ray["d"][bv], b["t"][bv]).max() <= 1e-5
    gap = parity.gap_mask(b["t"], b["second_t"]) & bv
    np.testing.assert_array_equal(ray["prim_id"][gap], b["prim_id"][gap])
    # silhouettes: the configured silhouette vertices (2D) / every mesh edge (3D), brute force in FP64
    sils = ext.configure_silhouette(V, F) if dim == 2 else ext.build_flat_silhouette_edges(V, F)
    sil = O.silhouette_baseline(dim, sils, x)
    b = O.bf_silhouette(dim, sils, x)
    bv = np.isfinite(b["d"])
    np.testing.assert_array_equal(sil["valid"].astype(bool), bv)
    assert parity.rel_err(sil["d"][bv], b["d"][bv]).max() <= parity.TOL_F64
    # boundary offset: a ray from the surface along -n, flagged on_boundary, does not hit its own
    # primitive at t ~ 0
    off = O.intersect_baseline(dim, V, F, cp["p"], -cp["n"], normal=cp["n"], on_boundary=np.ones(len(x), np.uint8))
    hit = off["valid"].astype(bool)
    assert (off["d"][hit] > 1e-6).all()
