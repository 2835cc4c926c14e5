"""CPU builder: bit-exact agreement with an independent Python restatement of the reference
builder (oracle/build_ref.py) on small inputs, and structural invariants on large ones."""
import numpy as np
import pytest

from gquery_amd import ext, synth
from oracle import build_ref

SQUARE_V = np.array([[0., 0.], [1., 0.], [0., 1.], [1., 1.]], np.float32)   # tests/test_bvh.py:7-8
SQUARE_E = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
PENTAGON_V = np.array([[0.0, 0.0], [1.0, 0.0], [1.5, 1.0], [0.5, 1.5], [-0.5, 1.0]], np.float32)  # examples/example_2d.py:7-20
PENTAGON_E = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 0]], np.int32)
TET_V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1.0, 0.0], [0.5, 0.5, 1.0]], np.float32)  # examples/example_3d.py:7-26
TET_F = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]], np.int32)


def _soup(n, dim, seed, degenerate=False):
    rng = np.random.default_rng(seed)
    V = rng.uniform(-1, 1, size=(n * dim, dim)).astype(np.float32)
    F = np.arange(n * dim, dtype=np.int32).reshape(n, dim)
    if degenerate:
        V[dim:2 * dim] = V[0:dim]          # duplicate primitive
        V[2 * dim:3 * dim] = V[2 * dim]    # zero-size primitive
        V[3 * dim + 1] = V[3 * dim]        # sliver
    return V, F


CASES = {
    "square": (SQUARE_V, SQUARE_E),
    "pentagon": (PENTAGON_V, PENTAGON_E),
    "star64": synth.star_polyline(64, seed=5),
    "star300": synth.star_polyline(300, seed=7),
    "tetrahedron": (TET_V, TET_F),
    "ico1": synth.displaced_icosphere(1, seed=3),
    "ico2": synth.displaced_icosphere(2, seed=4),
    "soup2d": _soup(120, 2, 1, True),
    "soup3d": _soup(90, 3, 2, True),
    "grid2d": (np.array([[i % 7, i // 7] for i in range(49)], np.float32),
               np.array([[i, i + 1] for i in range(48)], np.int32)),
}


def _ulp_close(a, b, ulps):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib) <= ulps


@pytest.mark.parametrize("name", sorted(CASES))
def test_builder_matches_restatement(name):
    V, F = CASES[name]
    dim = V.shape[1]
    s = (ext.SNCH if dim == 2 else ext.SNCH3D)(V, F)
    snode, sil, raw, bnode, prim = build_ref.build_snch(V, F)
    ns = 13 if dim == 2 else 16
    got_b = s.bvh_node_data().reshape(-1, 2 * dim + 3)
    got_s = s.node_data().reshape(-1, ns)
    assert got_b.shape == bnode.shape
    np.testing.assert_array_equal(got_b.view(np.int32), bnode.view(np.int32))
    np.testing.assert_array_equal(s.primitive_data().reshape(prim.shape).view(np.int32), prim.view(np.int32))
    np.testing.assert_array_equal(s.silhouette_data().reshape(sil.shape).view(np.int32), sil.view(np.int32))
    np.testing.assert_array_equal(s.raw_silhouette_data().reshape(raw.shape).view(np.int32), raw.view(np.int32))
    # all SNCH fields bit-exact except the half angle, whose acos comes from libm (float) in the
    # product and from a correctly rounded double acos here: allow 2 ulps.
    half = 3 * dim
    mask = np.ones(ns, bool)
    mask[half] = False
    np.testing.assert_array_equal(got_s[:, mask].view(np.int32), snode[:, mask].view(np.int32))
    assert _ulp_close(got_s[:, half], snode[:, half], 2).all()


def test_bvh_and_snch_share_tree():
    V, F = synth.displaced_icosphere(2, seed=9)
    b = ext.BVH3D(V, F)
    s = ext.SNCH3D(V, F)
    np.testing.assert_array_equal(b.node_data(), s.bvh_node_data())
    np.testing.assert_array_equal(b.primitive_data(), s.primitive_data())


def _check_invariants(s, V, F):
    dim = V.shape[1]
    info = s.info()
    nodes = s.bvh_node_data().reshape(-1, 2 * dim + 3)
    snodes = s.node_data().reshape(-1, 3 * dim + 7)
    prims = s.primitive_data().reshape(-1, dim * dim + 1)
    sil = s.silhouette_data().reshape(-1, 10 if dim == 2 else 19)
    n = nodes.shape[0]
    assert n == info["n_nodes"] and n == 2 * F.shape[0] - 1 or n <= 2 * F.shape[0] - 1
    # ordered primitives are a permutation of the input, with their own vertices
    ids = prims[:, -1].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(F.shape[0]))
    np.testing.assert_array_equal(prims[:, :-1].reshape(-1, dim, dim), V[F[ids]])
    covered = np.zeros(F.shape[0], np.int64)
    lo, hi = nodes[:, :dim], nodes[:, dim:2 * dim]
    cnt = nodes[:, 2 * dim + 2].astype(np.int64)
    off = nodes[:, 2 * dim].astype(np.int64)
    sec = nodes[:, 2 * dim + 1].astype(np.int64)
    for i in range(n):
        if cnt[i] > 0:
            sl = slice(off[i], off[i] + cnt[i])
            covered[sl] += 1
            pts = prims[sl, :-1].reshape(-1, dim)
            assert (pts >= lo[i]).all() and (pts <= hi[i]).all()
            assert sec[i] == 0
        else:
            l, r = i + 1, i + sec[i]
            assert i < r < n
            np.testing.assert_array_equal(lo[i], np.minimum(lo[l], lo[r]))
            np.testing.assert_array_equal(hi[i], np.maximum(hi[l], hi[r]))
    assert (covered == 1).all()
    # each leaf's silhouettes are exactly the vertices / edges of its primitives, deduplicated
    sil_off = snodes[:, 3 * dim + 5].astype(np.int64)
    sil_cnt = snodes[:, 3 * dim + 6].astype(np.int64)
    pid = sil[:, -1].astype(np.int64)
    for i in np.flatnonzero(cnt > 0)[:2000]:
        f = F[ids[off[i]:off[i] + cnt[i]]]
        if dim == 2:
            want = set(f.reshape(-1).tolist())
            got = pid[sil_off[i]:sil_off[i] + sil_cnt[i]]
            assert set(got.tolist()) == want and len(got) == len(want)
        else:
            got = sil[sil_off[i]:sil_off[i] + sil_cnt[i]]
            want = {tuple(sorted((int(a), int(b)))) for t in f for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
            have = {(int(r[13]), int(r[14])) for r in got}
            assert have == want and len(got) == len(want)
    # cones: a valid, non-degenerate cone contains the normal_0 of every silhouette below it
    half = snodes[:, 3 * dim]
    assert ((half == np.float32(-np.pi)) | (half >= 0)).all()


@pytest.mark.parametrize("which", ["star65536", "ico5", "soup3d_big"])
def test_builder_invariants_large(which):
    if which == "star65536":
        V, F = synth.star_polyline(65536)
    elif which == "ico5":
        V, F = synth.displaced_icosphere(5)
    else:
        V, F = _soup(3000, 3, 11, True)
    dim = V.shape[1]
    s = (ext.SNCH if dim == 2 else ext.SNCH3D)(V, F)
    _check_invariants(s, V, F)


def test_builder_deterministic():
    V, F = synth.displaced_icosphere(4, seed=2)
    a = ext.SNCH3D(V, F)
    b = ext.SNCH3D(V, F)
    for f in ("node_data", "primitive_data", "silhouette_data"):
        np.testing.assert_array_equal(getattr(a, f)(), getattr(b, f)())


def test_flat_silhouette_edges():
    V, F = synth.displaced_icosphere(1, seed=1)
    got = ext.build_flat_silhouette_edges(V, F).reshape(-1, 19)
    raw, _ = build_ref.silhouettes(V, F)
    np.testing.assert_array_equal(got, np.array(raw, np.float32))
    s = ext.SNCH3D(V, F)
    np.testing.assert_array_equal(got.reshape(-1), s.raw_silhouette_data())
    # closed genus-0 mesh: E = 3F/2, each edge has both faces
    assert got.shape[0] == 3 * F.shape[0] // 2


@pytest.mark.parametrize("bad", ["index", "nan", "empty", "dim"])
def test_builder_rejects_bad_input(bad):
    V, F = PENTAGON_V.copy(), PENTAGON_E.copy()
    if bad == "index":
        F[2, 1] = 99
    elif bad == "nan":
        V[1, 0] = np.nan
    elif bad == "empty":
        F = F[:0]
    with pytest.raises(ValueError):
        if bad == "dim":
            ext.BVH(np.zeros((4, 5), np.float32), F)
        else:
            ext.SNCH(V, F)
