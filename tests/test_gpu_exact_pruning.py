"""GPU: the engine's closest-point pruning is exact at any coordinate scale.

A traversal that never drops a primitive the float32 leaf test would accept returns, bit for bit,
the smallest float32 leaf distance over all primitives.  For meshes that is the GPU brute-force
baseline (Mesh.closest_point_baseline, mesh.py:92-118, which calls the same closest_point_triangle,
math.py:42-107); for polylines the oracle's float32 brute force with the traversal's own segment
formula (math.py:14-21), compared with the leaf distance of the segment the engine chose.  Scenes
are translated and scaled far from the unit cube, where float32 rounding of the leaf distance is
absolute in the coordinates; queries sit from 1e-6 to 0.3 of the scene size off the surface, plus
uniform ones.  Both node formats, with and without the
wave-per-query pass (a 6-step packet budget defers nearly every query to it).
"""
import numpy as np
import pytest
import torch

from gquery_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCALES = [((0.0, 0.0, 0.0), 1.0), ((1000.0, -700.0, 300.0), 1.0), ((0.0, 0.0, 0.0), 1e3),
          ((0.0, 0.0, 0.0), 1e-3), ((5e3, 0.0, 0.0), 10.0)]


def _near_surface(V, F, n, scale, seed):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, F.shape[0], n)
    P = V[F[f]].astype(np.float64)                       # (n, D, D)
    D = V.shape[1]
    w = rng.dirichlet(np.ones(D), n)
    on = (w[:, :, None] * P).sum(1)
    if D == 3:
        nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    else:
        e = P[:, 1] - P[:, 0]
        nrm = np.stack([e[:, 1], -e[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    h = scale * 10.0 ** rng.uniform(-6, np.log10(0.3), n) * rng.choice([-1.0, 1.0], n)
    return (on + h[:, None] * nrm).astype(np.float32)


def _queries(V, F, offset, scale, n, seed):
    D = V.shape[1]
    u = (synth.uniform_points(n // 2, D, seed=seed).astype(np.float64) * scale
         + np.asarray(offset[:D])).astype(np.float32)
    return np.concatenate([u, _near_surface(V, F, n - n // 2, scale, seed + 1)])


@pytest.mark.parametrize("offset,scale", SCALES)
@pytest.mark.parametrize("method", ["engine", "engine_fp32"])
@pytest.mark.parametrize("budget", [0, 6])
def test_exact_pruning_3d(gpu, offset, scale, method, budget):
    from gquery_amd.shapes.mesh import Mesh
    V, F = synth.displaced_icosphere(5, seed=7)
    V = (V.astype(np.float64) * scale + np.asarray(offset)).astype(np.float32)
    mesh = Mesh(V, F, device="cuda:0", packet_budget=budget)
    x = _queries(V, F, offset, scale, 65536, seed=31)
    print(f"[exact 3D] offset={offset} scale={scale} method={method} budget={budget}", flush=True)
    xt = torch.from_numpy(x).cuda()
    got = mesh.closest_point_bvh(xt, method=method)
    bf = mesh.closest_point_baseline(xt)
    torch.cuda.synchronize()
    gd, bd = got.d.cpu().numpy(), bf.d.cpu().numpy()
    assert got.valid.all().item() and bf.valid.all().item()
    bad = np.flatnonzero(gd.view(np.int32) != bd.view(np.int32))
    assert bad.size == 0, (bad.size, gd[bad[:4]], bd[bad[:4]])
    # the same triangle or an exactly tied one wins (exact ties are common: a nearest point on a
    # shared vertex is equally near on every triangle around it); the record is rebuilt from the
    # winning triangle with the leaf formula, so equal ids give bit-identical records
    same = got.prim_id.cpu().numpy() == bf.prim_id.cpu().numpy()
    assert same.mean() > 0.5
    for f in ("p", "uv"):
        a, b = getattr(got, f).cpu().numpy()[same], getattr(bf, f).cpu().numpy()[same]
        assert np.array_equal(a.view(np.int32), b.view(np.int32)), f
    np.testing.assert_allclose(got.n.cpu().numpy()[same], bf.n.cpu().numpy()[same], atol=1e-6)


@pytest.mark.parametrize("offset,scale", SCALES)
@pytest.mark.parametrize("method", ["engine", "engine_fp32"])
@pytest.mark.parametrize("budget", [0, 6])
def test_exact_pruning_2d(gpu, offset, scale, method, budget):
    from gquery_amd.shapes.polyline import Polyline
    V, F = synth.star_polyline(4096, seed=9)
    V = (V.astype(np.float64) * scale + np.asarray(offset[:2])).astype(np.float32)
    poly = Polyline(V, F, device="cuda:0", packet_budget=budget)
    x = _queries(V, F, offset, scale, 32768, seed=41)
    print(f"[exact 2D] offset={offset} scale={scale} method={method} budget={budget}", flush=True)
    got = poly.closest_point_bvh(torch.from_numpy(x).cuda(), method=method)
    torch.cuda.synchronize()
    ref = O.bf_closest_point_f32(2, poly.snch.primitive_data(), x)
    assert got.valid.all().item()
    # the 2D record's distance is LineSegment.closest_point's (line_segment.py:63-78), not the leaf
    # test's (math.py:14-21): compare the leaf distance of the segment the engine chose
    ld = O.leaf_distance(2, poly.snch.primitive_data(), got.prim_id.cpu().numpy(), x)
    bad = np.flatnonzero(ld.view(np.int32) != ref["d"].view(np.int32))
    assert bad.size == 0, (bad.size, ld[bad[:4]], ref["d"][bad[:4]])
