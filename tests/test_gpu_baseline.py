"""GPU: the brute-force (baseline) kernels and mesh normals through the C ABI, bit for bit against
the oracle's FP32 restatement (oracle/gq_oracle.c Part 3), and consistent with the BVH queries."""
import numpy as np
import pytest
import torch

from gquery_amd import ext, synth
from gquery_amd.shapes.mesh import Mesh
from gquery_amd.shapes.polyline import Polyline
from oracle import oracle as O
from tests import parity

pytestmark = pytest.mark.gpu

FIELDS = ("valid", "p", "n", "uv", "t", "d", "prim_id", "on_boundary")


def _np(r):
    torch.cuda.synchronize()
    return {f: getattr(r, f).cpu().numpy() for f in FIELDS if hasattr(r, f)}


def _same(got, ref, fields):
    for f in fields:
        if f not in got:
            continue
        g, r = np.asarray(got[f]), np.asarray(ref[f])
        if g.dtype == np.float32:
            assert np.array_equal(g.view(np.int32), r.reshape(g.shape).view(np.int32)), f
        else:
            assert np.array_equal(g.astype(np.int64), r.reshape(g.shape).astype(np.int64)), f


def _scene(name):
    if name == "star":
        V, F = synth.star_polyline(700, seed=12)
    elif name == "ostar":
        V, F = synth.star_polyline(700, seed=12)
        F = F[:-40]
    elif name == "ico":
        V, F = synth.displaced_icosphere(3, seed=13)
    else:   # open mesh: a band of the icosphere's faces removed
        V, F = synth.displaced_icosphere(3, seed=13)
        F = F[(V[F[:, 0], 2] < 0.3) | (V[F[:, 0], 2] > 0.6)]
    shape = (Polyline if V.shape[1] == 2 else Mesh)(V, F, device="cuda:0")
    return V, F, shape


@pytest.mark.parametrize("name", ["star", "ostar", "ico", "oico"])
def test_baselines_bit_exact(gpu, name):
    V, F, shape = _scene(name)
    dim = V.shape[1]
    n = 5000
    x = synth.uniform_points(n, dim, seed=41)
    v = synth.uniform_directions(n, dim, seed=42)
    r_arr = np.random.default_rng(43).uniform(0.05, 2.0, n).astype(np.float32)
    fn = O.normals(V, F)[0] if dim == 3 else None
    _same(_np(shape.closest_point_baseline(x)), O.closest_point_baseline(dim, V, F, x, fn), FIELDS)
    for r_max in (float("inf"), r_arr):
        got = _np(shape.intersect_baseline(x, v, r_max=torch.as_tensor(r_max) if np.ndim(r_max) else r_max))
        _same(got, O.intersect_baseline(dim, V, F, x, v, r_max), FIELDS)
        sils = shape.silhouettes.cpu().numpy()
        got = _np(shape.closest_silhouette_baseline(x, r_max=torch.as_tensor(r_max) if np.ndim(r_max) else r_max))
        _same(got, O.silhouette_baseline(dim, sils, x, r_max), FIELDS)
    # Polyline/Mesh.intersect: origins on the boundary move by -1e-5 n before the brute force
    cp = _np(shape.closest_point_baseline(x))
    flag = (np.arange(n) % 2).astype(bool)
    got = _np(shape.intersect(cp["p"], -cp["n"], n=cp["n"], on_boundary=torch.as_tensor(flag)))
    _same(got, O.intersect_baseline(dim, V, F, cp["p"], -cp["n"], normal=cp["n"], on_boundary=flag.astype(np.uint8)),
          FIELDS)


@pytest.mark.parametrize("name", ["star", "ico"])
def test_baselines_agree_with_bvh_queries(gpu, name):
    V, F, shape = _scene(name)
    dim = V.shape[1]
    x = synth.uniform_points(4000, dim, seed=51)
    v = synth.uniform_directions(4000, dim, seed=52)
    a, b = _np(shape.closest_point_baseline(x)), _np(shape.closest_point_bvh(x))
    assert a["valid"].all() and b["valid"].all()
    assert parity.rel_err(a["d"], b["d"]).max() <= parity.TOL_REF
    a, b = _np(shape.intersect_baseline(x, v)), _np(shape.intersect_bvh(x, v))
    assert np.array_equal(a["valid"], b["valid"])
    assert parity.rel_err(a["d"][a["valid"]], b["d"][b["valid"]]).max() <= parity.TOL_REF
    if dim == 3:   # the SNCH's silhouettes are the same edges (closed mesh)
        a, b = _np(shape.closest_silhouette_baseline(x)), _np(shape.closest_silhouette_snch(x))
        assert np.array_equal(a["valid"], b["valid"])
        assert parity.rel_err(a["d"][a["valid"]], b["d"][b["valid"]]).max() <= parity.TOL_REF


def test_normals_bit_exact(gpu):
    V, F, shape = _scene("ico")
    fn, vn = O.normals(V, F)
    torch.cuda.synchronize()
    assert np.array_equal(shape.face_normals.cpu().numpy().view(np.int32), fn.view(np.int32))
    assert np.array_equal(shape.normals.cpu().numpy().view(np.int32), vn.view(np.int32))


def test_shape_info_and_errors(gpu):
    V, F, shape = _scene("star")
    bf = shape._brute_force()
    info = bf.info()
    assert info["dim"] == 2 and info["n_prims"] == len(F) and info["n_silhouettes"] == len(shape.silhouettes)
    with pytest.raises(NotImplementedError):
        bf.normals()
    with pytest.raises(ValueError):
        Polyline(V, np.array([[0, 10 ** 6]], np.int32), device="cuda:0").closest_point_baseline(V[:3])
    empty = _np(shape.closest_point_baseline(np.zeros((0, 2), np.float32)))
    assert empty["d"].shape == (0,)
