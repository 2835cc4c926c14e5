"""GPU: the silhouette setups on the device (configure_silhouette for polylines,
build_flat_silhouette_edges for meshes) are bit-identical to the host restatements, on closed and
open shapes, shuffled inputs, repeated and non-manifold edges, degenerate segments and triangles,
and the full-size configuration scenes; invalid inputs fail as on the host."""
import ctypes as C

import numpy as np
import pytest
import torch

from gquery_amd import _lib, ext, synth

pytestmark = pytest.mark.gpu


def _both_2d(V, E):
    host = ext.configure_silhouette(V, E)
    dev = ext.configure_silhouette(torch.from_numpy(np.asarray(V, np.float32)).cuda(),
                                   torch.from_numpy(np.asarray(E, np.int32)).cuda())
    return host, dev.cpu().numpy()


def _both_3d(V, F):
    host = ext.build_flat_silhouette_edges(V, F)
    dev = ext.build_flat_silhouette_edges(torch.from_numpy(np.asarray(V, np.float32)).cuda(),
                                          torch.from_numpy(np.asarray(F, np.int32)).cuda())
    return host, dev.cpu().numpy()


def _same(host, dev):
    assert host.shape == dev.shape, (host.shape, dev.shape)
    assert host.tobytes() == dev.tobytes()   # bit-identical


def _polylines():
    rng = np.random.default_rng(11)
    V, E = synth.star_polyline(1024)
    yield "star1024", V, E
    yield "star65536", *synth.star_polyline(65536)
    yield "open", V, E[:-1]
    yield "shuffled", V, E[rng.permutation(len(E))]
    yield "reversed", V, E[:, ::-1].copy()
    # a vertex with two outgoing and two incoming edges: the last ones set next / prev
    yield "nonmanifold", V, np.concatenate([E, E[:40, ::-1], E[5:9]])
    # nearly collinear neighbours are dropped; a repeated vertex gives a zero-length segment (NaN
    # direction, dropped)
    Vc = np.array([[0, 0], [1, 0], [2, 1e-6], [3, 0], [3, 0], [4, 2], [0, 3]], np.float32)
    Ec = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [6, 0]], np.int32)
    yield "collinear", Vc, Ec
    # unused vertices and edges listed out of order
    Vu = np.concatenate([V, rng.random((50, 2), dtype=np.float32)])
    yield "unused", Vu, E[::-1].copy()
    # examples/example_2d.py's pentagon
    yield "pentagon", np.array([[0.0, 0.0], [1.0, 0.0], [1.5, 1.0], [0.5, 1.5], [-0.5, 1.0]], np.float32), \
        np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 0]], np.int32)


def _meshes():
    rng = np.random.default_rng(12)
    for s in (1, 4, 6):
        yield f"ico{s}", *synth.displaced_icosphere(s)
    V, F = synth.displaced_icosphere(4, seed=6)
    yield "open", V, F[V[F].mean(1)[:, 2] < 0.7]
    yield "shuffled", V, F[rng.permutation(len(F))]
    yield "rotated", V, np.roll(F, 1, axis=1).copy()
    # non-manifold: every edge of 300 identical triangles is shared 300 times
    yield "dup", np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.tile(np.array([[0, 1, 2]], np.int32), (300, 1))
    # degenerate triangles (a repeated vertex gives an edge (i, i))
    Fd = F.copy()
    Fd[::7, 1] = Fd[::7, 0]
    yield "degenerate", V, Fd
    # examples/example_3d.py's tetrahedron
    yield "tet", np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1.0, 0.0], [0.5, 0.5, 1.0]], np.float32), \
        np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]], np.int32)


@pytest.mark.parametrize("case", list(range(9)))
def test_configure_silhouette_device(gpu, case):
    name, V, E = list(_polylines())[case]
    print("polyline", name, len(V), len(E), flush=True)
    host, dev = _both_2d(V, E)
    _same(host, dev)
    if name == "star1024":
        assert len(host) > 0.5 * len(V)   # (at 65,536 segments every vertex is nearly collinear)


@pytest.mark.parametrize("case", list(range(9)))
def test_flat_silhouette_edges_device(gpu, case):
    name, V, F = list(_meshes())[case]
    print("mesh", name, len(V), len(F), flush=True)
    host, dev = _both_3d(V, F)
    _same(host, dev)
    if name.startswith("ico"):
        assert len(host) == 19 * (3 * len(F) // 2)   # closed mesh: E = 3F / 2


def test_full_size_mesh_setup(gpu):
    # the 327,680-triangle configuration mesh, and the Mesh constructor's own records
    V, F = synth.displaced_icosphere(7)
    host, dev = _both_3d(V, F)
    _same(host, dev)
    raw = ext.SNCH3D(V, F).raw_silhouette_data()
    assert raw.tobytes() == dev.tobytes()


def test_component_first_inputs(gpu):
    from gquery_amd.fwd import Array2, Array2i
    V, E = synth.star_polyline(300)
    dev = ext.configure_silhouette(Array2(torch.from_numpy(V.T.copy()).cuda()), Array2i(E.T))
    _same(ext.configure_silhouette(V, E), dev.cpu().numpy())


def _err():
    return _lib.lib().gq_last_error().decode(errors="replace")


def _raw(fn, V, I, out, n):
    L = _lib.lib()
    return getattr(L, fn)(V.data_ptr(), V.shape[0], I.data_ptr(), I.shape[0], None if out is None else out.data_ptr(),
                          C.byref(n), None)


@pytest.mark.parametrize("fn,dim", [("gq_configure_silhouette_device", 2), ("gq_build_flat_silhouette_edges_device", 3)])
def test_invalid_inputs(gpu, fn, dim):
    V, I = synth.star_polyline(64) if dim == 2 else synth.displaced_icosphere(1)
    Vt, It = torch.from_numpy(V).cuda(), torch.from_numpy(I.astype(np.int32)).cuda()
    n = C.c_int64(0)
    assert _raw(fn, Vt, It, None, n) == 0 and n.value > 0
    need = n.value
    # an output buffer one record short fails and reports the size
    out = torch.full((need,), -7.0, device="cuda")
    n = C.c_int64(need - 1)
    assert _raw(fn, Vt, It, out, n) == -1 and n.value == need
    assert "too small" in _err()
    assert (out == -7.0).all()
    # an index out of range and a non-finite coordinate fail as on the host
    bad = It.clone()
    bad[3, 1] = len(V)
    assert _raw(fn, Vt, bad, None, C.c_int64(0)) == -1 and "out of range" in _err()
    badv = Vt.clone()
    badv[2, 0] = float("nan")
    assert _raw(fn, badv, It, None, C.c_int64(0)) == -1 and "non-finite" in _err()
    # and the device is still fine
    host = ext.configure_silhouette(V, I) if dim == 2 else ext.build_flat_silhouette_edges(V, I)
    dev = ext.configure_silhouette(Vt, It) if dim == 2 else ext.build_flat_silhouette_edges(Vt, It)
    _same(host, dev.cpu().numpy())
