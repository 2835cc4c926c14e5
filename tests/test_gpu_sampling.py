"""GPU: BVH-guided boundary sampling (bvh3d.py:246-346 and its 2D form) through the C ABI against
the CPU oracle, bit for bit: hit flags, points, normals, uv / t, pdf, pmf, primitive ids and the
advanced PCG32 states (same float32 operations in the same order, integer generator)."""
import numpy as np
import pytest
import torch

from gquery_amd import synth
from gquery_amd.sampler import PCG32
from oracle import oracle as O
from tests.test_gpu_parity import scene

pytestmark = pytest.mark.gpu

FIELDS = ("p", "n", "uv", "pdf", "pmf", "prim_id")


def _u64(t):
    return t.cpu().numpy().view(np.uint64).copy()


def _check(shape, dim, x, R, g):
    """One call on the device and the oracle from the same generator states; returns the
    oracle's states after the call."""
    st0, inc = _u64(g.state), _u64(g.inc)
    hit, rec = shape.sample_boundary(x, R, g)
    torch.cuda.synchronize()
    ref, st_ref = O.sample_boundary(dim, shape.snch.bvh_node_data(), shape.snch.primitive_data(), x, R, st0, inc)
    assert np.array_equal(hit.cpu().numpy().astype(np.uint8), ref["hit"])
    got = {"p": rec.p, "n": rec.n, "uv": rec.uv if dim == 3 else rec.t, "pdf": rec.pdf, "pmf": rec.pmf,
           "prim_id": rec.prim_id}
    for f in FIELDS:
        a, b = got[f].cpu().numpy(), ref[f]
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b), f
    assert np.array_equal(_u64(g.state), st_ref)
    return ref


@pytest.mark.parametrize("radius", ["scalar", "array"])
@pytest.mark.parametrize("name", ["ico3", "ico5", "oico4", "star1024", "ostar1024"])
def test_sample_boundary_matches_oracle(gpu, name, radius):
    V, F, shape = scene(name)
    dim = V.shape[1]
    n = 8192
    x = synth.uniform_points(n, dim, seed=51)
    R = 0.3 if radius == "scalar" else np.random.default_rng(52).uniform(0.02, 1.5, n).astype(np.float32)
    g = PCG32(n, initstate=12345, device="cuda:0")
    ref = _check(shape, dim, x, R, g)
    assert 0.1 < ref["hit"].mean() < 1.0
    ref = _check(shape, dim, x, R, g)                  # the streams continue where they stopped
    assert ref["hit"].any()


def test_sample_boundary_full_size(gpu):
    """1M queries against the 327,680-triangle icosphere, every record against the oracle."""
    V, F = synth.displaced_icosphere(7)
    from gquery_amd.shapes.mesh import Mesh
    shape = Mesh(V, F, device="cuda:0")
    n = 1 << 20
    x = torch.from_numpy(synth.uniform_points(n, 3, seed=53)).cuda()
    g = PCG32(n, device="cuda:0")
    ref = _check(shape, 3, x.cpu().numpy(), 0.25, g)
    assert ref["hit"].mean() > 0.05                     # 11.5% of the cube's points are that close


@pytest.mark.parametrize("name", ["ico3", "star1024"])
def test_sample_boundary_edge_cases(gpu, name):
    V, F, shape = scene(name)
    dim = V.shape[1]
    x = synth.uniform_points(100, dim, seed=54)
    g = PCG32(100, device="cuda:0")
    st0 = _u64(g.state)
    hit, rec = shape.sample_boundary(x, -1.0, g)       # empty balls: nothing drawn, zero records
    assert not hit.any() and not rec.pdf.any() and not rec.p.any() and not rec.prim_id.any()
    assert np.array_equal(_u64(g.state), st0)
    _check(shape, dim, x, float("inf"), g)              # every primitive reachable
    hit, rec = shape.sample_boundary(np.zeros((0, dim), np.float32), 1.0, PCG32(0, device="cuda:0"))
    assert hit.numel() == 0 and rec.p.shape == (0, dim)
    with pytest.raises(ValueError):
        shape.sample_boundary(x, 1.0, PCG32(99, device="cuda:0"))
    with pytest.raises(ValueError):
        shape.sample_boundary(x, 1.0, PCG32(100, device="cpu"))


def test_generator_draws_on_the_device(gpu):
    """PCG32.next_uint32 / next_float32 run where the state lives (wrapping int64 tensor
    arithmetic) and continue the same streams the sampling kernel advances."""
    g = PCG32(4096, initstate=99, initseq=7, device="cuda:0")
    st, inc = _u64(g.state), _u64(g.inc)
    want, _ = O.pcg32_stream(st, inc, 4)
    for k in range(3):
        got = g.next_uint32()
        assert got.device.type == "cuda"
        assert np.array_equal(got.cpu().numpy().astype(np.uint32), want[:, k])
    f = g.next_float32().cpu().numpy()
    assert np.array_equal(f, ((want[:, 3] >> np.uint32(9)) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1))


def test_zero_area_leaf_gives_no_sample(gpu):
    """The kernel's handling of a leaf whose visible primitives all have zero area (no sample),
    bit for bit against the oracle, beside a leaf with a real triangle."""
    from gquery_amd.shapes.mesh import Mesh
    V = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [5, 5, 5], [6, 5, 5], [5, 6, 5]], np.float32)
    F = np.array([[0, 1, 2], [4, 5, 6]], np.int32)
    shape = Mesh(V, F, device="cuda:0")
    x = np.repeat(np.array([[1.0, 0.0, 0.1]], np.float32), 2048, 0)
    for R in (0.5, 100.0):
        ref = _check(shape, 3, x, R, PCG32(len(x), initstate=5, device="cuda:0"))
        assert (ref["hit"] == 0).all() if R == 0.5 else 0 < ref["hit"].mean() < 1
