"""GPU parity: every query path through the C ABI on the device, against the CPU oracle.

Plain kernels restate the reference traversal in the same arithmetic as the oracle, so they must
agree with it bit for bit (transcendentals in the SNCH cone test aside).  Engine kernels must meet
the acceptance rule of tests/parity.py against the oracle and the FP64 brute force."""
import numpy as np
import pytest

from gquery_amd import ext, synth
from oracle import oracle as O
from tests import parity
from tests.test_builder import PENTAGON_E, PENTAGON_V, TET_F, TET_V
from tests.test_oracle import TWO_TRI_F, TWO_TRI_HIT, TWO_TRI_RAY, TWO_TRI_V

pytestmark = pytest.mark.gpu

METHODS = ["plain", "engine_fp32", "engine"]

_SCENES = {}


def scene(name):
    """Cached (V, F, Mesh/Polyline) for a named synthetic scene."""
    if name not in _SCENES:
        from gquery_amd.shapes.mesh import Mesh
        from gquery_amd.shapes.polyline import Polyline
        if name.startswith("star"):
            V, F = synth.star_polyline(int(name[4:]))
            _SCENES[name] = (V, F, Polyline(V, F, device="cuda:0"))
        elif name.startswith("ostar"):
            # open polyline: the closing segment removed (end vertices with one neighbour)
            V, F = synth.star_polyline(int(name[5:]))
            F = F[:-1]
            _SCENES[name] = (V, F, Polyline(V, F, device="cuda:0"))
        elif name.startswith("oico"):
            # open mesh: a cap of faces removed (boundary edges with one adjacent face)
            V, F = synth.displaced_icosphere(int(name[4:]))
            F = F[V[F].mean(1)[:, 2] < 0.7]
            _SCENES[name] = (V, F, Mesh(V, F, device="cuda:0"))
        else:
            V, F = synth.displaced_icosphere(int(name[3:]))
            _SCENES[name] = (V, F, Mesh(V, F, device="cuda:0"))
    return _SCENES[name]


def _np(r, fields=("valid", "prim_id", "d", "p", "n", "uv", "t", "on_boundary")):
    import torch
    torch.cuda.synchronize()
    return {f: getattr(r, f).cpu().numpy() for f in fields if hasattr(r, f)}


def _oracle(kind, dim, s, x, v=None, r_max=float("inf")):
    if kind == "cp":
        return O.closest_point(dim, s.bvh_node_data(), s.primitive_data(), x)
    if kind == "ray":
        return O.intersect(dim, s.bvh_node_data(), s.primitive_data(), x, v, r_max)
    return O.silhouette(dim, s.node_data(), s.silhouette_data(), x, r_max)


def _query(shape, kind, method, x, v=None, r_max=float("inf")):
    if kind == "cp":
        return shape.closest_point_bvh(x, method=method)
    if kind == "ray":
        return shape.intersect_bvh(x, v, r_max, method=method)
    return shape.closest_silhouette_snch(x, r_max, method=method)


def _gap_from_bruteforce(kind, dim, s, V, F, x, v, ref=None):
    if kind == "cp":
        b = O.bf_closest_point(dim, V, F, x)
        return parity.gap_mask(b["d"], b["second_d"])
    if kind == "ray":
        b = O.bf_intersect(dim, V, F, x, v)
        return parity.gap_mask(b["t"], b["second_t"])
    b = O.bf_silhouette(dim, s.raw_silhouette_data(), x)
    gap = parity.gap_mask(b["d"], b["second_d"])
    if ref is not None:
        # On open meshes the reference cone test can prune the true nearest silhouette (its cones
        # are built from the zero-filled missing-face slots); where the reference answer is not the
        # brute-force one, the runner-up gap says nothing about ties among what it does find.
        gap &= np.abs(ref["d"] - b["d"]) <= 1e-6 * np.maximum(b["d"], 1.0)
    return gap


@pytest.mark.parametrize("name", ["star1024", "star4096", "ico3", "ico5", "ostar1024", "oico4"])
@pytest.mark.parametrize("kind", ["cp", "ray", "sil"])
@pytest.mark.parametrize("method", METHODS)
def test_parity_vs_oracle(gpu, name, kind, method):
    V, F, shape = scene(name)
    dim = V.shape[1]
    n = 4096
    x = synth.uniform_points(n, dim, seed=21)
    v = synth.uniform_directions(n, dim, seed=22)
    ref = _oracle(kind, dim, shape.snch, x, v)
    got = _np(_query(shape, kind, method, x, v))
    gap = _gap_from_bruteforce(kind, dim, shape.snch, V, F, x, v, ref)
    keep = np.ones(n, bool)
    if kind == "ray" and method != "plain" and dim == 3:
        # The engine re-evaluates grazing ray-triangle tests in FP64: there it must meet the FP64
        # bound, and the FP32 reference is compared only where it is itself that accurate.
        b = O.bf_intersect(dim, V, F, x, v)
        keep = parity.ray3d_vs_fp64(got, ref, b, gap)
    sub = lambda r: {k: a[keep] for k, a in r.items() if isinstance(a, np.ndarray) and len(a) == n}
    res = parity.summarize(f"{name}/{kind}/{method}", sub(got), sub(ref), gap[keep])
    assert res["ok"], parity.brief(res)
    if method == "plain" and kind != "sil":
        # same arithmetic as the oracle: bit-identical records
        assert np.array_equal(got["prim_id"], ref["prim_id"])
        assert np.array_equal(got["d"].view(np.int32), ref["d"].view(np.int32))
        assert np.array_equal(got["p"].view(np.int32), ref["p"].view(np.int32))


def _grazing_rays(n, seed=5):
    """A flat 8x8 grid of 128 triangles, rotated, and rays that meet it at grazing angles
    (direction slope 2e-5 to 3e-4 against the plane): Moller-Trumbore determinants of 1e-4 to
    1e-6 of their scale, where float32 t loses 2 to 5 digits."""
    g = np.linspace(-1.0, 1.0, 9)
    X, Y = np.meshgrid(g, g, indexing="ij")
    V = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    F = []
    for i in range(8):
        for j in range(8):
            a, b, c, d = i * 9 + j, (i + 1) * 9 + j, (i + 1) * 9 + j + 1, i * 9 + j + 1
            F += [[a, b, c], [a, c, d]]
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    slope = np.exp(rng.uniform(np.log(2e-5), np.log(3e-4), n))
    L = rng.uniform(0.1, 0.4, n)
    o = np.stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n), slope * L], 1)
    d = np.stack([np.cos(phi), np.sin(phi), -slope], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return ((V @ Q.T).astype(np.float32), np.array(F, np.int32), (o @ Q.T).astype(np.float32),
            (d @ Q.T).astype(np.float32))


@pytest.mark.parametrize("method", METHODS)
def test_grazing_rays_fp64_fallback(gpu, method):
    from gquery_amd.shapes.mesh import Mesh
    V, F, x, v = _grazing_rays(4096)
    shape = Mesh(V, F, device="cuda:0")
    got = _np(shape.intersect_bvh(x, v, method=method))
    b = O.bf_intersect(3, V, F, x, v)
    hit = np.isfinite(b["t"])
    assert hit.all()
    err = parity.rel_err(got["d"], b["t"])
    if method == "plain":
        # the reference's float32 arithmetic: most grazing hits are off by more than 1e-5
        assert (err > parity.TOL_REF).mean() > 0.5
        return
    assert np.array_equal(got["valid"].astype(bool), hit)
    assert err.max() <= parity.TOL_F64, err.max()
    gap = parity.gap_mask(b["t"], b["second_t"])
    assert np.array_equal(got["prim_id"][gap], b["prim_id"][gap])


@pytest.mark.parametrize("method", METHODS)
def test_known_answers_2d(gpu, method):
    import torch
    from gquery_amd.shapes.polyline import Polyline
    pl = Polyline(PENTAGON_V, PENTAGON_E, device="cuda:0")
    r = _np(pl.closest_point_bvh(np.array([[0.8, 0.8]], np.float32), method=method))
    assert r["valid"][0] and r["prim_id"][0] == 2
    np.testing.assert_allclose(r["p"][0], [1.02, 1.24], atol=1e-6)
    np.testing.assert_allclose(r["d"][0], np.sqrt(0.242), rtol=1e-6)
    np.testing.assert_allclose(r["t"][0], 0.48, atol=1e-6)                      # segment t
    np.testing.assert_allclose(r["n"][0], [0.5 / np.sqrt(1.25), 1 / np.sqrt(1.25)], atol=1e-6)
    r = _np(pl.intersect_bvh(np.array([[0.8, 0.8]], np.float32), np.array([[1.0, 1.0]], np.float32), method=method))
    assert r["valid"][0] and r["prim_id"][0] == 2 and r["on_boundary"][0]
    np.testing.assert_allclose(r["d"][0], 0.7 - 1 / 3, rtol=1e-6)
    np.testing.assert_allclose(r["p"][0], [0.8 + 0.7 - 1 / 3] * 2, atol=1e-6)
    assert r["t"][0] == -1.0                                                     # bvh.py:267-274
    np.testing.assert_allclose(r["n"][0], [0.5 / np.sqrt(1.25), 1 / np.sqrt(1.25)], atol=1e-6)
    r = _np(pl.closest_silhouette_snch(torch.tensor([[2.0, 0.5]]), method=method), ("valid", "prim_id", "d", "p"))
    assert r["valid"][0] and r["prim_id"][0] == 2
    np.testing.assert_allclose(r["d"][0], np.sqrt(0.5), rtol=1e-6)


@pytest.mark.parametrize("method", METHODS)
def test_known_answers_3d(gpu, method):
    from gquery_amd.shapes.mesh import Mesh
    m = Mesh(TET_V, TET_F, device="cuda:0")
    r = _np(m.closest_point_bvh(np.array([[0.6, 0.6, 0.6]], np.float32), method=method))
    assert r["valid"][0] and r["prim_id"][0] == 2
    np.testing.assert_allclose(r["d"][0], 0.2 / np.sqrt(21), rtol=2e-6)
    np.testing.assert_allclose(r["p"][0], np.array([59, 61, 62]) / 105, atol=1e-6)
    np.testing.assert_allclose(r["uv"][0], [2 / 7, 62 / 105], atol=1e-6)           # (bary_b, bary_c)
    np.testing.assert_allclose(r["n"][0], np.array([1, 0.5, 0.25]) / np.sqrt(1.3125), atol=1e-6)
    r = _np(m.intersect_bvh(np.array([[0.6, 0.6, 0.6]], np.float32), np.array([[1.0, 1.0, 1.0]], np.float32),
                            method=method))
    assert not r["valid"][0] and r["d"][0] == 0 and r["prim_id"][0] == 0
    assert not r["uv"].any() and not r["n"].any() and not r["on_boundary"][0]
    # a hit with hand-derived barycentrics: uv = (1 - v - w, v) = (0.625, 0.25) (triangle.py:48-73)
    m2 = Mesh(TWO_TRI_V, TWO_TRI_F, device="cuda:0")
    r = _np(m2.intersect_bvh(*TWO_TRI_RAY, method=method))
    assert r["valid"][0] and r["on_boundary"][0] and r["prim_id"][0] == TWO_TRI_HIT["prim_id"]
    np.testing.assert_allclose(r["d"][0], TWO_TRI_HIT["d"], rtol=1e-7)
    for f in ("p", "n", "uv"):
        np.testing.assert_allclose(r[f][0], TWO_TRI_HIT[f], atol=1e-7)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("dim", [2, 3])
def test_edge_cases(gpu, method, dim):
    """Empty batch, single query, queries on vertices / far away, finite and per-query r_max."""
    import torch
    name = "star1024" if dim == 2 else "ico3"
    V, F, shape = scene(name)
    s = shape.snch
    # empty batch
    r = shape.closest_point_bvh(np.zeros((0, dim), np.float32), method=method)
    assert r.d.shape[0] == 0
    # vertices, far points, a single query
    far = np.full((1, dim), 1e4, np.float32)
    x = np.concatenate([V[:50], far, -far, np.zeros((1, dim), np.float32)])
    ref = _oracle("cp", dim, s, x)
    got = _np(shape.closest_point_bvh(x, method=method))
    assert got["valid"].all()
    np.testing.assert_allclose(got["d"], ref["d"], rtol=1e-5, atol=1e-6)
    assert (got["d"][:50] == 0).all()
    # rays with a finite r_max (scalar and per query) and silhouettes with a finite r_max
    x = synth.uniform_points(2000, dim, seed=5)
    v = synth.uniform_directions(2000, dim, seed=6)
    rm = np.random.default_rng(7).uniform(0.05, 1.0, 2000).astype(np.float32)
    for r_max in (0.3, rm):
        ref = _oracle("ray", dim, s, x, v, r_max)
        got = _np(shape.intersect_bvh(x, v, torch.as_tensor(r_max), method=method))
        res = parity.summarize("ray r_max", got, ref, None)
        assert res["valid_mismatch"] == 0 and res["max_err_d"] < 1e-5, res
        ref = _oracle("sil", dim, s, x, None, r_max)
        got = _np(shape.closest_silhouette_snch(x, torch.as_tensor(r_max), method=method), ("valid", "prim_id", "d", "p"))
        res = parity.summarize("sil r_max", got, ref, None)
        assert res["valid_mismatch"] == 0 and res["max_err_d"] < 1e-5, res


@pytest.mark.parametrize("traversal", [1, 2])
def test_sil2d_cone_fast_path_matches_reference(gpu, traversal):
    """2D silhouettes: the engine decides the reference cone test without transcendentals away from
    its thresholds.  On the 65,536-segment star (where that test culls at 92% of nodes and its
    float32 half-angles make it miss some true silhouettes) the engine must find exactly what the
    plain kernel (the reference's float32 test everywhere) finds, except where the engine's FP64
    sign predicate overrules a float32 cross product that rounds to the wrong sign (nearly
    collinear neighbours): there the engine's answer must be the FP64-correct one."""
    import torch
    from gquery_amd.shapes.polyline import Polyline
    V, F = synth.star_polyline(65536)
    shape = Polyline(V, F, device="cuda:0", traversal=traversal)
    xh = synth.uniform_points(200000, 2, seed=77)
    x = torch.from_numpy(xh).cuda()
    got = _np(shape.closest_silhouette_snch(x, method="engine"), ("valid", "prim_id", "d"))
    ref = _np(shape.closest_silhouette_snch(x, method="plain"), ("valid", "prim_id", "d"))
    assert ref["valid"].mean() > 0.5
    diff, ok = parity.sil2d_differences_explained(shape.snch.silhouette_data(), xh, got, ref)
    assert len(diff) <= 20 and ok.all(), (len(diff), diff[~ok][:5])


@pytest.mark.parametrize("wave_front", [0, 40])
@pytest.mark.parametrize("budget", [1, 6])
@pytest.mark.parametrize("name", ["star4096", "ico5"])
@pytest.mark.parametrize("kind", ["cp", "sil"])
@pytest.mark.parametrize("method", ["engine_fp32", "engine"])
def test_wave_per_query_pass(gpu, wave_front, budget, name, kind, method):
    """A tiny packet step budget sends (nearly) every query through the wave-per-query pass, which
    starts from the packet pass's partial answer; a 40-entry first-pass frontier hands most of them
    on to the full-frontier pass, which restarts from the best answer so far.  Results must meet
    the same rule as above."""
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    V, F, base = scene(name)
    dim = V.shape[1]
    shape = (Polyline if dim == 2 else Mesh)(V, F, device="cuda:0", packet_budget=budget, traversal=1,
                                             wave_front=wave_front)
    x = synth.uniform_points(3000, dim, seed=31)
    for r_max in (float("inf"), 0.25):
        if kind == "cp" and r_max != float("inf"):
            continue
        ref = _oracle(kind, dim, shape.snch, x, None, r_max)
        got = _np(_query(shape, kind, method, x, None, r_max))
        gap = _gap_from_bruteforce(kind, dim, shape.snch, V, F, x, None, ref) if r_max == float("inf") else None
        res = parity.summarize(f"{name}/{kind}/{method}/budget{budget}/r{r_max}", got, ref, gap)
        if gap is None:   # ids and points only defined up to ties here: valid flags and distances
            assert res["valid_mismatch"] == 0 and res["max_err_d"] < parity.TOL_REF, res
        else:
            assert res["ok"], parity.brief(res)


@pytest.mark.parametrize("traversal", [1, 2])
@pytest.mark.parametrize("name", ["star4096", "ico5", "ostar1024", "oico4"])
@pytest.mark.parametrize("method", ["engine_fp32", "engine"])
def test_silhouette_traversal_modes(gpu, traversal, name, method):
    """Packet + wave pass (1) and lane per query (2) give the same silhouettes in 2D and 3D."""
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    V, F, base = scene(name)
    dim = V.shape[1]
    shape = (Polyline if dim == 2 else Mesh)(V, F, device="cuda:0", traversal=traversal)
    x = synth.uniform_points(3000, dim, seed=41)
    ref = _oracle("sil", dim, shape.snch, x)
    got = _np(_query(shape, "sil", method, x))
    gap = _gap_from_bruteforce("sil", dim, shape.snch, V, F, x, None, ref)
    res = parity.summarize(f"{name}/sil/{method}/traversal{traversal}", got, ref, gap)
    assert res["ok"], parity.brief(res)


@pytest.mark.parametrize("traversal", [1, 2])
@pytest.mark.parametrize("budget", [0, 1])
@pytest.mark.parametrize("name", ["ico5", "oico4", "star4096", "ostar1024"])
def test_quantized_silhouette_pairs(gpu, traversal, budget, name):
    """Silhouettes on the quantized pair records (scene option sil_pairs=1: 8-bit boxes in 3D,
    16-bit in 2D, int16 cones, reference cones from a table): the same rule against the oracle,
    and the same distances as the float32 pairs query by query (both layouts only ever keep nodes
    the reference keeps)."""
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    V, F, base = scene(name)
    dim = V.shape[1]
    x = synth.uniform_points(3000, dim, seed=43)
    ref = _oracle("sil", dim, base.snch, x)
    gap = _gap_from_bruteforce("sil", dim, base.snch, V, F, x, None, ref)
    out = {}
    for pairs in (0, 1):
        shape = (Mesh if dim == 3 else Polyline)(V, F, device="cuda:0", traversal=traversal, packet_budget=budget,
                                                 sil_pairs=pairs, structure=base.snch)
        out[pairs] = _np(_query(shape, "sil", "engine", x))
        res = parity.summarize(f"{name}/sil/pairs{pairs}/traversal{traversal}/budget{budget}", out[pairs], ref,
                               gap)
        assert res["ok"], parity.brief(res)
    assert np.array_equal(out[0]["valid"], out[1]["valid"])
    assert np.array_equal(out[0]["d"], out[1]["d"])
    assert np.array_equal(out[0]["prim_id"][gap], out[1]["prim_id"][gap])


@pytest.mark.parametrize("name", ["star4096", "ico4"])
@pytest.mark.parametrize("kind", ["cp", "ray", "sil"])
@pytest.mark.parametrize("method", ["engine_fp32", "engine"])
def test_sorted_batch(gpu, name, kind, method):
    """Batches of 8,192 queries and more go through the Morton-cell counting sort first (closest
    point and silhouettes; rays keep their input order)."""
    V, F, shape = scene(name)
    dim = V.shape[1]
    n = 20000
    x = synth.uniform_points(n, dim, seed=51)
    v = synth.uniform_directions(n, dim, seed=52)
    ref = _oracle(kind, dim, shape.snch, x, v)
    got = _np(_query(shape, kind, method, x, v))
    gap = _gap_from_bruteforce(kind, dim, shape.snch, V, F, x, v, ref)
    res = parity.summarize(f"{name}/{kind}/{method}/sorted", got, ref, gap)
    assert res["ok"], parity.brief(res)
