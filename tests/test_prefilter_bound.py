"""CPU: the closest-point pruning bounds of the engine never exceed what the float32 leaf test
computes, at any coordinate scale.

The engine skips a triangle when the squared float32 lower bound from its circumscribing disk
(centre m, unit normal n, radius r: h = (x - m).n, q = (x - m) - h n, h^2 + max(|q| - r, 0)^2)
exceeds (best (1 + 1e-6) + slack)^2; it skips a BVH4 child when its squared box distance or the
squared bound from its bounding cylinder does.  Each is exact only if, for every triangle below,
the widened bound stays at or above the bound computed for it: the float32 leaf distance
(closest_point_triangle, math.py:42-107) rebuilds the nearest point in absolute coordinates, so its
rounding grows with the coordinates, and the slack is kAbsSlack (|x|_1 + S_1) (engine.hip).  These
tests emulate every bound in numpy float32, operation by operation (square roots rounded up one
ulp, the hardware's error), on scenes translated and scaled far from the unit cube, against the
leaf distance computed by the oracle; an absolute 1e-6 margin (what the engine used before) fails
the same check.
"""
import numpy as np
import pytest

from gquery_amd import ext, synth
from oracle import oracle as O

K_ABS_SLACK = np.float32(32.0 * 2.0 ** -24)


def fma32(a, b, c):
    """float32 fma emulated in float64 (the product of two float32 values is exact in float64)."""
    return (np.asarray(a, np.float32).astype(np.float64) * np.asarray(b, np.float32).astype(np.float64)
            + np.asarray(c, np.float32).astype(np.float64)).astype(np.float32)


def sqrt_up(x):
    """float32 square root, one ulp up (v_sqrt_f32 is accurate to 1 ulp)."""
    r = np.sqrt(np.asarray(x, np.float32))
    return np.nextafter(r, np.float32(np.inf)).astype(np.float32)


def disk_lb2(pr, p):
    """The triangle prefilter's squared bound as the kernel evaluates it (tri_may_beat): pr are the
    engine's (n, 4, 4) float32 triangle records, p the (n, 3) float32 queries."""
    m, nrm, rad = pr[:, :3, 3], pr[:, 3, :3], pr[:, 3, 3]
    d = (p - m).astype(np.float32)
    h = (d[:, 0] * nrm[:, 0]).astype(np.float32)
    h = fma32(d[:, 1], nrm[:, 1], h)
    h = fma32(d[:, 2], nrm[:, 2], h)
    q = [fma32(-h, nrm[:, k], d[:, k]) for k in range(3)]
    q2 = (q[0] * q[0]).astype(np.float32)
    q2 = fma32(q[1], q[1], q2)
    q2 = fma32(q[2], q[2], q2)
    mm = np.maximum((sqrt_up(q2) - rad).astype(np.float32), np.float32(0))
    return fma32(h, h, (mm * mm).astype(np.float32))


def box_d2(lo, hi, p):
    """Squared float32 box distance: o = max(max(lo - p, p - hi), 0), d2 = fma chain."""
    o = np.maximum(np.maximum((lo - p).astype(np.float32), (p - hi).astype(np.float32)), np.float32(0))
    d2 = (o[:, 0] * o[:, 0]).astype(np.float32)
    d2 = fma32(o[:, 1], o[:, 1], d2)
    return fma32(o[:, 2], o[:, 2], d2)


def _scene(offset, scale):
    V, F = synth.displaced_icosphere(3, seed=2)
    V = (V.astype(np.float64) * scale + np.asarray(offset, np.float64)).astype(np.float32)
    return V, F


def _pairs(V, F, scale, offset, n=4000, k=12, seed=0):
    """Queries at distances from 1e-5 to 0.5 of the scene size off the surface, each paired with its
    k nearest triangles by centroid (the triangles a traversal actually compares)."""
    rng = np.random.default_rng(seed)
    f = rng.integers(0, F.shape[0], n)
    a, b, c = (V[F[f, j]].astype(np.float64) for j in range(3))
    w = rng.dirichlet([1, 1, 1], n)
    on = w[:, :1] * a + w[:, 1:2] * b + w[:, 2:] * c
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    h = scale * 10.0 ** rng.uniform(-5, np.log10(0.5), n) * rng.choice([-1.0, 1.0], n)
    p = (on + h[:, None] * nrm).astype(np.float32)
    cen = V[F].astype(np.float64).mean(1)
    d = np.linalg.norm(p[:, None, :].astype(np.float64) - cen[None], axis=2)
    near = np.argsort(d, axis=1)[:, :k]
    return np.repeat(p, k, axis=0), near.reshape(-1)


@pytest.mark.parametrize("offset,scale", [((0, 0, 0), 1.0), ((10, -7, 3), 1.0), ((100, 40, -70), 1.0),
                                          ((1000, -700, 300), 1.0), ((0, 0, 0), 1e3), ((0, 0, 0), 1e-3),
                                          ((5e3, 0, 0), 10.0)])
def test_pruning_bounds_hold(offset, scale):
    V, F = _scene(offset, scale)
    s = ext.SNCH3D(V, F)
    n_prims = F.shape[0]
    prims = s.engine_layout(4).view(np.float32).reshape(n_prims, 4, 4)
    ids = s.engine_layout(5).view(np.int32)
    slot_of = np.empty(n_prims, np.int64)
    slot_of[ids] = np.arange(n_prims)
    p, tri_id = _pairs(V, F, scale, offset)
    slot = slot_of[tri_id]
    tri = prims[slot, :3, :3]
    d32 = O.cp_triangle_pairs(p, tri)
    # S_1 from the scene's root box (the engine reads the padded BVH root box, which is larger)
    s1 = np.float32(np.abs(V).max(0).sum())
    slack = (K_ABS_SLACK * (np.abs(p).sum(1, dtype=np.float32) + s1)).astype(np.float32)

    bb = fma32(d32, np.float32(1.000001), slack)
    bound2 = (bb * bb).astype(np.float32)
    lb2 = disk_lb2(prims[slot], p)
    assert (lb2 <= bound2).all(), np.max(lb2 - bound2)

    lo, hi = tri.min(1), tri.max(1)
    d2 = box_d2(lo, hi, p)
    assert (d2 <= bound2).all()


def test_absolute_margin_alone_fails_away_from_the_unit_cube():
    """The check above has teeth: the former margin (1e-6 relative + 1e-6 absolute) drops triangles
    whose float32 leaf distance is below the bound once coordinates reach ~100."""
    V, F = _scene((100, 40, -70), 1.0)
    s = ext.SNCH3D(V, F)
    n_prims = F.shape[0]
    prims = s.engine_layout(4).view(np.float32).reshape(n_prims, 4, 4)
    ids = s.engine_layout(5).view(np.int32)
    slot_of = np.empty(n_prims, np.int64)
    slot_of[ids] = np.arange(n_prims)
    p, tri_id = _pairs(V, F, 1.0, (100, 40, -70))
    slot = slot_of[tri_id]
    d32 = O.cp_triangle_pairs(p, prims[slot, :3, :3])
    old = fma32(d32, np.float32(1.000001), np.float32(1e-6))
    assert (disk_lb2(prims[slot], p) > (old * old).astype(np.float32)).sum() > 0


def _child_paths(s, n_prims, which):
    """For every primitive slot, the (node, child) pairs on its path from the BVH4 root; the child
    box centres as the kernels compute them and the bounding-cylinder records of array `which`
    (14: about the quantized box centres, 15: about the float32 ones)."""
    f = s.engine_layout(1).view(np.uint8).reshape(-1, 128)
    child = f[:, 96:112].copy().view(np.uint32)
    nch = f[:, 112:116].copy().view(np.uint32).reshape(-1)
    if which == 15:
        blo = f[:, 0:48].copy().view(np.float32).reshape(-1, 3, 4)
        bhi = f[:, 48:96].copy().view(np.float32).reshape(-1, 3, 4)
        cen = (np.float32(0.5) * (blo + bhi).astype(np.float32)).astype(np.float32)
    else:
        # centre = fma(qlo + qhi, 2^(e - 1), origin), rounded once
        q = s.engine_layout(0).view(np.uint8).reshape(-1, 64)
        org = q[:, 0:12].copy().view(np.float32)
        ex = q[:, 12:15].copy().view(np.int8).astype(np.int64)
        hs = np.ldexp(1.0, ex - 1)[:, :, None]
        qs = q[:, 16:28].reshape(-1, 3, 4).astype(np.float64) + q[:, 28:40].reshape(-1, 3, 4).astype(np.float64)
        cen = (qs * hs + org[:, :, None]).astype(np.float32)
    big = s.engine_layout(9).view(np.uint32)
    raw = s.engine_layout(which).view(np.uint8).reshape(-1, 64)
    h16 = raw[:, :48].copy().view(np.uint16).astype(np.uint32) << 16
    ax = h16[:, :12].view(np.float32).reshape(-1, 3, 4)
    lo, hi, rad = (h16[:, 12 + 4 * k:16 + 4 * k].view(np.float32) for k in range(3))
    paths = [[] for _ in range(n_prims)]

    def walk(node, path):
        for j in range(int(nch[node])):
            ref = int(child[node, j])
            here = path + [(node, j)]
            if ref & 0x80000000:
                first, cnt = ref & 0xFFFFFF, (ref >> 24) & 0x7F
                if cnt == 127:
                    first, cnt = int(big[2 * first]), int(big[2 * first + 1])
                for t in range(first, first + cnt):
                    paths[t] = here
            else:
                walk(ref, here)

    walk(0, [])
    return paths, cen, ax, lo, hi, rad


def cyl_lb2(c, ax, lo, hi, rad, x):
    """cyl_lb2 (engine.hip) in float32, one row per (query, child); c the child box centre."""
    dv = (x - c).astype(np.float32)
    sx = (dv[:, 0] * ax[:, 0]).astype(np.float32)
    sx = fma32(dv[:, 1], ax[:, 1], sx)
    sx = fma32(dv[:, 2], ax[:, 2], sx)
    q2 = np.zeros(len(x), np.float32)
    for k in range(3):
        qk = fma32(-sx, ax[:, k], dv[:, k])
        q2 = fma32(qk, qk, q2)
    sd = np.maximum(np.maximum((lo - sx).astype(np.float32), (sx - hi).astype(np.float32)), np.float32(0))
    m = np.maximum((sqrt_up(q2) - rad).astype(np.float32), np.float32(0))
    return fma32(sd, sd, (m * m).astype(np.float32))


@pytest.mark.parametrize("which", [14, 15])
@pytest.mark.parametrize("offset,scale", [((0, 0, 0), 1.0), ((1000, -700, 300), 1.0), ((0, 0, 0), 1e3),
                                          ((0, 0, 0), 1e-3), ((5e3, 0, 0), 10.0)])
def test_cylinder_bounds_hold(offset, scale, which):
    """The child bounding-cylinder bound, evaluated in float32 as the kernels do (for both node
    formats' centres), never exceeds the widened bound of a triangle below the child: here each
    pair's triangle against every child on its path from the root."""
    V, F = _scene(offset, scale)
    s = ext.SNCH3D(V, F)
    n_prims = F.shape[0]
    prims = s.engine_layout(4).view(np.float32).reshape(n_prims, 4, 4)
    ids = s.engine_layout(5).view(np.int32)
    slot_of = np.empty(n_prims, np.int64)
    slot_of[ids] = np.arange(n_prims)
    p, tri_id = _pairs(V, F, scale, offset, n=1500, k=8)
    slot = slot_of[tri_id]
    d32 = O.cp_triangle_pairs(p, prims[slot, :3, :3])
    s1 = np.float32(np.abs(V).max(0).sum())
    slack = (K_ABS_SLACK * (np.abs(p).sum(1, dtype=np.float32) + s1)).astype(np.float32)
    bb = fma32(d32, np.float32(1.000001), slack)
    bound2 = (bb * bb).astype(np.float32)
    paths, cen, ax, lo, hi, rad = _child_paths(s, n_prims, which)
    depth = max(len(q) for q in paths)
    for level in range(depth):
        rows = np.array([i for i in range(len(p)) if len(paths[slot[i]]) > level])
        nj = np.array([paths[slot[i]][level] for i in rows])
        n_, j_ = nj[:, 0], nj[:, 1]
        lb2 = cyl_lb2(cen[n_, :, j_], ax[n_, :, j_], lo[n_, j_], hi[n_, j_], rad[n_, j_], p[rows])
        assert (lb2 <= bound2[rows]).all(), (level, np.max(lb2 - bound2[rows]))
