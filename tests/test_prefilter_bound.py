"""CPU: the closest-point pruning bounds of the engine never exceed what the float32 leaf test
computes, at any coordinate scale.

The engine skips a triangle when its plane-distance lower bound |n.x - w|, computed in float32 with
an fma chain from the stored plane, exceeds the current best distance by more than its rounding
slack; it skips a box when the squared float32 box distance exceeds (best (1 + 1e-6) + slack)^2.
Both are exact only if, for every triangle, the widened bound stays below the triangle's own float32
leaf distance (closest_point_triangle, math.py:42-107).  The float32 leaf distance rebuilds the
nearest point in absolute coordinates, so its rounding grows with the coordinates: the slack is
kAbsSlack (|x|_1 + S_1) (engine.hip).  This test emulates both bounds in numpy float32 on scenes
translated and scaled far from the unit cube and checks them against the leaf distance computed by
the oracle; an absolute 1e-6 margin (what the engine used before) fails the same check.
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


def plane_bound(pl, p):
    """fabsf(fmaf(pl.x, p.x, fmaf(pl.y, p.y, fmaf(pl.z, p.z, -pl.w)))) as the kernel evaluates it."""
    t = fma32(pl[:, 2], p[:, 2], -pl[:, 3])
    t = fma32(pl[:, 1], p[:, 1], t)
    t = fma32(pl[:, 0], p[:, 0], t)
    return np.abs(t)


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
    pl = prims[slot, 3, :]
    d32 = O.cp_triangle_pairs(p, tri)
    # S_1 from the scene's root box (the engine reads the padded BVH root box, which is larger)
    s1 = np.float32(np.abs(V).max(0).sum())
    slack = (K_ABS_SLACK * (np.abs(p).sum(1, dtype=np.float32) + s1)).astype(np.float32)

    lb = plane_bound(pl, p)
    keep = fma32(lb, np.float32(0.99999), -slack)
    assert (keep <= d32).all(), np.max(keep - d32)

    lo, hi = tri.min(1), tri.max(1)
    d2 = box_d2(lo, hi, p)
    bb = fma32(d32, np.float32(1.000001), slack)
    assert (d2 <= (bb * bb).astype(np.float32)).all()


def test_absolute_margin_alone_fails_away_from_the_unit_cube():
    """The check above has teeth: the former margin (1e-5 relative + 1e-6 absolute) drops triangles
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
    lb = plane_bound(prims[slot, 3, :], p)
    old_keep = (lb * np.float32(0.99999) - np.float32(1e-6)).astype(np.float32)
    assert (old_keep > d32).sum() > 0
