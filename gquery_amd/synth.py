"""Seeded synthetic scenes and query batches for the closest-point / ray / silhouette workloads.

Scenes:
  * star_polyline(N, seed): closed 2D star, vertices at theta_i = 2*pi*i/N with
    r = 1 + sum_{k=1..6} a_k sin(k*theta + phi_k), a_k ~ U(0, 0.3/k), phi_k ~ U(0, 2*pi);
    edges (i, (i+1) mod N).
  * displaced_icosphere(s, seed): icosahedron subdivided s times (midpoint split with shared
    midpoints, re-projected to the unit sphere), every vertex moved radially to
    n * (1 + 0.15 * sum_{k=1..8} c_k sin(f_k (n . u_k) + psi_k)) with f_k ~ U(1, 6),
    u_k uniform on S^2, c_k ~ U(0, 1)/8, psi_k ~ U(0, 2*pi).
Queries:
  * uniform_points(n, dim, seed): uniform in [-1.5, 1.5]^dim.
  * uniform_directions(n, dim, seed): uniform on S^1 (angle) or S^2 (normalised Gaussian).

All randomness comes from numpy Generator(PCG64(seed)); geometry uses seed 0, queries seed 1,
scene k of a multi-scene workload adds 1000*k to both.  Arithmetic is float64 and the result is
rounded once to float32, so the arrays are identical on every machine.
"""
from __future__ import annotations

import numpy as np

GEOMETRY_SEED = 0
QUERY_SEED = 1
DIRECTION_SEED = 2
SCENE_SEED_STRIDE = 1000


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def star_polyline(n: int, seed: int = GEOMETRY_SEED):
    """Closed star-shaped polyline with n segments -> (vertices float32 (n,2), edges int32 (n,2))."""
    if n < 3:
        raise ValueError("star polyline needs at least 3 segments")
    rng = _rng(seed)
    k = np.arange(1, 7, dtype=np.float64)
    a = rng.uniform(0.0, 1.0, size=6) * (0.3 / k)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=6)
    theta = 2.0 * np.pi * np.arange(n, dtype=np.float64) / n
    r = 1.0 + np.sum(a[None, :] * np.sin(k[None, :] * theta[:, None] + phi[None, :]), axis=1)
    v = np.stack([r * np.cos(theta), r * np.sin(theta)], axis=1).astype(np.float32)
    i = np.arange(n, dtype=np.int32)
    e = np.stack([i, (i + 1) % n], axis=1).astype(np.int32)
    return v, e


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
    ], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    return v, f


def _subdivide(v: np.ndarray, f: np.ndarray):
    """One 1->4 midpoint split.  New vertices are numbered in first-use order over the faces
    (edges a-b, b-c, c-a of face 0, then face 1, ...); face i becomes faces 4i..4i+3 =
    (a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)."""
    nv = v.shape[0]
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)  # (3F, 2), grouped by slot
    # order edges as face-major: face i slot 0,1,2
    nf = f.shape[0]
    order = (np.arange(3)[None, :] * nf + np.arange(nf)[:, None]).reshape(-1)
    e = e[order]
    key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # number unique edges by first appearance
    rank = np.empty_like(first)
    rank[np.argsort(first, kind="stable")] = np.arange(first.shape[0])
    mid_id = nv + rank[inv]
    lo = (uniq // nv)[np.argsort(rank)]
    hi = (uniq % nv)[np.argsort(rank)]
    m = 0.5 * (v[lo] + v[hi])
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    v2 = np.concatenate([v, m], axis=0)
    mid = mid_id.reshape(nf, 3)
    ab, bc, ca = mid[:, 0], mid[:, 1], mid[:, 2]
    a, b, c = f[:, 0], f[:, 1], f[:, 2]
    f2 = np.stack([
        np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
        np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], axis=1).reshape(-1, 3)
    return v2, f2


def icosphere(s: int):
    v, f = _icosahedron()
    for _ in range(s):
        v, f = _subdivide(v, f)
    return v, f


def displaced_icosphere(s: int, seed: int = GEOMETRY_SEED):
    """Radially displaced icosphere -> (vertices float32 (V,3), faces int32 (F,3)).  F = 20*4^s."""
    if s < 0 or s > 9:
        raise ValueError("subdivision level must be in [0, 9]")
    rng = _rng(seed)
    freq = rng.uniform(1.0, 6.0, size=8)
    u = rng.standard_normal(size=(8, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    c = rng.uniform(0.0, 1.0, size=8) / 8.0
    psi = rng.uniform(0.0, 2.0 * np.pi, size=8)
    n, f = icosphere(s)
    disp = np.sum(c[None, :] * np.sin(freq[None, :] * (n @ u.T) + psi[None, :]), axis=1)
    v = (n * (1.0 + 0.15 * disp)[:, None]).astype(np.float32)
    return v, f.astype(np.int32)


def icosphere_level_for(n_prims: float) -> int:
    """Subdivision level whose face count 20*4^s is closest to n_prims in log scale."""
    s = np.log(max(n_prims, 20.0) / 20.0) / np.log(4.0)
    return int(np.clip(np.round(s), 0, 9))


def uniform_points(n: int, dim: int, seed: int = QUERY_SEED, lo: float = -1.5, hi: float = 1.5):
    rng = _rng(seed)
    return rng.uniform(lo, hi, size=(n, dim)).astype(np.float32)


def uniform_directions(n: int, dim: int, seed: int = DIRECTION_SEED):
    rng = _rng(seed)
    if dim == 2:
        ang = rng.uniform(0.0, 2.0 * np.pi, size=n)
        d = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    elif dim == 3:
        d = rng.standard_normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    else:
        raise ValueError("dim must be 2 or 3")
    return d.astype(np.float32)


WORKLOAD_SEED = 7


def batched_workload(n_scenes: int = 64, scale: float = 1.0, seed: int = WORKLOAD_SEED):
    """Scene and batch specs of the uneven batched workload: half 2D star polylines, half 3D
    displaced icospheres, primitive counts log-uniform in [1k, 2M] (3D: the subdivision level
    nearest in log), one batch per scene with a size log-uniform in [4k, 4M] and a query kind
    uniform over closest point / ray / silhouette.  `scale` multiplies both ranges (for smaller
    runs).  Scene k uses geometry seed GEOMETRY_SEED + 1000 (k + 1) and query seed
    QUERY_SEED + 1000 (k + 1)."""
    rng = _rng(seed)
    dims = np.array([2] * (n_scenes // 2) + [3] * (n_scenes - n_scenes // 2))
    rng.shuffle(dims)
    specs = []
    for k in range(n_scenes):
        target = float(np.exp(rng.uniform(np.log(1e3), np.log(2e6)))) * scale
        nq = max(int(np.exp(rng.uniform(np.log(4e3), np.log(4e6))) * scale), 1)
        kind = ["cp", "ray", "sil"][int(rng.integers(0, 3))]
        if dims[k] == 3:
            levels = np.arange(1, 9)
            s = int(levels[np.argmin(np.abs(np.log(20.0 * 4.0 ** levels) - np.log(max(target, 20.0))))])
            spec = dict(dim=3, level=s, prims=20 * 4 ** s)
        else:
            spec = dict(dim=2, segments=max(int(round(target)), 3), prims=max(int(round(target)), 3))
        spec.update(geometry_seed=GEOMETRY_SEED + 1000 * (k + 1), query_seed=QUERY_SEED + 1000 * (k + 1),
                    queries=nq, kind=kind)
        specs.append(spec)
    return specs


def scene_geometry(spec):
    """(vertices, indices) of a workload scene spec."""
    if spec["dim"] == 3:
        return displaced_icosphere(spec["level"], seed=spec["geometry_seed"])
    return star_polyline(spec["segments"], seed=spec["geometry_seed"])


def served_workload(n_scenes: int = 16, n_requests: int = 400, scale: float = 1.0, seed: int = WORKLOAD_SEED + 1):
    """Resident scenes and request stream of the serving workload: scenes mixed 2D / 3D with
    primitive counts log-uniform in [10k, 2M]; requests name a scene uniformly, a kind with
    probabilities closest point 0.6 / ray 0.25 / silhouette 0.15, and a batch size log-uniform in
    [256, 262144].  Returns (scene specs, request specs)."""
    rng = _rng(seed)
    scenes = []
    for k in range(n_scenes):
        target = float(np.exp(rng.uniform(np.log(1e4), np.log(2e6)))) * scale
        if k % 2:
            levels = np.arange(1, 9)
            s = int(levels[np.argmin(np.abs(np.log(20.0 * 4.0 ** levels) - np.log(max(target, 20.0))))])
            spec = dict(dim=3, level=s, prims=20 * 4 ** s)
        else:
            spec = dict(dim=2, segments=max(int(round(target)), 3), prims=max(int(round(target)), 3))
        spec.update(geometry_seed=GEOMETRY_SEED + 1000 * (k + 101))
        scenes.append(spec)
    kinds = rng.choice(["cp", "ray", "sil"], size=n_requests, p=[0.6, 0.25, 0.15])
    reqs = []
    for i in range(n_requests):
        n = max(int(np.exp(rng.uniform(np.log(256), np.log(262144))) * scale), 1)
        reqs.append(dict(scene=int(rng.integers(0, n_scenes)), kind=str(kinds[i]), n=n,
                         query_seed=QUERY_SEED + 1000 * (i + 101)))
    return scenes, reqs
