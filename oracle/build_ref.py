"""TEST INFRASTRUCTURE ONLY -- pure-Python restatement of gQuery's BVH / SNCH builder for small
inputs, used to pin the product's C++ builder bit-for-bit.

Follows src/shapes/bvh.cpp:29-353 and src/shapes/snch.cpp:11-285 with numpy float32 scalar
arithmetic (every intermediate rounded to float32, no fused operations):
  * BoundingBox (include/shapes/bounding_box.h:9-72): expand(point) pads by +-FLT_EPSILON,
    min/max are std::min/std::max, surface_area is 2(e0+e1) / 2(e0e1+e0e2+e1e2) left to right,
    max_dimension is the first largest extent, offset divides by max(extent, FLT_EPSILON);
  * binned SAH with 12 buckets (bvh.cpp:29-91), split test 0.5 + cost/SA < n or n > 10
    (bvh.cpp:249-254), libstdc++ std::partition (two-ended swap) and std::nth_element for two
    elements (a swap when the second centroid is smaller);
  * Eigen fixed-size reductions: 3-term sums associate as c0 + (c1 + c2);
  * SNCH cones (snch.cpp:102-154) with std::acos of float.
Independent of the product code: it shares no source with gquery_amd/csrc/builder.cpp.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32
EPS = f32(np.finfo(np.float32).eps)
FMAX = f32(np.finfo(np.float32).max)
PI = f32(math.pi)


def _min(x, y):  # std::min(x, y)
    return y if y < x else x


def _max(x, y):  # std::max(x, y)
    return y if x < y else x


def _sum(c):
    if len(c) == 2:
        return f32(c[0] + c[1])
    return f32(c[0] + f32(c[1] + c[2]))


def _dot(a, b):
    return _sum([f32(a[k] * b[k]) for k in range(len(a))])


def _norm(a):
    return f32(np.sqrt(_dot(a, a)))


def _normalized(a):
    z = _dot(a, a)
    if z > 0:
        s = f32(np.sqrt(z))
        return [f32(v / s) for v in a]
    return list(a)


class Box:
    def __init__(self, D):
        self.lo = [FMAX] * D
        self.hi = [f32(-FMAX)] * D

    def copy(self):
        b = Box(len(self.lo))
        b.lo, b.hi = list(self.lo), list(self.hi)
        return b

    def expand_point(self, p):
        for k in range(len(self.lo)):
            self.lo[k] = _min(self.lo[k], f32(p[k] - EPS))
            self.hi[k] = _max(self.hi[k], f32(p[k] + EPS))

    def expand_box(self, o):
        for k in range(len(self.lo)):
            self.lo[k] = _min(self.lo[k], o.lo[k])
            self.hi[k] = _max(self.hi[k], o.hi[k])

    def centroid(self, k):
        return f32(f32(self.lo[k] + self.hi[k]) * f32(0.5))

    def surface_area(self):
        e = [f32(self.hi[k] - self.lo[k]) for k in range(len(self.lo))]
        with np.errstate(all="ignore"):
            if len(e) == 2:
                return f32(f32(2) * f32(e[0] + e[1]))
            return f32(f32(2) * f32(f32(f32(e[0] * e[1]) + f32(e[0] * e[2])) + f32(e[1] * e[2])))

    def max_dimension(self):
        e = [f32(self.hi[k] - self.lo[k]) for k in range(len(self.lo))]
        idx, m = 0, e[0]
        for k in range(1, len(e)):
            if e[k] > m:
                idx, m = k, e[k]
        return idx

    def offset(self, p, k):
        ext = _max(f32(self.hi[k] - self.lo[k]), EPS)
        return f32(f32(p - self.lo[k]) / ext)


def _bucket(cbox, prim, dim):
    b = int(f32(f32(12) * cbox.offset(prim[1].centroid(dim), dim)))
    return 11 if b == 12 else b


def _partition(a, pred):
    """libstdc++ std::partition for bidirectional iterators (two-ended swap); returns mid."""
    first, last = 0, len(a)
    while True:
        while True:
            if first == last:
                return first
            if pred(a[first]):
                first += 1
            else:
                break
        last -= 1
        while True:
            if first == last:
                return first
            if not pred(a[last]):
                last -= 1
            else:
                break
        a[first], a[last] = a[last], a[first]
        first += 1


def build_bvh(V, F, max_prims_in_node=10):
    """Returns (nodes, ordered) with nodes = list of dict(lo, hi, off, second, n) in DFS order."""
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.int64)
    D = V.shape[1]
    prims = []
    for i in range(F.shape[0]):
        b = Box(D)
        for k in range(D):
            b.expand_point([f32(c) for c in V[F[i, k]]])
        prims.append((i, b))
    nodes, ordered = [], []

    def leaf(node, items, box):
        nodes[node].update(lo=box.lo, hi=box.hi, off=len(ordered), n=len(items), second=0)
        ordered.extend(p[0] for p in items)

    def rec(items):
        node = len(nodes)
        nodes.append({})
        n = len(items)
        box = Box(D)
        for p in items:
            box.expand_box(p[1])
        if box.surface_area() == 0 or n == 1:
            leaf(node, items, box)
            return node
        cbox = Box(D)
        for p in items:
            cbox.expand_point([p[1].centroid(k) for k in range(D)])
        dim = cbox.max_dimension()
        if cbox.hi[dim] == cbox.lo[dim]:
            leaf(node, items, box)
            return node
        mid = n // 2
        if n <= 2:
            if items[1][1].centroid(dim) < items[0][1].centroid(dim):
                items[0], items[1] = items[1], items[0]
        else:
            cnt = [0] * 12
            bb = [Box(D) for _ in range(12)]
            for p in items:
                b = _bucket(cbox, p, dim)
                cnt[b] += 1
                bb[b].expand_box(p[1])
            costs = [f32(0)] * 11
            below, boxb = 0, Box(D)
            for i in range(11):
                boxb.expand_box(bb[i])
                below += cnt[i]
                with np.errstate(all="ignore"):
                    costs[i] = f32(f32(below) * boxb.surface_area())
            above, boxa = 0, Box(D)
            for i in range(10, -1, -1):
                boxa.expand_box(bb[i + 1])
                above += cnt[i + 1]
                with np.errstate(all="ignore"):
                    costs[i] = f32(costs[i] + f32(f32(above) * boxa.surface_area()))
            split, cost = 0, costs[0]
            for i in range(1, 11):
                if costs[i] < cost:
                    split, cost = i, costs[i]
            cost = f32(f32(0.5) + f32(cost / box.surface_area()))
            if n > max_prims_in_node or cost < f32(n):
                mid = _partition(items, lambda p: _bucket(cbox, p, dim) <= split)
            else:
                leaf(node, items, box)
                return node
        left = rec(items[:mid])
        right = rec(items[mid:])
        ib = Box(D)
        ib.expand_box(Box._from(nodes[left]))
        ib.expand_box(Box._from(nodes[right]))
        nodes[node].update(lo=ib.lo, hi=ib.hi, off=0, n=0, second=right - node)
        return node

    rec(prims)
    return nodes, ordered


def _box_from(d):
    b = Box(len(d["lo"]))
    b.lo, b.hi = list(d["lo"]), list(d["hi"])
    return b


Box._from = staticmethod(_box_from)


def bvh_node_data(nodes):
    return np.array([list(n["lo"]) + list(n["hi"]) + [n["off"], n["second"], n["n"]] for n in nodes], np.float32)


def primitive_data(V, F, ordered):
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.int64)
    rows = []
    for i in ordered:
        rows.append(list(V[F[i]].reshape(-1)) + [i])
    return np.array(rows, np.float32)


def silhouettes(V, F):
    """Raw silhouette records (snch.cpp:156-241) as lists of floats (10 / 19 per record)."""
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.int64)
    D = V.shape[1]
    if D == 2:
        recs = [[f32(0)] * 10 for _ in range(V.shape[0])]
        for i0, i1 in F:
            r0 = recs[i0]
            r0[2:4], r0[4:6] = list(V[i0]), list(V[i1])
            r0[7], r0[8], r0[9] = i0, i1, i0
            r1 = recs[i1]
            r1[0:2], r1[2:4] = list(V[i0]), list(V[i1])
            r1[6], r1[7], r1[9] = i0, i1, i1
        return recs, None
    emap = {}
    for f in F:
        for j in range(3):
            e = tuple(sorted((int(f[j]), int(f[(j + 1) % 3]))))
            if e not in emap:
                emap[e] = len(emap)
    recs = [[f32(0)] * 19 for _ in range(len(emap))]
    for i, f in enumerate(F):
        for j in range(3):
            i0, i1, i2 = int(f[j]), int(f[(j + 1) % 3]), int(f[(j + 2) % 3])
            v0, v1, v2 = list(V[i0]), list(V[i1]), list(V[i2])
            sw = i0 > i1
            if sw:
                i0, i1 = i1, i0
            e = emap[(i0, i1)]
            r = recs[e]
            r[3:6], r[6:9] = v0, v1
            s = 3 if sw else 0
            r[3 * s:3 * s + 3] = v2
            r[13], r[14] = i0, i1
            r[15 if sw else 12] = i2
            r[17 if sw else 16] = i
            r[18] = e
    return recs, emap


def _sil_geom(D, r):
    if D == 2:
        m = [r[0:2], r[2:4], r[4:6]]
        v0 = [f32(m[1][k] - m[0][k]) for k in range(2)]
        v1 = [f32(m[2][k] - m[1][k]) for k in range(2)]
        a, b = [v0[1], f32(-v0[0])], [v1[1], f32(-v1[0])]
        n = [f32(f32(f32(0) + a[k]) + b[k]) for k in range(2)]
        return _normalized(n), _normalized(a), _normalized(b), list(m[1])
    m = [r[0:3], r[3:6], r[6:9], r[9:12]]

    def cross(a, b):
        return [f32(f32(a[1] * b[2]) - f32(a[2] * b[1])), f32(f32(a[2] * b[0]) - f32(a[0] * b[2])),
                f32(f32(a[0] * b[1]) - f32(a[1] * b[0]))]
    n0 = cross([f32(m[2][k] - m[1][k]) for k in range(3)], [f32(m[0][k] - m[1][k]) for k in range(3)])
    n1 = cross([f32(m[0][k] - m[2][k]) for k in range(3)], [f32(m[1][k] - m[2][k]) for k in range(3)])
    n = [f32(f32(f32(0) + n0[k]) + n1[k]) for k in range(3)]
    cen = [f32(f32(0.5) * f32(m[1][k] + m[2][k])) for k in range(3)]
    return _normalized(n), _normalized(n0), _normalized(n1), cen


def build_snch(V, F, max_prims_in_node=10):
    """Returns (snch_node_data (N x 13/16), silhouette_data (M x 10/19), raw (K x 10/19))."""
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.int64)
    D = V.shape[1]
    nodes, ordered = build_bvh(V, F, max_prims_in_node)
    raw, emap = silhouettes(V, F)
    sil_off, sil_cnt, osil = [0] * len(nodes), [0] * len(nodes), []
    for i, nd in enumerate(nodes):
        if nd["n"] == 0:
            continue
        start, seen = len(osil), set()
        for j in range(nd["n"]):
            f = F[ordered[nd["off"] + j]]
            if D == 2:
                keys = [int(f[0]), int(f[1])]
            else:
                keys = [emap[tuple(sorted((int(f[k]), int(f[(k + 1) % 3]))))] for k in range(3)]
            for key in keys:
                if key not in seen:
                    seen.add(key)
                    osil.append(key)
        sil_off[i], sil_cnt[i] = start, len(osil) - start
    geom = [_sil_geom(D, r) for r in raw]
    cones = [None] * len(nodes)

    def rec(start, end):
        nd = nodes[start]
        cen = [f32(f32(nd["lo"][k] + nd["hi"][k]) * f32(0.5)) for k in range(D)]
        axis, half, radius = [f32(0)] * D, PI, f32(0)
        anyv = False
        for i in range(start, end):
            for j in range(sil_cnt[i]):
                g = geom[osil[sil_off[i] + j]]
                axis = [f32(axis[k] + g[0][k]) for k in range(D)]
                radius = _max(radius, _norm([f32(g[3][k] - cen[k]) for k in range(D)]))
                anyv = True
        if not anyv:
            half = f32(-PI)
        else:
            an = _norm(axis)
            if an > EPS:
                axis = [f32(a / an) for a in axis]
                half = f32(0)
                for i in range(start, end):
                    for j in range(sil_cnt[i]):
                        g = geom[osil[sil_off[i] + j]]
                        x0 = min(max(_dot(axis, g[1]), f32(-1)), f32(1))
                        x1 = min(max(_dot(axis, g[2]), f32(-1)), f32(1))
                        a0 = f32(math.acos(float(x0)))
                        a1 = f32(math.acos(float(x1)))
                        half = _max(half, _max(a0, a1))
        cones[start] = (axis, half, radius)
        if nd["n"] == 0:
            rec(start + 1, start + nd["second"])
            rec(start + nd["second"], end)

    rec(0, len(nodes))
    rows = []
    for i, nd in enumerate(nodes):
        axis, half, radius = cones[i]
        rows.append(list(nd["lo"]) + list(nd["hi"]) + list(axis) + [half, radius, nd["off"] if nd["n"] else 0,
                                                                     nd["second"], nd["n"], sil_off[i], sil_cnt[i]])
    sil = np.array([raw[k] for k in osil], np.float32)
    return np.array(rows, np.float32), sil, np.array(raw, np.float32), bvh_node_data(nodes), primitive_data(V, F, ordered)
