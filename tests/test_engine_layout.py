"""Engine layout built on the host (CPU-only checks): every reference in range, every primitive and
silhouette reachable exactly once, decoded quantized boxes containing the true geometry, and the
exact-cull cones containing every predicate normal below them."""
import numpy as np
import pytest

from gquery_amd import ext, synth

LEAF = 0x80000000


def qnode_dtype(D):
    if D == 3:
        return np.dtype([("origin", "<f4", 3), ("exp", "i1", 3), ("nchild", "u1"), ("qlo", "u1", (3, 4)),
                         ("qhi", "u1", (3, 4)), ("child", "<u4", 4), ("pad", "u1", 8)])
    return np.dtype([("origin", "<f4", 2), ("exp", "i1", 2), ("nchild", "u1"), ("pad0", "u1"),
                     ("qlo", "<u2", (2, 4)), ("qhi", "<u2", (2, 4)), ("child", "<u4", 4), ("pad", "u1", 4)])


def fnode_dtype(D):
    return np.dtype([("lo", "<f4", (D, 4)), ("hi", "<f4", (D, 4)), ("child", "<u4", 4), ("nchild", "<u4"),
                     ("pad", "u1", 128 - (2 * D * 16 + 20))])


def spq_dtype(D):
    if D == 3:
        return np.dtype([("origin", "<f4", 3), ("exp", "i1", 3), ("flags", "u1"), ("qlo", "u1", (2, 3)),
                         ("qhi", "u1", (2, 3)), ("axis", "<i2", (2, 3)), ("sin", "<i2", 2),
                         ("child", "<u4", 2), ("ref_base", "<i4"), ("ext_mid", "<i2", 2), ("ext_half", "u1", 2),
                         ("ext_perp", "u1", 2)])
    return np.dtype([("origin", "<f4", 2), ("exp", "i1", 2), ("flags", "u1"), ("pad0", "u1"),
                     ("qlo", "<u2", (2, 2)), ("qhi", "<u2", (2, 2)), ("axis", "<i2", (2, 2)), ("sin", "<i2", 2),
                     ("child", "<u4", 2), ("ref_base", "<i4"), ("ext_mid", "<i2", 2), ("ext_half", "u1", 2),
                     ("ext_perp", "u1", 2), ("pad", "u1", 4)])


REFCONE = np.dtype([("axis", "<f4", 3), ("half", "<f4"), ("radius", "<f4"), ("lo", "<f4", 3), ("hi", "<f4", 3),
                    ("sin", "<f4"), ("cos", "<f4"), ("pad", "<f4", 3)])
REFCONE2 = np.dtype([("lo", "<f4", 2), ("hi", "<f4", 2), ("axis", "<f4", 2), ("half", "<f4"), ("radius", "<f4"),
                     ("sin", "<f4"), ("cos", "<f4"), ("pad", "<f4", 2)])
SEED = np.dtype([("lo", "<f4", 3), ("ref", "<u4"), ("hi", "<f4", 3), ("pad", "<f4")])


def spf_dtype(D):
    if D == 2:
        return np.dtype([("lo", "<f4", (2, 2)), ("hi", "<f4", (2, 2)), ("cone_axis", "<f4", (2, 2)),
                         ("half", "<f4", 2), ("radius", "<f4", 2), ("child", "<u4", 2), ("ref", "<i4", 2),
                         ("cull", "<i4", 2), ("sin", "<f4", 2), ("cos", "<f4", 2), ("ecos", "<f4", 2),
                         ("esin", "<f4", 2), ("eaxis", "<f4", (2, 2)), ("elo", "<f4", 2), ("ehi", "<f4", 2),
                         ("eperp", "<f4", 2), ("pad", "u1", 32)])
    return np.dtype([("lo", "<f4", (2, D)), ("hi", "<f4", (2, D)), ("axis0", "<f4", 3), ("cos0", "<f4"),
                     ("sin0", "<f4"), ("axis1", "<f4", 3), ("cos1", "<f4"), ("sin1", "<f4"),
                     ("child", "<u4", 2), ("ref", "<i4", 2), ("slo", "<f4", 2), ("shi", "<f4", 2),
                     ("sperp", "<f4", 2)])


def decode(q, e, origin):
    """fma(q, 2^e, origin) rounded once to float32 (exact in float64 for these magnitudes)."""
    return (q.astype(np.float64) * np.ldexp(1.0, e.astype(np.int64)) + origin.astype(np.float64)).astype(np.float32)


def leaf_range(ref, big):
    first, count = ref & 0xFFFFFF, (ref >> 24) & 0x7F
    if count == 127:
        return int(big[2 * first]), int(big[2 * first + 1])
    return int(first), int(count)


SCENES = {
    "star300": lambda: synth.star_polyline(300, seed=3),
    "star8192": lambda: synth.star_polyline(8192, seed=4),
    "ico1": lambda: synth.displaced_icosphere(1),
    "ico4": lambda: synth.displaced_icosphere(4, seed=6),
    "dup3d": lambda: (np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (300, 1)),
                      np.arange(900, dtype=np.int32).reshape(300, 3)),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bvh4_layout(name):
    V, F = SCENES[name]()
    D = V.shape[1]
    s = (ext.SNCH if D == 2 else ext.SNCH3D)(V, F)
    scal = s.engine_layout(10).view(np.uint32)
    root, n_nodes = scal[0], scal[2]
    q = s.engine_layout(0).view(qnode_dtype(D))
    f = s.engine_layout(1).view(fnode_dtype(D))
    prims = s.engine_layout(4).view(np.float32).reshape(-1, 4)
    ids = s.engine_layout(5).view(np.int32)
    big = s.engine_layout(9).view(np.uint32)
    assert len(q) == len(f) == n_nodes and root == 0
    n_prims = F.shape[0]
    assert np.array_equal(np.sort(ids), np.arange(n_prims))
    pts = prims.reshape(n_prims, 4, 4)[:, :3, :3] if D == 3 else prims.reshape(n_prims, 2, 2)
    np.testing.assert_array_equal(pts, V[F[ids]])
    if D == 3:
        # circumscribing disk: centre m in the free lanes, |n| <= 1, every vertex within rad of the
        # line m + t n and within float32 rounding of the plane through m
        pr = prims.reshape(n_prims, 4, 4).astype(np.float64)
        m, nrm, rad = pr[:, :3, 3], pr[:, 3, :3], pr[:, 3, 3]
        assert (np.linalg.norm(nrm, axis=1) <= 1.0).all()
        np.testing.assert_allclose(m, pts.astype(np.float64).mean(1), rtol=1e-6, atol=1e-7)
        for k in range(3):
            dv = pts[:, k].astype(np.float64) - m
            h = (dv * nrm).sum(1)
            lat = np.linalg.norm(dv - h[:, None] * nrm, axis=1)
            assert (lat <= rad).all()
            assert (np.abs(h) <= 4e-7 * (1.0 + np.abs(m).sum(1))).all()
    covered = np.zeros(n_prims, np.int64)
    seen = np.zeros(n_nodes, np.int64)

    def subtree_slots(ref):
        if ref & LEAF:
            a, c = leaf_range(ref, big)
            return list(range(a, a + c))
        out = []
        nd = f[ref]
        for j in range(int(nd["nchild"])):
            out += subtree_slots(int(nd["child"][j]))
        return out

    stack = [0]
    while stack:
        i = stack.pop()
        seen[i] += 1
        qn, fn = q[i], f[i]
        assert qn["nchild"] == fn["nchild"] and 1 <= qn["nchild"] <= 4
        for j in range(int(fn["nchild"])):
            ref = int(fn["child"][j])
            assert ref == int(qn["child"][j])
            if ref & LEAF:
                a, c = leaf_range(ref, big)
                assert c >= 1 and a + c <= n_prims
                covered[a:a + c] += 1
            else:
                assert ref < n_nodes
                stack.append(ref)
            slots = subtree_slots(ref)
            geo = pts[slots].reshape(-1, D)
            lo = np.array([fn["lo"][k][j] for k in range(D)], np.float32)
            hi = np.array([fn["hi"][k][j] for k in range(D)], np.float32)
            assert (geo >= lo).all() and (geo <= hi).all()
            qlo = decode(qn["qlo"][:, j], qn["exp"], qn["origin"])
            qhi = decode(qn["qhi"][:, j], qn["exp"], qn["origin"])
            assert (qlo <= lo).all() and (qhi >= hi).all(), (i, j, qlo, lo, qhi, hi)
            # the quantized box is tight to one quantization step
            step = np.ldexp(1.0, qn["exp"].astype(np.int64))
            assert ((lo - qlo) <= step * 1.0001 + 1e-30).all() and ((qhi - hi) <= step * 1.0001 + 1e-30).all()
    assert (covered == 1).all() and (seen == 1).all()


@pytest.mark.parametrize("name", ["star300", "star8192", "ico1", "ico4"])
def test_snch_pair_layout(name):
    V, F = SCENES[name]()
    D = V.shape[1]
    s = (ext.SNCH if D == 2 else ext.SNCH3D)(V, F)
    scal = s.engine_layout(10).view(np.uint32)
    sroot, n_pairs = scal[1], scal[3]
    pf = s.engine_layout(3).view(spf_dtype(D))
    pq = s.engine_layout(2).view(spq_dtype(D))
    refs = s.engine_layout(8 if D == 3 else 11).view(REFCONE if D == 3 else REFCONE2)
    if D == 2:
        assert s.engine_layout(8).size == 0
        # SNCH node records of the reference (box, cone axis, half-angle, radius) in node order;
        # children of pair records are matched to them through the subtree primitive ranges below
        snd = s.node_data().reshape(-1, 13)
        node_keys = {r.tobytes() for r in np.ascontiguousarray(snd[:, :8], np.float32)}
    sils = s.engine_layout(6).view(np.float32).reshape(-1, 12 if D == 3 else 8)
    sil_ids = s.engine_layout(7).view(np.int32)
    big = s.engine_layout(9).view(np.uint32)
    ref_sil = s.silhouette_data().reshape(-1, 19 if D == 3 else 10)
    assert len(pf) == n_pairs and sroot == 0 and len(pq) == n_pairs
    np.testing.assert_array_equal(sil_ids, ref_sil[:, -1].astype(np.int32))
    if D == 3:
        np.testing.assert_array_equal(sils[:, 0:3], ref_sil[:, 3:6])   # a = m1
        np.testing.assert_array_equal(sils[:, 3:6], ref_sil[:, 6:9])   # b = m2
        seg = np.stack([sils[:, 0:3], sils[:, 3:6]], 1)
        n0, n1 = sils[:, 6:9], sils[:, 9:12]
    else:
        seg = np.stack([sils[:, 0:2], sils[:, 2:4]], 1)
    covered = np.zeros(len(sils), np.int64)

    def slots(ref):
        if ref & LEAF:
            a, c = leaf_range(ref, big)
            return list(range(a, a + c))
        return slots(int(pf[ref]["child"][0])) + slots(int(pf[ref]["child"][1]))

    stack = [0]
    while stack:
        i = stack.pop()
        for j in range(2):
            ref = int(pf[i]["child"][j])
            assert ref == int(pq[i]["child"][j])
            if ref & LEAF:
                a, c = leaf_range(ref, big)
                assert a + c <= len(sils)
                covered[a:a + c] += 1
            else:
                assert ref < n_pairs
                stack.append(ref)
            sl = slots(ref)
            lo, hi = pf[i]["lo"][j], pf[i]["hi"][j]
            if D == 2:
                # the record carries the reference node's own float32 box and cone
                geo = sils[sl, 2:4]   # the silhouette vertex b of each (a, b, c) record
                assert (geo >= lo).all() and (geo <= hi).all()
                rec = np.concatenate([lo, hi, pf[i]["cone_axis"][j], [pf[i]["half"][j], pf[i]["radius"][j]]])
                assert rec.astype(np.float32).tobytes() in node_keys, (i, j)
                half = pf[i]["half"][j]
                assert pf[i]["ref"][j] == (0 if 0.0 <= half <= np.float32(np.pi / 2) else -1)
                assert abs(pf[i]["sin"][j] - np.sin(np.float64(half))) <= 1e-7
                assert abs(pf[i]["cos"][j] - np.cos(np.float64(half))) <= 1e-7
                # exact-cull cone over the vertex predicate's segment normals, and the vertices'
                # extent along / across its axis about the float32 box centre
                ecos = pf[i]["ecos"][j]
                assert (ecos > 1.5) == bool(pf[i]["cull"][j])
                if 0 < ecos < 1.5 and sl:
                    ax = pf[i]["eaxis"][j].astype(np.float64)
                    A, B, Cc = (sils[sl, k:k + 2].astype(np.float64) for k in (0, 2, 4))
                    n1 = np.stack([-(B - A)[:, 1], (B - A)[:, 0]], 1)
                    n2 = np.stack([-(Cc - B)[:, 1], (Cc - B)[:, 0]], 1)
                    ok = np.isfinite(n1).all(1) & np.isfinite(n2).all(1)
                    c1 = (n1[ok] @ ax) / np.linalg.norm(n1[ok], axis=1)
                    c2 = (n2[ok] @ ax) / np.linalg.norm(n2[ok], axis=1)
                    sg = np.sign(c1 + c2)
                    assert (sg * c1 >= ecos - 1e-6).all() and (sg * c2 >= ecos - 1e-6).all()
                    cen = (np.float32(0.5) * (lo + hi)).astype(np.float64)
                    d = B - cen
                    al = d @ ax
                    pe = np.abs(d[:, 0] * ax[1] - d[:, 1] * ax[0])
                    assert pf[i]["elo"][j] <= al.min() and al.max() <= pf[i]["ehi"][j] and pe.max() <= pf[i]["eperp"][j]
                # quantized twin: the reference cone moves to the table, the exact cull is kept
                check_qpair(pq[i], j, lo, hi, geo.astype(np.float64), None if not (0 < ecos < 1.5) else
                            (pf[i]["eaxis"][j], ecos, *_normals2d(sils, sl)), ecos > 1.5)
                has = bool(pq[i]["flags"] & (1 << j))
                assert has == (pf[i]["ref"][j] == 0)
                if has:
                    rc = refs[qref(pq[i], j)]
                    np.testing.assert_array_equal(rc["lo"], lo)
                    np.testing.assert_array_equal(rc["hi"], hi)
                    np.testing.assert_array_equal(rc["axis"], pf[i]["cone_axis"][j])
                    assert rc["half"] == pf[i]["half"][j] and rc["radius"] == pf[i]["radius"][j]
                    assert rc["sin"] == pf[i]["sin"][j] and rc["cos"] == pf[i]["cos"][j]
                continue
            if D == 3:
                has = bool(pq[i]["flags"] & (1 << j))
                assert has == (pf[i]["ref"][j] >= 0)
                if has:
                    assert qref(pq[i], j) == pf[i]["ref"][j]
                    rc = refs[pf[i]["ref"][j]]
                    np.testing.assert_array_equal(rc["lo"], lo)
                    np.testing.assert_array_equal(rc["hi"], hi)
                    assert 0.0 <= rc["half"] <= np.float32(np.pi / 2)
                    assert abs(rc["sin"] - np.sin(np.float64(rc["half"]))) <= 1e-7
                    assert abs(rc["cos"] - np.cos(np.float64(rc["half"]))) <= 1e-7
            if D == 3 and sl:
                geo = seg[sl].reshape(-1, 3)
                assert (geo >= lo).all() and (geo <= hi).all()
                # the exact cull's extents: every edge endpoint below, relative to the float32 box
                # centre, projects onto the cone axis within [slo, shi] and lies within sperp of the
                # axis line; each bound is attained to within rounding
                cen = (np.float32(0.5) * (lo + hi)).astype(np.float64)
                ax = pf[i]["axis0" if j == 0 else "axis1"].astype(np.float64)
                d = geo.astype(np.float64) - cen
                al = d @ ax
                pe = np.linalg.norm(d - al[:, None] * ax, axis=1)
                slo, shi, sperp = (float(pf[i][k][j]) for k in ("slo", "shi", "sperp"))
                scale = np.abs(d).max() + 1e-30
                assert slo <= al.min() and al.max() <= shi and pe.max() <= sperp
                assert al.min() - slo <= 1e-6 * scale and shi - al.max() <= 1e-6 * scale
                assert sperp - pe.max() <= 1e-6 * scale
                cos_t = pf[i]["cos0" if j == 0 else "cos1"]
                if 0 < cos_t < 1.5:
                    ax = pf[i]["axis0" if j == 0 else "axis1"].astype(np.float64)
                    a0 = n0[sl].astype(np.float64) @ ax
                    a1 = n1[sl].astype(np.float64) @ ax
                    a0 /= np.linalg.norm(n0[sl].astype(np.float64), axis=1)
                    a1 /= np.linalg.norm(n1[sl].astype(np.float64), axis=1)
                    sg = np.sign(a0 + a1)
                    assert (sg * a0 >= cos_t - 1e-6).all() and (sg * a1 >= cos_t - 1e-6).all()
                check_qpair(pq[i], j, lo, hi, geo.astype(np.float64),
                            (ax, cos_t, n0[sl], n1[sl]) if 0 < cos_t < 1.5 else None, cos_t > 1.5)
    assert (covered == 1).all()


def qref(rec, j):
    """Reference-cone table index of child j of a quantized pair (consecutive per pair)."""
    return int(rec["ref_base"]) + (int(rec["flags"]) & 1 if j == 1 else 0)


def _normals2d(sils, sl):
    A, B, Cc = (sils[sl, k:k + 2].astype(np.float64) for k in (0, 2, 4))
    n1 = np.stack([-(B - A)[:, 1], (B - A)[:, 0], np.zeros(len(sl))], 1)
    n2 = np.stack([-(Cc - B)[:, 1], (Cc - B)[:, 0], np.zeros(len(sl))], 1)
    ok = np.isfinite(n1).all(1) & np.isfinite(n2).all(1)
    return n1[ok], n2[ok]


def check_qpair(rec, j, lo, hi, pts, cone, none):
    """Child j of a quantized pair record against its float32 child: the decoded box holds the float
    box; the decoded cone (int16 axis, rounded-up sine) holds every predicate normal pair the float
    cone holds and is at most 2e-4 rad wider; the extents about the decoded centre, scaled by the
    decoded half-diagonal, hold every silhouette point.  `cone` = (float axis, cos, n0, n1) when the
    float cone can cull."""
    D = len(lo)
    qlo = decode(rec["qlo"][j], rec["exp"], rec["origin"])
    qhi = decode(rec["qhi"][j], rec["exp"], rec["origin"])
    assert (qlo <= lo).all() and (qhi >= hi).all()
    assert bool(rec["flags"] & (4 << j)) == bool(none)
    qsin = rec["sin"][j] / 32767.0
    qcos = 0.0 if rec["sin"][j] >= 32767 else np.sqrt(1.0 - qsin * qsin)
    a = rec["axis"][j].astype(np.float64) / 32767.0
    u = a / np.linalg.norm(a) if np.linalg.norm(a) > 0 else a
    if cone is None:
        assert none or qcos == 0.0
    elif qcos > 0:
        ax, cos_t, n0, n1 = cone
        assert qcos <= cos_t + 1e-6   # no narrower than the float cone
        assert np.arccos(qcos) <= np.arccos(min(float(cos_t), 1.0)) + 2e-4
        u3 = np.zeros(3)
        u3[:D] = u
        n0, n1 = np.asarray(n0, np.float64), np.asarray(n1, np.float64)
        c0 = n0 @ u3 / np.linalg.norm(n0, axis=1)
        c1 = n1 @ u3 / np.linalg.norm(n1, axis=1)
        sg = np.sign(c0 + c1)
        assert (sg * c0 >= qcos - 1e-6).all() and (sg * c1 >= qcos - 1e-6).all()
    if none or len(pts) == 0:
        return
    cen = (np.float32(0.5) * (qlo + qhi)).astype(np.float64)
    h = (np.float32(0.5) * (qhi - qlo)).astype(np.float64)
    R = np.sqrt((h * h).sum())
    d = pts.reshape(-1, D) - cen
    al = d @ u
    pe = np.linalg.norm(d - al[:, None] * u, axis=1)
    tol = 1e-6 * (R + 1e-30)
    mid, hw = rec["ext_mid"][j] / 32767.0 * R, (rec["ext_half"][j] / 255.0) ** 2 * R
    assert mid - hw <= al.min() + tol and al.max() <= mid + hw + tol
    assert pe.max() <= rec["ext_perp"][j] / 255.0 * R + tol
    # tight: the half-width to one step of its square-root code, the middle to R / 65534
    need = max(mid - al.min(), al.max() - mid)
    assert np.sqrt(hw / R) - np.sqrt(max(need, 0.0) / R) <= 1.01 / 255.0 + 1e-6
    assert abs(mid - 0.5 * (al.min() + al.max())) <= 0.51 * R / 32767.0 + tol


@pytest.mark.parametrize("name", ["star300", "star8192", "ico1", "ico4"])
def test_seed_cuts(name):
    """The wave-per-query passes' seed cuts: each partitions the tree's primitives / silhouettes
    below the root (every one under exactly one seed), carries its node's float32 box, and the SNCH
    cut never passes below a child whose reference cone test can cull."""
    V, F = SCENES[name]()
    D = V.shape[1]
    s = (ext.SNCH if D == 2 else ext.SNCH3D)(V, F)
    big = s.engine_layout(9).view(np.uint32)
    f = s.engine_layout(1).view(fnode_dtype(D))
    pf = s.engine_layout(3).view(spf_dtype(D))
    for which, n_items in ((12, F.shape[0]), (13, len(s.engine_layout(7).view(np.int32)))):
        seeds = s.engine_layout(which).view(SEED)
        assert len(seeds) <= 64
        if len(seeds) == 0:
            # no cut: a child of the root pair has a reference cone that can cull
            assert which == 13 and (pf[0]["ref"] >= 0).any()
            continue
        # box of every reachable ref, and whether it sits below a reference cone
        boxes, below_cone = {}, {}
        stack = [(0, False)]
        while stack:
            i, cone = stack.pop()
            if which == 12:
                nd = f[i]
                kids = [(int(nd["child"][j]), [nd["lo"][k][j] for k in range(D)], [nd["hi"][k][j] for k in range(D)], False)
                        for j in range(int(nd["nchild"]))]
            else:
                kids = [(int(pf[i]["child"][j]), pf[i]["lo"][j], pf[i]["hi"][j], int(pf[i]["ref"][j]) >= 0)
                        for j in range(2)]
            for ref, lo, hi, rc in kids:
                boxes[ref] = (np.float32(lo), np.float32(hi))
                below_cone[ref] = cone or rc
                if not ref & LEAF:
                    stack.append((ref, cone or rc))
        covered = np.zeros(n_items, np.int64)

        def cover(ref):
            if ref & LEAF:
                a, c = leaf_range(ref, big)
                covered[a:a + c] += 1
            elif which == 12:
                for j in range(int(f[ref]["nchild"])):
                    cover(int(f[ref]["child"][j]))
            else:
                cover(int(pf[ref]["child"][0]))
                cover(int(pf[ref]["child"][1]))

        for e in seeds:
            ref = int(e["ref"])
            lo, hi = boxes[ref]
            np.testing.assert_array_equal(e["lo"][:D], lo)
            np.testing.assert_array_equal(e["hi"][:D], hi)
            assert not below_cone[ref]
            cover(ref)
        assert (covered == 1).all()
        if which == 12 and F.shape[0] > 4096:
            assert len(seeds) > 16   # the cut goes a few levels down on a large tree


def halves(raw16):
    """Upper halves of float32 patterns, widened."""
    return (raw16.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bvh4_child_cylinders(name):
    """Child bounding cylinders (CBound), both arrays: axis length <= 1; every vertex below child j
    lies in the slab [lo_j, hi_j] along the axis and within rad_j of the axis line through the child
    box's centre as the kernel computes it (fma(qlo + qhi, 2^(e - 1), origin) for the quantized
    nodes, 0.5f * (lo + hi) of the float32 box for the other array), checked in float64; tight to
    the 16-bit rounding."""
    V, F = SCENES[name]()
    D = V.shape[1]
    s = (ext.SNCH if D == 2 else ext.SNCH3D)(V, F)
    q = s.engine_layout(0).view(qnode_dtype(D))
    f = s.engine_layout(1).view(fnode_dtype(D))
    big = s.engine_layout(9).view(np.uint32)
    n_prims = F.shape[0]
    prims = s.engine_layout(4).view(np.float32).reshape(-1, 4)
    pts = prims.reshape(n_prims, 4, 4)[:, :3, :3] if D == 3 else prims.reshape(n_prims, 2, 2)

    def slots(ref):
        if ref & LEAF:
            a, c = leaf_range(ref, big)
            return np.arange(a, a + c)
        nd = f[ref]
        return np.concatenate([slots(int(nd["child"][j])) for j in range(int(nd["nchild"]))])

    for which in (14, 15):
        raw = s.engine_layout(which).view(np.uint8).reshape(-1, 64)
        assert len(raw) == len(f)
        h16 = raw[:, :8 * D + 24].copy().view(np.uint16)
        ax = halves(h16[:, :4 * D]).reshape(-1, D, 4)
        lo, hi, rad = (halves(h16[:, 4 * D + 4 * k:4 * D + 4 * k + 4]) for k in range(3))
        for i in range(len(f)):
            for j in range(int(f[i]["nchild"])):
                if which == 14:
                    # fma(qlo + qhi, 2^(e - 1), origin), rounded once to float32
                    qs = q[i]["qlo"][:, j].astype(np.int64) + q[i]["qhi"][:, j].astype(np.int64)
                    c = decode(qs, q[i]["exp"].astype(np.int64) - 1, q[i]["origin"]).astype(np.float64)
                else:
                    blo = np.array([f[i]["lo"][k][j] for k in range(D)], np.float32)
                    bhi = np.array([f[i]["hi"][k][j] for k in range(D)], np.float32)
                    c = (np.float32(0.5) * (blo + bhi)).astype(np.float64)
                u = ax[i, :, j]
                assert np.linalg.norm(u) <= 1.0
                dv = pts[slots(int(f[i]["child"][j]))].reshape(-1, D).astype(np.float64) - c
                t = dv @ u
                r = np.linalg.norm(dv - t[:, None] * u[None], axis=1)
                assert lo[i, j] <= t.min() and t.max() <= hi[i, j] and r.max() <= rad[i, j], (which, i, j)
                tol = 2.0 ** -7
                assert t.min() - lo[i, j] <= tol * abs(t.min()) + 1e-30
                assert hi[i, j] - t.max() <= tol * abs(t.max()) + 1e-30
                assert rad[i, j] - r.max() <= tol * r.max() + 1e-30
