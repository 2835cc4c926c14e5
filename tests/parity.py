"""Shared parity checks (the acceptance rule for every query path).

A result matches a reference answer when:
  * valid flags agree;
  * the primitive index agrees wherever the runner-up gap exceeds GAP_REL relative
    (ties and near-ties may legitimately resolve to another primitive);
  * distances and points agree within TOL_REF relative (to an FP32 reference) or TOL_F64
    (to the FP64 brute force), with an absolute floor of the same size for values near zero.
"""
import numpy as np

GAP_REL = 1e-5   # runner-up gap above which the primitive index must match
TOL_REF = 1e-5   # vs the FP32 restatement
TOL_F64 = 1e-6   # vs the FP64 oracle (absolute for |x| <= 1, relative above)


def rel_err(a, b, floor=1.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def gap_mask(best, second):
    best = np.asarray(best, np.float64)
    second = np.asarray(second, np.float64)
    with np.errstate(invalid="ignore"):
        gap = second - best
    return np.where(np.isfinite(best), gap > GAP_REL * np.maximum(np.abs(best), 1e-30), False) | (
        np.isfinite(best) & ~np.isfinite(second))


RECORD_FIELDS = ("p", "n", "uv", "t")   # compared where present in both records


def summarize(name, got, ref, gap=None, tol=TOL_REF, vec_fields=RECORD_FIELDS, scalar_fields=("d",)):
    """Returns a dict of mismatch counts and max errors; `gap` masks queries where ids must match.

    Every field of `vec_fields` present in both records (p, n, uv: the 3D barycentrics, t: the
    2D segment parameter t) is compared where the ids must match; `on_boundary` must be equal on
    every query both call valid."""
    valid_g = np.asarray(got["valid"]).astype(bool)
    valid_r = np.asarray(ref["valid"]).astype(bool)
    out = {"name": name, "n": int(valid_r.size), "valid_mismatch": int((valid_g != valid_r).sum())}
    both = valid_g & valid_r
    m = both if gap is None else both & gap
    out["id_checked"] = int(m.sum())
    out["id_mismatch"] = int((np.asarray(got["prim_id"])[m] != np.asarray(ref["prim_id"])[m]).sum())
    for f in scalar_fields:
        e = rel_err(np.asarray(got[f])[both], np.asarray(ref[f])[both])
        out[f"max_err_{f}"] = float(e.max()) if e.size else 0.0
    for f in vec_fields:
        if f not in got or f not in ref:
            continue
        # points, normals and barycentrics are only defined up to ties: compare them where the ids
        # must match
        g = np.asarray(got[f]).reshape(valid_r.size, -1)[m]
        r = np.asarray(ref[f]).reshape(valid_r.size, -1)[m]
        e = rel_err(g, r)
        out[f"max_err_{f}"] = float(e.max()) if e.size else 0.0
    if "on_boundary" in got and "on_boundary" in ref:
        ob_g = np.asarray(got["on_boundary"]).astype(bool)[both]
        ob_r = np.asarray(ref["on_boundary"]).astype(bool)[both]
        out["on_boundary_mismatch"] = int((ob_g != ob_r).sum())
    out["fields"] = [f for f in vec_fields if f in got and f in ref]
    out["ok"] = (out["valid_mismatch"] == 0 and out["id_mismatch"] == 0 and out.get("on_boundary_mismatch", 0) == 0
                 and all(out[k] <= tol for k in out if k.startswith("max_err_")))
    return out


def ray3d_vs_fp64(got, ref, b, gap):
    """3D rays through the engine, which re-evaluates grazing ray-triangle tests in FP64 (elsewhere
    it runs the reference's float32 arithmetic).  The engine's hit distance t is held to the FP64
    brute force `b` on every hit; its barycentrics uv, where the first hit is not a near tie and the
    ids agree, equal either the reference's float32 ones (within TOL_REF) or the FP64 ones (within
    TOL_F64: the FP64 band).  Float32 barycentrics carry an error of about 2^-24 |x - a| / |edge|,
    which on small triangles exceeds TOL_F64 at any angle, so FP64 alone cannot pin them.  Returns
    the queries on which the FP32 reference is itself accurate (no hit, or t within 0.1 TOL_REF and
    uv within 0.5 TOL_REF of FP64), the ones the other fields are compared on."""
    hit = np.isfinite(b["t"]) & np.asarray(got["valid"]).astype(bool)
    assert rel_err(got["d"][hit], b["t"][hit]).max(initial=0.0) <= TOL_F64, "t vs FP64"
    m = hit & gap & (np.asarray(got["prim_id"]) == b["prim_id"]) & (np.asarray(ref["prim_id"]) == b["prim_id"])
    e_ref = rel_err(got["uv"][m], ref["uv"][m]).max(1, initial=0.0)
    e_f64 = rel_err(got["uv"][m], b["uv"][m]).max(1, initial=0.0)
    assert ((e_ref <= TOL_REF) | (e_f64 <= TOL_F64)).all(), "uv vs FP32 reference and FP64"
    with np.errstate(invalid="ignore"):
        acc = (rel_err(ref["d"], b["t"]) <= 0.1 * TOL_REF) & (rel_err(ref["uv"], b["uv"]).max(1) <= 0.5 * TOL_REF)
    return ~np.isfinite(b["t"]) | acc


def brief(res, tol=TOL_REF):
    """The failing entries of a summarize() result, for assertion messages."""
    keep = {"name", "n", "id_checked"}
    return {k: v for k, v in res.items()
            if k in keep or (k.endswith("mismatch") and v) or (k.startswith("max_err_") and v > tol)}


def sil2d_fp64_predicate(silhouette_data, ids, x):
    """FP64 silhouette-vertex predicate cross(b-a, x-a) * cross(c-b, x-b) < 0 of the vertices `ids`
    (the 2D records' prim_id) for query points x, from the flat 2D silhouette records (stride 10:
    a, b, c, neighbour indices, vertex id)."""
    sil = np.asarray(silhouette_data, np.float64).reshape(-1, 10)
    row = {int(v): k for k, v in enumerate(sil[:, 9])}
    r = np.array([row[int(i)] for i in ids], np.int64)
    a, b, c = sil[r, 0:2], sil[r, 2:4], sil[r, 4:6]
    X = np.asarray(x, np.float64).reshape(-1, 2)
    c0 = (b[:, 0] - a[:, 0]) * (X[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (X[:, 0] - a[:, 0])
    c1 = (c[:, 0] - b[:, 0]) * (X[:, 1] - b[:, 1]) - (c[:, 1] - b[:, 1]) * (X[:, 0] - b[:, 0])
    return c0 * c1 < 0.0


def sil2d_differences_explained(silhouette_data, x, eng, ref):
    """Queries where the engine's 2D silhouette differs from the reference float32 one, and whether
    every difference is the FP64 predicate overruling a float32 sign: the engine's nearer vertex
    passes in FP64, or the reference's nearer vertex fails in FP64."""
    ev, rv = np.asarray(eng["valid"]).astype(bool), np.asarray(ref["valid"]).astype(bool)
    ed, rd = np.asarray(eng["d"]), np.asarray(ref["d"])
    diff = np.where((ev != rv) | (ev & rv & (ed != rd)))[0]
    eng_nearer = ev[diff] & (~rv[diff] | (ed[diff] < rd[diff]))
    ok = np.ones(len(diff), bool)
    x = np.asarray(x).reshape(-1, 2)
    i = diff[eng_nearer]
    if len(i):
        ok[eng_nearer] = sil2d_fp64_predicate(silhouette_data, np.asarray(eng["prim_id"])[i], x[i])
    i = diff[~eng_nearer]
    if len(i):
        ok[~eng_nearer] = ~sil2d_fp64_predicate(silhouette_data, np.asarray(ref["prim_id"])[i], x[i])
    return diff, ok
