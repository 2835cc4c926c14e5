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


def summarize(name, got, ref, gap=None, tol=TOL_REF, vec_fields=("p",), scalar_fields=("d",)):
    """Returns a dict of mismatch counts and max errors; `gap` masks queries where ids must match."""
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
        # points (and normals) are only defined up to ties: compare them where the ids must match
        g = np.asarray(got[f]).reshape(valid_r.size, -1)[m]
        r = np.asarray(ref[f]).reshape(valid_r.size, -1)[m]
        e = rel_err(g, r)
        out[f"max_err_{f}"] = float(e.max()) if e.size else 0.0
    out["ok"] = (out["valid_mismatch"] == 0 and out["id_mismatch"] == 0
                 and all(out[k] <= tol for k in out if k.startswith("max_err_")))
    return out


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
