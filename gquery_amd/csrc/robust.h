// FP64 re-evaluation of the float32 decisions that rounding can flip (engine kernels only; the
// plain kernels keep the reference's float32 arithmetic throughout).
//
//  * Möller–Trumbore (math.py:110-128, triangle.py:48-73) when the determinant is below 0.05 of
//    the product of the edge and direction lengths: a grazing ray, where float32 t, v and w lose
//    digits.
//  * The silhouette-edge sign predicate (silhouette_edge.py:7-12) and the silhouette-vertex sign
//    predicate (silhouette_vertex.py:26-37) when either factor is within its float32 rounding
//    bound of zero.
// Outside those bands the float32 result is returned unchanged, so the engine agrees with the
// reference bit for bit there; inside them it returns the FP64 evaluation of the same formulas.
#pragma once

#include "gq_device.h"

namespace gqd {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(const V3 &v) { return {(double)v.x, (double)v.y, (double)v.z}; }
__device__ __forceinline__ D3 dsub(const D3 &a, const D3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double ddot(const D3 &a, const D3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dcross(const D3 &a, const D3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Ray x + t·d against triangle abc with the reference's acceptance rules: |det| > 2^-24,
// v in [0, 1], w >= 0, v + w <= 1, t >= 0.  Returns t (or +inf) and the barycentric v, w.
__device__ __forceinline__ float ray_triangle_robust(V3 x, V3 d, V3 a, V3 b, V3 c, float &vv, float &ww) {
    const V3 v1 = b - a, v2_ = c - a;
    const V3 p = cross(d, v2_);
    const float det = dot(v1, p);
    // |det| = |v1|·|d|·|v2|·|sin| terms.  The float32 t carries a relative error of about
    // 2e-8 · scale / |det| (measured against FP64 on grazing rays over a flat grid: 1.6e-5 at
    // |det| = 1e-3 scale, 6.8e-7 at 0.03-0.1 scale), so below 0.05 of the scale the test runs in
    // FP64 and the float32 path keeps t within ~4e-7 of the FP64 value.
    const float scale = (fabsf(v1.x) + fabsf(v1.y) + fabsf(v1.z)) * (fabsf(d.x) + fabsf(d.y) + fabsf(d.z)) *
                        (fabsf(v2_.x) + fabsf(v2_.y) + fabsf(v2_.z));
    vv = 0.0f;
    ww = 0.0f;
    if (fabsf(det) > 0.05f * scale) {
        if (!(fabsf(det) > kDrEps)) return __builtin_inff();
        const float inv_det = 1.0f / det;
        const V3 s = x - a;
        const float v = dot(s, p) * inv_det;
        if (!((v >= 0.0f) & (v <= 1.0f))) return __builtin_inff();
        const V3 q = cross(s, v1);
        const float w = dot(d, q) * inv_det;
        if (!((w >= 0.0f) & (v + w <= 1.0f))) return __builtin_inff();
        const float t = dot(v2_, q) * inv_det;
        if (!(t >= 0.0f)) return __builtin_inff();
        vv = v;
        ww = w;
        return t;
    }
    const D3 A = d3(a), Dd = d3(d), e1 = dsub(d3(b), A), e2 = dsub(d3(c), A);
    const D3 P = dcross(Dd, e2);
    const double det64 = ddot(e1, P);
    if (!(fabs(det64) > (double)kDrEps)) return __builtin_inff();
    const double inv = 1.0 / det64;
    const D3 S = dsub(d3(x), A);
    const double v = ddot(S, P) * inv;
    if (!((v >= 0.0) & (v <= 1.0))) return __builtin_inff();
    const D3 Q = dcross(S, e1);
    const double w = ddot(Dd, Q) * inv;
    if (!((w >= 0.0) & (v + w <= 1.0))) return __builtin_inff();
    const double t = ddot(e2, Q) * inv;
    if (!(t >= 0.0)) return __builtin_inff();
    vv = (float)v;
    ww = (float)w;
    return (float)t;
}

// Silhouette-edge predicate dot(pt − x, n0) · dot(pt − x, n1) < 0, pt the nearest point of the
// segment ab to x.  The float32 factors carry an absolute error below ~1e-6 of Σ|vd_i·n_i| plus
// the error of pt itself (~1e-7 of the coordinates); inside that band both factors are
// recomputed in FP64 from a, b, x and the stored normals.
__device__ __forceinline__ bool sil_edge_pred_robust(V3 x, V3 a, V3 b, V3 pt, V3 n0, V3 n1) {
    const V3 vd = pt - x;
    const float d0 = dot(vd, n0), d1 = dot(vd, n1);
    const float mag = fabsf(x.x) + fabsf(x.y) + fabsf(x.z) + fabsf(a.x) + fabsf(a.y) + fabsf(a.z) +
                      fabsf(b.x) + fabsf(b.y) + fabsf(b.z);
    const float e0 = 2e-6f * (fabsf(vd.x * n0.x) + fabsf(vd.y * n0.y) + fabsf(vd.z * n0.z)) + 1e-6f * mag;
    const float e1 = 2e-6f * (fabsf(vd.x * n1.x) + fabsf(vd.y * n1.y) + fabsf(vd.z * n1.z)) + 1e-6f * mag;
    if (fabsf(d0) > e0 && fabsf(d1) > e1) return d0 * d1 < 0.0f;
    const D3 A = d3(a), X = d3(x), ab = dsub(d3(b), A);
    const double l2 = ddot(ab, ab);
    double t = l2 > 0.0 ? ddot(dsub(X, A), ab) / l2 : 0.0;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    const D3 vd64 = {A.x + t * ab.x - X.x, A.y + t * ab.y - X.y, A.z + t * ab.z - X.z};
    return ddot(vd64, d3(n0)) * ddot(vd64, d3(n1)) < 0.0;
}

// Silhouette-vertex predicate cross(b − a, x − a) · cross(c − b, x − b) < 0 with the same
// treatment (each 2D cross product is exact in FP64 from float32 inputs).
__device__ __forceinline__ bool sil_vertex_pred_robust(V2 x, V2 a, V2 b, V2 c) {
    const float c0 = cross2(b - a, x - a), c1 = cross2(c - b, x - b);
    const V2 u = b - a, w = x - a, u1 = c - b, w1 = x - b;
    const float e0 = 2e-6f * (fabsf(u.x * w.y) + fabsf(u.y * w.x)) + 1e-7f * (fabsf(a.x) + fabsf(a.y) + fabsf(x.x) + fabsf(x.y)) * (fabsf(u.x) + fabsf(u.y));
    const float e1 = 2e-6f * (fabsf(u1.x * w1.y) + fabsf(u1.y * w1.x)) + 1e-7f * (fabsf(b.x) + fabsf(b.y) + fabsf(x.x) + fabsf(x.y)) * (fabsf(u1.x) + fabsf(u1.y));
    if (fabsf(c0) > e0 && fabsf(c1) > e1) return c0 * c1 < 0.0f;
    const double ax = a.x, ay = a.y, bx = b.x, by = b.y, cx = c.x, cy = c.y, px = x.x, py = x.y;
    const double C0 = (bx - ax) * (py - ay) - (by - ay) * (px - ax);
    const double C1 = (cx - bx) * (py - by) - (cy - by) * (px - bx);
    return C0 * C1 < 0.0;
}

} // namespace gqd
