/* gq_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of gQuery's traversal path, used by tests/ as the parity checker and by
 * bench.py as the CPU baseline.  Nothing in the product (gquery_amd/) links, loads or calls it.
 *
 * Part 1 (FP32): the six reference traversals over the flat float32 arrays of the reference
 * wire format, following the reference line by line:
 *   or_closest_point    BVH.closest_point     gquery/shapes/bvh.py:127-198
 *                       BVH3D.closest_point   gquery/shapes/bvh3d.py:108-174
 *   or_intersect        BVH.intersect         gquery/shapes/bvh.py:200-276
 *                       BVH3D.intersect       gquery/shapes/bvh3d.py:176-239
 *   or_silhouette       SNCH.closest_silhouette   gquery/shapes/snch.py:106-164
 *                       SNCH3D.closest_silhouette gquery/shapes/snch3d.py:143-203
 *   leaf tests / predicates: gquery/core/math.py, gquery/shapes/{line_segment,triangle,
 *   silhouette_vertex,silhouette_edge}.py
 * DrJit arithmetic forms are restated as: dot = fma chain, cross = fma(a.y,b.z,-(a.z*b.y)),
 * lerp(a,b,t) = fma(b,t,fma(-a,t,a)), normalize = v * (1/sqrt(dot)), every other expression in
 * source order; compiled with -ffp-contract=off.  Which ulp DrJit itself produces is not known
 * offline: those bits are "parity unpinned", absorbed by the 1e-5 tolerance of the tests.
 * Each query also reports its work: popped nodes, child-node tests and leaf tests, the counts
 * the algorithmic-byte figure is built from.
 *
 * Part 2 (FP64): brute force over every primitive / silhouette, independent of any tree: nearest
 * primitive and runner-up distance (closest point), first and second hit (rays), nearest valid
 * silhouette and runner-up (silhouettes).  These pin the FP32 results to the mathematics.
 *
 * Part 2b (FP32): brute force with the traversal's own leaf test (exact target for conservative
 * pruning at any coordinate scale), and the leaf distance of point/triangle pairs.
 * Part 3 (FP32): the brute-force *_baseline queries and mesh normals (gquery/shapes/polyline.py,
 * mesh.py).  Part 4: BVH-guided boundary sampling (gquery/shapes/bvh3d.py:246-346) and the PCG32
 * generator it draws from.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_STACK 256

typedef struct { float x, y, z; } v3;

/* ---- FP32 helpers ------------------------------------------------------------------------- */
static inline float dot2f(const float *a, const float *b) { return fmaf(a[1], b[1], a[0] * b[0]); }
static inline float dot3f(const float *a, const float *b) {
    return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}
static inline float dotf(int D, const float *a, const float *b) { return D == 2 ? dot2f(a, b) : dot3f(a, b); }
static inline float normf(int D, const float *a) { return sqrtf(dotf(D, a, a)); }
static inline void cross3f(const float *a, const float *b, float *o) {
    o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}
static inline void normalizef(int D, float *a) {
    float s = 1.0f / sqrtf(dotf(D, a, a));
    for (int k = 0; k < D; ++k) a[k] = a[k] * s;
}
static inline float clampf_(float x, float lo, float hi) { return fmaxf(fminf(x, hi), lo); }
static inline float cross2f(const float *a, const float *b) { return a[0] * b[1] - a[1] * b[0]; } /* math.py:9-11 */

static const float EPS24 = 5.9604644775390625e-08f; /* dr.epsilon(Float32) */
static const float HALF_PI = 1.57079637050628662109375f;

/* BoundingBox.distance (bvh.py:22-27) */
static inline void box_dist(int D, const float *lo, const float *hi, const float *p, float *dmin, float *dmax) {
    float a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        float u = lo[k] - p[k], v = p[k] - hi[k];
        a[k] = fmaxf(fmaxf(u, v), 0.0f);
        b[k] = fminf(u, v);
    }
    *dmin = normf(D, a);
    *dmax = normf(D, b);
}

/* BoundingBox.intersect (bvh.py:33-50) */
static inline int box_isect(int D, const float *lo, const float *hi, const float *x, const float *v,
                            float r_max, float *tmin) {
    float tn = -INFINITY, tf = INFINITY;
    for (int k = 0; k < D; ++k) {
        float t0 = (lo[k] - x[k]) / v[k], t1 = (hi[k] - x[k]) / v[k];
        float n = fminf(t0, t1), f = fmaxf(t0, t1);
        tn = (k == 0) ? n : fmaxf(tn, n);
        tf = (k == 0) ? f : fminf(tf, f);
    }
    float tnm = fmaxf(tn, 0.0f), tfm = fminf(tf, r_max);
    if (tnm > tfm) { *tmin = INFINITY; return 0; }
    *tmin = tnm;
    return 1;
}

/* closest_point_line_segment (math.py:14-21) */
static inline float cp_segment(int D, const float *x, const float *a, const float *b, float *p) {
    float u[3], w[3], q[3];
    for (int k = 0; k < D; ++k) { u[k] = b[k] - a[k]; w[k] = x[k] - a[k]; }
    float t = clampf_(dotf(D, w, u) / dotf(D, u, u), 0.0f, 1.0f);
    for (int k = 0; k < D; ++k) { p[k] = fmaf(b[k], t, fmaf(-a[k], t, a[k])); q[k] = p[k] - x[k]; }
    return normf(D, q);
}

/* closest_point_triangle (math.py:42-107): regions a, b, c, ab, ac, bc, face; first match wins */
static float cp_triangle(const float *p, const float *a, const float *b, const float *c, float *pt, float *uvo) {
    float ab[3], ac[3], ax[3], bx[3], cx[3], uv[2], q[3];
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ax[k] = p[k] - a[k]; }
    float d1 = dot3f(ab, ax), d2 = dot3f(ac, ax);
    if ((d1 <= 0.0f) && (d2 <= 0.0f)) {
        memcpy(pt, a, 12); uv[0] = 1.0f; uv[1] = 0.0f; goto done;
    }
    for (int k = 0; k < 3; ++k) bx[k] = p[k] - b[k];
    float d3 = dot3f(ab, bx), d4 = dot3f(ac, bx);
    if ((d3 >= 0.0f) && (d4 <= d3)) {
        memcpy(pt, b, 12); uv[0] = 0.0f; uv[1] = 1.0f; goto done;
    }
    for (int k = 0; k < 3; ++k) cx[k] = p[k] - c[k];
    float d5 = dot3f(ab, cx), d6 = dot3f(ac, cx);
    if ((d6 >= 0.0f) && (d5 <= d6)) {
        memcpy(pt, c, 12); uv[0] = 0.0f; uv[1] = 0.0f; goto done;
    }
    float vc = d1 * d4 - d3 * d2;
    if ((vc <= 0.0f) && (d1 >= 0.0f) && (d3 <= 0.0f)) {
        float v = d1 / (d1 - d3);
        for (int k = 0; k < 3; ++k) pt[k] = a[k] + ab[k] * v;
        uv[0] = 1.0f - v; uv[1] = v; goto done;
    }
    float vb = d5 * d2 - d1 * d6;
    if ((vb <= 0.0f) && (d2 >= 0.0f) && (d6 <= 0.0f)) {
        float w = d2 / (d2 - d6);
        for (int k = 0; k < 3; ++k) pt[k] = a[k] + ac[k] * w;
        uv[0] = 1.0f - w; uv[1] = 0.0f; goto done;
    }
    float va = d3 * d6 - d5 * d4;
    if ((va <= 0.0f) && ((d4 - d3) >= 0.0f) && ((d5 - d6) >= 0.0f)) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int k = 0; k < 3; ++k) pt[k] = b[k] + (c[k] - b[k]) * w;
        uv[0] = 0.0f; uv[1] = 1.0f - w; goto done;
    }
    {
        float denom = 1.0f / (va + vb + vc);
        float v = vb * denom, w = vc * denom;
        for (int k = 0; k < 3; ++k) pt[k] = a[k] + ab[k] * v + ac[k] * w;
        uv[0] = 1.0f - v - w; uv[1] = v;
    }
done:
    uvo[0] = uv[1];
    uvo[1] = 1.0f - uv[0] - uv[1];
    for (int k = 0; k < 3; ++k) q[k] = p[k] - pt[k];
    return normf(3, q);
}

/* ray_triangle_intersect (math.py:110-128) with r_max = inf; also returns v, w for the record
 * (Triangle.ray_intersect, triangle.py:48-73). */
static float ray_tri(const float *x, const float *d, const float *a, const float *b, const float *c,
                     float *vo, float *wo) {
    float v1[3], v2[3], p[3], s[3], q[3];
    for (int k = 0; k < 3; ++k) { v1[k] = b[k] - a[k]; v2[k] = c[k] - a[k]; }
    cross3f(d, v2, p);
    float det = dot3f(v1, p);
    if (fabsf(det) > EPS24) {
        float inv = 1.0f / det;
        for (int k = 0; k < 3; ++k) s[k] = x[k] - a[k];
        float v = dot3f(s, p) * inv;
        if ((v >= 0.0f) && (v <= 1.0f)) {
            cross3f(s, v1, q);
            float w = dot3f(d, q) * inv;
            if ((w >= 0.0f) && (v + w <= 1.0f)) {
                float t = dot3f(v2, q) * inv;
                if (t >= 0.0f) { *vo = v; *wo = w; return t; }
            }
        }
    }
    return INFINITY;
}

/* ray_intersection (math.py:31-39) */
static float ray_seg(const float *x, const float *v, const float *a, const float *b) {
    float u[2] = {b[0] - a[0], b[1] - a[1]}, w[2] = {x[0] - a[0], x[1] - a[1]};
    float d = cross2f(v, u);
    float s = cross2f(v, w) / d;
    float t = cross2f(u, w) / d;
    return ((t > 0.0f) && (s >= 0.0f) && (s <= 1.0f)) ? t : INFINITY;
}

/* SNCHNode.overlaps / SNCHNode3D.overlaps (snch.py:15-60, snch3d.py:50-95) */
static inline __attribute__((always_inline)) int snch_overlap(int D, const float *nd, const float *x, float r_max, float *dmin) {
    const float *lo = nd, *hi = nd + D, *axis = nd + 2 * D;
    const float half = nd[3 * D], radius = nd[3 * D + 1];
    float dmax;
    box_dist(D, lo, hi, x, dmin, &dmax);
    int ov = *dmin <= r_max;
    if ((half >= 0.0f) && ov) {
        ov = (half > HALF_PI) && (*dmin < EPS24);
        if (!ov) {
            float c[3], va[3];
            for (int k = 0; k < D; ++k) { c[k] = (lo[k] + hi[k]) / 2.0f; va[k] = c[k] - x[k]; }
            float l = normf(D, va);
            for (int k = 0; k < D; ++k) va[k] = va[k] / l;
            float ang = acosf(clampf_(dotf(D, axis, va), -1.0f, 1.0f));
            ov = (HALF_PI >= ang - half) && (HALF_PI <= ang + half);
            if (!ov) {
                if (l > radius) {
                    float vh = asinf(radius / l);
                    float hs = half + vh;
                    ov = (hs > HALF_PI) || ((HALF_PI >= ang - hs) && (HALF_PI <= ang + hs));
                } else {
                    float e[3] = {0, 0, 0}, av[3] = {0, 0, 0};
                    for (int k = 0; k < D; ++k) { e[k] = hi[k] - c[k]; av[k] = fabsf(va[k]); }
                    float dd = dotf(D, e, av);
                    float s = l - dd;
                    ov = s < 0.0f;
                    if (!ov) {
                        if (D == 2) {
                            /* project_to_plane (math.py:24-28) */
                            float bb[2] = {fabsf(-va[1]), fabsf(va[0])};
                            dd = fabsf(dot2f(e, bb));
                        } else {
                            /* compute_orthonomal_basis + project_to_plane_3D (snch3d.py:11-30) */
                            float sign = copysignf(1.0f, va[2]);
                            float a = -1.0f / (sign + va[2]);
                            float b = va[0] * va[1] * a;
                            float b1[3] = {fabsf(1.0f + sign * va[0] * va[0] * a), fabsf(sign * b), fabsf(-sign * va[0])};
                            float b2[3] = {fabsf(b), fabsf(sign + va[1] * va[1] * a), fabsf(-va[1])};
                            float r1 = dot3f(e, b1), r2 = dot3f(e, b2);
                            dd = sqrtf(r1 * r1 + r2 * r2);
                        }
                        float vh = atan2f(dd, s);
                        float hs = half + vh;
                        ov = (hs > HALF_PI) || ((HALF_PI >= ang - hs) && (HALF_PI <= ang + hs));
                    }
                }
            }
        }
    }
    return ov;
}

/* Output arrays of the FP32 restatement; any pointer may be NULL. */
typedef struct {
    uint8_t *valid;
    float *p, *n, *uv, *d;
    int32_t *prim_id;
    uint8_t *on_boundary;
    int32_t *visits, *child_tests, *leaf_tests; /* work counters per query */
} or_out;

static void zero_out(int D, const or_out *o, int64_t i, int full) {
    if (o->valid) o->valid[i] = 0;
    if (o->p) for (int k = 0; k < D; ++k) o->p[D * i + k] = 0.0f;
    if (full) {
        if (o->n) for (int k = 0; k < D; ++k) o->n[D * i + k] = 0.0f;
        if (o->uv) for (int k = 0; k < D - 1; ++k) o->uv[(D - 1) * i + k] = 0.0f;
        if (o->on_boundary) o->on_boundary[i] = 0;
    }
    if (o->d) o->d[i] = 0.0f;
    if (o->prim_id) o->prim_id[i] = 0;
}

static void counters(const or_out *o, int64_t i, int visits, int child, int leaf) {
    if (o->visits) o->visits[i] = visits;
    if (o->child_tests) o->child_tests[i] = child;
    if (o->leaf_tests) o->leaf_tests[i] = leaf;
}

/* Stack push helper shared by the three traversals: far child first, near child last. */
#define PUSH2(li, ri, dl, dr_)                                                               \
    do {                                                                                     \
        int closer_ = (li), other_ = (ri);                                                   \
        float d0_ = (dl), d1_ = (dr_);                                                       \
        if (d1_ < d0_) { closer_ = (ri); other_ = (li); float t_ = d0_; d0_ = d1_; d1_ = t_; } \
        if (sp + 2 >= OR_STACK) { overflow = 1; break; }                                     \
        si[++sp] = other_; sd[sp] = d1_;                                                     \
        si[++sp] = closer_; sd[sp] = d0_;                                                    \
    } while (0)
#define PUSH1(ii, dd)                                        \
    do {                                                     \
        if (sp + 1 >= OR_STACK) { overflow = 1; break; }     \
        si[++sp] = (ii); sd[sp] = (dd);                      \
    } while (0)

/* Returns the number of queries whose stack would have overflowed (never expected). */
static inline __attribute__((always_inline)) int64_t or_closest_point_d(const int D, const float *nodes, const float *prims, const float *x, int64_t n,
                         or_out o) {
    const int NS = 2 * D + 3, PS = D == 2 ? 5 : 10;
    int64_t overflows = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : overflows)
    for (int64_t i = 0; i < n; ++i) {
        const float *p = x + D * i;
        float r_max = INFINITY;
        int idx = -1, si[OR_STACK], sp = 0, overflow = 0;
        float sd[OR_STACK];
        int visits = 0, child = 0, leaf = 0;
        si[0] = 0; sd[0] = r_max;
        while (sp >= 0 && !overflow) {
            int ni = si[sp];
            float cur = sd[sp];
            --sp;
            if (!(cur <= r_max)) continue;
            const float *nd = nodes + (int64_t)NS * ni;
            ++visits;
            int nref = (int)nd[2 * D + 2];
            if (nref > 0) {
                int off = (int)nd[2 * D];
                for (int j = 0; j < nref; ++j) {
                    const float *pr = prims + (int64_t)PS * (off + j);
                    float pt[3], uv[2], dd;
                    ++leaf;
                    if (D == 2) dd = cp_segment(2, p, pr, pr + 2, pt);
                    else dd = cp_triangle(p, pr, pr + 3, pr + 6, pt, uv);
                    if (dd < r_max) { r_max = dd; idx = off + j; }
                }
            } else {
                int li = ni + 1, ri = ni + (int)nd[2 * D + 1];
                const float *ln = nodes + (int64_t)NS * li, *rn = nodes + (int64_t)NS * ri;
                float dl, xl, dr, xr;
                child += 2;
                box_dist(D, ln, ln + D, p, &dl, &xl);
                int hl = dl <= r_max;
                r_max = fminf(r_max, xl);
                box_dist(D, rn, rn + D, p, &dr, &xr);
                int hr = dr <= r_max;
                r_max = fminf(r_max, xr);
                if (hl && hr) PUSH2(li, ri, dl, dr);
                else if (hl) PUSH1(li, dl);
                else if (hr) PUSH1(ri, dr);
            }
        }
        overflows += overflow;
        counters(&o, i, visits, child, leaf);
        if (idx < 0) { zero_out(D, &o, i, 1); continue; }
        const float *pr = prims + (int64_t)PS * idx;
        if (D == 2) {
            /* LineSegment.closest_point (line_segment.py:63-78) */
            float pa[2] = {p[0] - pr[0], p[1] - pr[1]}, ba[2] = {pr[2] - pr[0], pr[3] - pr[1]};
            float t[2] = {ba[0], ba[1]};
            normalizef(2, t);
            float h = clampf_(dot2f(pa, ba) / dot2f(ba, ba), 0.0f, 1.0f);
            float ip[2] = {pr[0] + ba[0] * h, pr[1] + ba[1] * h};
            float q[2] = {ip[0] - p[0], ip[1] - p[1]};
            if (o.p) { o.p[2 * i] = ip[0]; o.p[2 * i + 1] = ip[1]; }
            if (o.n) { o.n[2 * i] = t[1]; o.n[2 * i + 1] = -t[0]; }
            if (o.uv) o.uv[i] = h;
            if (o.d) o.d[i] = normf(2, q);
            if (o.prim_id) o.prim_id[i] = (int)pr[4];
        } else {
            /* Triangle.closest_point (triangle.py:75-81) */
            float pt[3], uv[2], e1[3], e2[3], nn[3];
            float dd = cp_triangle(p, pr, pr + 3, pr + 6, pt, uv);
            for (int k = 0; k < 3; ++k) { e1[k] = pr[3 + k] - pr[k]; e2[k] = pr[6 + k] - pr[k]; }
            cross3f(e1, e2, nn);
            normalizef(3, nn);
            if (o.p) memcpy(o.p + 3 * i, pt, 12);
            if (o.n) memcpy(o.n + 3 * i, nn, 12);
            if (o.uv) memcpy(o.uv + 2 * i, uv, 8);
            if (o.d) o.d[i] = dd;
            if (o.prim_id) o.prim_id[i] = (int)pr[9];
        }
        if (o.valid) o.valid[i] = 1;
    }
    return overflows;
}

int64_t or_closest_point(int D, const float *nodes, const float *prims, const float *x, int64_t n, or_out o) {
    return D == 2 ? or_closest_point_d(2, nodes, prims, x, n, o) : or_closest_point_d(3, nodes, prims, x, n, o);
}

static inline __attribute__((always_inline)) int64_t or_intersect_d(const int D, const float *nodes, const float *prims, const float *x, const float *vdir,
                     const float *r_max_arr, float r_max_scalar, int64_t n, or_out o) {
    const int NS = 2 * D + 3, PS = D == 2 ? 5 : 10;
    int64_t overflows = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : overflows)
    for (int64_t i = 0; i < n; ++i) {
        const float *px = x + D * i, *pv = vdir + D * i;
        float r_max = r_max_arr ? r_max_arr[i] : r_max_scalar;
        float d_min = r_max; /* 2D best hit; the 2D traversal never lowers r_max (bvh.py:206-240) */
        float troot;
        int idx = -1, overflow = 0, visits = 0, child = 0, leaf = 0;
        if (box_isect(D, nodes, nodes + D, px, pv, r_max, &troot)) {
            int si[OR_STACK], sp = 0;
            float sd[OR_STACK];
            si[0] = 0; sd[0] = troot;
            while (sp >= 0 && !overflow) {
                int ni = si[sp];
                float cur = sd[sp];
                --sp;
                if (!(cur <= r_max)) continue;
                const float *nd = nodes + (int64_t)NS * ni;
                ++visits;
                int nref = (int)nd[2 * D + 2];
                if (nref > 0) {
                    int off = (int)nd[2 * D];
                    for (int j = 0; j < nref; ++j) {
                        const float *pr = prims + (int64_t)PS * (off + j);
                        ++leaf;
                        if (D == 2) {
                            float dd = ray_seg(px, pv, pr, pr + 2);
                            if (dd < d_min) { d_min = dd; idx = off + j; }
                        } else {
                            float vv, ww;
                            float dd = ray_tri(px, pv, pr, pr + 3, pr + 6, &vv, &ww);
                            if (dd < r_max) { r_max = dd; idx = off + j; }
                        }
                    }
                } else {
                    int li = ni + 1, ri = ni + (int)nd[2 * D + 1];
                    const float *ln = nodes + (int64_t)NS * li, *rn = nodes + (int64_t)NS * ri;
                    float t0, t1;
                    child += 2;
                    int h0 = box_isect(D, ln, ln + D, px, pv, r_max, &t0);
                    int h1 = box_isect(D, rn, rn + D, px, pv, r_max, &t1);
                    if (h0 && h1) PUSH2(li, ri, t0, t1);
                    else if (h0) PUSH1(li, t0);
                    else if (h1) PUSH1(ri, t1);
                }
            }
        }
        overflows += overflow;
        counters(&o, i, visits, child, leaf);
        if (idx < 0) { zero_out(D, &o, i, 1); continue; }
        const float *pr = prims + (int64_t)PS * idx;
        if (D == 2) {
            /* bvh.py:267-274 */
            float t[2] = {pr[2] - pr[0], pr[3] - pr[1]};
            float nn[2] = {t[1], -t[0]};
            normalizef(2, nn);
            if (o.valid) o.valid[i] = 1;
            if (o.p) { o.p[2 * i] = px[0] + pv[0] * d_min; o.p[2 * i + 1] = px[1] + pv[1] * d_min; }
            if (o.n) { o.n[2 * i] = nn[0]; o.n[2 * i + 1] = nn[1]; }
            if (o.uv) o.uv[i] = -1.0f;
            if (o.d) o.d[i] = d_min;
            if (o.prim_id) o.prim_id[i] = (int)pr[4];
            if (o.on_boundary) o.on_boundary[i] = 1;
        } else {
            /* Triangle.ray_intersect (triangle.py:48-73) with r_max = the hit distance */
            float vv = 0, ww = 0;
            float t = ray_tri(px, pv, pr, pr + 3, pr + 6, &vv, &ww);
            if (!(t <= r_max)) { zero_out(3, &o, i, 1); continue; }
            float v1[3], v2[3], nn[3];
            for (int k = 0; k < 3; ++k) { v1[k] = pr[3 + k] - pr[k]; v2[k] = pr[6 + k] - pr[k]; }
            cross3f(v1, v2, nn);
            normalizef(3, nn);
            if (o.valid) o.valid[i] = 1;
            if (o.p) for (int k = 0; k < 3; ++k) o.p[3 * i + k] = pr[k] + v1[k] * vv + v2[k] * ww;
            if (o.n) memcpy(o.n + 3 * i, nn, 12);
            if (o.uv) { o.uv[2 * i] = 1.0f - vv - ww; o.uv[2 * i + 1] = vv; }
            if (o.d) o.d[i] = t;
            if (o.prim_id) o.prim_id[i] = (int)pr[9];
            if (o.on_boundary) o.on_boundary[i] = 1;
        }
    }
    return overflows;
}

int64_t or_intersect(int D, const float *nodes, const float *prims, const float *x, const float *vdir, const float *r_max_arr, float r_max_scalar, int64_t n, or_out o) {
    return D == 2 ? or_intersect_d(2, nodes, prims, x, vdir, r_max_arr, r_max_scalar, n, o) : or_intersect_d(3, nodes, prims, x, vdir, r_max_arr, r_max_scalar, n, o);
}

static inline __attribute__((always_inline)) int64_t or_silhouette_d(const int D, const float *snodes, const float *sils, const float *x,
                      const float *r_max_arr, float r_max_scalar, int64_t n, or_out o) {
    const int NS = D == 2 ? 13 : 16, SS = D == 2 ? 10 : 19;
    int64_t overflows = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : overflows)
    for (int64_t i = 0; i < n; ++i) {
        const float *px = x + D * i;
        float r_max = r_max_arr ? r_max_arr[i] : r_max_scalar;
        float droot, dmx, best_p[3] = {0, 0, 0}, best_d = 0.0f;
        int found = 0, best_id = 0, overflow = 0, visits = 0, child = 0, leaf = 0;
        box_dist(D, snodes, snodes + D, px, &droot, &dmx);
        if (droot <= r_max) {
            int si[OR_STACK], sp = 0;
            float sd[OR_STACK];
            si[0] = 0; sd[0] = droot;
            while (sp >= 0 && !overflow) {
                int ni = si[sp];
                float cur = sd[sp];
                --sp;
                if (!(cur <= r_max)) continue;
                const float *nd = snodes + (int64_t)NS * ni;
                ++visits;
                if ((int)nd[3 * D + 4] > 0) {
                    int off = (int)nd[3 * D + 5], cnt = (int)nd[3 * D + 6];
                    for (int j = 0; j < cnt; ++j) {
                        const float *sl = sils + (int64_t)SS * (off + j);
                        ++leaf;
                        if (D == 2) {
                            /* SilhouetteVertex.closest_silhouette (silhouette_vertex.py:39-53) */
                            const float *a = sl, *b = sl + 2, *c = sl + 4;
                            float xb[2] = {px[0] - b[0], px[1] - b[1]};
                            float dd = normf(2, xb);
                            if (((int)sl[6] != -1) && ((int)sl[8] != -1) && (dd < r_max)) {
                                float ba[2] = {b[0] - a[0], b[1] - a[1]}, xa[2] = {px[0] - a[0], px[1] - a[1]};
                                float cb[2] = {c[0] - b[0], c[1] - b[1]}, xb2[2] = {px[0] - b[0], px[1] - b[1]};
                                if (cross2f(ba, xa) * cross2f(cb, xb2) < 0.0f) {
                                    r_max = fminf(r_max, dd);
                                    found = 1; best_d = dd; best_id = (int)sl[9];
                                    best_p[0] = b[0]; best_p[1] = b[1];
                                }
                            }
                        } else {
                            /* SilhouetteEdge.closest_silhouette (silhouette_edge.py:49-64) */
                            const float *c = sl, *a = sl + 3, *b = sl + 6, *e = sl + 9;
                            float pt[3];
                            float dd = cp_segment(3, px, a, b, pt);
                            if (dd < r_max) {
                                int sil = ((int)sl[12] == -1) || ((int)sl[15] == -1);
                                if (!sil) {
                                    float ba[3], ca[3], ea[3], n0[3], n1[3], vd[3];
                                    for (int k = 0; k < 3; ++k) {
                                        ba[k] = b[k] - a[k]; ca[k] = c[k] - a[k]; ea[k] = e[k] - a[k];
                                        vd[k] = pt[k] - px[k];
                                    }
                                    cross3f(ba, ca, n0);
                                    normalizef(3, n0);
                                    cross3f(ea, ba, n1);
                                    normalizef(3, n1);
                                    sil = dot3f(vd, n0) * dot3f(vd, n1) < 0.0f;
                                }
                                if (sil) {
                                    r_max = fminf(r_max, dd);
                                    found = 1; best_d = dd; best_id = (int)sl[18];
                                    memcpy(best_p, pt, 12);
                                }
                            }
                        }
                    }
                } else {
                    int li = ni + 1, ri = ni + (int)nd[3 * D + 3];
                    float d0, d1;
                    child += 2;
                    int h0 = snch_overlap(D, snodes + (int64_t)NS * li, px, r_max, &d0);
                    int h1 = snch_overlap(D, snodes + (int64_t)NS * ri, px, r_max, &d1);
                    if (h0 && h1) PUSH2(li, ri, d0, d1);
                    else if (h0) PUSH1(li, d0);
                    else if (h1) PUSH1(ri, d1);
                }
            }
        }
        overflows += overflow;
        counters(&o, i, visits, child, leaf);
        if (!found) { zero_out(D, &o, i, 0); continue; }
        if (o.valid) o.valid[i] = 1;
        if (o.p) for (int k = 0; k < D; ++k) o.p[D * i + k] = best_p[k];
        if (o.d) o.d[i] = best_d;
        if (o.prim_id) o.prim_id[i] = best_id;
    }
    return overflows;
}

int64_t or_silhouette(int D, const float *snodes, const float *sils, const float *x, const float *r_max_arr, float r_max_scalar, int64_t n, or_out o) {
    return D == 2 ? or_silhouette_d(2, snodes, sils, x, r_max_arr, r_max_scalar, n, o) : or_silhouette_d(3, snodes, sils, x, r_max_arr, r_max_scalar, n, o);
}

/* ---- Part 2: FP64 brute force -------------------------------------------------------------- */

static inline double dotd(int D, const double *a, const double *b) {
    double s = 0.0;
    for (int k = 0; k < D; ++k) s += a[k] * b[k];
    return s;
}

/* Exact-arithmetic-style closest point on a segment in double. */
static double cp_segment_d(int D, const double *x, const double *a, const double *b, double *p) {
    double u[3], w[3], q[3];
    for (int k = 0; k < D; ++k) { u[k] = b[k] - a[k]; w[k] = x[k] - a[k]; }
    double uu = dotd(D, u, u);
    double t = uu > 0.0 ? dotd(D, w, u) / uu : 0.0;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    for (int k = 0; k < D; ++k) { p[k] = a[k] + u[k] * t; q[k] = p[k] - x[k]; }
    return sqrt(dotd(D, q, q));
}

/* Closest point on a triangle in double (Ericson region test, textbook form). */
static double cp_triangle_d(const double *p, const double *a, const double *b, const double *c, double *pt) {
    double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
    double d1 = dotd(3, ab, ap), d2 = dotd(3, ac, ap);
    if (d1 <= 0 && d2 <= 0) { memcpy(pt, a, 24); goto done; }
    for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
    double d3 = dotd(3, ab, bp), d4 = dotd(3, ac, bp);
    if (d3 >= 0 && d4 <= d3) { memcpy(pt, b, 24); goto done; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        for (int k = 0; k < 3; ++k) pt[k] = a[k] + v * ab[k];
        goto done;
    }
    for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
    double d5 = dotd(3, ab, cp), d6 = dotd(3, ac, cp);
    if (d6 >= 0 && d5 <= d6) { memcpy(pt, c, 24); goto done; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        for (int k = 0; k < 3; ++k) pt[k] = a[k] + w * ac[k];
        goto done;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int k = 0; k < 3; ++k) pt[k] = b[k] + w * (c[k] - b[k]);
        goto done;
    }
    {
        double denom = 1.0 / (va + vb + vc);
        double v = vb * denom, w = vc * denom;
        for (int k = 0; k < 3; ++k) pt[k] = a[k] + ab[k] * v + ac[k] * w;
    }
done:
    for (int k = 0; k < 3; ++k) q[k] = p[k] - pt[k];
    return sqrt(dotd(3, q, q));
}

static inline void load_d(int D, const float *V, int32_t idx, double *o) {
    for (int k = 0; k < D; ++k) o[k] = (double)V[(int64_t)D * idx + k];
}

/* Nearest primitive by brute force in double: best distance/point/id and the runner-up
 * distance over the other primitives. */
void or_bf_closest_point_f64(int D, const float *V, const int32_t *F, int64_t nf, const float *x, int64_t n,
                             double *best_d, double *best_p, int32_t *best_id, double *second_d) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        double p[3], a[3], b[3], c[3], pt[3], bp[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) p[k] = x[D * i + k];
        double b1 = INFINITY, b2 = INFINITY;
        int32_t id = -1;
        for (int64_t f = 0; f < nf; ++f) {
            double dd;
            load_d(D, V, F[D * f], a);
            load_d(D, V, F[D * f + 1], b);
            if (D == 2) dd = cp_segment_d(2, p, a, b, pt);
            else { load_d(3, V, F[3 * f + 2], c); dd = cp_triangle_d(p, a, b, c, pt); }
            if (dd < b1) { b2 = b1; b1 = dd; id = (int32_t)f; memcpy(bp, pt, sizeof bp); }
            else if (dd < b2) b2 = dd;
        }
        best_d[i] = b1; second_d[i] = b2; best_id[i] = id;
        for (int k = 0; k < D; ++k) best_p[D * i + k] = bp[k];
    }
}

/* First ray hit by brute force in double (exact Moller-Trumbore / 2D cross test, no epsilon):
 * hit distance, id and the second-smallest hit distance. */
void or_bf_intersect_f64(int D, const float *V, const int32_t *F, int64_t nf, const float *x, const float *vdir,
                         int64_t n, double *best_t, int32_t *best_id, double *second_t, double *best_uv) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        double o[3], v[3], a[3], b[3], c[3];
        for (int k = 0; k < D; ++k) { o[k] = x[D * i + k]; v[k] = vdir[D * i + k]; }
        double t1 = INFINITY, t2 = INFINITY, uv1[2] = {0.0, 0.0}, uv[2] = {0.0, 0.0};
        int32_t id = -1;
        for (int64_t f = 0; f < nf; ++f) {
            double t = INFINITY;
            load_d(D, V, F[D * f], a);
            load_d(D, V, F[D * f + 1], b);
            if (D == 2) {
                double u[2] = {b[0] - a[0], b[1] - a[1]}, w[2] = {o[0] - a[0], o[1] - a[1]};
                double den = v[0] * u[1] - v[1] * u[0];
                if (den != 0.0) {
                    double s = (v[0] * w[1] - v[1] * w[0]) / den;
                    double tt = (u[0] * w[1] - u[1] * w[0]) / den;
                    if (tt > 0 && s >= 0 && s <= 1) t = tt;
                }
            } else {
                load_d(3, V, F[3 * f + 2], c);
                double e1[3], e2[3], pv[3], s[3], q[3];
                for (int k = 0; k < 3; ++k) { e1[k] = b[k] - a[k]; e2[k] = c[k] - a[k]; s[k] = o[k] - a[k]; }
                pv[0] = v[1] * e2[2] - v[2] * e2[1];
                pv[1] = v[2] * e2[0] - v[0] * e2[2];
                pv[2] = v[0] * e2[1] - v[1] * e2[0];
                double det = dotd(3, e1, pv);
                if (det != 0.0) {
                    double inv = 1.0 / det;
                    double uu = dotd(3, s, pv) * inv;
                    if (uu >= 0 && uu <= 1) {
                        q[0] = s[1] * e1[2] - s[2] * e1[1];
                        q[1] = s[2] * e1[0] - s[0] * e1[2];
                        q[2] = s[0] * e1[1] - s[1] * e1[0];
                        double ww = dotd(3, v, q) * inv;
                        if (ww >= 0 && uu + ww <= 1) {
                            double tt = dotd(3, e2, q) * inv;
                            /* barycentrics of the hit as Triangle.ray_intersect reports them
                             * (triangle.py:69): (1 - v - w, v) */
                            if (tt >= 0) { t = tt; uv[0] = 1.0 - uu - ww; uv[1] = uu; }
                        }
                    }
                }
            }
            if (t < t1) { t2 = t1; t1 = t; id = (int32_t)f; uv1[0] = uv[0]; uv1[1] = uv[1]; }
            else if (t < t2) t2 = t;
        }
        best_t[i] = t1; best_id[i] = id; second_t[i] = t2;
        if (best_uv) { best_uv[2 * i] = uv1[0]; best_uv[2 * i + 1] = uv1[1]; }
    }
}

/* Nearest valid silhouette by brute force in double over the raw silhouette records (one per
 * vertex in 2D, per edge in 3D; flat layout of silhouette_vertex.h / silhouette_edge.h).  A 2D
 * vertex missing a neighbour is never a silhouette (silhouette_vertex.py:46); a 3D edge missing a
 * face always is (silhouette_edge.py:54). */
void or_bf_silhouette_f64(int D, const float *raw, int64_t ns, const float *x, int64_t n,
                          double *best_d, double *best_p, int32_t *best_id, double *second_d) {
    const int SS = D == 2 ? 10 : 19;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        double p[3], bp[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) p[k] = x[D * i + k];
        double b1 = INFINITY, b2 = INFINITY;
        int32_t id = -1;
        for (int64_t s = 0; s < ns; ++s) {
            const float *sl = raw + (int64_t)SS * s;
            double dd = INFINITY, pt[3];
            if (D == 2) {
                double a[2] = {sl[0], sl[1]}, b[2] = {sl[2], sl[3]}, c[2] = {sl[4], sl[5]};
                double c0 = (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]);
                double c1 = (c[0] - b[0]) * (p[1] - b[1]) - (c[1] - b[1]) * (p[0] - b[0]);
                if ((int)sl[6] != -1 && (int)sl[8] != -1 && c0 * c1 < 0.0) {
                    dd = sqrt((p[0] - b[0]) * (p[0] - b[0]) + (p[1] - b[1]) * (p[1] - b[1]));
                    pt[0] = b[0]; pt[1] = b[1];
                }
            } else {
                double c[3] = {sl[0], sl[1], sl[2]}, a[3] = {sl[3], sl[4], sl[5]},
                       b[3] = {sl[6], sl[7], sl[8]}, e[3] = {sl[9], sl[10], sl[11]};
                double dseg = cp_segment_d(3, p, a, b, pt);
                double ba[3], ca[3], ea[3], n0[3], n1[3], vd[3];
                for (int k = 0; k < 3; ++k) { ba[k] = b[k] - a[k]; ca[k] = c[k] - a[k]; ea[k] = e[k] - a[k]; vd[k] = pt[k] - p[k]; }
                n0[0] = ba[1] * ca[2] - ba[2] * ca[1]; n0[1] = ba[2] * ca[0] - ba[0] * ca[2]; n0[2] = ba[0] * ca[1] - ba[1] * ca[0];
                n1[0] = ea[1] * ba[2] - ea[2] * ba[1]; n1[1] = ea[2] * ba[0] - ea[0] * ba[2]; n1[2] = ea[0] * ba[1] - ea[1] * ba[0];
                if ((int)sl[12] == -1 || (int)sl[15] == -1 || dotd(3, vd, n0) * dotd(3, vd, n1) < 0.0) dd = dseg;
            }
            if (dd < b1) { b2 = b1; b1 = dd; id = (int32_t)(int)sl[SS - 1]; memcpy(bp, pt, sizeof bp); }
            else if (dd < b2) b2 = dd;
        }
        best_d[i] = b1; second_d[i] = b2; best_id[i] = id;
        for (int k = 0; k < D; ++k) best_p[D * i + k] = bp[k];
    }
}

/* ---- Part 2b (FP32): brute force with the traversal's own leaf test ----------------------
 * The smallest float32 leaf distance over every primitive of the reference flat array (stride
 * 5 segments / 10 triangles, the leaf tests of bvh.py:156-160 / bvh3d.py:123-128, math.py:14-21 /
 * math.py:42-107).  A traversal whose pruning never drops a primitive the float32 leaf test would
 * accept returns exactly this distance, bit for bit, at any coordinate scale. */
void or_bf_closest_point_f32(int D, const float *prims, int64_t np, const float *x, int64_t n, float *best_d,
                             int32_t *best_id) {
    const int PS = D == 2 ? 5 : 10;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        const float *p = x + D * i;
        float b = INFINITY, pt[3], uv[2];
        int32_t id = -1;
        for (int64_t j = 0; j < np; ++j) {
            const float *pr = prims + PS * j;
            const float dd = D == 3 ? cp_triangle(p, pr, pr + 3, pr + 6, pt, uv) : cp_segment(2, p, pr, pr + 2, pt);
            if (dd < b) { b = dd; id = (int32_t)pr[PS - 1]; }
        }
        best_d[i] = b;
        best_id[i] = id;
    }
}

/* The float32 leaf distance from query i to the primitive in row rows[i] of the flat array. */
void or_leaf_distance(int D, const float *prims, const int64_t *rows, const float *x, int64_t n, float *d) {
    const int PS = D == 2 ? 5 : 10;
    for (int64_t i = 0; i < n; ++i) {
        const float *pr = prims + PS * rows[i];
        float pt[3], uv[2];
        d[i] = D == 3 ? cp_triangle(x + 3 * i, pr, pr + 3, pr + 6, pt, uv) : cp_segment(2, x + 2 * i, pr, pr + 2, pt);
    }
}

/* The float32 leaf distance of n (point, triangle) pairs (closest_point_triangle, math.py:42-107):
 * p[3n], tri[9n] (a, b, c).  For checks of lower bounds against what the leaf test computes. */
void or_cp_triangle_pairs(const float *p, const float *tri, int64_t n, float *d) {
    for (int64_t i = 0; i < n; ++i) {
        float pt[3], uv[2];
        d[i] = cp_triangle(p + 3 * i, tri + 9 * i, tri + 9 * i + 3, tri + 9 * i + 6, pt, uv);
    }
}

/* ---- Part 3 (FP32): brute-force baselines and mesh normals --------------------------------
 *   or_normals        Mesh.configure_normal            gquery/shapes/mesh.py:35-50
 *   or_cp_baseline    Polyline/Mesh.closest_point_baseline  polyline.py:176-209, mesh.py:92-118
 *   or_ray_baseline   Polyline/Mesh.intersect_baseline (and .intersect's boundary offset)
 *                                                      polyline.py:125-170, mesh.py:53-85
 *   or_sil_baseline   Polyline/Mesh.closest_silhouette_baseline  polyline.py:212-223, mesh.py:124-136
 * The loops visit primitives in index order and keep the first strictly nearer one. */

void or_normals(const float *V, int64_t nv, const int32_t *F, int64_t nf, float *fn, float *vn) {
    for (int64_t f = 0; f < nf; ++f) {
        const float *a = V + 3 * (int64_t)F[3 * f], *b = V + 3 * (int64_t)F[3 * f + 1], *c = V + 3 * (int64_t)F[3 * f + 2];
        float u[3], w[3];
        for (int k = 0; k < 3; ++k) { u[k] = b[k] - a[k]; w[k] = c[k] - a[k]; }
        cross3f(u, w, fn + 3 * f);
        normalizef(3, fn + 3 * f);
    }
    /* dr.scatter_reduce(Add) over indices.x, then .y, then .z (mesh.py:44-49), into zeros */
    memset(vn, 0, (size_t)nv * 3 * sizeof(float));
    for (int slot = 0; slot < 3; ++slot)
        for (int64_t f = 0; f < nf; ++f)
            for (int k = 0; k < 3; ++k) vn[3 * (int64_t)F[3 * f + slot] + k] = vn[3 * (int64_t)F[3 * f + slot] + k] + fn[3 * f + k];
    for (int64_t v = 0; v < nv; ++v) normalizef(3, vn + 3 * v);
}

void or_cp_baseline(int D, const float *V, const int32_t *F, int64_t nf, const float *fn, const float *x,
                    int64_t n, or_out o) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float *p = x + D * i;
        float d_min = INFINITY, pt[3], uv[2];
        int idx = -1;
        for (int64_t j = 0; j < nf; ++j) {
            float d;
            if (D == 2) {
                const float *a = V + 2 * (int64_t)F[2 * j], *b = V + 2 * (int64_t)F[2 * j + 1];
                float pa[2] = {p[0] - a[0], p[1] - a[1]}, ba[2] = {b[0] - a[0], b[1] - a[1]};
                float h = clampf_(dot2f(pa, ba) / dot2f(ba, ba), 0.0f, 1.0f);
                float q[2] = {pa[0] - ba[0] * h, pa[1] - ba[1] * h};
                d = normf(2, q);
            } else {
                d = cp_triangle(p, V + 3 * (int64_t)F[3 * j], V + 3 * (int64_t)F[3 * j + 1], V + 3 * (int64_t)F[3 * j + 2], pt, uv);
            }
            if (d < d_min) { d_min = d; idx = (int)j; }
        }
        if (idx < 0) { zero_out(D, &o, i, 1); continue; }
        if (D == 2) {
            const float *a = V + 2 * (int64_t)F[2 * idx], *b = V + 2 * (int64_t)F[2 * idx + 1];
            float pa[2] = {p[0] - a[0], p[1] - a[1]}, ba[2] = {b[0] - a[0], b[1] - a[1]};
            float h = clampf_(dot2f(pa, ba) / dot2f(ba, ba), 0.0f, 1.0f);
            float nn[2] = {ba[1], -ba[0]};
            normalizef(2, nn);
            if (o.p) for (int k = 0; k < 2; ++k) o.p[2 * i + k] = fmaf(b[k], h, fmaf(-a[k], h, a[k]));
            if (o.n) { o.n[2 * i] = nn[0]; o.n[2 * i + 1] = nn[1]; }
            if (o.uv) o.uv[i] = h;
        } else {
            cp_triangle(p, V + 3 * (int64_t)F[3 * idx], V + 3 * (int64_t)F[3 * idx + 1], V + 3 * (int64_t)F[3 * idx + 2], pt, uv);
            if (o.p) memcpy(o.p + 3 * i, pt, 12);
            if (o.n) memcpy(o.n + 3 * i, fn + 3 * (int64_t)idx, 12);
            if (o.uv) memcpy(o.uv + 2 * i, uv, 8);
        }
        if (o.d) o.d[i] = d_min;
        if (o.prim_id) o.prim_id[i] = idx;
        if (o.valid) o.valid[i] = 1;
    }
}

void or_ray_baseline(int D, const float *V, const int32_t *F, int64_t nf, const float *x, const float *vdir,
                     const float *normal, const uint8_t *on_boundary, const float *r_max_arr, float r_max_scalar,
                     int64_t n, or_out o) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        float p[3] = {0, 0, 0};
        const float *v = vdir + D * i;
        for (int k = 0; k < D; ++k) p[k] = x[D * i + k];
        if (on_boundary && on_boundary[i])   /* p - 1e-5 * n (polyline.py:129, mesh.py:57) */
            for (int k = 0; k < D; ++k) p[k] = p[k] - normal[D * i + k] * 1e-5f;
        const float r_max = r_max_arr ? r_max_arr[i] : r_max_scalar;
        float d_min = r_max, vv = 0, ww = 0;
        int idx = -1;
        for (int64_t j = 0; j < nf; ++j) {
            float d;
            if (D == 2) d = ray_seg(p, v, V + 2 * (int64_t)F[2 * j], V + 2 * (int64_t)F[2 * j + 1]);
            else d = ray_tri(p, v, V + 3 * (int64_t)F[3 * j], V + 3 * (int64_t)F[3 * j + 1], V + 3 * (int64_t)F[3 * j + 2], &vv, &ww);
            if (d < d_min) { d_min = d; idx = (int)j; }
        }
        if (idx < 0) { zero_out(D, &o, i, 1); continue; }
        if (D == 2) {
            const float *a = V + 2 * (int64_t)F[2 * idx], *b = V + 2 * (int64_t)F[2 * idx + 1];
            float ab[2] = {b[0] - a[0], b[1] - a[1]};
            normalizef(2, ab);
            if (o.p) for (int k = 0; k < 2; ++k) o.p[2 * i + k] = p[k] + v[k] * d_min;
            if (o.n) { o.n[2 * i] = ab[1]; o.n[2 * i + 1] = -ab[0]; }
            if (o.uv) o.uv[i] = -1.0f;
            if (o.d) o.d[i] = d_min;
        } else {
            const float *a = V + 3 * (int64_t)F[3 * idx], *b = V + 3 * (int64_t)F[3 * idx + 1], *c = V + 3 * (int64_t)F[3 * idx + 2];
            float t = ray_tri(p, v, a, b, c, &vv, &ww), v1[3], v2[3], nn[3];
            for (int k = 0; k < 3; ++k) { v1[k] = b[k] - a[k]; v2[k] = c[k] - a[k]; }
            cross3f(v1, v2, nn);
            normalizef(3, nn);
            if (o.p) for (int k = 0; k < 3; ++k) o.p[3 * i + k] = a[k] + v1[k] * vv + v2[k] * ww;
            if (o.n) memcpy(o.n + 3 * i, nn, 12);
            if (o.uv) { o.uv[2 * i] = 1.0f - vv - ww; o.uv[2 * i + 1] = vv; }
            if (o.d) o.d[i] = t;
        }
        if (o.prim_id) o.prim_id[i] = idx;
        if (o.on_boundary) o.on_boundary[i] = 1;
        if (o.valid) o.valid[i] = 1;
    }
}

void or_sil_baseline(int D, const float *sils, int64_t ns, const float *x, const float *r_max_arr, float r_max_scalar,
                     int64_t n, or_out o) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float *p = x + D * i;
        const float r_max = r_max_arr ? r_max_arr[i] : r_max_scalar;
        float d_min = r_max, best[3] = {0, 0, 0};
        int id = -1;
        for (int64_t j = 0; j < ns; ++j) {
            if (D == 2) {
                /* SilhouetteVertex.closest_silhouette (silhouette_vertex.py:39-53) */
                const float *r = sils + 10 * j, *a = r, *b = r + 2, *c = r + 4;
                float q[2] = {p[0] - b[0], p[1] - b[1]};
                float d = normf(2, q);
                if (((int)r[6] != -1) && ((int)r[8] != -1) && d < r_max) {
                    float ba[2] = {b[0] - a[0], b[1] - a[1]}, xa[2] = {p[0] - a[0], p[1] - a[1]};
                    float cb[2] = {c[0] - b[0], c[1] - b[1]}, xb[2] = {p[0] - b[0], p[1] - b[1]};
                    if (cross2f(ba, xa) * cross2f(cb, xb) < 0.0f && d < d_min) {
                        d_min = d; best[0] = b[0]; best[1] = b[1]; id = (int)r[9];
                    }
                }
            } else {
                /* SilhouetteEdge.closest_silhouette (silhouette_edge.py:49-64), flat edge records */
                const float *r = sils + 19 * j, *c = r, *a = r + 3, *b = r + 6, *e = r + 9;
                float pt[3];
                float d = cp_segment(3, p, a, b, pt);
                if (d < r_max) {
                    int sil = ((int)r[12] == -1) || ((int)r[15] == -1);
                    if (!sil) {
                        float u[3], w[3], n0[3], n1[3], vd[3];
                        for (int k = 0; k < 3; ++k) { u[k] = b[k] - a[k]; w[k] = c[k] - a[k]; }
                        cross3f(u, w, n0); normalizef(3, n0);
                        for (int k = 0; k < 3; ++k) { u[k] = e[k] - a[k]; w[k] = b[k] - a[k]; }
                        cross3f(u, w, n1); normalizef(3, n1);
                        for (int k = 0; k < 3; ++k) vd[k] = pt[k] - p[k];
                        sil = dot3f(vd, n0) * dot3f(vd, n1) < 0.0f;
                    }
                    if (sil && d < d_min) { d_min = d; memcpy(best, pt, 12); id = (int)r[18]; }
                }
            }
        }
        if (id < 0) { zero_out(D, &o, i, 0); continue; }
        if (o.valid) o.valid[i] = 1;
        if (o.p) for (int k = 0; k < D; ++k) o.p[D * i + k] = best[k];
        if (o.d) o.d[i] = d_min;
        if (o.prim_id) o.prim_id[i] = id;
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- Part 4: BVH-guided boundary sampling ---------------------------------------------------
 *
 * or_sample_boundary follows BVH3D.sample_boundary / _sample_boundary (gquery/shapes/bvh3d.py:
 * 246-346): one root-to-leaf walk per query.  At an interior node each child whose box overlaps
 * the ball (x, R) gets weight 1 (branch_traversal_weight is 1, bvh3d.py:240-243); one uniform
 * number picks a child with probability weight / total, which multiplies the pmf, and r_max
 * becomes the chosen child's farthest-corner distance.  At the leaf, every primitive that is
 * "visible" (r_max <= R: the whole leaf box lies in the ball, or the primitive meets the ball,
 * Triangle.sphere_intersect triangle.py:44-47) joins a one-pass weighted reservoir with weight =
 * surface area (triangle.py:23-25), drawing one number per visible primitive; the pmf gains
 * area / total.  Then Triangle.sample_point (triangle.py:27-42) draws two numbers:
 * u1 = sqrt(r0), u = 1 - u1, v = r1 * u1, w = 1 - u - v, p = a*u + b*v + c*w, pdf = 2 / |cross|.
 * Record: p, n (unit face normal), uv = (u, v), pdf = pmf * pdf_point, pmf, prim_id = original
 * index; a query whose root box misses the ball, or with no visible primitive, gets hit = 0 and
 * an all-zero record.
 *
 * Where the reference cannot run as written, the meaning it states is taken:
 *   - `prim.sorted_index` (bvh3d.py:303; no such field on Triangle) is the leaf-order slot of the
 *     primitive, reference_offset + j, which is what self.primitives[...] indexes;
 *   - `type=prim.type` (bvh3d.py:265) is dropped: BoundarySamplingRecord3D has no such field;
 *   - a leaf whose visible primitives all have zero area yields no sample (the reference would
 *     gather primitive -1).
 * 2D (stride-7 nodes, stride-5 segments) is the same walk with LineSegment.surface_area =
 * |b - a|, sphere_intersect by closest_point_line_segment and LineSegment.sample_point
 * (line_segment.py:23-36: t = r0, p = a + (b - a) t, pdf = 1 / |b - a|, n = normal()); the
 * reference has these pieces and BoundarySamplingRecord (primitive.py:51-58) but no 2D walk.
 *
 * Random numbers: DrJit's PCG32 (drjit.auto.ad.PCG32, gquery/core/fwd.py:5; DrJit is neither
 * pinned by the reference nor installed here).  Its generator is O'Neill's pcg32 (XSH RR 64/32):
 * state' = state * 6364136223846793005 + inc, output = rotr32(((old >> 18) ^ old) >> 27,
 * old >> 59); next_float32 = float bits ((out >> 9) | 0x3f800000) - 1; seed(initstate, initseq):
 * inc = (initseq << 1) | 1, state = 0, step, state += initstate, step.  The per-lane stream of a
 * PCG32 of size n is initseq + lane.  Each query carries its own (state, inc); the state is
 * advanced in place, as the DrJit sampler object is.
 *
 * Pins (tests/test_sampling_oracle.py): the generator step against the published pcg32
 * known-answer vector; the walk's pmf against float64 enumeration of every branch; frequencies
 * (chi-square); unbiasedness of 1/pdf for the FP64 brute-force area; point placement and density.
 * Parity unpinned: the per-lane seeding convention (stream initseq + lane) -- DrJit is absent and
 * no reference test or fixture fixes it; callers passing their own states do not depend on it.
 */
static inline uint32_t pcg32_next(uint64_t *state, uint64_t inc) {
    const uint64_t old = *state;
    *state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static inline float pcg32_float(uint64_t *state, uint64_t inc) {
    const uint32_t b = (pcg32_next(state, inc) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &b, 4);
    return f - 1.0f;
}

void or_pcg32_seed(int64_t n, uint64_t initstate, uint64_t initseq, uint64_t *state, uint64_t *inc) {
    for (int64_t i = 0; i < n; ++i) {
        inc[i] = ((initseq + (uint64_t)i) << 1u) | 1u;
        state[i] = 0;
        pcg32_next(&state[i], inc[i]);
        state[i] += initstate;
        pcg32_next(&state[i], inc[i]);
    }
}

/* k outputs of next_uint32 per lane, row-major (lane, k); states advanced in place */
void or_pcg32_stream(uint64_t *state, const uint64_t *inc, int64_t n, int64_t k, uint32_t *out) {
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < k; ++j) out[i * k + j] = pcg32_next(&state[i], inc[i]);
}

typedef struct { void *hit, *p, *n, *uv, *pdf, *pmf, *prim_id; } or_sample_out;

static inline float prim_area(int D, const float *pr) {
    if (D == 3) {
        float ab[3], ac[3], c[3];
        for (int k = 0; k < 3; ++k) { ab[k] = pr[3 + k] - pr[k]; ac[k] = pr[6 + k] - pr[k]; }
        cross3f(ab, ac, c);
        return normf(3, c) / 2.0f;
    }
    float ba[2] = {pr[2] - pr[0], pr[3] - pr[1]};
    return normf(2, ba);
}

static inline int prim_in_ball(int D, const float *pr, const float *x, float R) {
    float pt[3], uv[2];
    const float d = D == 3 ? cp_triangle(x, pr, pr + 3, pr + 6, pt, uv) : cp_segment(2, x, pr, pr + 2, pt);
    return d <= R;
}

void or_sample_boundary(int D, const float *nodes, const float *prims, const float *x, const float *R_arr,
                        float R_scalar, uint64_t *state, const uint64_t *inc, int64_t n, or_sample_out o) {
    const int NS = D == 3 ? 9 : 7, PS = D == 3 ? 10 : 5;
    uint8_t *hit_o = (uint8_t *)o.hit;
    float *p_o = (float *)o.p, *n_o = (float *)o.n, *uv_o = (float *)o.uv, *pdf_o = (float *)o.pdf,
          *pmf_o = (float *)o.pmf;
    int32_t *id_o = (int32_t *)o.prim_id;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        const float *xi = x + D * i;
        const float R = R_arr ? R_arr[i] : R_scalar;
        uint64_t st = state[i];
        const uint64_t in = inc[i];
        int hit = 0, slot = -1;
        float pdf = 1.0f, dmin, dmax;
        box_dist(D, nodes, nodes + D, xi, &dmin, &dmax);
        if (dmin <= R) {
            float r_max = INFINITY;
            int64_t node = 0;
            for (;;) {
                const float *nd = nodes + (int64_t)NS * node;
                const int n_refs = (int)nd[2 * D + 2];
                if (n_refs > 0) {
                    const int off = (int)nd[2 * D];
                    float total = 0.0f;
                    for (int j = 0; j < n_refs; ++j) {
                        const float *pr = prims + (int64_t)PS * (off + j);
                        const float area = prim_area(D, pr);
                        if ((r_max <= R) || prim_in_ball(D, pr, xi, R)) {
                            hit = 1;
                            total += area;
                            const float sel = area / total;
                            const float u = pcg32_float(&st, in);
                            if (u < sel) slot = off + j;
                        }
                    }
                    if (total > 0.0f) pdf *= prim_area(D, prims + (int64_t)PS * slot) / total;
                    break;
                }
                const int64_t left = node + 1, right = node + (int64_t)(int)nd[2 * D + 1];
                float dmin0, dmax0, dmin1, dmax1;
                const float *b0 = nodes + (int64_t)NS * left, *b1 = nodes + (int64_t)NS * right;
                box_dist(D, b0, b0 + D, xi, &dmin0, &dmax0);
                box_dist(D, b1, b1 + D, xi, &dmin1, &dmax1);
                const float w0 = dmin0 <= R ? 1.0f : 0.0f, w1 = dmin1 <= R ? 1.0f : 0.0f;
                const float tw = w0 + w1;
                if (!(tw > 0.0f)) break;
                const float p0 = w0 / tw, p1 = 1.0f - p0;
                const float u = pcg32_float(&st, in);
                if (u < p0) { node = left; pdf *= p0; r_max = dmax0; }
                else { node = right; pdf *= p1; r_max = dmax1; }
            }
        }
        if (slot < 0) hit = 0;
        if (hit_o) hit_o[i] = (uint8_t)hit;
        float P[3] = {0, 0, 0}, N[3] = {0, 0, 0}, UV[2] = {0, 0}, pdf_rec = 0.0f, pmf_rec = 0.0f;
        int32_t id = 0;
        if (hit) {
            const float *pr = prims + (int64_t)PS * slot;
            if (D == 3) {
                const float u1 = sqrtf(pcg32_float(&st, in));
                const float u2 = pcg32_float(&st, in);
                const float u = 1.0f - u1, v = u2 * u1, w = 1.0f - u - v;
                float ab[3], ac[3], c[3];
                for (int k = 0; k < 3; ++k) {
                    P[k] = pr[k] * u + pr[3 + k] * v + pr[6 + k] * w;
                    ab[k] = pr[3 + k] - pr[k];
                    ac[k] = pr[6 + k] - pr[k];
                }
                cross3f(ab, ac, c);
                const float pdf_pt = 2.0f / normf(3, c);
                memcpy(N, c, 12);
                normalizef(3, N);
                UV[0] = u;
                UV[1] = v;
                pdf_rec = pdf * pdf_pt;
                id = (int32_t)pr[9];
            } else {
                const float t = pcg32_float(&st, in);
                const float ba[2] = {pr[2] - pr[0], pr[3] - pr[1]};
                P[0] = pr[0] + ba[0] * t;
                P[1] = pr[1] + ba[1] * t;
                const float pdf_pt = 1.0f / normf(2, ba);
                N[0] = ba[1];
                N[1] = -ba[0];
                normalizef(2, N);
                UV[0] = t;
                pdf_rec = pdf * pdf_pt;
                id = (int32_t)pr[4];
            }
            pmf_rec = pdf;
        }
        if (p_o) for (int k = 0; k < D; ++k) p_o[D * i + k] = P[k];
        if (n_o) for (int k = 0; k < D; ++k) n_o[D * i + k] = N[k];
        if (uv_o) for (int k = 0; k < D - 1; ++k) uv_o[(D - 1) * i + k] = UV[k];
        if (pdf_o) pdf_o[i] = pdf_rec;
        if (pmf_o) pmf_o[i] = pmf_rec;
        if (id_o) id_o[i] = id;
        state[i] = st;
    }
}
