// Device-side vector math and leaf tests shared by the traversal kernels.
//
// Arithmetic forms follow the reference's DrJit expressions (gquery/core/math.py):
//   dot      = fma chain  a.x*b.x -> fma(a.y,b.y,.) -> fma(a.z,b.z,.)   (drjit dot)
//   cross    = fma(a.y, b.z, -(a.z*b.y)), ...                            (drjit cross / fmsub)
//   lerp     = fma(b, t, fma(-a, t, a))                                  (drjit lerp)
//   normalize= v * (1 / sqrt(dot(v, v)))                                 (drjit normalize / rsqrt)
//   clamp    = max(min(x, hi), lo)
// Every other expression is written in the order it appears in the reference Python.  Translation
// units that must reproduce the CPU restatement bit-for-bit are compiled with -ffp-contract=off;
// sqrt and division are the correctly rounded forms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gqd {

struct V2 { float x, y; };
struct V3 { float x, y, z; };

__device__ __forceinline__ V2 v2(float x, float y) { return {x, y}; }
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ V2 operator/(V2 a, float s) { return {a.x / s, a.y / s}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }

__device__ __forceinline__ float dot(V2 a, V2 b) { return fmaf(a.y, b.y, a.x * b.x); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ float norm(V2 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float norm(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V2 normalize(V2 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
// math.py:9-11 (explicit expression, no fma)
__device__ __forceinline__ float cross2(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
__device__ __forceinline__ V2 vabs(V2 a) { return {fabsf(a.x), fabsf(a.y)}; }
__device__ __forceinline__ V3 vabs(V3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ V2 vmin(V2 a, V2 b) { return {fminf(a.x, b.x), fminf(a.y, b.y)}; }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ V2 vmax(V2 a, V2 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ V2 vmax(V2 a, float s) { return {fmaxf(a.x, s), fmaxf(a.y, s)}; }
__device__ __forceinline__ V3 vmax(V3 a, float s) { return {fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }
__device__ __forceinline__ float hmax(V2 a) { return fmaxf(a.x, a.y); }
__device__ __forceinline__ float hmax(V3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
__device__ __forceinline__ float hmin(V2 a) { return fminf(a.x, a.y); }
__device__ __forceinline__ float hmin(V3 a) { return fminf(fminf(a.x, a.y), a.z); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fmaxf(fminf(x, hi), lo); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(b, t, fmaf(-a, t, a)); }
__device__ __forceinline__ V2 lerp(V2 a, V2 b, float t) { return {lerpf(a.x, b.x, t), lerpf(a.y, b.y, t)}; }
__device__ __forceinline__ V3 lerp(V3 a, V3 b, float t) {
    return {lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t)};
}

constexpr float kDrEps = 5.9604644775390625e-08f;  // dr.epsilon(Float32) = 2^-24
constexpr float kHalfPi = 1.57079637050628662109375f;  // float(pi / 2)

template <int D> struct VT;
template <> struct VT<2> { using T = V2; };
template <> struct VT<3> { using T = V3; };

__device__ __forceinline__ V2 ld2(const float *p) { return {p[0], p[1]}; }
__device__ __forceinline__ V3 ld3(const float *p) { return {p[0], p[1], p[2]}; }
template <int D> __device__ __forceinline__ typename VT<D>::T ldv(const float *p);
template <> __device__ __forceinline__ V2 ldv<2>(const float *p) { return ld2(p); }
template <> __device__ __forceinline__ V3 ldv<3>(const float *p) { return ld3(p); }
__device__ __forceinline__ void st(float *p, V2 a) { p[0] = a.x; p[1] = a.y; }
__device__ __forceinline__ void st(float *p, V3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

// BoundingBox.distance (bvh.py:22-27, bvh3d.py:19-24): d_min to the box, d_max to its farthest corner.
template <typename V>
__device__ __forceinline__ void box_distance(V lo, V hi, V p, float &d_min, float &d_max) {
    V u = lo - p;
    V w = p - hi;
    d_min = norm(vmax(vmax(u, w), 0.0f));
    d_max = norm(vmin(u, w));
}

// BoundingBox.intersect (bvh.py:33-50, bvh3d.py:30-47): slab test, returns hit and t_min.
template <typename V>
__device__ __forceinline__ bool box_intersect(V lo, V hi, V x, V v, float r_max, float &t_min) {
    V t0 = (lo - x);
    V t1 = (hi - x);
    if constexpr (sizeof(V) == sizeof(V2)) {
        t0 = {t0.x / v.x, t0.y / v.y};
        t1 = {t1.x / v.x, t1.y / v.y};
    } else {
        t0 = {t0.x / v.x, t0.y / v.y, t0.z / v.z};
        t1 = {t1.x / v.x, t1.y / v.y, t1.z / v.z};
    }
    V tn = vmin(t0, t1);
    V tf = vmax(t0, t1);
    float t_near_max = fmaxf(hmax(tn), 0.0f);
    float t_far_min = fminf(hmin(tf), r_max);
    if (t_near_max > t_far_min) {
        t_min = __builtin_inff();
        return false;
    }
    t_min = t_near_max;
    return true;
}

// closest_point_line_segment (math.py:14-21)
template <typename V>
__device__ __forceinline__ float closest_point_segment(V x, V a, V b, V &p) {
    V u = b - a;
    V w = x - a;
    float t = clampf(dot(w, u) / dot(u, u), 0.0f, 1.0f);
    p = lerp(a, b, t);
    return norm(p - x);
}

// closest_point_triangle (math.py:42-107): Ericson's region test, first matching region wins.
// Returns the distance; pt and uv = (bary_b, bary_c) as the reference returns them.
__device__ __forceinline__ float closest_point_triangle(V3 p, V3 a, V3 b, V3 c, V3 &pt, V2 &uvo) {
    V3 ab = b - a;
    V3 ac = c - a;
    V3 ax = p - a;
    float d1 = dot(ab, ax);
    float d2 = dot(ac, ax);
    V2 uv;
    do {
        if ((d1 <= 0.0f) & (d2 <= 0.0f)) { pt = a; uv = v2(1.0f, 0.0f); break; }
        V3 bx = p - b;
        float d3 = dot(ab, bx);
        float d4 = dot(ac, bx);
        if ((d3 >= 0.0f) & (d4 <= d3)) { pt = b; uv = v2(0.0f, 1.0f); break; }
        V3 cx = p - c;
        float d5 = dot(ab, cx);
        float d6 = dot(ac, cx);
        if ((d6 >= 0.0f) & (d5 <= d6)) { pt = c; uv = v2(0.0f, 0.0f); break; }
        float vc = d1 * d4 - d3 * d2;
        if ((vc <= 0.0f) & (d1 >= 0.0f) & (d3 <= 0.0f)) {
            float v = d1 / (d1 - d3);
            pt = a + ab * v;
            uv = v2(1.0f - v, v);
            break;
        }
        float vb = d5 * d2 - d1 * d6;
        if ((vb <= 0.0f) & (d2 >= 0.0f) & (d6 <= 0.0f)) {
            float w = d2 / (d2 - d6);
            pt = a + ac * w;
            uv = v2(1.0f - w, 0.0f);
            break;
        }
        float va = d3 * d6 - d5 * d4;
        if ((va <= 0.0f) & ((d4 - d3) >= 0.0f) & ((d5 - d6) >= 0.0f)) {
            float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            pt = b + (c - b) * w;
            uv = v2(0.0f, 1.0f - w);
            break;
        }
        float denom = 1.0f / (va + vb + vc);
        float v = vb * denom;
        float w = vc * denom;
        pt = a + ab * v + ac * w;
        uv = v2(1.0f - v - w, v);
    } while (0);
    uvo = v2(uv.y, 1.0f - uv.x - uv.y);
    return norm(p - pt);
}

// ray_triangle_intersect (math.py:110-128) with r_max = inf: t of the hit, +inf on a miss.
__device__ __forceinline__ float ray_triangle_t(V3 x, V3 d, V3 a, V3 b, V3 c) {
    V3 v1 = b - a;
    V3 v2_ = c - a;
    V3 p = cross(d, v2_);
    float det = dot(v1, p);
    float r = __builtin_inff();
    if (fabsf(det) > kDrEps) {
        float inv_det = 1.0f / det;
        V3 s = x - a;
        float v = dot(s, p) * inv_det;
        if ((v >= 0.0f) & (v <= 1.0f)) {
            V3 q = cross(s, v1);
            float w = dot(d, q) * inv_det;
            if ((w >= 0.0f) & (v + w <= 1.0f)) {
                float t = dot(v2_, q) * inv_det;
                if (t >= 0.0f) r = t;
            }
        }
    }
    return r;
}

// ray_intersection (math.py:31-39): 2D ray vs segment, t or +inf.
__device__ __forceinline__ float ray_segment_t(V2 x, V2 v, V2 a, V2 b) {
    V2 u = b - a;
    V2 w = x - a;
    float d = cross2(v, u);
    float s = cross2(v, w) / d;
    float t = cross2(u, w) / d;
    bool valid = (t > 0.0f) & (s >= 0.0f) & (s <= 1.0f);
    return valid ? t : __builtin_inff();
}

// project_to_plane (math.py:24-28)
__device__ __forceinline__ float project_to_plane(V2 n, V2 e) {
    V2 b = v2(-n.y, n.x);
    return fabsf(dot(e, vabs(b)));
}
// compute_orthonomal_basis + project_to_plane_3D (snch3d.py:11-30)
__device__ __forceinline__ float project_to_plane(V3 n, V3 e) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    V3 b1 = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    V3 b2 = v3(b, sign + n.y * n.y * a, -n.y);
    float r1 = dot(e, vabs(b1));
    float r2 = dot(e, vabs(b2));
    return sqrtf(r1 * r1 + r2 * r2);
}

// SNCHNode.overlaps / SNCHNode3D.overlaps (snch.py:15-60, snch3d.py:50-95): box test, then the
// normal-cone / view-cone test.  Returns overlap; d_min is the box distance.
template <typename V>
__device__ __forceinline__ bool snch_overlaps(V lo, V hi, V axis, float half, float radius, V x,
                                              float r_max, float &d_min) {
    float d_max;
    box_distance(lo, hi, x, d_min, d_max);
    bool ov = d_min <= r_max;
    if ((half >= 0.0f) & ov) {
        ov = (half > kHalfPi) & (d_min < kDrEps);
        if (!ov) {
            V c = (lo + hi) / 2.0f;
            V va = c - x;
            float l = norm(va);
            va = va / l;
            float ang = acosf(clampf(dot(axis, va), -1.0f, 1.0f));
            ov = (kHalfPi >= ang - half) & (kHalfPi <= ang + half);
            if (!ov) {
                if (l > radius) {
                    float vh = asinf(radius / l);
                    float hs = half + vh;
                    float mn = ang - hs, mx = ang + hs;
                    ov = (hs > kHalfPi) | ((kHalfPi >= mn) & (kHalfPi <= mx));
                } else {
                    V e = hi - c;
                    float dd = dot(e, vabs(va));
                    float s = l - dd;
                    ov = s < 0.0f;
                    if (!ov) {
                        dd = project_to_plane(va, e);
                        float vh = atan2f(dd, s);
                        float hs = half + vh;
                        float mn = ang - hs, mx = ang + hs;
                        ov = (hs > kHalfPi) | ((kHalfPi >= mn) & (kHalfPi <= mx));
                    }
                }
            }
        }
    }
    return ov;
}

} // namespace gqd
