// Brute-force ("baseline") queries and mesh normals: the reference's *_baseline methods, which test
// every primitive or silhouette of the shape for every query, and Mesh.configure_normal.
//   Polyline.closest_point_baseline / intersect_baseline / closest_silhouette_baseline
//     (gquery/shapes/polyline.py:125-223)
//   Mesh.closest_point_baseline / intersect_baseline / closest_silhouette_baseline, configure_normal
//     (gquery/shapes/mesh.py:35-50, 53-136)
// One lane per query; the primitive index is wave-uniform, so every lane reads the same record
// and the loads are broadcast.  Float32 arithmetic in the reference's order (compiled with
// -ffp-contract=off), strict `<` over increasing primitive index: the first nearest primitive wins.
#include "gq_device.h"
#include "gq_internal.h"

namespace gqa {
namespace {

using namespace gqd;

constexpr int kBlockB = 256;

__device__ __forceinline__ int f2i(float f) { return (int)f; }

template <int D>
__device__ __forceinline__ void zero_rec(const gq_records &o, int64_t i, bool cp_or_ray) {
    if (o.valid) o.valid[i] = 0;
    if (o.p) for (int k = 0; k < D; ++k) o.p[D * i + k] = 0.0f;
    if (cp_or_ray) {
        if (o.n) for (int k = 0; k < D; ++k) o.n[D * i + k] = 0.0f;
        if (o.uv) for (int k = 0; k < D - 1; ++k) o.uv[(D - 1) * i + k] = 0.0f;
        if (o.on_boundary) o.on_boundary[i] = 0;
    }
    if (o.d) o.d[i] = 0.0f;
    if (o.prim_id) o.prim_id[i] = 0;
}

// Mesh.configure_normal (mesh.py:35-50): face normals normalize(cross(b - a, c - a)), then the
// vertex normals normalize(sum of the incident face normals) in the reference's scatter order.
__global__ void __launch_bounds__(kBlockB) k_face_normals(ShapeView s, float *fn) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s.n_prims) return;
    const V3 a = ld3(s.V + 3 * (int64_t)s.F[3 * f]);
    const V3 b = ld3(s.V + 3 * (int64_t)s.F[3 * f + 1]);
    const V3 c = ld3(s.V + 3 * (int64_t)s.F[3 * f + 2]);
    st(fn + 3 * f, normalize(cross(b - a, c - a)));
}

__global__ void __launch_bounds__(kBlockB) k_vertex_normals(ShapeView s, const float *fn, float *vn) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= s.n_vertices) return;
    V3 sum = v3(0.0f, 0.0f, 0.0f);
    for (int32_t k = s.vf_start[v]; k < s.vf_start[v + 1]; ++k) sum = sum + ld3(fn + 3 * (int64_t)s.vf_list[k]);
    st(vn + 3 * v, normalize(sum));
}

// closest_point_baseline (polyline.py:176-209, mesh.py:92-118)
template <int D>
__global__ void __launch_bounds__(kBlockB) k_cp_baseline(ShapeView s, QueryArgs q) {
    using V = typename VT<D>::T;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < q.n;
    const V x = ldv<D>(q.x + D * (live ? i : 0));
    float d_min = __builtin_inff();
    int idx = -1;
    for (int64_t j = 0; j < s.n_prims; ++j) {
        const int32_t *f = s.F + D * j;
        float d;
        if constexpr (D == 2) {
            const V2 a = ld2(s.V + 2 * (int64_t)f[0]), b = ld2(s.V + 2 * (int64_t)f[1]);
            const V2 pa = x - a, ba = b - a;
            const float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
            d = norm(pa - ba * h);
        } else {
            const V3 a = ld3(s.V + 3 * (int64_t)f[0]), b = ld3(s.V + 3 * (int64_t)f[1]), c = ld3(s.V + 3 * (int64_t)f[2]);
            V3 pt;
            V2 uv;
            d = closest_point_triangle(x, a, b, c, pt, uv);
        }
        if (d < d_min) {
            d_min = d;
            idx = (int)j;
        }
    }
    if (!live) return;
    const gq_records &o = q.out;
    if (idx < 0) {
        zero_rec<D>(o, i, true);
        return;
    }
    const int32_t *f = s.F + D * (int64_t)idx;
    if constexpr (D == 2) {
        const V2 a = ld2(s.V + 2 * (int64_t)f[0]), b = ld2(s.V + 2 * (int64_t)f[1]);
        const V2 pa = x - a, ba = b - a;
        const float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        if (o.p) st(o.p + 2 * i, lerp(a, b, h));
        if (o.n) st(o.n + 2 * i, normalize(v2(ba.y, -ba.x)));
        if (o.uv) o.uv[i] = h;
    } else {
        const V3 a = ld3(s.V + 3 * (int64_t)f[0]), b = ld3(s.V + 3 * (int64_t)f[1]), c = ld3(s.V + 3 * (int64_t)f[2]);
        V3 pt;
        V2 uv;
        closest_point_triangle(x, a, b, c, pt, uv);
        if (o.p) st(o.p + 3 * i, pt);
        if (o.n) st(o.n + 3 * i, ld3(s.face_normals + 3 * (int64_t)idx));
        if (o.uv) st(o.uv + 2 * i, uv);
    }
    if (o.d) o.d[i] = d_min;
    if (o.prim_id) o.prim_id[i] = idx;
    if (o.valid) o.valid[i] = 1;
}

// intersect_baseline (polyline.py:138-170, mesh.py:61-85); Polyline/Mesh.intersect move origins on
// the boundary by -1e-5 n first (polyline.py:125-131, mesh.py:53-59).
template <int D>
__global__ void __launch_bounds__(kBlockB) k_ray_baseline(ShapeView s, QueryArgs q, const float *normal,
                                                          const uint8_t *on_boundary) {
    using V = typename VT<D>::T;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < q.n;
    const int64_t qi = live ? i : 0;
    V x = ldv<D>(q.x + D * qi);
    const V v = ldv<D>(q.v + D * qi);
    if (on_boundary && on_boundary[qi]) x = x - ldv<D>(normal + D * qi) * 1e-5f;
    const float r_max = q.r_max ? q.r_max[qi] : q.r_max_scalar;
    float d_min = r_max;
    int idx = -1;
    for (int64_t j = 0; j < s.n_prims; ++j) {
        const int32_t *f = s.F + D * j;
        float d;
        if constexpr (D == 2) {
            d = ray_segment_t(x, v, ld2(s.V + 2 * (int64_t)f[0]), ld2(s.V + 2 * (int64_t)f[1]));
        } else {
            // Triangle.ray_intersect accepts t <= r_max; t < d_min <= r_max implies it
            d = ray_triangle_t(x, v, ld3(s.V + 3 * (int64_t)f[0]), ld3(s.V + 3 * (int64_t)f[1]),
                               ld3(s.V + 3 * (int64_t)f[2]));
        }
        if (d < d_min) {
            d_min = d;
            idx = (int)j;
        }
    }
    if (!live) return;
    const gq_records &o = q.out;
    if (idx < 0) {
        zero_rec<D>(o, i, true);
        return;
    }
    const int32_t *f = s.F + D * (int64_t)idx;
    if constexpr (D == 2) {
        const V2 a = ld2(s.V + 2 * (int64_t)f[0]), b = ld2(s.V + 2 * (int64_t)f[1]);
        const V2 ab = normalize(b - a);
        if (o.p) st(o.p + 2 * i, x + v * d_min);
        if (o.n) st(o.n + 2 * i, v2(ab.y, -ab.x));
        if (o.uv) o.uv[i] = -1.0f;
        if (o.d) o.d[i] = d_min;
    } else {
        // Triangle.ray_intersect (triangle.py:48-73) of the winning triangle
        const V3 a = ld3(s.V + 3 * (int64_t)f[0]), b = ld3(s.V + 3 * (int64_t)f[1]), c = ld3(s.V + 3 * (int64_t)f[2]);
        const V3 v1 = b - a, v2_ = c - a;
        const V3 pp = cross(v, v2_);
        const float inv_det = 1.0f / dot(v1, pp);
        const V3 sv = x - a;
        const float vv = dot(sv, pp) * inv_det;
        const V3 qq = cross(sv, v1);
        const float ww = dot(v, qq) * inv_det;
        if (o.p) st(o.p + 3 * i, a + v1 * vv + v2_ * ww);
        if (o.n) st(o.n + 3 * i, normalize(cross(v1, v2_)));
        if (o.uv) st(o.uv + 2 * i, v2(1.0f - vv - ww, vv));
        if (o.d) o.d[i] = dot(v2_, qq) * inv_det;
    }
    if (o.prim_id) o.prim_id[i] = idx;
    if (o.on_boundary) o.on_boundary[i] = 1;
    if (o.valid) o.valid[i] = 1;
}

// closest_silhouette_baseline (polyline.py:212-223, mesh.py:124-136): every silhouette vertex
// (Polyline.configure_silhouette records, stride 10) or edge (build_flat_silhouette_edges records,
// stride 19) through SilhouetteVertex / SilhouetteEdge.closest_silhouette.
template <int D>
__global__ void __launch_bounds__(kBlockB) k_sil_baseline(ShapeView s, QueryArgs q) {
    using V = typename VT<D>::T;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < q.n;
    const int64_t qi = live ? i : 0;
    const V x = ldv<D>(q.x + D * qi);
    const float r_max = q.r_max ? q.r_max[qi] : q.r_max_scalar;
    float d_min = r_max;
    V best_p{};
    int best_id = -1;
    for (int64_t j = 0; j < s.n_sils; ++j) {
        if constexpr (D == 2) {
            // SilhouetteVertex.closest_silhouette (silhouette_vertex.py:39-53)
            const float *r = s.sils + 10 * j;
            const V2 a = ld2(r), b = ld2(r + 2), c = ld2(r + 4);
            const float d = norm(x - b);
            if ((f2i(r[6]) != -1) & (f2i(r[8]) != -1) & (d < r_max) & (d < d_min)) {
                if (cross2(b - a, x - a) * cross2(c - b, x - b) < 0.0f) {
                    d_min = d;
                    best_p = b;
                    best_id = f2i(r[9]);
                }
            }
        } else {
            // SilhouetteEdge.closest_silhouette (silhouette_edge.py:49-64)
            const float *r = s.sils + 19 * j;
            const V3 c = ld3(r), a = ld3(r + 3), b = ld3(r + 6), e = ld3(r + 9);
            V3 pt;
            const float d = closest_point_segment(x, a, b, pt);
            if ((d < r_max) & (d < d_min)) {
                bool sil = (f2i(r[12]) == -1) | (f2i(r[15]) == -1);
                if (!sil) {
                    const V3 n0 = normalize(cross(b - a, c - a));
                    const V3 n1 = normalize(cross(e - a, b - a));
                    const V3 vd = pt - x;
                    sil = dot(vd, n0) * dot(vd, n1) < 0.0f;
                }
                if (sil) {
                    d_min = d;
                    best_p = pt;
                    best_id = f2i(r[18]);
                }
            }
        }
    }
    if (!live) return;
    const gq_records &o = q.out;
    if (best_id < 0) {
        zero_rec<D>(o, i, false);
        return;
    }
    if (o.valid) o.valid[i] = 1;
    if (o.p) st(o.p + D * i, best_p);
    if (o.d) o.d[i] = d_min;
    if (o.prim_id) o.prim_id[i] = best_id;
}

template <int D>
hipError_t launch_baseline_d(QueryKind kind, const ShapeView &s, const QueryArgs &q, const float *normal,
                             const uint8_t *on_boundary, hipStream_t st) {
    const unsigned grid = (unsigned)((q.n + kBlockB - 1) / kBlockB);
    static const char *names[2][3] = {
        {"baseline_closest_point_2d", "baseline_intersect_2d", "baseline_silhouette_2d"},
        {"baseline_closest_point_3d", "baseline_intersect_3d", "baseline_silhouette_3d"}};
    prof_begin(st);
    if (kind == kClosestPoint) hipLaunchKernelGGL(k_cp_baseline<D>, dim3(grid), dim3(kBlockB), 0, st, s, q);
    else if (kind == kIntersect)
        hipLaunchKernelGGL(k_ray_baseline<D>, dim3(grid), dim3(kBlockB), 0, st, s, q, normal, on_boundary);
    else hipLaunchKernelGGL(k_sil_baseline<D>, dim3(grid), dim3(kBlockB), 0, st, s, q);
    prof_end(st, names[D - 2][kind]);
    return hipGetLastError();
}

} // namespace

hipError_t launch_normals(const ShapeView &s, float *face_normals, float *vertex_normals, hipStream_t st) {
    if (s.dim != 3 || s.n_prims == 0) return hipSuccess;
    hipLaunchKernelGGL(k_face_normals, dim3((unsigned)((s.n_prims + kBlockB - 1) / kBlockB)), dim3(kBlockB), 0, st,
                       s, face_normals);
    hipLaunchKernelGGL(k_vertex_normals, dim3((unsigned)((s.n_vertices + kBlockB - 1) / kBlockB)), dim3(kBlockB), 0,
                       st, s, (const float *)face_normals, vertex_normals);
    return hipGetLastError();
}

hipError_t launch_baseline(QueryKind kind, const ShapeView &s, const QueryArgs &q, const float *normal,
                           const uint8_t *on_boundary, hipStream_t st) {
    if (q.n <= 0) return hipSuccess;
    return s.dim == 2 ? launch_baseline_d<2>(kind, s, q, normal, on_boundary, st)
                      : launch_baseline_d<3>(kind, s, q, normal, on_boundary, st);
}

} // namespace gqa
