// Plain FP32 traversal kernels: one lane per query, the reference's flat AoS float32 node and
// primitive records, a 64-entry private stack, the reference visit order and pruning rules.
// These are the faithful GPU restatement of gquery/shapes/{bvh,bvh3d,snch,snch3d}.py and the
// baseline the engine kernels are measured against.  Compiled with -ffp-contract=off.
#include "gq_device.h"
#include "gq_internal.h"

namespace gqa {
namespace {

using namespace gqd;

constexpr int kStack = 64;  // dr.alloc_local(TraversalStack, 64) (bvh.py:132)

template <int D> struct L;  // record strides of the reference layout
template <> struct L<2> { static constexpr int node = 7, prim = 5, snode = 13, sil = 10; };
template <> struct L<3> { static constexpr int node = 9, prim = 10, snode = 16, sil = 19; };

__device__ __forceinline__ int f2i(float f) { return (int)f; }

template <int D>
__device__ __forceinline__ void zero_record(const gq_records &o, int64_t i, bool cp_or_ray) {
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

// ---------------------------------------------------------------------------------------------
// closest point: BVH.closest_point (bvh.py:127-198), BVH3D.closest_point (bvh3d.py:108-174)
template <int D>
__global__ void __launch_bounds__(256) k_closest_point(PlainView s, QueryArgs q) {
    using V = typename VT<D>::T;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const V p = ldv<D>(q.x + D * i);
    float r_max = __builtin_inff();
    int idx = -1;
    int st_i[kStack];
    float st_d[kStack];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = r_max;
    while (sp >= 0) {
        const int ni = st_i[sp];
        const float cur = st_d[sp];
        --sp;
        if (cur <= r_max) {
            const float *nd = s.bvh_nodes + (int64_t)L<D>::node * ni;
            const int n_refs = f2i(nd[2 * D + 2]);
            if (n_refs > 0) {
                const int off = f2i(nd[2 * D]);
                for (int j = 0; j < n_refs; ++j) {
                    const float *pr = s.prims + (int64_t)L<D>::prim * (off + j);
                    float dd;
                    if constexpr (D == 2) {
                        V pt;
                        dd = closest_point_segment(p, ld2(pr), ld2(pr + 2), pt);
                    } else {
                        V3 pt;
                        V2 uv;
                        dd = closest_point_triangle(p, ld3(pr), ld3(pr + 3), ld3(pr + 6), pt, uv);
                    }
                    if (dd < r_max) {
                        r_max = dd;
                        idx = off + j;
                    }
                }
            } else {
                const int li = ni + 1;
                const int ri = ni + f2i(nd[2 * D + 1]);
                const float *ln = s.bvh_nodes + (int64_t)L<D>::node * li;
                const float *rn = s.bvh_nodes + (int64_t)L<D>::node * ri;
                float dminl, dmaxl, dminr, dmaxr;
                box_distance(ldv<D>(ln), ldv<D>(ln + D), p, dminl, dmaxl);
                const bool hl = dminl <= r_max;
                r_max = fminf(r_max, dmaxl);
                box_distance(ldv<D>(rn), ldv<D>(rn + D), p, dminr, dmaxr);
                const bool hr = dminr <= r_max;
                r_max = fminf(r_max, dmaxr);
                if (hl & hr) {
                    int closer = li, other = ri;
                    if (dminr < dminl) {
                        closer = ri; other = li;
                        float t = dminl; dminl = dminr; dminr = t;
                    }
                    if (sp + 2 < kStack) {
                        st_i[++sp] = other; st_d[sp] = dminr;
                        st_i[++sp] = closer; st_d[sp] = dminl;
                    }
                } else if (hl) {
                    if (sp + 1 < kStack) { st_i[++sp] = li; st_d[sp] = dminl; }
                } else if (hr) {
                    if (sp + 1 < kStack) { st_i[++sp] = ri; st_d[sp] = dminr; }
                }
            }
        }
    }
    const gq_records &o = q.out;
    if (idx < 0) {
        zero_record<D>(o, i, true);
        return;
    }
    const float *pr = s.prims + (int64_t)L<D>::prim * idx;
    if constexpr (D == 2) {
        // LineSegment.closest_point (line_segment.py:63-78)
        V2 a = ld2(pr), b = ld2(pr + 2);
        V2 pa = p - a, ba = b - a;
        V2 t = normalize(ba);
        V2 nn = v2(t.y, -t.x);
        float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        V2 ip = a + ba * h;
        float dd = norm(ip - p);
        if (o.p) st(o.p + 2 * i, ip);
        if (o.n) st(o.n + 2 * i, nn);
        if (o.uv) o.uv[i] = h;
        if (o.d) o.d[i] = dd;
        if (o.prim_id) o.prim_id[i] = f2i(pr[4]);
    } else {
        // Triangle.closest_point (triangle.py:75-81)
        V3 a = ld3(pr), b = ld3(pr + 3), c = ld3(pr + 6);
        V3 pt;
        V2 uv;
        float dd = closest_point_triangle(p, a, b, c, pt, uv);
        V3 nn = normalize(cross(b - a, c - a));
        if (o.p) st(o.p + 3 * i, pt);
        if (o.n) st(o.n + 3 * i, nn);
        if (o.uv) st(o.uv + 2 * i, uv);
        if (o.d) o.d[i] = dd;
        if (o.prim_id) o.prim_id[i] = f2i(pr[9]);
    }
    if (o.valid) o.valid[i] = 1;
}

// ---------------------------------------------------------------------------------------------
// ray: BVH.intersect (bvh.py:200-276), BVH3D.intersect (bvh3d.py:176-239)
template <int D>
__global__ void __launch_bounds__(256) k_intersect(PlainView s, QueryArgs q) {
    using V = typename VT<D>::T;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const V x = ldv<D>(q.x + D * i);
    const V v = ldv<D>(q.v + D * i);
    float r_max = q.r_max ? q.r_max[i] : q.r_max_scalar;
    const gq_records &o = q.out;
    float t_root;
    const bool root_hit = box_intersect(ldv<D>(s.bvh_nodes), ldv<D>(s.bvh_nodes + D), x, v, r_max, t_root);
    int idx = -1;
    float d_min = r_max;  // 2D: best hit distance; r_max stays fixed (bvh.py:206, Q4)
    if (root_hit) {
        int st_i[kStack];
        float st_d[kStack];
        int sp = 0;
        st_i[0] = 0;
        st_d[0] = t_root;
        while (sp >= 0) {
            const int ni = st_i[sp];
            const float cur = st_d[sp];
            --sp;
            if (cur <= r_max) {
                const float *nd = s.bvh_nodes + (int64_t)L<D>::node * ni;
                const int n_refs = f2i(nd[2 * D + 2]);
                if (n_refs > 0) {
                    const int off = f2i(nd[2 * D]);
                    for (int j = 0; j < n_refs; ++j) {
                        const float *pr = s.prims + (int64_t)L<D>::prim * (off + j);
                        if constexpr (D == 2) {
                            float dd = ray_segment_t(x, v, ld2(pr), ld2(pr + 2));
                            if (dd < d_min) { d_min = dd; idx = off + j; }
                        } else {
                            float dd = ray_triangle_t(x, v, ld3(pr), ld3(pr + 3), ld3(pr + 6));
                            if (dd < r_max) { r_max = dd; idx = off + j; }
                        }
                    }
                } else {
                    const int li = ni + 1;
                    const int ri = ni + f2i(nd[2 * D + 1]);
                    const float *ln = s.bvh_nodes + (int64_t)L<D>::node * li;
                    const float *rn = s.bvh_nodes + (int64_t)L<D>::node * ri;
                    float t0, t1;
                    const bool h0 = box_intersect(ldv<D>(ln), ldv<D>(ln + D), x, v, r_max, t0);
                    const bool h1 = box_intersect(ldv<D>(rn), ldv<D>(rn + D), x, v, r_max, t1);
                    if (h0 & h1) {
                        int closer = li, other = ri;
                        if (t1 < t0) {
                            closer = ri; other = li;
                            float t = t0; t0 = t1; t1 = t;
                        }
                        if (sp + 2 < kStack) {
                            st_i[++sp] = other; st_d[sp] = t1;
                            st_i[++sp] = closer; st_d[sp] = t0;
                        }
                    } else if (h0) {
                        if (sp + 1 < kStack) { st_i[++sp] = li; st_d[sp] = t0; }
                    } else if (h1) {
                        if (sp + 1 < kStack) { st_i[++sp] = ri; st_d[sp] = t1; }
                    }
                }
            }
        }
    }
    if (idx < 0) {
        zero_record<D>(o, i, true);
        return;
    }
    const float *pr = s.prims + (int64_t)L<D>::prim * idx;
    if constexpr (D == 2) {
        // record of bvh.py:267-274
        V2 a = ld2(pr), b = ld2(pr + 2);
        V2 t = b - a;
        V2 nn = normalize(v2(t.y, -t.x));
        if (o.valid) o.valid[i] = 1;
        if (o.p) st(o.p + 2 * i, x + v * d_min);
        if (o.n) st(o.n + 2 * i, nn);
        if (o.uv) o.uv[i] = -1.0f;
        if (o.d) o.d[i] = d_min;
        if (o.prim_id) o.prim_id[i] = f2i(pr[4]);
        if (o.on_boundary) o.on_boundary[i] = 1;
    } else {
        // Triangle.ray_intersect (triangle.py:48-73) with r_max = the hit distance
        V3 a = ld3(pr), b = ld3(pr + 3), c = ld3(pr + 6);
        V3 v1 = b - a, v2_ = c - a;
        V3 pp = cross(v, v2_);
        float det = dot(v1, pp);
        bool ok = false;
        float vv = 0.0f, ww = 0.0f, t = 0.0f;
        if (fabsf(det) > kDrEps) {
            float inv_det = 1.0f / det;
            V3 sv = x - a;
            vv = dot(sv, pp) * inv_det;
            if ((vv >= 0.0f) & (vv <= 1.0f)) {
                V3 qq = cross(sv, v1);
                ww = dot(v, qq) * inv_det;
                if ((ww >= 0.0f) & (vv + ww <= 1.0f)) {
                    t = dot(v2_, qq) * inv_det;
                    ok = (t >= 0.0f) & (t <= r_max);
                }
            }
        }
        if (!ok) {
            zero_record<D>(o, i, true);
            return;
        }
        if (o.valid) o.valid[i] = 1;
        if (o.p) st(o.p + 3 * i, a + v1 * vv + v2_ * ww);
        if (o.n) st(o.n + 3 * i, normalize(cross(v1, v2_)));
        if (o.uv) st(o.uv + 2 * i, v2(1.0f - vv - ww, vv));
        if (o.d) o.d[i] = t;
        if (o.prim_id) o.prim_id[i] = f2i(pr[9]);
        if (o.on_boundary) o.on_boundary[i] = 1;
    }
}

// ---------------------------------------------------------------------------------------------
// silhouette: SNCH.closest_silhouette (snch.py:106-164), SNCH3D.closest_silhouette (snch3d.py:143-203)
template <int D>
__global__ void __launch_bounds__(256) k_silhouette(PlainView s, QueryArgs q) {
    using V = typename VT<D>::T;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const V x = ldv<D>(q.x + D * i);
    float r_max = q.r_max ? q.r_max[i] : q.r_max_scalar;
    const gq_records &o = q.out;
    constexpr int NS = L<D>::snode;
    float d_root, dmx;
    box_distance(ldv<D>(s.snch_nodes), ldv<D>(s.snch_nodes + D), x, d_root, dmx);
    bool found = false;
    V best_p{};
    float best_d = 0.0f;
    int best_id = 0;
    if (d_root <= r_max) {
        int st_i[kStack];
        float st_d[kStack];
        int sp = 0;
        st_i[0] = 0;
        st_d[0] = d_root;
        while (sp >= 0) {
            const int ni = st_i[sp];
            const float cur = st_d[sp];
            --sp;
            if (cur <= r_max) {
                const float *nd = s.snch_nodes + (int64_t)NS * ni;
                const int n_prims = f2i(nd[3 * D + 4]);
                if (n_prims > 0) {
                    const int off = f2i(nd[3 * D + 5]);
                    const int cnt = f2i(nd[3 * D + 6]);
                    for (int j = 0; j < cnt; ++j) {
                        const float *sl = s.sils + (int64_t)L<D>::sil * (off + j);
                        if constexpr (D == 2) {
                            // SilhouetteVertex.closest_silhouette (silhouette_vertex.py:39-53)
                            V2 a = ld2(sl), b = ld2(sl + 2), c = ld2(sl + 4);
                            float dd = norm(x - b);
                            if ((f2i(sl[6]) != -1) & (f2i(sl[8]) != -1) & (dd < r_max)) {
                                if (cross2(b - a, x - a) * cross2(c - b, x - b) < 0.0f) {
                                    r_max = fminf(r_max, dd);
                                    found = true; best_p = b; best_d = dd; best_id = f2i(sl[9]);
                                }
                            }
                        } else {
                            // SilhouetteEdge.closest_silhouette (silhouette_edge.py:49-64)
                            V3 c = ld3(sl), a = ld3(sl + 3), b = ld3(sl + 6), e = ld3(sl + 9);
                            V3 pt;
                            float dd = closest_point_segment(x, a, b, pt);
                            if (dd < r_max) {
                                bool sil = (f2i(sl[12]) == -1) | (f2i(sl[15]) == -1);
                                if (!sil) {
                                    V3 n0 = normalize(cross(b - a, c - a));
                                    V3 n1 = normalize(cross(e - a, b - a));
                                    V3 vd = pt - x;
                                    sil = dot(vd, n0) * dot(vd, n1) < 0.0f;
                                }
                                if (sil) {
                                    r_max = fminf(r_max, dd);
                                    found = true; best_p = pt; best_d = dd; best_id = f2i(sl[18]);
                                }
                            }
                        }
                    }
                } else {
                    const int li = ni + 1;
                    const int ri = ni + f2i(nd[3 * D + 3]);
                    const float *ln = s.snch_nodes + (int64_t)NS * li;
                    const float *rn = s.snch_nodes + (int64_t)NS * ri;
                    float d0, d1;
                    const bool h0 = snch_overlaps(ldv<D>(ln), ldv<D>(ln + D), ldv<D>(ln + 2 * D),
                                                  ln[3 * D], ln[3 * D + 1], x, r_max, d0);
                    const bool h1 = snch_overlaps(ldv<D>(rn), ldv<D>(rn + D), ldv<D>(rn + 2 * D),
                                                  rn[3 * D], rn[3 * D + 1], x, r_max, d1);
                    if (h0 & h1) {
                        int closer = li, other = ri;
                        if (d1 < d0) {
                            closer = ri; other = li;
                            float t = d0; d0 = d1; d1 = t;
                        }
                        if (sp + 2 < kStack) {
                            st_i[++sp] = other; st_d[sp] = d1;
                            st_i[++sp] = closer; st_d[sp] = d0;
                        }
                    } else if (h0) {
                        if (sp + 1 < kStack) { st_i[++sp] = li; st_d[sp] = d0; }
                    } else if (h1) {
                        if (sp + 1 < kStack) { st_i[++sp] = ri; st_d[sp] = d1; }
                    }
                }
            }
        }
    }
    if (!found) {
        zero_record<D>(o, i, false);
        return;
    }
    if (o.valid) o.valid[i] = 1;
    if (o.p) st(o.p + D * i, best_p);
    if (o.d) o.d[i] = best_d;
    if (o.prim_id) o.prim_id[i] = best_id;
}

template <int D> hipError_t launch_dim(QueryKind kind, const PlainView &s, const QueryArgs &q, hipStream_t st) {
    const int block = 256;
    const unsigned grid = (unsigned)((q.n + block - 1) / block);
    static const char *names[2][3] = {
        {"plain_closest_point_2d", "plain_intersect_2d", "plain_silhouette_2d"},
        {"plain_closest_point_3d", "plain_intersect_3d", "plain_silhouette_3d"}};
    prof_begin(st);
    if (kind == kClosestPoint) hipLaunchKernelGGL(k_closest_point<D>, dim3(grid), dim3(block), 0, st, s, q);
    else if (kind == kIntersect) hipLaunchKernelGGL(k_intersect<D>, dim3(grid), dim3(block), 0, st, s, q);
    else hipLaunchKernelGGL(k_silhouette<D>, dim3(grid), dim3(block), 0, st, s, q);
    prof_end(st, names[D - 2][kind]);
    return hipGetLastError();
}

} // namespace

hipError_t launch_plain(QueryKind kind, const PlainView &s, const QueryArgs &q, hipStream_t st) {
    if (q.n <= 0) return hipSuccess;
    return s.dim == 2 ? launch_dim<2>(kind, s, q, st) : launch_dim<3>(kind, s, q, st);
}

} // namespace gqa
