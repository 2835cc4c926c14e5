// Engine kernels: Morton-sorted queries, 4-wide BVH (float32 or quantized child boxes) for the
// closest-point and ray queries, SNCH pair records for the silhouette query.
//
// Every traversal is one lane per query over the sorted order (lanes of a wavefront hold nearby
// queries, so they walk the same nodes), with a per-lane stack whose first kLdsDepth entries live
// in LDS and whose deeper entries spill to private memory.  Children of a node are visited nearest
// first.  Pruning is conservative: a node is skipped only when no primitive in it can beat the
// current answer, so results equal the reference traversal's up to exact ties:
//  * closest point: box distance^2 against a bound = min(best^2, MinMaxDist^2 of every child box
//    seen so far), MinMaxDist widened by the quantization step and the builder's padding;
//  * ray: slab test with the entry/exit distances widened by 4 ulp;
//  * silhouette: the reference's box and normal-cone tests (the cone test is evaluated with the
//    reference's float32 cone values wherever it can cull) plus a cull of subtrees that provably
//    contain no silhouette for the query (every predicate normal of the subtree lies in a cone whose
//    dot products with all view directions into the subtree share one sign).
// Leaf tests and output records use the reference formulas (gq_device.h).
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "gq_device.h"
#include "gq_internal.h"
#include "wide.h"

namespace gqa {
namespace {

using namespace gqd;

constexpr int kBlock = 256;
constexpr int kLdsDepth = 16;
constexpr int kSpillDepth = 96;
constexpr int kStackCap = kLdsDepth + kSpillDepth;
constexpr int kSortMin = 8192;   // smaller batches are traversed in input order
constexpr float kInf = __builtin_inff();

struct EngineParams {
    const void *nodes;        // BVH4 (FNode / QNode) or SNCH pairs (SPairF / SPairQ)
    const float4 *prims;
    const int32_t *ids;
    const RefCone *ref_cones;
    const uint32_t *big;
    uint32_t root;
    float root_lo[3], root_hi[3];
    const float *x, *v, *r_max;
    float r_max_scalar;
    const uint32_t *perm;
    int n;
    gq_records out;
};

// ---- per-lane traversal stack: LDS for the top kLdsDepth entries' slots, private spill below ----
struct Stack {
    uint2 *lds;                // this lane's column: slot s at lds[s * kBlock]
    uint2 spill[kSpillDepth];
    int sp;
    __device__ __forceinline__ void push(uint32_t ref, float key) {
        if (sp >= kStackCap) return;  // unreachable: engine_supported() bounds the depth
        uint2 e = make_uint2(ref, __float_as_uint(key));
        if (sp < kLdsDepth) lds[sp * kBlock] = e;
        else spill[sp - kLdsDepth] = e;
        ++sp;
    }
    // Pops entries until one with key <= limit; returns false when the stack is empty.
    __device__ __forceinline__ bool pop(float limit, uint32_t &ref) {
        while (sp > 0) {
            --sp;
            uint2 e = sp < kLdsDepth ? lds[sp * kBlock] : spill[sp - kLdsDepth];
            if (__uint_as_float(e.y) <= limit) {
                ref = e.x;
                return true;
            }
        }
        return false;
    }
};

template <typename T> __device__ __forceinline__ T load_rec(const T *p) {
    static_assert(sizeof(T) % 16 == 0, "records are multiples of 16 bytes");
    union U {
        T t;
        uint4 v[sizeof(T) / 16];
        __device__ U() {}
    } u;
    const uint4 *s = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 16); ++k) u.v[k] = s[k];
    return u.t;
}

__device__ __forceinline__ void leaf_range(uint32_t ref, const uint32_t *big, uint32_t &first, uint32_t &count) {
    first = ref_first(ref);
    count = ref_count(ref);
    if (count == 127) {
        count = big[2 * first + 1];
        first = big[2 * first];
    }
}

// Builder padding of every box face: FLT_EPSILON plus rounding of the padded coordinate.
__device__ __forceinline__ float pad_of(float c) { return 2.5e-7f * (1.0f + fabsf(c)); }

// Child boxes of a BVH4 node, decoded; delta[k] bounds how far a child face may lie outside the
// nearest primitive point (quantization step + builder padding).
template <int D, bool Q> struct Node4 {
    float lo[D][4], hi[D][4], delta[D];
    uint32_t child[4];
    int n;
};

template <int D, bool Q>
__device__ __forceinline__ void load_node4(const void *base, uint32_t idx, Node4<D, Q> &nd) {
    if constexpr (Q) {
        const QNode<D> q = load_rec(reinterpret_cast<const QNode<D> *>(base) + idx);
        nd.n = q.nchild;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float s = ldexpf(1.0f, (int)q.exp[k]);
            nd.delta[k] = s + pad_of(q.origin[k]) + pad_of(fmaf((float)QBits<D>::levels, s, q.origin[k]));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                nd.lo[k][j] = fmaf((float)q.qlo[k][j], s, q.origin[k]);
                nd.hi[k][j] = fmaf((float)q.qhi[k][j], s, q.origin[k]);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) nd.child[j] = q.child[j];
    } else {
        const FNode<D> f = load_rec(reinterpret_cast<const FNode<D> *>(base) + idx);
        nd.n = (int)f.nchild;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            nd.delta[k] = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                nd.lo[k][j] = f.lo[k][j];
                nd.hi[k][j] = f.hi[k][j];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) nd.child[j] = f.child[j];
    }
}

__device__ __forceinline__ void cswap(float *k, uint32_t *v, int a, int b) {
    const bool s = k[b] < k[a];
    const float ka = k[a], kb = k[b];
    const uint32_t va = v[a], vb = v[b];
    k[a] = s ? kb : ka;
    k[b] = s ? ka : kb;
    v[a] = s ? vb : va;
    v[b] = s ? va : vb;
}
__device__ __forceinline__ void sort4(float *k, uint32_t *v) {
    cswap(k, v, 0, 1);
    cswap(k, v, 2, 3);
    cswap(k, v, 0, 2);
    cswap(k, v, 1, 3);
    cswap(k, v, 1, 2);
}

template <int D> __device__ __forceinline__ float comp(const typename VT<D>::T &p, int k) {
    if constexpr (D == 2) return k == 0 ? p.x : p.y;
    else return k == 0 ? p.x : (k == 1 ? p.y : p.z);
}

template <int D> __device__ __forceinline__ uint32_t query_index(const EngineParams &P, int i) {
    return P.perm ? P.perm[i] : (uint32_t)i;
}

// ---------------------------------------------------------------------------------------------
// closest point
template <int D, bool Q>
__global__ void __launch_bounds__(kBlock) k_closest_point(EngineParams P) {
    using V = typename VT<D>::T;
    __shared__ uint2 lds[kLdsDepth * kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t qi = query_index<D>(P, i);
    const V p = ldv<D>(P.x + (size_t)D * qi);
    float pc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pc[k] = comp<D>(p, k);
    Stack stk;
    stk.lds = lds + threadIdx.x;
    stk.sp = 0;
    float best = kInf, bound2 = kInf;
    int best_slot = -1;
    uint32_t cur = P.root;
    bool go = true;
    while (go) {
        if (!ref_is_leaf(cur)) {
            Node4<D, Q> nd;
            load_node4<D, Q>(P.nodes, cur, nd);
            float key[4];
            uint32_t ref[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float d2 = 0.0f, fsum = 0.0f, far2[D], near2[D];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float a = nd.lo[k][j] - pc[k], b = pc[k] - nd.hi[k][j];
                    const float o = fmaxf(fmaxf(a, b), 0.0f);
                    d2 = fmaf(o, o, d2);
                    const float fa = fabsf(a), fb = fabsf(b);
                    const float nr = fminf(fa, fb) + nd.delta[k];
                    const float fr = fmaxf(fa, fb);
                    far2[k] = fr * fr;
                    near2[k] = nr * nr;
                    fsum += far2[k];
                }
                float ub2 = kInf;
#pragma unroll
                for (int k = 0; k < D; ++k) ub2 = fminf(ub2, fsum - far2[k] + near2[k]);
                const bool valid = j < nd.n;
                if (valid) bound2 = fminf(bound2, ub2 * 1.000002f);
                key[j] = valid ? d2 : kInf;
                ref[j] = nd.child[j];
            }
            sort4(key, ref);
            int nh = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) nh += (key[j] <= bound2) & (key[j] < kInf);
            if (nh == 0) {
                go = stk.pop(bound2, cur);
                continue;
            }
#pragma unroll
            for (int j = 3; j >= 1; --j)
                if (j < nh) stk.push(ref[j], key[j]);
            cur = ref[0];
        } else {
            uint32_t first, count;
            leaf_range(cur, P.big, first, count);
            for (uint32_t s = first; s < first + count; ++s) {
                float dd;
                if constexpr (D == 3) {
                    const float4 a = P.prims[3 * s], b = P.prims[3 * s + 1], c = P.prims[3 * s + 2];
                    V3 pt;
                    V2 uv;
                    dd = closest_point_triangle(p, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), pt, uv);
                } else {
                    const float4 ab = P.prims[s];
                    V2 pt;
                    dd = closest_point_segment(p, v2(ab.x, ab.y), v2(ab.z, ab.w), pt);
                }
                if (dd < best) {
                    best = dd;
                    best_slot = (int)s;
                    bound2 = fminf(bound2, dd * dd * 1.000002f);
                }
            }
            go = stk.pop(bound2, cur);
        }
    }
    const gq_records &o = P.out;
    if (best_slot < 0) {
        if (o.valid) o.valid[qi] = 0;
        if (o.p) for (int k = 0; k < D; ++k) o.p[(size_t)D * qi + k] = 0.0f;
        if (o.n) for (int k = 0; k < D; ++k) o.n[(size_t)D * qi + k] = 0.0f;
        if (o.uv) for (int k = 0; k < D - 1; ++k) o.uv[(size_t)(D - 1) * qi + k] = 0.0f;
        if (o.d) o.d[qi] = 0.0f;
        if (o.prim_id) o.prim_id[qi] = 0;
        return;
    }
    const int s = best_slot;
    if constexpr (D == 3) {
        // Triangle.closest_point (triangle.py:75-81)
        const float4 a4 = P.prims[3 * s], b4 = P.prims[3 * s + 1], c4 = P.prims[3 * s + 2];
        const V3 a = v3(a4.x, a4.y, a4.z), b = v3(b4.x, b4.y, b4.z), c = v3(c4.x, c4.y, c4.z);
        V3 pt;
        V2 uv;
        const float dd = closest_point_triangle(p, a, b, c, pt, uv);
        if (o.p) st(o.p + 3 * (size_t)qi, pt);
        if (o.n) st(o.n + 3 * (size_t)qi, normalize(cross(b - a, c - a)));
        if (o.uv) st(o.uv + 2 * (size_t)qi, uv);
        if (o.d) o.d[qi] = dd;
    } else {
        // LineSegment.closest_point (line_segment.py:63-78)
        const float4 ab = P.prims[s];
        const V2 a = v2(ab.x, ab.y), b = v2(ab.z, ab.w);
        const V2 pa = p - a, ba = b - a;
        const V2 t = normalize(ba);
        const float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        const V2 ip = a + ba * h;
        if (o.p) st(o.p + 2 * (size_t)qi, ip);
        if (o.n) st(o.n + 2 * (size_t)qi, v2(t.y, -t.x));
        if (o.uv) o.uv[qi] = h;
        if (o.d) o.d[qi] = norm(ip - p);
    }
    if (o.prim_id) o.prim_id[qi] = P.ids[s];
    if (o.valid) o.valid[qi] = 1;
}

// ---------------------------------------------------------------------------------------------
// ray
template <int D, bool Q>
__global__ void __launch_bounds__(kBlock) k_intersect(EngineParams P) {
    using V = typename VT<D>::T;
    __shared__ uint2 lds[kLdsDepth * kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t qi = query_index<D>(P, i);
    const V x = ldv<D>(P.x + (size_t)D * qi);
    const V v = ldv<D>(P.v + (size_t)D * qi);
    float xc[D], inv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        xc[k] = comp<D>(x, k);
        inv[k] = 1.0f / comp<D>(v, k);
    }
    float best = P.r_max ? P.r_max[qi] : P.r_max_scalar;
    const float r_max0 = best;
    Stack stk;
    stk.lds = lds + threadIdx.x;
    stk.sp = 0;
    int best_slot = -1;
    uint32_t cur = P.root;
    bool go = true;
    constexpr float kLo = 1.0f - 4.8e-7f, kHi = 1.0f + 4.8e-7f;
    while (go) {
        if (!ref_is_leaf(cur)) {
            Node4<D, Q> nd;
            load_node4<D, Q>(P.nodes, cur, nd);
            float key[4];
            uint32_t ref[4];
            const float lim = best * kHi;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float tn = 0.0f, tf = lim;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float t0 = (nd.lo[k][j] - xc[k]) * inv[k];
                    const float t1 = (nd.hi[k][j] - xc[k]) * inv[k];
                    tn = fmaxf(tn, fminf(t0, t1));
                    tf = fminf(tf, fmaxf(t0, t1));
                }
                const float tfw = tf >= 0.0f ? tf * kHi : tf * kLo;
                const bool hit = (j < nd.n) & (tn * kLo <= tfw);
                key[j] = hit ? tn * kLo : kInf;
                ref[j] = nd.child[j];
            }
            sort4(key, ref);
            int nh = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) nh += (key[j] <= lim) & (key[j] < kInf);
            if (nh == 0) {
                go = stk.pop(best * kHi, cur);
                continue;
            }
#pragma unroll
            for (int j = 3; j >= 1; --j)
                if (j < nh) stk.push(ref[j], key[j]);
            cur = ref[0];
        } else {
            uint32_t first, count;
            leaf_range(cur, P.big, first, count);
            for (uint32_t s = first; s < first + count; ++s) {
                float t;
                if constexpr (D == 3) {
                    const float4 a = P.prims[3 * s], b = P.prims[3 * s + 1], c = P.prims[3 * s + 2];
                    t = ray_triangle_t(x, v, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z));
                } else {
                    const float4 ab = P.prims[s];
                    t = ray_segment_t(x, v, v2(ab.x, ab.y), v2(ab.z, ab.w));
                }
                if (t < best) {
                    best = t;
                    best_slot = (int)s;
                }
            }
            go = stk.pop(best * kHi, cur);
        }
    }
    const gq_records &o = P.out;
    bool ok = best_slot >= 0;
    const int s = best_slot;
    if (ok) {
        if constexpr (D == 3) {
            // Triangle.ray_intersect (triangle.py:48-73) with r_max = the hit distance
            const float4 a4 = P.prims[3 * s], b4 = P.prims[3 * s + 1], c4 = P.prims[3 * s + 2];
            const V3 a = v3(a4.x, a4.y, a4.z), b = v3(b4.x, b4.y, b4.z), c = v3(c4.x, c4.y, c4.z);
            const V3 v1 = b - a, v2_ = c - a;
            const V3 pp = cross(v, v2_);
            const float det = dot(v1, pp);
            float vv = 0.0f, ww = 0.0f, t = 0.0f;
            ok = false;
            if (fabsf(det) > kDrEps) {
                const float inv_det = 1.0f / det;
                const V3 sv = x - a;
                vv = dot(sv, pp) * inv_det;
                if ((vv >= 0.0f) & (vv <= 1.0f)) {
                    const V3 qq = cross(sv, v1);
                    ww = dot(v, qq) * inv_det;
                    if ((ww >= 0.0f) & (vv + ww <= 1.0f)) {
                        t = dot(v2_, qq) * inv_det;
                        ok = (t >= 0.0f) & (t <= best);
                    }
                }
            }
            if (ok) {
                if (o.p) st(o.p + 3 * (size_t)qi, a + v1 * vv + v2_ * ww);
                if (o.n) st(o.n + 3 * (size_t)qi, normalize(cross(v1, v2_)));
                if (o.uv) st(o.uv + 2 * (size_t)qi, v2(1.0f - vv - ww, vv));
                if (o.d) o.d[qi] = t;
            }
        } else {
            // record of bvh.py:267-274
            const float4 ab = P.prims[s];
            const V2 tt = v2(ab.z - ab.x, ab.w - ab.y);
            if (o.p) st(o.p + 2 * (size_t)qi, x + v * best);
            if (o.n) st(o.n + 2 * (size_t)qi, normalize(v2(tt.y, -tt.x)));
            if (o.uv) o.uv[qi] = -1.0f;
            if (o.d) o.d[qi] = best;
        }
    }
    (void)r_max0;
    if (!ok) {
        if (o.valid) o.valid[qi] = 0;
        if (o.p) for (int k = 0; k < D; ++k) o.p[(size_t)D * qi + k] = 0.0f;
        if (o.n) for (int k = 0; k < D; ++k) o.n[(size_t)D * qi + k] = 0.0f;
        if (o.uv) for (int k = 0; k < D - 1; ++k) o.uv[(size_t)(D - 1) * qi + k] = 0.0f;
        if (o.d) o.d[qi] = 0.0f;
        if (o.prim_id) o.prim_id[qi] = 0;
        if (o.on_boundary) o.on_boundary[qi] = 0;
        return;
    }
    if (o.prim_id) o.prim_id[qi] = P.ids[s];
    if (o.on_boundary) o.on_boundary[qi] = 1;
    if (o.valid) o.valid[qi] = 1;
}

// ---------------------------------------------------------------------------------------------
// silhouette
struct PairChild {
    float lo[3], hi[3];
    float axis[3], cos_t, sin_t;
    uint32_t child;
    int32_t ref_cone;
};

template <int D, bool Q>
__device__ __forceinline__ void load_pair(const void *base, uint32_t idx, PairChild (&c)[2]) {
    if constexpr (Q) {
        const SPairQ<D> q = load_rec(reinterpret_cast<const SPairQ<D> *>(base) + idx);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float s = ldexpf(1.0f, (int)q.exp[k]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                c[j].lo[k] = fmaf((float)q.qlo[j][k], s, q.origin[k]);
                c[j].hi[k] = fmaf((float)q.qhi[j][k], s, q.origin[k]);
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            constexpr float kS = 1.0f / 32767.0f;
            float a0 = q.cone_axis[j][0] * kS, a1 = q.cone_axis[j][1] * kS, a2 = q.cone_axis[j][2] * kS;
            float l = a0 * a0 + a1 * a1 + a2 * a2;
            float r = l > 0.0f ? 1.0f / sqrtf(l) : 0.0f;
            c[j].axis[0] = a0 * r;
            c[j].axis[1] = a1 * r;
            c[j].axis[2] = a2 * r;
            c[j].cos_t = q.cone_cs[j][0] == 32767 && q.cone_cs[j][1] == -32767 ? 2.0f : q.cone_cs[j][0] * kS;
            c[j].sin_t = q.cone_cs[j][1] * kS;
            c[j].child = q.child[j];
            c[j].ref_cone = q.ref_cone[j];
        }
    } else {
        const SPairF<D> f = load_rec(reinterpret_cast<const SPairF<D> *>(base) + idx);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                c[j].lo[k] = f.lo[j][k];
                c[j].hi[k] = f.hi[j][k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) c[j].axis[k] = f.cone[j].axis[k];
            c[j].cos_t = f.cone[j].cos_t;
            c[j].sin_t = f.cone[j].sin_t;
            c[j].child = f.child[j];
            c[j].ref_cone = f.ref_cone[j];
        }
    }
}

// True when every silhouette predicate in the child is provably false for query x: the normals
// lie within the cone and the dot product of every view direction into the box with every cone
// normal has one sign (bounded away from zero by a margin covering float32 rounding).
template <int D>
__device__ __forceinline__ bool exact_cull(const PairChild &c, const float *xc) {
    if (c.cos_t > 1.5f) return true;     // no silhouette can ever pass in this subtree
    if (!(c.cos_t > 0.0f)) return false; // cone too wide to separate anything
    if constexpr (D == 3) {
        float w[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            w[k] = 0.5f * (c.lo[k] + c.hi[k]) - xc[k];
            h[k] = 0.5f * (c.hi[k] - c.lo[k]);
        }
        const float R = sqrtf(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        const float a = w[0] * c.axis[0] + w[1] * c.axis[1] + w[2] * c.axis[2];
        const float cx = w[1] * c.axis[2] - w[2] * c.axis[1];
        const float cy = w[2] * c.axis[0] - w[0] * c.axis[2];
        const float cz = w[0] * c.axis[1] - w[1] * c.axis[0];
        const float b = sqrtf(cx * cx + cy * cy + cz * cz);
        const float m = R + 1e-5f * (fabsf(a) + b + R) + 1e-6f;
        return (a * c.cos_t - b * c.sin_t > m) || (a * c.cos_t + b * c.sin_t < -m);
    } else {
        return false;
    }
}

template <int D, bool Q>
__global__ void __launch_bounds__(kBlock) k_silhouette(EngineParams P) {
    using V = typename VT<D>::T;
    __shared__ uint2 lds[kLdsDepth * kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t qi = query_index<D>(P, i);
    const V x = ldv<D>(P.x + (size_t)D * qi);
    float xc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = comp<D>(x, k);
    float r = P.r_max ? P.r_max[qi] : P.r_max_scalar;
    bool found = false;
    V best_p{};
    float best_d = 0.0f;
    int best_slot = 0;
    float d_root, d_unused;
    box_distance(ldv<D>(P.root_lo), ldv<D>(P.root_hi), x, d_root, d_unused);
    if (d_root <= r) {
        Stack stk;
        stk.lds = lds + threadIdx.x;
        stk.sp = 0;
        uint32_t cur = P.root;
        bool go = true;
        while (go) {
            if (!ref_is_leaf(cur)) {
                PairChild c[2];
                load_pair<D, Q>(P.nodes, cur, c);
                float key[2];
                bool hit[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float dm, dx;
                    box_distance(ldv<D>(c[j].lo), ldv<D>(c[j].hi), x, dm, dx);
                    bool h = dm <= r;
                    if (h && c[j].ref_cone >= 0) {
                        const RefCone &rc = P.ref_cones[c[j].ref_cone];
                        h = snch_overlaps(ldv<D>(rc.lo), ldv<D>(rc.hi), ldv<D>(rc.axis), rc.half, rc.radius, x, r, dm);
                    }
                    if (h && exact_cull<D>(c[j], xc)) h = false;
                    hit[j] = h;
                    key[j] = dm;
                }
                if (hit[0] & hit[1]) {
                    const int near = key[1] < key[0] ? 1 : 0;
                    stk.push(c[1 - near].child, key[1 - near]);
                    cur = c[near].child;
                } else if (hit[0] | hit[1]) {
                    cur = hit[0] ? c[0].child : c[1].child;
                } else {
                    go = stk.pop(r, cur);
                }
            } else {
                uint32_t first, count;
                leaf_range(cur, P.big, first, count);
                for (uint32_t s = first; s < first + count; ++s) {
                    if constexpr (D == 3) {
                        const float4 s0 = P.prims[3 * s], s1 = P.prims[3 * s + 1], s2 = P.prims[3 * s + 2];
                        const V3 a = v3(s0.x, s0.y, s0.z), b = v3(s0.w, s1.x, s1.y);
                        V3 pt;
                        const float dd = closest_point_segment(x, a, b, pt);
                        if (dd < r) {
                            const V3 n0 = v3(s1.z, s1.w, s2.x), n1 = v3(s2.y, s2.z, s2.w);
                            const V3 vd = pt - x;
                            if (dot(vd, n0) * dot(vd, n1) < 0.0f) {
                                r = dd;
                                found = true;
                                best_p = pt;
                                best_d = dd;
                                best_slot = (int)s;
                            }
                        }
                    } else {
                        const float4 s0 = P.prims[2 * s], s1 = P.prims[2 * s + 1];
                        const V2 a = v2(s0.x, s0.y), b = v2(s0.z, s0.w), cc = v2(s1.x, s1.y);
                        const float dd = norm(x - b);
                        if (dd < r && cross2(b - a, x - a) * cross2(cc - b, x - b) < 0.0f) {
                            r = dd;
                            found = true;
                            best_p = b;
                            best_d = dd;
                            best_slot = (int)s;
                        }
                    }
                }
                go = stk.pop(r, cur);
            }
        }
    }
    const gq_records &o = P.out;
    if (!found) {
        if (o.valid) o.valid[qi] = 0;
        if (o.p) for (int k = 0; k < D; ++k) o.p[(size_t)D * qi + k] = 0.0f;
        if (o.d) o.d[qi] = 0.0f;
        if (o.prim_id) o.prim_id[qi] = 0;
        return;
    }
    if (o.valid) o.valid[qi] = 1;
    if (o.p) st(o.p + (size_t)D * qi, best_p);
    if (o.d) o.d[qi] = best_d;
    if (o.prim_id) o.prim_id[qi] = P.ids[best_slot];
}

// ---------------------------------------------------------------------------------------------
// Morton ordering of the queries
__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ void k_bounds_init(uint32_t *b) {
    const int t = threadIdx.x;
    if (t < 3) b[t] = 0xFFFFFFFFu;        // min
    else if (t < 6) b[t] = 0u;            // max
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_bounds(const float *x, int n, uint32_t *b) {
    uint32_t mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { mn[k] = 0xFFFFFFFFu; mx[k] = 0u; }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float f = x[(size_t)D * i + k];
            if (isfinite(f)) {
                const uint32_t u = f2ord(f);
                mn[k] = min(mn[k], u);
                mx[k] = max(mx[k], u);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            atomicMin(&b[k], mn[k]);
            atomicMax(&b[3 + k], mx[k]);
        }
    }
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {  // 16 bits -> every second bit
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_morton(const float *x, int n, const uint32_t *b, uint32_t *keys,
                                                    uint32_t *vals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    constexpr float kCells = D == 3 ? 1023.0f : 65535.0f;
    uint32_t code = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float lo = ord2f(b[k]), hi = ord2f(b[3 + k]);
        const float ext = hi - lo;
        float t = ext > 0.0f ? (x[(size_t)D * i + k] - lo) / ext : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);  // NaN -> 0
        const uint32_t c = (uint32_t)(t * kCells);
        code |= (D == 3 ? spread3(c) : spread2(c)) << k;
    }
    keys[i] = code;
    vals[i] = (uint32_t)i;
}

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

template <int D, bool Q>
hipError_t launch_traversal(QueryKind kind, const EngineParams &P, hipStream_t st) {
    const unsigned grid = (unsigned)((P.n + kBlock - 1) / kBlock);
    static const char *names[2][3] = {
        {"engine_closest_point_2d", "engine_intersect_2d", "engine_silhouette_2d"},
        {"engine_closest_point_3d", "engine_intersect_3d", "engine_silhouette_3d"}};
    prof_begin(st);
    if (kind == kClosestPoint) hipLaunchKernelGGL((k_closest_point<D, Q>), dim3(grid), dim3(kBlock), 0, st, P);
    else if (kind == kIntersect) hipLaunchKernelGGL((k_intersect<D, Q>), dim3(grid), dim3(kBlock), 0, st, P);
    else hipLaunchKernelGGL((k_silhouette<D, Q>), dim3(grid), dim3(kBlock), 0, st, P);
    prof_end(st, names[D - 2][kind]);
    return hipGetLastError();
}

// Sorts the batch by the Morton code of the query points.  The buffers come from the
// stream-ordered allocator and are released on the same stream after the traversal, so calls on
// different streams never share scratch memory.
template <int D>
hipError_t sort_queries(const float *x, int n, hipStream_t st, const uint32_t **perm, void **buf) {
    *perm = nullptr;
    *buf = nullptr;
    if (n < kSortMin) return hipSuccess;
    size_t temp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0,
                                             D == 3 ? 30 : 32, st);
    if (e != hipSuccess) return e;
    const size_t nb = align256((size_t)n * 4);
    const size_t need = 256 + 4 * nb + align256(temp);
    if ((e = hipMallocAsync(buf, need, st)) != hipSuccess) return e;
    uint8_t *base = (uint8_t *)*buf;
    uint32_t *bounds = (uint32_t *)base;
    uint32_t *k_in = (uint32_t *)(base + 256), *k_out = (uint32_t *)(base + 256 + nb);
    uint32_t *v_in = (uint32_t *)(base + 256 + 2 * nb), *v_out = (uint32_t *)(base + 256 + 3 * nb);
    void *tmp = base + 256 + 4 * nb;
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, bounds);
    const unsigned rgrid = (unsigned)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_bounds<D>), dim3(rgrid), dim3(kBlock), 0, st, x, n, bounds);
    hipLaunchKernelGGL((k_morton<D>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, x, n, bounds, k_in, v_in);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(tmp, temp, k_in, k_out, v_in, v_out, (size_t)n, 0, D == 3 ? 30 : 32, st);
    if (e != hipSuccess) return e;
    *perm = v_out;
    return hipSuccess;
}

template <int D>
hipError_t launch_dim(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q,
                      EngineWorkspace &ws, hipStream_t st) {
    EngineParams P{};
    P.x = q.x;
    P.v = q.v;
    P.r_max = q.r_max;
    P.r_max_scalar = q.r_max_scalar;
    P.n = (int)q.n;
    P.out = q.out;
    P.big = w.big_leaves;
    if (kind == kSilhouette) {
        P.nodes = quantized ? w.qsnodes : w.snodes;
        P.prims = w.sils;
        P.ids = w.sil_ids;
        P.ref_cones = w.ref_cones;
        P.root = w.sroot_ref;
        for (int k = 0; k < 3; ++k) { P.root_lo[k] = w.sroot_lo[k]; P.root_hi[k] = w.sroot_hi[k]; }
    } else {
        P.nodes = quantized ? w.qnodes : w.nodes;
        P.prims = w.prims;
        P.ids = w.prim_ids;
        P.root = w.root_ref;
    }
    (void)ws;
    void *buf = nullptr;
    hipError_t e = sort_queries<D>(q.x, P.n, st, &P.perm, &buf);
    if (e == hipSuccess)
        e = quantized ? launch_traversal<D, true>(kind, P, st) : launch_traversal<D, false>(kind, P, st);
    if (buf) {
        hipError_t f = hipFreeAsync(buf, st);
        if (e == hipSuccess) e = f;
    }
    return e;
}

} // namespace

hipError_t EngineWorkspace::ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (buf) {
        // earlier calls on this scene may still read the old buffer
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
    }
    release();
    hipError_t e = hipMalloc(&buf, need);
    if (e == hipSuccess) bytes = need;
    return e;
}

void EngineWorkspace::release() {
    if (buf) hipFree(buf);
    buf = nullptr;
    bytes = 0;
}

bool engine_supported(QueryKind kind, const WideView &w) {
    if (kind == kSilhouette) return w.bin_depth + 2 <= kStackCap;
    return 3 * w.wide_depth + 4 <= kStackCap;
}

hipError_t launch_engine(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q,
                         EngineWorkspace &ws, hipStream_t st) {
    if (q.n <= 0) return hipSuccess;
    return w.dim == 2 ? launch_dim<2>(kind, w, quantized, q, ws, st) : launch_dim<3>(kind, w, quantized, q, ws, st);
}

} // namespace gqa
