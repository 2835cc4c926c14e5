// Engine kernels: Morton-sorted queries, 4-wide BVH (float32 or quantized child boxes) for the
// closest-point and ray queries, SNCH pair records for the silhouette query.
//
// Every traversal is one lane per query over the sorted order (lanes of a wavefront hold nearby
// queries, so they walk the same nodes), with a per-lane stack whose first kLdsDepth entries live
// in LDS and whose deeper entries spill to private memory.  Children of a node are visited nearest
// first.  Pruning is conservative: a node is skipped only when no primitive in it can beat the
// current answer, so results equal the reference traversal's up to exact ties:
//  * closest point: squared box distance against best^2 (inflated by 2e-6 relative so float
//    rounding never drops a box holding an equally near primitive);
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
#include "robust.h"
#include "wide.h"

namespace gqa {
namespace {

using namespace gqd;

constexpr int kBlock = 256;     // packet kernels: 4 wavefronts per block
constexpr int kRayBlock = 64;   // lane-per-query ray kernel: one wavefront per block
constexpr int kRayStackMax = 96;  // lane-stack entries: 96 * 64 lanes * 8 B = 48 KB of LDS per block
constexpr int kWavePassBlocks = 1024;  // persistent grid of the wave-per-query pass (4 per CU)
// Packet steps per wavefront before its queries are deferred to the wave-per-query pass, by query
// kind and dimension (measured optimum for 4-primitive leaves; a scene option overrides it).
constexpr int kPacketBudgetCp3 = 1500, kPacketBudgetCp2 = 1000, kPacketBudgetSil = 1000;
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
    int stack_depth;          // lane-stack entries (ray kernel)
    int budget;               // packet steps before a wave hands its queries to the wave-per-query pass (0 = never)
    uint32_t *defer_count;    // [0] deferred queries, [1] work counter of the wave-per-query pass
    uint4 *deferred;          // (query index, best distance bits, best slot, 0)
    int sil_lane;             // silhouettes: lane-per-query kernel instead of packet + wave pass
    gq_records out;
};

// ---- lane-per-query stack (rays): LDS only, depth chosen per launch from the tree depth, one
// column per lane of a [depth][blockDim] dynamic shared array.  No private spill area, so the
// kernels need no scratch memory and their occupancy is set by LDS alone.
#define GQ_LANE_STACK_DECL                     \
    extern __shared__ uint2 lane_stack[];      \
    const int lane_col = threadIdx.x;          \
    const int lane_stride = blockDim.x;        \
    int sp = 0;
#define GQ_LPUSH(REF, KEY)                                                                       \
    do {                                                                                         \
        if (sp < P.stack_depth) {                                                                \
            lane_stack[sp * lane_stride + lane_col] = make_uint2((REF), __float_as_uint(KEY));  \
            ++sp;                                                                                \
        }                                                                                        \
    } while (0)
#define GQ_LPOP(LIMIT, CUR, GO)                                                                  \
    do {                                                                                         \
        GO = false;                                                                              \
        while (sp > 0) {                                                                         \
            --sp;                                                                                \
            const uint2 e_ = lane_stack[sp * lane_stride + lane_col];                            \
            if (__uint_as_float(e_.y) <= (LIMIT)) {                                              \
                CUR = e_.x;                                                                      \
                GO = true;                                                                       \
                break;                                                                           \
            }                                                                                    \
        }                                                                                        \
    } while (0)

// ---- wave-coherent ("packet") traversal helpers: the 64 lanes of a wavefront hold 64 nearby
// queries (Morton order) and walk the tree together; the stack holds node references only and
// is shared by the wave, so it lives in a few LDS words; node and primitive records are read at
// wave-uniform addresses (scalar loads).
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniformf(float v) { return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v))); }
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint4 ld16(const void *p, int k) { return reinterpret_cast<const uint4 *>(p)[k]; }
__device__ __forceinline__ float ubyte(uint32_t w, int j) { return (float)((w >> (8 * j)) & 0xFFu); }
__device__ __forceinline__ float ushort(uint32_t w, int j) { return (float)((w >> (16 * j)) & 0xFFFFu); }
__device__ __forceinline__ int sbyte(uint32_t w, int j) { return (int)(int8_t)((w >> (8 * j)) & 0xFFu); }
__device__ __forceinline__ uint32_t comp4(const uint4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// A record as 32-bit words, loaded with 16-byte loads; fields are read at constant byte offsets.
template <int NW> struct Words {
    uint32_t w[NW];
    __device__ __forceinline__ void load(const void *p) {
#pragma unroll
        for (int k = 0; k < NW / 4; ++k) {
            const uint4 v = ld16(p, k);
            w[4 * k] = v.x; w[4 * k + 1] = v.y; w[4 * k + 2] = v.z; w[4 * k + 3] = v.w;
        }
    }
    __device__ __forceinline__ uint32_t u32(int off) const { return w[off / 4]; }
    __device__ __forceinline__ float f32(int off) const { return __uint_as_float(w[off / 4]); }
    __device__ __forceinline__ uint32_t u8(int off) const { return (w[off / 4] >> (8 * (off % 4))) & 0xFFu; }
    __device__ __forceinline__ int s8(int off) const { return (int)(int8_t)u8(off); }
    __device__ __forceinline__ uint32_t u16(int off) const { return (w[off / 4] >> (8 * (off % 4))) & 0xFFFFu; }
    __device__ __forceinline__ int s16(int off) const { return (int)(int16_t)u16(off); }
};

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
    float lo[D][4], hi[D][4], step[D];   // step: quantization step per axis (0 for float32 boxes)
    uint32_t child[4];
    int n;
};

template <int D, bool Q>
__device__ __forceinline__ void load_node4(const void *base, uint32_t idx, Node4<D, Q> &nd) {
    if constexpr (Q && D == 3) {
        // QNode<3>: [origin xyz | exp xyz, nchild] [qlo x,y,z | qhi x] [qhi y,z | child 0,1] [child 2,3 | pad]
        const uint8_t *p = reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64;
        const uint4 w0 = ld16(p, 0), w1 = ld16(p, 1), w2 = ld16(p, 2), w3 = ld16(p, 3);
        const float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
        nd.n = (int)(w0.w >> 24);
        const uint32_t qlo[3] = {w1.x, w1.y, w1.z}, qhi[3] = {w1.w, w2.x, w2.y};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float sc = ldexpf(1.0f, sbyte(w0.w, k));
            nd.step[k] = sc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                nd.lo[k][j] = fmaf(ubyte(qlo[k], j), sc, org[k]);
                nd.hi[k][j] = fmaf(ubyte(qhi[k], j), sc, org[k]);
            }
        }
        nd.child[0] = w2.z; nd.child[1] = w2.w; nd.child[2] = w3.x; nd.child[3] = w3.y;
    } else if constexpr (Q && D == 2) {
        // QNode<2>: [origin xy | exp xy, nchild, pad | qlo x0 x1] [qlo x2 x3, y0..y3 | qhi x0 x1] ...
        const uint8_t *p = reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64;
        const uint4 w0 = ld16(p, 0), w1 = ld16(p, 1), w2 = ld16(p, 2), w3 = ld16(p, 3);
        const float org[2] = {__uint_as_float(w0.x), __uint_as_float(w0.y)};
        nd.n = (int)((w0.z >> 16) & 0xFFu);
        // 16-bit fields from byte 12: qlo[2][4] then qhi[2][4] (32 bytes, 8 words)
        const uint32_t q[8] = {w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float sc = ldexpf(1.0f, sbyte(w0.z, k));
            nd.step[k] = sc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                nd.lo[k][j] = fmaf(ushort(q[k * 2 + j / 2], j & 1), sc, org[k]);
                nd.hi[k][j] = fmaf(ushort(q[4 + k * 2 + j / 2], j & 1), sc, org[k]);
            }
        }
        nd.child[0] = w2.w; nd.child[1] = w3.x; nd.child[2] = w3.y; nd.child[3] = w3.z;
    } else {
        // FNode<D>: lo[D][4], hi[D][4] as 2D float4, then child[4], nchild
        const uint8_t *p = reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 128;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const uint4 l = ld16(p, k), h = ld16(p, D + k);
            nd.step[k] = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                nd.lo[k][j] = __uint_as_float(comp4(l, j));
                nd.hi[k][j] = __uint_as_float(comp4(h, j));
            }
        }
        const uint4 c = ld16(p, 2 * D), m = ld16(p, 2 * D + 1);
        nd.child[0] = c.x; nd.child[1] = c.y; nd.child[2] = c.z; nd.child[3] = c.w;
        nd.n = (int)m.x;
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

// Output record of a closest-point query from the winning primitive slot (reference formulas).
template <int D>
__device__ __forceinline__ void cp_record(const EngineParams &P, uint32_t qi, const typename VT<D>::T &p, int best_slot) {
    using V = typename VT<D>::T;
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
        const float4 a4 = P.prims[4 * s], b4 = P.prims[4 * s + 1], c4 = P.prims[4 * s + 2];
        const V3 a = v3(a4.x, a4.y, a4.z), b = v3(b4.x, b4.y, b4.z), c = v3(c4.x, c4.y, c4.z);
        V3 pt;
        V2 uv;
        const float dd = closest_point_triangle(p, a, b, c, pt, uv);
        if (o.p) gqd::st(o.p + 3 * (size_t)qi, pt);
        if (o.n) gqd::st(o.n + 3 * (size_t)qi, normalize(cross(b - a, c - a)));
        if (o.uv) gqd::st(o.uv + 2 * (size_t)qi, uv);
        if (o.d) o.d[qi] = dd;
    } else {
        // LineSegment.closest_point (line_segment.py:63-78)
        const float4 ab = P.prims[s];
        const V2 a = v2(ab.x, ab.y), b = v2(ab.z, ab.w);
        const V2 pa = p - a, ba = b - a;
        const V2 t = normalize(ba);
        const float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        const V2 ip = a + ba * h;
        if (o.p) gqd::st(o.p + 2 * (size_t)qi, ip);
        if (o.n) gqd::st(o.n + 2 * (size_t)qi, v2(t.y, -t.x));
        if (o.uv) o.uv[qi] = h;
        if (o.d) o.d[qi] = norm(ip - p);
    }
    if (o.prim_id) o.prim_id[qi] = P.ids[s];
    if (o.valid) o.valid[qi] = 1;
}

// Silhouette-edge predicate on an engine record (s1, s2 hold n0 and n1; a, b the edge, pt its
// nearest point to x): a boundary edge is a silhouette everywhere (silhouette_edge.py:54), else
// is_silhouette_edge (silhouette_edge.py:7-12) with the FP64 fallback near the sign change.
__device__ __forceinline__ bool sil_edge_test(const V3 &x, const V3 &a, const V3 &b, const V3 &pt,
                                              const float4 &s1, const float4 &s2) {
    if (__float_as_uint(s2.y) == kSilAlwaysTag) return true;
    return sil_edge_pred_robust(x, a, b, pt, v3(s1.z, s1.w, s2.x), v3(s2.y, s2.z, s2.w));
}

// ---------------------------------------------------------------------------------------------
// closest point (packet)
constexpr int kPacketDepth = 128;   // wave stack entries (node references); bounded by engine_supported()

template <int D> __device__ __forceinline__ void packet_centroid(bool live, const float *pc, float *cen) {
    const float nl = fmaxf(wave_sum(live ? 1.0f : 0.0f), 1.0f);
#pragma unroll
    for (int k = 0; k < D; ++k) cen[k] = uniformf(wave_sum(live ? pc[k] : 0.0f) / nl);
}

// Hands the live lanes of a wavefront that exceeded its step budget to the wave-per-query pass:
// query index, current bound and best slot so far, appended with one atomic per wavefront.
__device__ __forceinline__ void defer_wave(const EngineParams &P, bool live, uint32_t qi, float bound, int slot) {
    const unsigned long long m = __ballot(live);
    const uint32_t nlive = (uint32_t)__popcll(m);
    // every lane takes part in the atomic (lane 0 adds the count) so no lane-dependent branch
    // precedes the return
    const uint32_t base = uniform(atomicAdd(P.defer_count, (threadIdx.x & 63) == 0 ? nlive : 0u));
    if (live) {
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull));
        P.deferred[base + rank] = make_uint4(qi, __float_as_uint(bound), (uint32_t)slot, 0u);
    }
}

template <int D, bool Q>
__global__ void __launch_bounds__(kBlock) k_closest_point(EngineParams P) {
    using V = typename VT<D>::T;
    __shared__ uint32_t wstack[kBlock / 64][kPacketDepth];
    uint32_t *wst = wstack[threadIdx.x >> 6];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < P.n;
    const uint32_t qi = live ? query_index<D>(P, i) : 0u;
    V p;
    if (live) p = ldv<D>(P.x + (size_t)D * qi);
    else p = ldv<D>(P.x);   // any valid point; the lane's result is discarded
    float pc[D], cen[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pc[k] = comp<D>(p, k);
    packet_centroid<D>(live, pc, cen);
    float best = kInf, bound2 = live ? kInf : -1.0f;   // a dead lane never opens anything
    int best_slot = -1;
    int sp = 0;
    int steps = 0;
    uint32_t cur = uniform(P.root);
    for (;;) {
        if (P.budget > 0 && ++steps > P.budget) {
            // Too much work for one wavefront (queries deep inside a closed surface see most of it
            // at nearly one distance): hand the queries, with the answer so far as the starting
            // bound, to the wave-per-query pass so they spread over many wavefronts.
            defer_wave(P, live, qi, best, best_slot);
            return;
        }
        if (!ref_is_leaf(cur)) {
            Node4<D, Q> nd;
            load_node4<D, Q>(P.nodes, cur, nd);
            float key[4];
            uint32_t ref[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float d2 = 0.0f, c2 = 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float o = fmaxf(fmaxf(nd.lo[k][j] - pc[k], pc[k] - nd.hi[k][j]), 0.0f);
                    d2 = fmaf(o, o, d2);
                    const float oc = fmaxf(fmaxf(nd.lo[k][j] - cen[k], cen[k] - nd.hi[k][j]), 0.0f);
                    c2 = fmaf(oc, oc, c2);
                }
                const bool any = __ballot((j < nd.n) && (d2 <= bound2)) != 0ull;
                key[j] = any ? uniformf(c2) : kInf;
                ref[j] = nd.child[j];
            }
            sort4(key, ref);
            int nh = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) nh += key[j] < kInf;
            if (nh > 0) {
#pragma unroll
                for (int j = 3; j >= 1; --j)
                    if (j < nh && sp < kPacketDepth) wst[sp++] = ref[j];
                cur = uniform(ref[0]);
                continue;
            }
        } else {
            uint32_t first, count;
            leaf_range(cur, P.big, first, count);
            first = uniform(first);
            count = uniform(count);
            for (uint32_t s = first; s < first + count; ++s) {
                float dd;
                if constexpr (D == 3) {
                    const float4 a = P.prims[4 * s], b = P.prims[4 * s + 1], c = P.prims[4 * s + 2];
                    const float4 pl = P.prims[4 * s + 3];
                    // plane-distance lower bound, widened for rounding: skip the exact test when no
                    // lane can get strictly closer through this triangle
                    const float lb = fabsf(fmaf(pl.x, pc[0], fmaf(pl.y, pc[1], fmaf(pl.z, pc[2], -pl.w))));
                    if (__ballot(lb * 0.99999f - 1e-6f < best) == 0ull) continue;
                    V3 pt;
                    V2 uv;
                    dd = closest_point_triangle(p, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), pt, uv);
                } else {
                    const float4 ab = P.prims[s];
                    V2 pt;
                    dd = closest_point_segment(p, v2(ab.x, ab.y), v2(ab.z, ab.w), pt);
                }
                if (live && (dd < best || (dd == best && (int)s < best_slot))) {
                    best = dd;
                    best_slot = (int)s;
                    bound2 = dd * dd * 1.000002f;
                }
            }
        }
        if (sp == 0) break;
        cur = uniform(wst[--sp]);
    }
    if (!live) return;
    cp_record<D>(P, qi, p, best_slot);
}

// ---------------------------------------------------------------------------------------------
// closest point, wave per query: finishes the queries the packet pass deferred.  The 64 lanes
// expand the top 64 frontier entries per round and append the opened children in lane order, so
// the frontier stays sorted by depth; every append at depth d happens after all older depth-d
// entries were taken, hence at most 256 entries per depth level and 256 * (levels) in total.
// Leaf candidates are reduced across the wave each round.
constexpr int kWaveBlock = 128;
constexpr int kFront = 2560;

__device__ __forceinline__ int wave_exclusive_scan(int v, int &total) {
    const int lane = threadIdx.x & 63;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

template <int D, bool Q>
__global__ void __launch_bounds__(kWaveBlock) k_closest_point_wave(EngineParams P) {
    using V = typename VT<D>::T;
    __shared__ uint2 front_all[kWaveBlock / 64][kFront];
    uint2 *front = front_all[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const uint32_t total = uniform(P.defer_count[0]);
    for (;;) {
        // All lanes take part in the fetch (lane 0 adds 1, the others 0) and later all lanes store
        // the same record: no lane-dependent branch inside this loop, so the compiler cannot split
        // the wavefront across iterations.
        const uint32_t k = uniform(atomicAdd(P.defer_count + 1, lane == 0 ? 1u : 0u));
        if (k >= total) break;
        const uint4 item = P.deferred[k];
        const uint32_t qi = uniform(item.x);
        float best = uniformf(__uint_as_float(item.y));
        int best_slot = (int)uniform(item.z);
        const V p = ldv<D>(P.x + (size_t)D * qi);
        float pc[D];
#pragma unroll
        for (int j = 0; j < D; ++j) pc[j] = comp<D>(p, j);
        float bound2 = best < kInf ? best * best * 1.000002f : kInf;
        int size = 1;
        front[0] = make_uint2(P.root, 0u);
        while (size > 0) {
            const int take = size < 64 ? size : 64;
            uint2 e = make_uint2(0u, __float_as_uint(kInf));
            if (lane < take) e = front[size - take + lane];
            size -= take;
            uint2 ent[4] = {};
            bool hit[4] = {false, false, false, false};
            float cand = kInf;
            int cand_slot = -1;
            if (lane < take && __uint_as_float(e.y) <= bound2) {
                if (!ref_is_leaf(e.x)) {
                    Node4<D, Q> nd;
                    load_node4<D, Q>(P.nodes, e.x, nd);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float d2 = 0.0f;
#pragma unroll
                        for (int c = 0; c < D; ++c) {
                            const float o = fmaxf(fmaxf(nd.lo[c][j] - pc[c], pc[c] - nd.hi[c][j]), 0.0f);
                            d2 = fmaf(o, o, d2);
                        }
                        hit[j] = j < nd.n && d2 <= bound2;
                        ent[j] = make_uint2(nd.child[j], __float_as_uint(d2));
                    }
                } else {
                    uint32_t first, count;
                    leaf_range(e.x, P.big, first, count);
                    for (uint32_t s = first; s < first + count; ++s) {
                        float dd;
                        if constexpr (D == 3) {
                            const float4 pl = P.prims[4 * s + 3];
                            const float lb = fabsf(fmaf(pl.x, pc[0], fmaf(pl.y, pc[1], fmaf(pl.z, pc[2], -pl.w))));
                            if (lb * 0.99999f - 1e-6f >= best) continue;
                            const float4 a = P.prims[4 * s], b = P.prims[4 * s + 1], c = P.prims[4 * s + 2];
                            V3 pt;
                            V2 uv;
                            dd = closest_point_triangle(p, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), pt, uv);
                        } else {
                            const float4 ab = P.prims[s];
                            V2 pt;
                            dd = closest_point_segment(p, v2(ab.x, ab.y), v2(ab.z, ab.w), pt);
                        }
                        if (dd < cand || (dd == cand && (int)s < cand_slot)) {
                            cand = dd;
                            cand_slot = (int)s;
                        }
                    }
                }
            }
            // wave minimum of the leaf candidates (ties to the lower slot, so every lane agrees)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const float oc = __shfl_xor(cand, o, 64);
                const int os = __shfl_xor(cand_slot, o, 64);
                if (oc < cand || (oc == cand && os < cand_slot)) {
                    cand = oc;
                    cand_slot = os;
                }
            }
            cand = uniformf(cand);
            cand_slot = (int)uniform((uint32_t)cand_slot);
            if (cand < best || (cand == best && cand_slot < best_slot)) {
                best = cand;
                best_slot = cand_slot;
                bound2 = cand * cand * 1.000002f;
            }
            // append the opened children
            int tot;
            const int nh = (int)hit[0] + (int)hit[1] + (int)hit[2] + (int)hit[3];
            int pos = size + wave_exclusive_scan(nh, tot);
            tot = (int)uniform((uint32_t)tot);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (hit[j] && pos < kFront) front[pos] = ent[j];
                pos += (int)hit[j];
            }
            size = size + tot < kFront ? size + tot : kFront;
        }
        cp_record<D>(P, qi, p, best_slot);
    }
}

// ---------------------------------------------------------------------------------------------
// ray (lane per query)
template <int D, bool Q>
__global__ void __launch_bounds__(kRayBlock) k_intersect(EngineParams P) {
    using V = typename VT<D>::T;
    GQ_LANE_STACK_DECL
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
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
                GQ_LPOP(best * kHi, cur, go);
                continue;
            }
#pragma unroll
            for (int j = 3; j >= 1; --j)
                if (j < nh) GQ_LPUSH(ref[j], key[j]);
            cur = ref[0];
        } else {
            uint32_t first, count;
            leaf_range(cur, P.big, first, count);
            for (uint32_t s = first; s < first + count; ++s) {
                float t;
                if constexpr (D == 3) {
                    const float4 a = P.prims[4 * s], b = P.prims[4 * s + 1], c = P.prims[4 * s + 2];
                    float bv, bw;
                    t = ray_triangle_robust(x, v, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), bv, bw);
                } else {
                    const float4 ab = P.prims[s];
                    t = ray_segment_t(x, v, v2(ab.x, ab.y), v2(ab.z, ab.w));
                }
                if (t < best || (t == best && (int)s < best_slot)) {
                    best = t;
                    best_slot = (int)s;
                }
            }
            GQ_LPOP(best * kHi, cur, go);
        }
    }
    const gq_records &o = P.out;
    bool ok = best_slot >= 0;
    const int s = best_slot;
    if (ok) {
        if constexpr (D == 3) {
            // Triangle.ray_intersect (triangle.py:48-73) with r_max = the hit distance
            const float4 a4 = P.prims[4 * s], b4 = P.prims[4 * s + 1], c4 = P.prims[4 * s + 2];
            const V3 a = v3(a4.x, a4.y, a4.z), b = v3(b4.x, b4.y, b4.z), c = v3(c4.x, c4.y, c4.z);
            const V3 v1 = b - a, v2_ = c - a;
            float vv, ww;
            const float t = ray_triangle_robust(x, v, a, b, c, vv, ww);   // as in the traversal
            ok = t <= best;
            if (ok) {
                if (o.p) gqd::st(o.p + 3 * (size_t)qi, a + v1 * vv + v2_ * ww);
                if (o.n) gqd::st(o.n + 3 * (size_t)qi, normalize(cross(v1, v2_)));
                if (o.uv) gqd::st(o.uv + 2 * (size_t)qi, v2(1.0f - vv - ww, vv));
                if (o.d) o.d[qi] = t;
            }
        } else {
            // record of bvh.py:267-274
            const float4 ab = P.prims[s];
            const V2 tt = v2(ab.z - ab.x, ab.w - ab.y);
            if (o.p) gqd::st(o.p + 2 * (size_t)qi, x + v * best);
            if (o.n) gqd::st(o.n + 2 * (size_t)qi, normalize(v2(tt.y, -tt.x)));
            if (o.uv) o.uv[qi] = -1.0f;
            if (o.d) o.d[qi] = best;
        }
    }
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
// silhouette (packet)
struct PairChild {
    float lo[3], hi[3];
    float axis[3], cos_t, sin_t;
    uint32_t child;
    int32_t ref_cone;
};

template <int D, bool Q>
__device__ __forceinline__ void load_pair(const void *base, uint32_t idx, PairChild (&c)[2]) {
    if constexpr (Q) {
        // SPairQ<D> (64 B): origin, exp | child boxes | cone_axis int16 @28 | cone_cs int16 @40 |
        // child @48 | ref_cone @56
        Words<16> r;
        r.load(reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float org = r.f32(4 * k);
            const float sc = ldexpf(1.0f, r.s8(4 * D + k));
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if constexpr (D == 3) {
                    c[j].lo[k] = fmaf((float)r.u8(16 + 3 * j + k), sc, org);
                    c[j].hi[k] = fmaf((float)r.u8(22 + 3 * j + k), sc, org);
                } else {
                    c[j].lo[k] = fmaf((float)r.u16(12 + 2 * (2 * j + k)), sc, org);
                    c[j].hi[k] = fmaf((float)r.u16(20 + 2 * (2 * j + k)), sc, org);
                }
            }
        }
        constexpr float kS = 1.0f / 32767.0f;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float a0 = r.s16(28 + 6 * j) * kS, a1 = r.s16(30 + 6 * j) * kS, a2 = r.s16(32 + 6 * j) * kS;
            const float l = a0 * a0 + a1 * a1 + a2 * a2;
            const float rn = l > 0.0f ? 1.0f / sqrtf(l) : 0.0f;
            c[j].axis[0] = a0 * rn;
            c[j].axis[1] = a1 * rn;
            c[j].axis[2] = a2 * rn;
            const int cs0 = r.s16(40 + 4 * j), cs1 = r.s16(42 + 4 * j);
            c[j].cos_t = (cs0 == 32767 && cs1 == -32767) ? 2.0f : cs0 * kS;
            c[j].sin_t = cs1 * kS;
            c[j].child = r.u32(48 + 4 * j);
            c[j].ref_cone = (int32_t)r.u32(56 + 4 * j);
        }
    } else {
        // SPairF<D> (128 B): lo[2][D] @0, hi[2][D] @8D, cones (axis3, cos, sin) @16D, child @16D+40,
        // ref_cone @16D+48
        Words<32> r;
        r.load(reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 128);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                c[j].lo[k] = r.f32(4 * (j * D + k));
                c[j].hi[k] = r.f32(8 * D + 4 * (j * D + k));
            }
            const int co = 16 * D + 20 * j;
#pragma unroll
            for (int k = 0; k < 3; ++k) c[j].axis[k] = r.f32(co + 4 * k);
            c[j].cos_t = r.f32(co + 12);
            c[j].sin_t = r.f32(co + 16);
            c[j].child = r.u32(16 * D + 40 + 4 * j);
            c[j].ref_cone = (int32_t)r.u32(16 * D + 48 + 4 * j);
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
    __shared__ uint32_t wstack[kBlock / 64][kPacketDepth];
    __shared__ unsigned long long wmask[kBlock / 64][kPacketDepth];
    uint32_t *wst = wstack[threadIdx.x >> 6];
    unsigned long long *wmk = wmask[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < P.n;
    const uint32_t qi = live ? query_index<D>(P, i) : 0u;
    V x;
    if (live) x = ldv<D>(P.x + (size_t)D * qi);
    else x = ldv<D>(P.x);
    float xc[3] = {0.0f, 0.0f, 0.0f}, cen[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = comp<D>(x, k);
    packet_centroid<D>(live, xc, cen);
    float r = live ? (P.r_max ? P.r_max[qi] : P.r_max_scalar) : -1.0f;
    bool found = false;
    V best_p{};
    float best_d = 0.0f;
    int best_slot = 0;
    float d_root, d_unused;
    box_distance(ldv<D>(P.root_lo), ldv<D>(P.root_hi), x, d_root, d_unused);
    // a lane whose query misses the root box takes no part (reference: snch.py:111-112)
    if (!(d_root <= r)) r = -1.0f;
    // `act`: this lane's own reference-order traversal reaches the current node.  Children are
    // opened for the whole wave when any lane needs them, but a lane only tests the leaves it
    // would reach itself, so every query sees exactly the silhouettes the reference sees (the
    // reference cone test is not conservative for every mesh, so extra leaves could change it).
    bool act = r >= 0.0f;
    if (__ballot(act) != 0ull) {
        int sp = 0;
        int steps = 0;
        uint32_t cur = uniform(P.root);
        for (;;) {
            if (P.budget > 0 && ++steps > P.budget) {
                // interior queries see no silhouette and little culling: finish them one per wave
                defer_wave(P, live, qi, r, found ? best_slot : -1);
                return;
            }
            if (!ref_is_leaf(cur)) {
                PairChild c[2];
                load_pair<D, Q>(P.nodes, cur, c);
                bool any[2], hh[2];
                unsigned long long hm[2];
                float key[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float dm, dx;
                    box_distance(ldv<D>(c[j].lo), ldv<D>(c[j].hi), x, dm, dx);
                    bool h = act && dm <= r;
                    if (h && c[j].ref_cone >= 0) {
                        const RefCone &rc = P.ref_cones[c[j].ref_cone];
                        h = snch_overlaps(ldv<D>(rc.lo), ldv<D>(rc.hi), ldv<D>(rc.axis), rc.half, rc.radius, x, r, dm);
                    }
                    if (h && exact_cull<D>(c[j], xc)) h = false;
                    hh[j] = h;
                    hm[j] = __ballot(h);
                    any[j] = hm[j] != 0ull;
                    float c2 = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const float oc = fmaxf(fmaxf(c[j].lo[k] - cen[k], cen[k] - c[j].hi[k]), 0.0f);
                        c2 = fmaf(oc, oc, c2);
                    }
                    key[j] = uniformf(c2);
                }
                if (any[0] && any[1]) {
                    const int near = key[1] < key[0] ? 1 : 0;
                    if (sp < kPacketDepth) {
                        wst[sp] = c[1 - near].child;
                        wmk[sp] = hm[1 - near];
                        ++sp;
                    }
                    cur = uniform(c[near].child);
                    act = near ? hh[1] : hh[0];
                    continue;
                }
                if (any[0] || any[1]) {
                    cur = uniform(any[0] ? c[0].child : c[1].child);
                    act = any[0] ? hh[0] : hh[1];
                    continue;
                }
            } else {
                uint32_t first, count;
                leaf_range(cur, P.big, first, count);
                first = uniform(first);
                count = uniform(count);
                for (uint32_t s = first; s < first + count; ++s) {
                    if constexpr (D == 3) {
                        const float4 s0 = P.prims[3 * s], s1 = P.prims[3 * s + 1], s2 = P.prims[3 * s + 2];
                        const V3 a = v3(s0.x, s0.y, s0.z), b = v3(s0.w, s1.x, s1.y);
                        V3 pt;
                        const float dd = closest_point_segment(x, a, b, pt);
                        if (act && (dd < r || (dd == r && found && (int)s < best_slot))) {
                            if (sil_edge_test(x, a, b, pt, s1, s2)) {
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
                        if (act && (dd < r || (dd == r && found && (int)s < best_slot)) &&
                            sil_vertex_pred_robust(x, a, b, cc)) {
                            r = dd;
                            found = true;
                            best_p = b;
                            best_d = dd;
                            best_slot = (int)s;
                        }
                    }
                }
            }
            if (sp == 0) break;
            --sp;
            cur = uniform(wst[sp]);
            act = ((wmk[sp] >> lane) & 1ull) != 0ull;
        }
    }
    if (!live) return;
    const gq_records &o = P.out;
    if (!found) {
        if (o.valid) o.valid[qi] = 0;
        if (o.p) for (int k = 0; k < D; ++k) o.p[(size_t)D * qi + k] = 0.0f;
        if (o.d) o.d[qi] = 0.0f;
        if (o.prim_id) o.prim_id[qi] = 0;
        return;
    }
    if (o.valid) o.valid[qi] = 1;
    if (o.p) gqd::st(o.p + (size_t)D * qi, best_p);
    if (o.d) o.d[qi] = best_d;
    if (o.prim_id) o.prim_id[qi] = P.ids[best_slot];
}

// Output record of a closest-silhouette query from the winning silhouette slot (-1: none).  The
// point and distance are recomputed exactly as the leaf test computed them.
template <int D>
__device__ __forceinline__ void sil_record(const EngineParams &P, uint32_t qi, const typename VT<D>::T &x,
                                           int best_slot) {
    using V = typename VT<D>::T;
    const gq_records &o = P.out;
    if (best_slot < 0) {
        if (o.valid) o.valid[qi] = 0;
        if (o.p) for (int j = 0; j < D; ++j) o.p[(size_t)D * qi + j] = 0.0f;
        if (o.d) o.d[qi] = 0.0f;
        if (o.prim_id) o.prim_id[qi] = 0;
        return;
    }
    const uint32_t s = (uint32_t)best_slot;
    V bp;
    float bd;
    if constexpr (D == 3) {
        const float4 s0 = P.prims[3 * s], s1 = P.prims[3 * s + 1];
        bd = closest_point_segment(x, v3(s0.x, s0.y, s0.z), v3(s0.w, s1.x, s1.y), bp);
    } else {
        const float4 s0 = P.prims[2 * s];
        bp = v2(s0.z, s0.w);
        bd = norm(x - bp);
    }
    if (o.valid) o.valid[qi] = 1;
    if (o.p) gqd::st(o.p + (size_t)D * qi, bp);
    if (o.d) o.d[qi] = bd;
    if (o.prim_id) o.prim_id[qi] = P.ids[s];
}

// ---------------------------------------------------------------------------------------------
// closest silhouette, lane per query: the same node tests as the packet pass, one query per lane
// over the Morton order with a per-lane LDS stack (one entry per binary level at most).  Cheaper
// than the packet pass when per-query work is small and culling is effective (2D silhouettes).
template <int D, bool Q>
__global__ void __launch_bounds__(kRayBlock) k_silhouette_lane(EngineParams P) {
    using V = typename VT<D>::T;
    GQ_LANE_STACK_DECL
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t qi = query_index<D>(P, i);
    const V x = ldv<D>(P.x + (size_t)D * qi);
    float xc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = comp<D>(x, k);
    float r = P.r_max ? P.r_max[qi] : P.r_max_scalar;
    int best_slot = -1;
    float d_root, d_unused;
    box_distance(ldv<D>(P.root_lo), ldv<D>(P.root_hi), x, d_root, d_unused);
    bool go = d_root <= r;   // reference: a query that misses the root box is not traversed
    uint32_t cur = P.root;
    while (go) {
        if (!ref_is_leaf(cur)) {
            PairChild c[2];
            load_pair<D, Q>(P.nodes, cur, c);
            bool h[2];
            float key[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float dm, dx;
                box_distance(ldv<D>(c[j].lo), ldv<D>(c[j].hi), x, dm, dx);
                bool hj = dm <= r;
                if (hj && c[j].ref_cone >= 0) {
                    const RefCone &rc = P.ref_cones[c[j].ref_cone];
                    float dr;
                    hj = snch_overlaps(ldv<D>(rc.lo), ldv<D>(rc.hi), ldv<D>(rc.axis), rc.half, rc.radius, x, r, dr);
                }
                if (hj && exact_cull<D>(c[j], xc)) hj = false;
                h[j] = hj;
                key[j] = dm;
            }
            if (h[0] && h[1]) {
                const int near = key[1] < key[0] ? 1 : 0;
                GQ_LPUSH(c[1 - near].child, key[1 - near]);
                cur = c[near].child;
                continue;
            }
            if (h[0] || h[1]) {
                cur = h[0] ? c[0].child : c[1].child;
                continue;
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
                    if (dd < r || (dd == r && best_slot >= 0 && (int)s < best_slot)) {
                        if (sil_edge_test(x, a, b, pt, s1, s2)) {
                            r = dd;
                            best_slot = (int)s;
                        }
                    }
                } else {
                    const float4 s0 = P.prims[2 * s], s1 = P.prims[2 * s + 1];
                    const V2 a = v2(s0.x, s0.y), b = v2(s0.z, s0.w), cc = v2(s1.x, s1.y);
                    const float dd = norm(x - b);
                    if ((dd < r || (dd == r && best_slot >= 0 && (int)s < best_slot)) &&
                        sil_vertex_pred_robust(x, a, b, cc)) {
                        r = dd;
                        best_slot = (int)s;
                    }
                }
            }
        }
        GQ_LPOP(r, cur, go);
    }
    sil_record<D>(P, qi, x, best_slot);
}

// ---------------------------------------------------------------------------------------------
// closest silhouette, wave per query: finishes the queries the packet pass deferred, with the same
// node tests (reference box and cone test, exact cone cull).  Frontier as in the closest-point
// pass: at most 128 entries per binary level.
constexpr int kFrontSil = 3072;

template <int D, bool Q>
__global__ void __launch_bounds__(kWaveBlock) k_silhouette_wave(EngineParams P) {
    using V = typename VT<D>::T;
    __shared__ uint2 front_all[kWaveBlock / 64][kFrontSil];
    uint2 *front = front_all[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const uint32_t total = uniform(P.defer_count[0]);
    for (;;) {
        // All lanes take part in the fetch (lane 0 adds 1, the others 0) and later all lanes store
        // the same record: no lane-dependent branch inside this loop, so the compiler cannot split
        // the wavefront across iterations.
        const uint32_t k = uniform(atomicAdd(P.defer_count + 1, lane == 0 ? 1u : 0u));
        if (k >= total) break;
        const uint4 item = P.deferred[k];
        const uint32_t qi = uniform(item.x);
        float r = uniformf(__uint_as_float(item.y));
        int best_slot = (int)uniform(item.z);
        const V x = ldv<D>(P.x + (size_t)D * qi);
        float xc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < D; ++j) xc[j] = comp<D>(x, j);
        int size = 0;
        if (r >= 0.0f) {
            front[0] = make_uint2(P.root, 0u);
            size = 1;
        }
        while (size > 0) {
            const int take = size < 64 ? size : 64;
            uint2 e = make_uint2(0u, __float_as_uint(kInf));
            if (lane < take) e = front[size - take + lane];
            size -= take;
            // opened children live in fixed slots (no dynamically indexed register arrays)
            uint2 ch0 = make_uint2(0u, 0u), ch1 = make_uint2(0u, 0u);
            bool h0 = false, h1 = false;
            float cand = kInf;
            int cand_slot = -1;
            if (lane < take && __uint_as_float(e.y) <= r) {
                if (!ref_is_leaf(e.x)) {
                    PairChild c[2];
                    load_pair<D, Q>(P.nodes, e.x, c);
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
                        const uint2 ent = make_uint2(c[j].child, __float_as_uint(dm));
                        if (j == 0) { h0 = h; ch0 = ent; }
                        else { h1 = h; ch1 = ent; }
                    }
                } else {
                    uint32_t first, count;
                    leaf_range(e.x, P.big, first, count);
                    for (uint32_t s = first; s < first + count; ++s) {
                        if constexpr (D == 3) {
                            const float4 s0 = P.prims[3 * s], s1 = P.prims[3 * s + 1], s2 = P.prims[3 * s + 2];
                            const V3 a = v3(s0.x, s0.y, s0.z), b = v3(s0.w, s1.x, s1.y);
                            V3 pt;
                            const float dd = closest_point_segment(x, a, b, pt);
                            if ((dd < r || (dd == r && best_slot >= 0 && (int)s < best_slot)) &&
                                (dd < cand || (dd == cand && (int)s < cand_slot))) {
                                if (sil_edge_test(x, a, b, pt, s1, s2)) {
                                    cand = dd;
                                    cand_slot = (int)s;
                                }
                            }
                        } else {
                            const float4 s0 = P.prims[2 * s], s1 = P.prims[2 * s + 1];
                            const V2 a = v2(s0.x, s0.y), b = v2(s0.z, s0.w), cc = v2(s1.x, s1.y);
                            const float dd = norm(x - b);
                            if ((dd < r || (dd == r && best_slot >= 0 && (int)s < best_slot)) &&
                                (dd < cand || (dd == cand && (int)s < cand_slot)) &&
                                sil_vertex_pred_robust(x, a, b, cc)) {
                                cand = dd;
                                cand_slot = (int)s;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const float oc = __shfl_xor(cand, o, 64);
                const int os = __shfl_xor(cand_slot, o, 64);
                if (oc < cand || (oc == cand && os < cand_slot)) {
                    cand = oc;
                    cand_slot = os;
                }
            }
            cand = uniformf(cand);
            cand_slot = (int)uniform((uint32_t)cand_slot);
            if (cand < r || (cand == r && best_slot >= 0 && cand_slot >= 0 && cand_slot < best_slot)) {
                r = cand;
                best_slot = cand_slot;
            }
            int tot;
            const int nh = (int)h0 + (int)h1;
            int pos = size + wave_exclusive_scan(nh, tot);
            tot = (int)uniform((uint32_t)tot);
            if (h0 && pos < kFrontSil) front[pos] = ch0;
            pos += (int)h0;
            if (h1 && pos < kFrontSil) front[pos] = ch1;
            size = size + tot < kFrontSil ? size + tot : kFrontSil;
        }
        sil_record<D>(P, qi, x, best_slot);
    }
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

__global__ void k_zero_u32(uint32_t *b, int n) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) b[t] = 0u;
}

constexpr int kBoundsBlocks = 256;

// Per-block partial bounds (order-preserving uint encoding) into part[block][6]; no atomics, so
// the result does not depend on scheduling and the pass can be recorded in a graph.
template <int D>
__global__ void __launch_bounds__(kBlock) k_bounds(const float *x, int n, uint32_t *part) {
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
    __shared__ uint32_t red[kBlock / 64][6];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        red[w][k] = k < D ? mn[k] : 0xFFFFFFFFu;   // every lane stores the same wave value
        red[w][3 + k] = k < D ? mx[k] : 0u;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int j = 1; j < kBlock / 64; ++j) v = k < 3 ? min(v, red[j][k]) : max(v, red[j][k]);
        part[blockIdx.x * 6 + k] = v;
    }
}

// Final bounds from the per-block partials: b[0..2] = min, b[3..5] = max.
__global__ void k_bounds_final(const uint32_t *part, int nblocks, uint32_t *b) {
    const int k = threadIdx.x;
    if (k >= 6) return;
    uint32_t v = k < 3 ? 0xFFFFFFFFu : 0u;
    for (int j = 0; j < nblocks; ++j) v = k < 3 ? min(v, part[j * 6 + k]) : max(v, part[j * 6 + k]);
    b[k] = v;
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
    static const char *names[2][3] = {
        {"engine_closest_point_2d", "engine_intersect_2d", "engine_silhouette_2d"},
        {"engine_closest_point_3d", "engine_intersect_3d", "engine_silhouette_3d"}};
    prof_begin(st);
    if (kind == kIntersect) {
        const unsigned grid = (unsigned)((P.n + kRayBlock - 1) / kRayBlock);
        const size_t lds = (size_t)P.stack_depth * kRayBlock * sizeof(uint2);
        hipLaunchKernelGGL((k_intersect<D, Q>), dim3(grid), dim3(kRayBlock), lds, st, P);
    } else {
        const unsigned grid = (unsigned)((P.n + kBlock - 1) / kBlock);
        if (kind == kClosestPoint) {
            if (P.budget > 0) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, P.defer_count, 2);
            hipLaunchKernelGGL((k_closest_point<D, Q>), dim3(grid), dim3(kBlock), 0, st, P);
            if (P.budget > 0)
                hipLaunchKernelGGL((k_closest_point_wave<D, Q>), dim3(kWavePassBlocks), dim3(kWaveBlock), 0, st, P);
        } else if (P.sil_lane) {
            const unsigned lgrid = (unsigned)((P.n + kRayBlock - 1) / kRayBlock);
            const size_t lds = (size_t)P.stack_depth * kRayBlock * sizeof(uint2);
            hipLaunchKernelGGL((k_silhouette_lane<D, Q>), dim3(lgrid), dim3(kRayBlock), lds, st, P);
        } else {
            if (P.budget > 0) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, P.defer_count, 2);
            hipLaunchKernelGGL((k_silhouette<D, Q>), dim3(grid), dim3(kBlock), 0, st, P);
            if (P.budget > 0)
                hipLaunchKernelGGL((k_silhouette_wave<D, Q>), dim3(kWavePassBlocks), dim3(kWaveBlock), 0, st, P);
        }
    }
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
    const size_t need = 256 + align256(kBoundsBlocks * 6 * 4) + 4 * nb + align256(temp);
    if ((e = hipMallocAsync(buf, need, st)) != hipSuccess) return e;
    uint8_t *base = (uint8_t *)*buf;
    uint32_t *bounds = (uint32_t *)base;
    uint32_t *part = (uint32_t *)(base + 256);
    base += 256 + align256(kBoundsBlocks * 6 * 4);
    uint32_t *k_in = (uint32_t *)base, *k_out = (uint32_t *)(base + nb);
    uint32_t *v_in = (uint32_t *)(base + 2 * nb), *v_out = (uint32_t *)(base + 3 * nb);
    void *tmp = base + 4 * nb;
    const int rgrid = (int)std::min<int64_t>(kBoundsBlocks, (n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_bounds<D>), dim3(rgrid), dim3(kBlock), 0, st, x, n, part);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, st, part, rgrid, bounds);
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
    P.stack_depth = 3 * w.wide_depth + 4;
    if (kind == kSilhouette) {
        // lane per query wins where culling leaves little work per query (2D); 3D interior queries
        // need the packet pass and its wave-per-query tail
        P.sil_lane = (w.traversal == 2 || (w.traversal == 0 && D == 2)) && w.bin_depth + 2 <= kRayStackMax;
        if (P.sil_lane) P.stack_depth = w.bin_depth + 2;
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
    void *buf = nullptr, *dbuf = nullptr;
    P.budget = 0;
    hipError_t e = sort_queries<D>(q.x, P.n, st, &P.perm, &buf);
    const bool defer = (kind == kClosestPoint && 256 * (w.wide_depth + 1) <= kFront) ||
                       (kind == kSilhouette && !P.sil_lane && 128 * (w.bin_depth + 2) <= kFrontSil);
    if (e == hipSuccess && defer && w.packet_budget >= 0) {
        e = hipMallocAsync(&dbuf, 256 + (size_t)P.n * sizeof(uint4), st);
        if (e == hipSuccess) {
            P.budget = w.packet_budget > 0 ? w.packet_budget
                       : kind == kSilhouette  ? kPacketBudgetSil
                       : D == 3               ? kPacketBudgetCp3
                                              : kPacketBudgetCp2;
            P.defer_count = (uint32_t *)dbuf;
            P.deferred = (uint4 *)((uint8_t *)dbuf + 256);
        }
    }
    if (e == hipSuccess)
        e = quantized ? launch_traversal<D, true>(kind, P, st) : launch_traversal<D, false>(kind, P, st);
    if (dbuf) {
        hipError_t f = hipFreeAsync(dbuf, st);
        if (e == hipSuccess) e = f;
    }
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
    if (kind == kSilhouette) return w.bin_depth + 2 <= kPacketDepth;
    if (kind == kClosestPoint) return 3 * w.wide_depth + 4 <= kPacketDepth;
    return 3 * w.wide_depth + 4 <= kRayStackMax;
}

hipError_t launch_engine(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q,
                         EngineWorkspace &ws, hipStream_t st) {
    if (q.n <= 0) return hipSuccess;
    return w.dim == 2 ? launch_dim<2>(kind, w, quantized, q, ws, st) : launch_dim<3>(kind, w, quantized, q, ws, st);
}

} // namespace gqa
