// Engine kernels: Morton-sorted queries, 4-wide BVH (float32 or 8/16-bit quantized child boxes)
// for the closest-point and ray queries, SNCH pair records for the silhouette query.
//
// Kernels run over a job table (one EngineParams per query batch) and a tile table (workgroup ->
// job, first query), so one launch serves a single call or a whole batch of scenes:
//  * closest point and 3D silhouettes: a packet pass (the 64 lanes of a wavefront walk the tree
//    together: wave-uniform node fetches, a per-wave stack in LDS, a child opened when any lane
//    needs it) with a step budget; wavefronts over budget defer their live queries to a persistent
//    wave-per-query pass (one wavefront expands 64 frontier entries of one query per round);
//  * rays and 2D silhouettes: one lane per query with a per-lane stack in LDS.
// Pruning is conservative: a node is skipped only when no primitive in it can beat the current
// answer, so results equal the reference traversal's up to exact ties (broken to the lowest
// primitive slot everywhere, so a query's result does not depend on its neighbours):
//  * closest point: squared box distance against (best (1 + 1e-6) + slack)^2, where the slack
//    scales with the coordinates, so float rounding never drops a box holding a primitive whose
//    float32 distance is equal or nearer;
//  * ray: slab test with a reciprocal direction widened so it never drops a box the reference's
//    division-based test keeps;
//  * silhouette: the reference's box and normal-cone tests (the cone test is evaluated with the
//    reference's float32 cone values wherever it can cull) plus a cull of subtrees that provably
//    contain no silhouette for the query (every predicate normal of the subtree lies in a cone whose
//    dot products with all view directions into the subtree share one sign).
// Leaf tests and output records use the reference formulas (gq_device.h); robust.h re-evaluates in
// FP64 the few float32 decisions rounding can flip.

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "gq_device.h"
#include "gq_internal.h"
#include "robust.h"
#include "wide.h"

namespace gqa {
namespace {

using namespace gqd;

constexpr int kBlock = 256;     // sort kernels: 4 wavefronts per block
// Packet kernels: one wavefront per block.  A block's wave slots and LDS are released only when its
// slowest wavefront ends, and packet wavefronts differ in length by up to 20x: with 4-wavefront
// blocks 3D closest point took 7.16 ms on 1M queries, with 1-wavefront blocks 6.93 ms (2-wavefront
// blocks 7.06 ms); 3D silhouettes 17.2 ms either way.
constexpr int kPacketBlock = 64;
constexpr int kRayBlock = 64;   // lane-per-query ray kernel: one wavefront per block
constexpr int kRayStackMax = 96;  // ray stack entries per lane (768 B of private memory)
constexpr int kWavePassBlocks = 2048;  // persistent grid of the full-frontier wave-per-query pass
constexpr int kXcds = 8;            // MI355X: 8 XCDs of 32 CUs, each with its own L2
// Packet steps per wavefront before its queries are deferred to the wave-per-query pass, by query
// kind and dimension (measured optimum on 1M uniform queries: 3D closest point on 327,680 triangles
// with 24-primitive leaves 2.66 ms at 300 vs 2.70 at 350, 2.78 at 250 and 3.08 at 200; 2D closest
// point on 65,536 segments with 16 1.96 ms at 300 vs 2.05 at 450 and 2.25 at 200; 3D silhouettes, with near-last appends in the wave pass,
// 29.05 ms at 128 vs 29.07 at 64, 29.30 at 250 and 31.47 at 750: the wave pass is the better
// place for nearly all of their work; 2D silhouettes in packet mode (65,536 segments) 3.28 ms at
// 750 vs 3.85 at 2,000 and 8.93 at 128: a 2D query has too little work for a whole wavefront).
// A scene or job option overrides it.
constexpr int kPacketBudgetCp3 = 300, kPacketBudgetCp2 = 300, kPacketBudgetSil3 = 128, kPacketBudgetSil2 = 750;
// A launch with few queries leaves most of the GPU idle, and then a packet wavefront's walk -- up to
// its budget of dependent node fetches -- is the latency of the whole call, while the wave-per-query
// pass puts 64 lanes on each query.  So below a launch size the default budget drops: measured on
// the 327,680-triangle icosphere with 24-primitive leaves (3D closest point, ms per call, budget 16
// vs the default): 300 queries 0.57 vs 1.18, 16,384 0.72 vs 1.44, 65,536 1.40 vs 1.70, 262,144
// 3.60 vs 1.79; 65,536-segment star with 16-primitive leaves: 16,384 0.45 vs 1.05, 65,536 0.99 vs
// 1.08, 262,144 3.32 vs 1.38; 3D silhouettes at budget 1: 300 0.42 vs 0.25, 65,536 1.53 vs 1.36.
// Below 16,384 queries 8 steps do better still (300 queries 0.32 vs 0.37 ms, 8,000 0.57 vs 0.60;
// 32,768: 1.00 vs 0.79, so 16 from there).
constexpr int kSmallBudgetCp = 16, kTinyBudgetCp = 8, kSmallBudgetSil3 = 1;
constexpr int64_t kTinyLaunchCp = 1 << 14;
constexpr int64_t kSmallLaunchCp3 = 1 << 17, kSmallLaunchCp2 = 1 << 17, kSmallLaunchSil3 = 1 << 18;
// When other launches share the device (gq_job.shared; a serving loop's stream slots) the GPU is not
// idle and total work decides throughput and queueing: shared groups are budgeted as full launches
// (served workload, 2 + 1 stream slots: 140 vs 131 Mq/s saturated, p99 at 100 Mq/s offered 33.4 vs
// 43.5 ms).
constexpr int64_t kFullLaunch = kSmallLaunchSil3;

// The budgets above are the optimum for scenes of about 65k-330k primitives.  The best budget grows
// with the tree (a packet needs more steps to finish) and with the query density rho = queries per
// primitive (denser packets share more of their walk).  Measured with one-job launches of uniform
// queries (displaced icospheres s = 3..8, stars of 1k-1M segments; 262k, 1M and 4M queries; ms):
//   3D closest point: s <= 5 (<= 20k triangles) 128 beats 350 by 5-25% (s5, 1M: 1.14 vs 1.21);
//     s6 (82k) 200 (1M: 1.63 vs 1.75 at 300); s7 300, and 450 at rho >= 8 (4M: 6.83 vs 7.11);
//     s8 (1.3M triangles) 1,000 beats 350 (1M: 6.35 vs 8.56, 4M: 14.2 vs 17.0).
//   2D closest point: <= 16k segments 128 beats 300 (16k, 262k queries: 0.68 vs 0.93); 65k
//     segments 300, and 450 at rho >= 32 (4M: 4.82 vs 4.89); 262k segments 700 (1M: 4.17 vs 6.34
//     at 300); from 512k segments with rho >= 1, 1,000 (512k segments, 1M: 6.56 vs 10.8; 1M
//     segments, 4M: 23.7 vs 49.5).
//   3D silhouettes: with rho >= 128 nearly every packet finishes and budget 3,000 wins (s4, 1M:
//     4.96 vs 11.95; s5, 4M: 26.3 vs 46.8); for 24 <= rho < 128 lane per query wins (s4 at 262k
//     queries 2.51 vs 3.34, s5 at 1M 9.35 vs 11.98, s6 at 4M 42.9 vs 48.2); below that 128.
//   2D silhouettes: lane per query keeps its stack in LDS, which a deep tree makes costly: from
//     393k segments packets win (512k segments, 1M queries: 8.86 at 3,000 vs 11.26 per lane; 1M
//     segments, 262k queries: 8.11 at 64 vs 27.4), with 3,000 at rho >= 1 and 128 below.
constexpr int kLongBudget = 3000;
constexpr int kPacketBudgetSmallTree = 128;
constexpr int kPacketBudgetLargeTree = 1000;
constexpr int kPacketBudgetMidTree = 200, kPacketBudgetCp2Large = 700;
constexpr int64_t kCp3SmallTree = 1 << 16, kCp3MidTree = 1 << 17, kCp3LargeTree = 1 << 19;
constexpr int64_t kCp2SmallTree = 1 << 14, kCp2MidTree = 1 << 17, kCp2LargeTree = 1 << 19;
constexpr double kCp3DenseRho = 8.0, kCp2DenseRho = 32.0;
constexpr int64_t kSil2PacketTree = 3 << 17;
constexpr double kSil3LaneRho = 24.0, kSil3DenseRho = 128.0;

// Default packet budget of a job of n queries on a scene of `prims` primitives, in a launch group
// of `group_q` queries (no job or scene override).
int default_budget(int kind, int dim, int64_t group_q, int64_t n, int64_t prims) {
    const double rho = (double)n / (double)std::max<int64_t>(prims, 1);
    if (kind == kSilhouette) {
        if (dim == 3) {
            if (group_q < kSmallLaunchSil3) return kSmallBudgetSil3;
            return rho >= kSil3DenseRho ? kLongBudget : kPacketBudgetSil3;
        }
        if (prims >= kSil2PacketTree) return rho >= 1.0 ? kLongBudget : kPacketBudgetSmallTree;
        return kPacketBudgetSil2;
    }
    if (group_q < kTinyLaunchCp) return kTinyBudgetCp;
    if (dim == 3) {
        if (group_q < kSmallLaunchCp3) return kSmallBudgetCp;
        if (prims < kCp3SmallTree) return kPacketBudgetSmallTree;
        if (prims < kCp3MidTree) return kPacketBudgetMidTree;
        if (prims < kCp3LargeTree) return rho >= kCp3DenseRho ? kPacketBudgetCp3 * 3 / 2 : kPacketBudgetCp3;
        return kPacketBudgetLargeTree;
    }
    if (group_q < kSmallLaunchCp2) return kSmallBudgetCp;
    if (prims <= kCp2SmallTree) return kPacketBudgetSmallTree;
    if (prims < kCp2MidTree) return rho >= kCp2DenseRho ? kPacketBudgetCp2 * 3 / 2 : kPacketBudgetCp2;
    if (prims < kCp2LargeTree) return kPacketBudgetCp2Large;
    return rho >= 1.0 ? kPacketBudgetLargeTree : kPacketBudgetCp2;
}

// Default silhouette traversal of a job (no job or scene override): lane per query for 2D scenes
// below kSil2PacketTree primitives and for full-size 3D launches at a middle query density,
// packets otherwise (measurements above).
bool default_sil_lane(int dim, int64_t n, int64_t prims) {
    if (dim == 2) return prims < kSil2PacketTree;
    const double rho = (double)n / (double)std::max<int64_t>(prims, 1);
    return n >= kSmallLaunchSil3 && rho >= kSil3LaneRho && rho < kSil3DenseRho;
}
constexpr int kSortMin = 8192;   // smaller batches are traversed in input order
constexpr float kInf = __builtin_inff();
// Absolute rounding slack per unit of coordinate magnitude: 32 float32 ulps of 1.  The float32
// distance to a triangle or segment rebuilds the nearest point in absolute coordinates, so its
// rounding error grows with |x| and the scene's coordinates, not with the distance itself; every
// pruning test that must never drop a primitive the float32 leaf test would accept widens its bound
// by kAbsSlack * (|x|_1 + S_1), where S_1 bounds |p|_1 over the scene (from its root box).  The
// error terms: the prefilter's fma chain and stored plane (4 ulps of |x|_1 + S_1), the nearest
// point's reconstruction (12 ulps of S_1), the exact cull's view vector and its cross product
// (about 20 ulps), each well inside 32.
constexpr float kAbsSlack = 32.0f * 5.9604644775390625e-08f;

struct EngineParams {
    const void *nodes;        // BVH4 (FNode / QNode) or SNCH pairs (SPairF / SPairQ)
    const void *cyl;          // closest point: CBound per BVH4 node (the node format's own)
    const float4 *prims;
    const int32_t *ids;
    const RefCone *ref_cones;
    const uint32_t *big;
    uint32_t root;
    float root_lo[3], root_hi[3];
    const float *x, *v, *r_max;
    float r_max_scalar;
    uint32_t *perm;           // Morton-cell order of the queries (null: input order)
    uint32_t *cell, *counts, *part, *bounds;   // counting-sort scratch (per job)
    int cell_bits;            // Morton cell bits per axis
    int n_cells;              // 2^(cell_bits * dimension)
    int cell_shift;           // cell = fine Morton code >> cell_shift
    uint32_t job;             // index in the job table (carried by deferred entries)
    int n;
    int stack_depth;          // lane-stack entries (ray kernel)
    int budget;               // packet steps before a wave hands its queries to the wave-per-query pass (0 = never)
    uint32_t *defer_count;    // [0] deferred queries, [1] work counter of the wave-per-query pass
    uint4 *deferred;          // (query index, best distance bits, best slot, job)
    int sil_lane;             // silhouettes: lane-per-query kernel instead of packet + wave pass
    const Seed *seed;         // seed cut the first wave-per-query pass starts from (n_seed = 0: the root)
    int n_seed;
    gq_records out;
};

// Device view of a job: the same fields with every pointer in the global address space.  Pointers
// read from a job table are generic to the compiler, which would access them with flat
// instructions; the explicit cast of each one makes every access a global load or store.
#define GQ_G __attribute__((address_space(1)))
struct GRecords {
    GQ_G uint8_t *valid;
    GQ_G float *p, *n, *uv, *d;
    GQ_G int32_t *prim_id;
    GQ_G uint8_t *on_boundary;
};
struct GP {
    const GQ_G uint8_t *nodes;
    const GQ_G uint8_t *cyl;
    const GQ_G float4 *prims;
    const GQ_G int32_t *ids;
    const GQ_G RefCone *ref_cones;
    const GQ_G uint32_t *big;
    uint32_t root;
    float root_lo[3], root_hi[3];
    const GQ_G float *x, *v, *r_max;
    float r_max_scalar;
    const GQ_G uint32_t *perm;
    uint32_t job;
    int n;
    int stack_depth;
    int budget;
    GQ_G uint32_t *defer_count;
    GQ_G uint4 *deferred;
    const GQ_G Seed *seed;
    int n_seed;
    GRecords out;
};
template <typename T> __device__ __forceinline__ GQ_G T *gcast(T *p) { return (GQ_G T *)p; }
typedef float gq_v4f __attribute__((ext_vector_type(4)));
typedef unsigned int gq_v4u __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg(const GQ_G float4 *p) {
    const gq_v4f v = *(const GQ_G gq_v4f *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg(GQ_G uint4 *p, uint4 v) { *(GQ_G gq_v4u *)p = gq_v4u{v.x, v.y, v.z, v.w}; }
// Global-pointer forms of the generic helpers (the cast back to a generic pointer keeps a global
// source, so the accesses stay global instructions after inlining).
template <int D> __device__ __forceinline__ typename VT<D>::T ldv(const GQ_G float *p) { return gqd::ldv<D>((const float *)p); }
template <int D> __device__ __forceinline__ typename VT<D>::T ldv(const float *p) { return gqd::ldv<D>(p); }
__device__ __forceinline__ void st(GQ_G float *p, const V3 &v) { gqd::st((float *)p, v); }
__device__ __forceinline__ void st(GQ_G float *p, const V2 &v) { gqd::st((float *)p, v); }

__device__ __forceinline__ RefCone ldg(const GQ_G RefCone *p) {
    const GQ_G float4 *f = (const GQ_G float4 *)p;
    const float4 a = ldg(f), b = ldg(f + 1), c = ldg(f + 2), d = ldg(f + 3);
    RefCone r;
    r.axis[0] = a.x; r.axis[1] = a.y; r.axis[2] = a.z;
    r.half = a.w; r.radius = b.x;
    r.lo[0] = b.y; r.lo[1] = b.z; r.lo[2] = b.w;
    r.hi[0] = c.x; r.hi[1] = c.y; r.hi[2] = c.z;
    r.sin_h = c.w; r.cos_h = d.x;
    return r;
}


__device__ __forceinline__ GP gview(const EngineParams &in) {
    GP P;
    P.nodes = gcast((const uint8_t *)in.nodes);
    P.cyl = gcast((const uint8_t *)in.cyl);
    P.prims = gcast(in.prims);
    P.ids = gcast(in.ids);
    P.ref_cones = gcast(in.ref_cones);
    P.big = gcast(in.big);
    P.root = in.root;
    for (int k = 0; k < 3; ++k) { P.root_lo[k] = in.root_lo[k]; P.root_hi[k] = in.root_hi[k]; }
    P.x = gcast(in.x);
    P.v = gcast(in.v);
    P.r_max = gcast(in.r_max);
    P.r_max_scalar = in.r_max_scalar;
    P.perm = gcast((const uint32_t *)in.perm);
    P.job = in.job;
    P.n = in.n;
    P.stack_depth = in.stack_depth;
    P.budget = in.budget;
    P.defer_count = gcast(in.defer_count);
    P.deferred = gcast(in.deferred);
    P.seed = gcast(in.seed);
    P.n_seed = in.n_seed;
    P.out.valid = gcast(in.out.valid);
    P.out.p = gcast(in.out.p);
    P.out.n = gcast(in.out.n);
    P.out.uv = gcast(in.out.uv);
    P.out.d = gcast(in.out.d);
    P.out.prim_id = gcast(in.out.prim_id);
    P.out.on_boundary = gcast(in.out.on_boundary);
    return P;
}

// Absolute rounding slack of a query at xc against the job's scene (kAbsSlack above).
template <int D> __device__ __forceinline__ float coord_slack(const GP &P, const float *xc) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < D; ++k) s += fabsf(xc[k]) + fmaxf(fabsf(P.root_lo[k]), fabsf(P.root_hi[k]));
    return kAbsSlack * s;
}
// Squared pruning bound for a closest-point answer at distance d: a child or triangle whose
// squared lower bound exceeds this cannot hold a point whose float32 distance is below or equal to d.
__device__ __forceinline__ float cp_bound2(float d, float slack) {
    const float b = fmaf(d, 1.000001f, slack);
    return b * b;
}

// Lower bound on the distance from p to triangle abc that decides whether the exact leaf test can
// matter: its circumscribing disk (centre m in the free lanes of a, b, c; unit normal n and radius
// in pl), sqrt(h^2 + max(|q| - rad, 0)^2) with h = (p - m).n and q = (p - m) - h n, squared, against
// bound2.  Near the true distance wherever the query is not right above the triangle: a query at
// distance d from a smooth surface sees many triangles whose box or supporting plane is within d,
// and the disk rejects nearly all of them (on a 327,680-triangle sphere-like mesh, 4.7 triangles per
// uniform query pass at the final distance, against 239 for the box and plane together).
__device__ __forceinline__ bool tri_may_beat(const float *pc, const float4 &a, const float4 &b, const float4 &c,
                                             const float4 &pl, float bound2) {
    const float d0 = pc[0] - a.w, d1 = pc[1] - b.w, d2 = pc[2] - c.w;
    const float h = fmaf(d2, pl.z, fmaf(d1, pl.y, d0 * pl.x));
    const float q0 = fmaf(-h, pl.x, d0), q1 = fmaf(-h, pl.y, d1), q2 = fmaf(-h, pl.z, d2);
    const float m = fmaxf(__builtin_amdgcn_sqrtf(fmaf(q2, q2, fmaf(q1, q1, q0 * q0))) - pl.w, 0.0f);
    return fmaf(h, h, m * m) <= bound2;
}

// ---- lane-per-query stacks.  In LDS: depth chosen per launch from the tree depth, one column per
// lane of a [depth][blockDim] dynamic shared array.  In private (scratch) memory: a per-lane array
// of the largest bound, so LDS does not limit how many wavefronts a CU holds; only the top entries
// are touched and they stay in the L1 cache.  Measured on 1M queries: 3D rays against 327,680
// triangles 0.50 ms private vs 0.68 ms in LDS (2 wavefronts per SIMD); 3D silhouettes per lane
// 44.3 vs 51.9 ms (20,480 triangles: 15.7 vs 16.2 ms); 2D silhouettes (65,536 segments) 2.82 ms
// private vs 2.73 in LDS.  GQ_PRIVATE_STACK_DECL(N) declares a private stack of N entries;
// GQ_LPUSH / GQ_LPOP work on either kind through lane_stack, lane_col and lane_stride.
#define GQ_PRIVATE_STACK_DECL(N)               \
    uint2 lane_stack[N];                       \
    const int lane_col = 0;                    \
    const int lane_stride = 1;                 \
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

// Child box centres and half-extents of a BVH4 node, for the closest-point tests.  Quantized
// nodes form both from the integer sum and difference of a child's quantized corners (scalar
// integer work): centre = fma(qlo + qhi, 2^(e - 1), origin), half-extent = (qhi - qlo) * 2^(e - 1);
// float32 nodes take 0.5f * (lo + hi) and 0.5f * (hi - lo).  The host builds the bounding
// cylinders about exactly these centres.
template <int D> struct NodeC4 {
    float c[D][4], h[D][4];
    uint32_t child[4];
    int n;
};
template <int D, bool Q>
__device__ __forceinline__ void load_node_c(const void *base, uint32_t idx, NodeC4<D> &nd) {
    if constexpr (Q && D == 3) {
        const uint8_t *p = reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64;
        const uint4 w0 = ld16(p, 0), w1 = ld16(p, 1), w2 = ld16(p, 2), w3 = ld16(p, 3);
        const float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
        nd.n = (int)(w0.w >> 24);
        const uint32_t qlo[3] = {w1.x, w1.y, w1.z}, qhi[3] = {w1.w, w2.x, w2.y};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float hs = ldexpf(1.0f, sbyte(w0.w, k) - 1);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t a = (qlo[k] >> (8 * j)) & 0xFFu, b = (qhi[k] >> (8 * j)) & 0xFFu;
                nd.c[k][j] = fmaf((float)(a + b), hs, org[k]);
                nd.h[k][j] = (float)(b - a) * hs;
            }
        }
        nd.child[0] = w2.z; nd.child[1] = w2.w; nd.child[2] = w3.x; nd.child[3] = w3.y;
    } else if constexpr (Q && D == 2) {
        const uint8_t *p = reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64;
        const uint4 w0 = ld16(p, 0), w1 = ld16(p, 1), w2 = ld16(p, 2), w3 = ld16(p, 3);
        const float org[2] = {__uint_as_float(w0.x), __uint_as_float(w0.y)};
        nd.n = (int)((w0.z >> 16) & 0xFFu);
        const uint32_t q[8] = {w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float hs = ldexpf(1.0f, sbyte(w0.z, k) - 1);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t a = (q[k * 2 + j / 2] >> (16 * (j & 1))) & 0xFFFFu;
                const uint32_t b = (q[4 + k * 2 + j / 2] >> (16 * (j & 1))) & 0xFFFFu;
                nd.c[k][j] = fmaf((float)(a + b), hs, org[k]);
                nd.h[k][j] = (float)(b - a) * hs;
            }
        }
        nd.child[0] = w2.w; nd.child[1] = w3.x; nd.child[2] = w3.y; nd.child[3] = w3.z;
    } else {
        Node4<D, Q> b;
        load_node4<D, Q>(base, idx, b);
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                nd.c[k][j] = 0.5f * (b.lo[k][j] + b.hi[k][j]);
                nd.h[k][j] = 0.5f * (b.hi[k][j] - b.lo[k][j]);
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) nd.child[j] = b.child[j];
        nd.n = b.n;
    }
}

// Bounding cylinders of a BVH4 node's children (CBound<D>, the first 48 bytes of its 64): every
// field is the upper half of a float32 pattern, widened with a shift (scalar instructions when the
// record is wave-uniform).
template <int D> struct Cyl4 {
    float ax[D][4], lo[4], hi[4], rad[4];
};
template <int D>
__device__ __forceinline__ void load_cyl4(const void *base, uint32_t idx, Cyl4<D> &cy) {
    Words<12> w;
    w.load(reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64);
    auto half = [&](int i) {
        const uint32_t word = w.w[i >> 1];
        return __uint_as_float((i & 1) ? (word & 0xFFFF0000u) : (word << 16));
    };
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int k = 0; k < D; ++k) cy.ax[k][j] = half(4 * k + j);
        cy.lo[j] = half(4 * D + j);
        cy.hi[j] = half(4 * D + 4 + j);
        cy.rad[j] = half(4 * D + 8 + j);
    }
}
// Squared lower bound on the distance from x to anything below child j: x against the child's
// bounding cylinder about its box centre c (the same float32 centre the host used), s = (x - c).u
// the position along the axis (returned for the ordering key).  In 2D the child box is tested as
// well: a run of a polyline's segments can bend through a wide angle, and the box then prunes what
// its cylinder cannot (65,536-segment star: 2.25 ms with both tests, 3.22 with the cylinder alone;
// in 3D the box adds nothing, 3.47 vs 3.53 ms).
template <int D>
__device__ __forceinline__ float cyl_lb2(const NodeC4<D> &nd, const Cyl4<D> &cy, int j, const float *x, float &s) {
    float dv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) dv[k] = x[k] - nd.c[k][j];
    s = dv[0] * cy.ax[0][j];
#pragma unroll
    for (int k = 1; k < D; ++k) s = fmaf(dv[k], cy.ax[k][j], s);
    float q2 = 0.0f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float q = fmaf(-s, cy.ax[k][j], dv[k]);
        q2 = fmaf(q, q, q2);
    }
    const float sd = fmaxf(fmaxf(cy.lo[j] - s, s - cy.hi[j]), 0.0f);
    const float m = fmaxf(__builtin_amdgcn_sqrtf(q2) - cy.rad[j], 0.0f);
    const float lb2 = fmaf(sd, sd, m * m);
    if constexpr (D == 2) {
        float d2 = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float o = fmaxf(fabsf(dv[k]) - nd.h[k][j], 0.0f);
            d2 = fmaf(o, o, d2);
        }
        return fmaxf(lb2, d2);
    }
    return lb2;
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

template <int D> __device__ __forceinline__ uint32_t query_index(const GP &P, int i) {
    return P.perm ? P.perm[i] : (uint32_t)i;
}

// Output record of a closest-point query from the winning primitive slot (reference formulas).
template <int D>
__device__ __forceinline__ void cp_record(const GP &P, uint32_t qi, const typename VT<D>::T &p, int best_slot) {
    using V = typename VT<D>::T;
    const GRecords &o = P.out;
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
        const float4 a4 = ldg(P.prims + (4 * s)), b4 = ldg(P.prims + (4 * s + 1)), c4 = ldg(P.prims + (4 * s + 2));
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
        const float4 ab = ldg(P.prims + (s));
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
__device__ __forceinline__ void defer_wave(const GP &P, bool live, uint32_t qi, float bound, int slot) {
    const unsigned long long m = __ballot(live);
    const uint32_t nlive = (uint32_t)__popcll(m);
    // every lane takes part in the atomic (lane 0 adds the count) so no lane-dependent branch
    // precedes the return
    const uint32_t base = uniform(atomicAdd((uint32_t *)P.defer_count, (threadIdx.x & 63) == 0 ? nlive : 0u));
    if (live) {
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull));
        stg(P.deferred + base + rank, make_uint4(qi, __float_as_uint(bound), (uint32_t)slot, P.job));
    }
}

// Number of set bits of a wave mask in the lanes below this one (mbcnt): with one ballot per
// child slot, the append position of a lane's children is the sum over slots, and the round's
// total the sum of the masks' popcounts -- no shuffles through LDS.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Packet leaf work for triangles, queued.  The wave opens a child when any lane needs it, but a
// lane only takes part below a child whose bounds it passed itself (its bit in the child's lane
// mask, which the wave stack carries with the reference), and with the cylinder bounds few lanes
// need any one leaf (on the 327,680-triangle icosphere a wavefront of uniform queries visits about
// 190 leaves while each of its queries needs about 7).  So a leaf visit appends one (triangle, lane)
// pair per triangle and needing lane to an LDS queue, and the queue is drained a full round at a
// time: lane t takes pair t, reads the owner's query and its bound at the drain from LDS and the
// triangle from memory, runs the disk prefilter and the exact test, and folds (distance, slot) into
// the owner's LDS minimum (one 64-bit key: distance bits, which order like unsigned integers for
// d >= 0, then slot, so ties go to the lowest slot exactly as a sequential scan with that rule).
// Each lane then merges its minimum into its answer.  A drain later than the leaf's visit only
// delays bound updates (more tests, never fewer), and the result does not depend on the order of
// the tests.  Drain threshold, measured on 1M uniform queries (3D: 327,680-triangle icosphere; 2D:
// 65,536-segment star): 64 pairs 3.72 / 2.40 ms, 128 3.56 / 2.35, 256 3.57 / 2.30, 448 3.63 / 2.34
// (draining at once while some lane has no answer yet was no better: 3.61 / 2.31 at 256).
constexpr int kQueueCap = 512;      // queued pairs at most
constexpr int kQueueRound = 256;    // a drain starts once this many pairs are queued
struct PacketLds {
    float q[64][3];                 // the lanes' query points (2D: the first two)
    float b2[64];                   // the lanes' squared pruning bounds at the drain
    unsigned long long mn[64];      // per-lane minimum of a drain: (distance bits << 32) | slot
    uint32_t pair[kQueueCap];       // (triangle slot << 6) | lane
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D>
__device__ __forceinline__ void cp_drain(const GP &P, PacketLds &L, int &qn, bool live, float slack, float &best,
                                         int &best_slot, float &bound2) {
    const int lane = threadIdx.x & 63;
    L.b2[lane] = bound2;
    L.mn[lane] = ~0ull;
    wave_sync_lds();
    for (int b0 = 0; b0 < qn; b0 += 64) {
        const int k = b0 + lane;
        if (k < qn) {
            const uint32_t e = L.pair[k];
            const uint32_t owner = e & 63u, s = e >> 6;
            const float q[3] = {L.q[owner][0], L.q[owner][1], L.q[owner][2]};
            if constexpr (D == 3) {
                const float4 a = ldg(P.prims + (4 * s)), b = ldg(P.prims + (4 * s + 1)), c = ldg(P.prims + (4 * s + 2));
                const float4 pl = ldg(P.prims + (4 * s + 3));
                if (tri_may_beat(q, a, b, c, pl, L.b2[owner])) {
                    V3 pt;
                    V2 uv;
                    const float dd = closest_point_triangle(v3(q[0], q[1], q[2]), v3(a.x, a.y, a.z), v3(b.x, b.y, b.z),
                                                            v3(c.x, c.y, c.z), pt, uv);
                    atomicMin(&L.mn[owner], ((unsigned long long)__float_as_uint(dd) << 32) | s);
                }
            } else {
                // a segment's exact test costs about what a prefilter would
                const float4 ab = ldg(P.prims + (s));
                V2 pt;
                const float dd = closest_point_segment(v2(q[0], q[1]), v2(ab.x, ab.y), v2(ab.z, ab.w), pt);
                atomicMin(&L.mn[owner], ((unsigned long long)__float_as_uint(dd) << 32) | s);
            }
        }
    }
    wave_sync_lds();
    const unsigned long long mk = L.mn[lane];
    if (live && mk != ~0ull) {
        const float dd = __uint_as_float((uint32_t)(mk >> 32));
        const int s = (int)(uint32_t)mk;
        if (dd < best || (dd == best && s < best_slot)) {
            best = dd;
            best_slot = s;
            bound2 = cp_bound2(dd, slack);
        }
    }
    qn = 0;
}

// Appends the pairs of a leaf's triangles [first, first + count) with the lanes in `act`, draining
// the queue whenever the next piece would not fit; pairs of one triangle are adjacent.
template <int D>
__device__ __forceinline__ void cp_enqueue(const GP &P, PacketLds &L, int &qn, uint32_t first, uint32_t count,
                                           bool act, bool live, float slack, float &best, int &best_slot,
                                           float &bound2) {
    const int lane = threadIdx.x & 63;
    const unsigned long long m = __ballot(act);
    const int np = __popcll(m);
    if (np == 0) return;
    const int rank = lanes_below(m);
    const uint32_t per = (uint32_t)(kQueueCap / np);   // triangles whose pairs fit an empty queue
    for (uint32_t c0 = first; c0 < first + count;) {
        const uint32_t ch = min(per, first + count - c0);
        if (qn + (int)ch * np > kQueueCap) cp_drain<D>(P, L, qn, live, slack, best, best_slot, bound2);
        if (act)
            for (uint32_t t = 0; t < ch; ++t) L.pair[qn + (int)t * np + rank] = ((c0 + t) << 6) | (uint32_t)lane;
        qn += (int)ch * np;
        c0 += ch;
    }
    if (qn >= kQueueRound) cp_drain<D>(P, L, qn, live, slack, best, best_slot, bound2);
}

template <int D, bool Q>
__device__ __forceinline__ void cp_packet(const GP &P, int base) {
    using V = typename VT<D>::T;
    __shared__ uint32_t wstack[kPacketBlock / 64][kPacketDepth];
    __shared__ unsigned long long wmask[kPacketBlock / 64][kPacketDepth];
    __shared__ PacketLds wlds[kPacketBlock / 64];
    uint32_t *wst = wstack[threadIdx.x >> 6];
    unsigned long long *wmk = wmask[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int i = base + threadIdx.x;
    const bool live = i < P.n;
    const uint32_t qi = live ? query_index<D>(P, i) : 0u;
    V p;
    if (live) p = ldv<D>(P.x + (size_t)D * qi);
    else p = ldv<D>(P.x);   // any valid point; the lane's result is discarded
    float pc[D], cen[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pc[k] = comp<D>(p, k);
    packet_centroid<D>(live, pc, cen);
    const float slack = coord_slack<D>(P, pc);
    float best = kInf, bound2 = live ? kInf : -1.0f;   // a dead lane never opens anything
    int best_slot = -1;
    int sp = 0;
    int steps = 0;
    int qn = 0;   // queued leaf pairs
#pragma unroll
    for (int k = 0; k < D; ++k) wlds[threadIdx.x >> 6].q[lane][k] = pc[k];
    // `act`: this lane passed the bounds of every node on the way to the current one
    bool act = live;
    uint32_t cur = uniform(P.root);
    for (;;) {
        if (P.budget > 0 && ++steps > P.budget) {
            // Too much work for one wavefront (queries deep inside a closed surface see most of it
            // at nearly one distance): hand the queries, with the answer so far as the starting
            // bound, to the wave-per-query pass so they spread over many wavefronts.
            if (qn > 0) cp_drain<D>(P, wlds[threadIdx.x >> 6], qn, live, slack, best, best_slot, bound2);
            defer_wave(P, live, qi, best, best_slot);
            return;
        }
        if (!ref_is_leaf(cur)) {
            NodeC4<D> nd;
            load_node_c<D, Q>((const void *)P.nodes, cur, nd);
            Cyl4<D> cy;
            load_cyl4<D>((const void *)P.cyl, cur, cy);
            float key[4];
            uint32_t ref[4];
            // children any acting lane must visit: within its bound of the child's bounding cylinder
            // (the child box adds nothing to it: the same node visits per query on the
            // 327,680-triangle icosphere); the ordering key (distance from the packet centroid) is
            // computed only when more than one is open (3D: 7.41 vs 7.45 ms; float32 nodes 7.18 vs
            // 7.32 ms)
            unsigned long long hm[4];
            int nh = 0;
            int only = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float sx;
                const float lb2 = cyl_lb2<D>(nd, cy, j, pc, sx);
                hm[j] = __ballot(act && (j < nd.n) && (lb2 <= bound2));
                nh += (int)(hm[j] != 0ull);
                ref[j] = nd.child[j];
                if (hm[j] != 0ull) only = j;
            }
            if (nh == 1) {
                cur = uniform(ref[only]);
                act = (hm[only] >> lane) & 1ull;
                continue;
            }
            if (nh > 1) {
                uint32_t slot[4] = {0u, 1u, 2u, 3u};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    key[j] = kInf;
                    if (hm[j] != 0ull) {
                        // box distance of the centroid, or its distance behind the cylinder's slab
                        // (the side opposite the normals) where larger: the slab's near side would
                        // order the tilted patches in front of a convex surface first
                        float c2 = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; ++k) {
                            const float oc = fmaxf(fabsf(cen[k] - nd.c[k][j]) - nd.h[k][j], 0.0f);
                            c2 = fmaf(oc, oc, c2);
                        }
                        float sc;
                        (void)cyl_lb2<D>(nd, cy, j, cen, sc);
                        const float behind = fmaxf(cy.lo[j] - sc, 0.0f);
                        key[j] = uniformf(fmaxf(c2, behind * behind));
                    }
                }
                sort4(key, slot);
#pragma unroll
                for (int j = 3; j >= 1; --j)
                    if (j < nh && sp < kPacketDepth) {
                        wst[sp] = ref[slot[j]];
                        wmk[sp] = hm[slot[j]];
                        ++sp;
                    }
                cur = uniform(ref[slot[0]]);
                act = (hm[slot[0]] >> lane) & 1ull;
                continue;
            }
        } else {
            uint32_t first, count;
            leaf_range(cur, (const uint32_t *)P.big, first, count);
            first = uniform(first);
            count = uniform(count);
            cp_enqueue<D>(P, wlds[threadIdx.x >> 6], qn, first, count, act, live, slack, best, best_slot, bound2);
        }
        if (sp == 0) break;
        --sp;
        cur = uniform(wst[sp]);
        act = (wmk[sp] >> lane) & 1ull;
    }
    if (qn > 0) cp_drain<D>(P, wlds[threadIdx.x >> 6], qn, live, slack, best, best_slot, bound2);
    if (!live) return;
    cp_record<D>(P, qi, p, best_slot);
}

// ---------------------------------------------------------------------------------------------
// closest point, wave per query: finishes the queries the packet pass deferred.  The 64 lanes
// expand the top 64 frontier entries per round and append the opened children in lane order, so
// the frontier stays sorted by depth; every append at depth d happens after all older depth-d
// entries were taken, hence at most 256 entries per depth level and 256 * (levels) in total.
// Leaf candidates are reduced across the wave each round.
constexpr int kWaveBlock = 64;      // one wavefront per block: the frontier is sized per launch
constexpr int kFrontMax = 8192;     // frontier entries at most (64 KB of LDS)
// The wave-per-query pass runs twice: first with a small frontier, so that LDS does not limit how
// many wavefronts a CU holds, then, for the queries whose frontier would outgrow it, a second pass
// whose frontier has the proven bound.  A handed-on query restarts from the root with its best
// answer so far as its bound, which gives the same result.  Sizes measured on 1M uniform queries
// against the 327,680-triangle icosphere: closest point 8.04 ms at 1,024 entries (no hand-offs)
// vs 8.84 ms at 512 (1,540 hand-offs) and 9.07 ms with the full frontier only; silhouettes (peak
// frontier 437) 40.7 ms at 512-768 entries vs 42.8 ms at 1,024 (fewer resident wavefronts).
constexpr int kFrontSmallCp = 1024, kFrontSmallSil = 640;
// Silhouette small-frontier pass: children whose key is within kSilNear x the round's smallest are
// appended last, so they are expanded first.  On the workload above, first measured at 34.6 ms
// with 1.25 (35.8 at 1.5, 37.5 at 2.0; three or four groups no better) vs 40.1 ms with lane-order
// appends; with the tighter exact cull the optimum moved: 15.8 ms at 1.1 (15.6 at 1.05, 15.9 at
// 1.02, 17.1 at 1.0, 16.8 at 1.25, 17.8 at 1.5; 20,480 triangles: 11.7-11.9 ms at 1.0-1.1).  The
// same grouping in the closest-point pass was slower (7.57-7.69 ms vs 7.47 ms).
constexpr float kSilNear = 1.1f;

// Wave-wide minima through DPP lane moves (register to register, no LDS): swap within pairs and
// quads, mirror half-rows and rows (every lane of a 16-lane row then holds the row's minimum), then
// broadcast lane 15 into rows 1 and 3 and lane 31 into rows 2 and 3; lane 63 ends with the
// minimum of the wave, which readlane hands to every lane.  Lanes a row mask leaves out keep their
// own value (`old` = the value itself).  ArgMin orders (distance, slot) lexicographically, so ties
// go to the lower slot.
template <int CTRL, int ROWS>
__device__ __forceinline__ void argmin_step(float &c, int &s) {
    const float oc = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(c), __float_as_int(c), CTRL, ROWS, 0xF, false));
    const int os = __builtin_amdgcn_update_dpp(s, s, CTRL, ROWS, 0xF, false);
    if (oc < c || (oc == c && os < s)) {
        c = oc;
        s = os;
    }
}
__device__ __forceinline__ void wave_argmin(float &c, int &s) {
    argmin_step<0xB1, 0xF>(c, s);    // quad_perm [1,0,3,2]
    argmin_step<0x4E, 0xF>(c, s);    // quad_perm [2,3,0,1]
    argmin_step<0x141, 0xF>(c, s);   // row_half_mirror
    argmin_step<0x140, 0xF>(c, s);   // row_mirror
    argmin_step<0x142, 0xA>(c, s);   // row_bcast:15 into rows 1, 3
    argmin_step<0x143, 0xC>(c, s);   // row_bcast:31 into rows 2, 3
    c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 63));
    s = __builtin_amdgcn_readlane(s, 63);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float min_step(float v) {
    return fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWS, 0xF, false)));
}
__device__ __forceinline__ float wave_min(float v) {
    v = min_step<0xB1, 0xF>(v);
    v = min_step<0x4E, 0xF>(v);
    v = min_step<0x141, 0xF>(v);
    v = min_step<0x140, 0xF>(v);
    v = min_step<0x142, 0xA>(v);
    v = min_step<0x143, 0xC>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}


// Exclusive prefix sum of v over the wavefront's lanes; total receives the sum over all 64.
__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
    const int lane = threadIdx.x & 63;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// Appends one query (index, bound, best slot, job) to a hand-off list: every lane takes part in the
// atomic (lane 0 adds 1) and stores the same entry, so no lane-dependent branch is involved.
__device__ __forceinline__ void wave_handoff(uint4 *list, uint32_t *count, uint4 entry) {
    const uint32_t o = uniform(atomicAdd(count, (threadIdx.x & 63) == 0 ? 1u : 0u));
    stg(gcast(list) + o, entry);
}

// The first wave-per-query pass starts from the scene's seed cut instead of the root: lane k puts
// seed k on the frontier, keyed as an opened child would be (closest points: squared box distance;
// silhouettes: box distance).  Skipping the tests of the seeds' ancestors is exact, as the cut
// never passes below a node whose reference cone test can cull and the box tests and exact culls
// it skips are conservative.  The full-frontier pass keeps the root, where its frontier bound holds.
template <int D, bool Squared>
__device__ __forceinline__ int seed_front(const GP &P, const float *xc, uint2 *front) {
    const int lane = threadIdx.x & 63;
    const int n = P.n_seed;
    if (lane < n) {
        const GQ_G float4 *sp = (const GQ_G float4 *)(P.seed + lane);
        const float4 a = ldg(sp), b = ldg(sp + 1);
        float key;
        if constexpr (Squared) {
            const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
            key = 0.0f;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const float o = fmaxf(fmaxf(lo[c] - xc[c], xc[c] - hi[c]), 0.0f);
                key = fmaf(o, o, key);
            }
        } else {
            float dx;
            if constexpr (D == 3) box_distance(v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(xc[0], xc[1], xc[2]), key, dx);
            else box_distance(v2(a.x, a.y), v2(b.x, b.y), v2(xc[0], xc[1]), key, dx);
        }
        front[lane] = make_uint2(__float_as_uint(a.w), __float_as_uint(key));
    }
    return n;
}

template <int D, bool Q>
__device__ __forceinline__ void cp_wave_item(const GP &P, const uint4 &item, uint2 *front, int cap, uint4 *ovf,
                                             uint32_t *ovf_n) {
    using V = typename VT<D>::T;
    const int lane = threadIdx.x & 63;
    {
        const uint32_t qi = uniform(item.x);
        float best = uniformf(__uint_as_float(item.y));
        int best_slot = (int)uniform(item.z);
        const V p = ldv<D>(P.x + (size_t)D * qi);
        float pc[D];
#pragma unroll
        for (int j = 0; j < D; ++j) pc[j] = comp<D>(p, j);
        const float slack = coord_slack<D>(P, pc);
        float bound2 = best < kInf ? cp_bound2(best, slack) : kInf;
        int size = 1;
        bool overflow = false;
        if (ovf != nullptr && P.n_seed > 0 && P.n_seed <= cap) size = seed_front<D, true>(P, pc, front);
        else front[0] = make_uint2(P.root, 0u);
        while (size > 0) {
            const int take = size < 64 ? size : 64;
            uint2 e = make_uint2(0u, __float_as_uint(kInf));
            if (lane < take) e = front[size - take + lane];
            size -= take;
            uint2 ent[4] = {};
            bool hit[4] = {false, false, false, false};
            float cand = kInf;
            int cand_slot = -1;
            uint32_t lfirst = 0, lcount = 0;
            if (lane < take && __uint_as_float(e.y) <= bound2) {
                if (!ref_is_leaf(e.x)) {
                    NodeC4<D> nd;
                    load_node_c<D, Q>((const void *)P.nodes, e.x, nd);
                    Cyl4<D> cy;
                    load_cyl4<D>((const void *)P.cyl, e.x, cy);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float sx;
                        const float lb2 = cyl_lb2<D>(nd, cy, j, pc, sx);
                        hit[j] = j < nd.n && lb2 <= bound2;
                        // the entry's key is re-tested against the bound when it is taken
                        ent[j] = make_uint2(nd.child[j], __float_as_uint(lb2));
                    }
                } else {
                    leaf_range(e.x, (const uint32_t *)P.big, lfirst, lcount);
                }
            }
            // the round's leaf primitives, one per lane: their slots are listed in the free part of
            // the frontier (entries [size, cap) were taken into registers above), so a leaf's
            // primitives are tested side by side instead of one after another by its lane
            auto test = [&](uint32_t s) {
                float dd;
                if constexpr (D == 3) {
                    const float4 a = ldg(P.prims + (4 * s)), b = ldg(P.prims + (4 * s + 1)), c = ldg(P.prims + (4 * s + 2));
                    const float4 pl = ldg(P.prims + (4 * s + 3));
                    if (!tri_may_beat(pc, a, b, c, pl, bound2)) return;
                    V3 pt;
                    V2 uv;
                    dd = closest_point_triangle(p, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), pt, uv);
                } else {
                    const float4 ab = ldg(P.prims + (s));
                    V2 pt;
                    dd = closest_point_segment(p, v2(ab.x, ab.y), v2(ab.z, ab.w), pt);
                }
                if (dd < cand || (dd == cand && (int)s < cand_slot)) {
                    cand = dd;
                    cand_slot = (int)s;
                }
            };
            if (__ballot(lcount > 0u) != 0ull) {
                int ltot;
                const int loff = wave_excl_scan((int)lcount, ltot);
                if (ltot <= 2 * (cap - size)) {
                    uint32_t *list = reinterpret_cast<uint32_t *>(front + size);
                    for (uint32_t t = 0; t < lcount; ++t) list[loff + (int)t] = lfirst + t;
                    wave_sync_lds();
                    for (int b0 = 0; b0 < ltot; b0 += 64)
                        if (b0 + lane < ltot) test(list[b0 + lane]);
                    wave_sync_lds();   // the appends below reuse this space
                } else {
                    for (uint32_t s = lfirst; s < lfirst + lcount; ++s) test(s);
                }
            }
            // wave minimum of the leaf candidates (ties to the lower slot, so every lane agrees)
            wave_argmin(cand, cand_slot);
            if (cand < best || (cand == best && cand_slot < best_slot)) {
                best = cand;
                best_slot = cand_slot;
                bound2 = cp_bound2(cand, slack);
            }
            // append the opened children, in lane order and child order within a lane
            int tot = 0, pos = size;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const unsigned long long b = __ballot(hit[j]);
                tot += __popcll(b);
                pos += lanes_below(b);
            }
            if (ovf != nullptr && size + tot > cap) {   // wave-uniform: hand the query on
                overflow = true;
                break;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (hit[j] && pos < cap) front[pos] = ent[j];
                pos += (int)hit[j];
            }
            size = size + tot < cap ? size + tot : cap;
        }
        if (overflow) {
            wave_handoff(ovf, ovf_n, make_uint4(qi, __float_as_uint(best), (uint32_t)best_slot, item.w));
            return;
        }
        cp_record<D>(P, qi, p, best_slot);
    }
}

// ---------------------------------------------------------------------------------------------
// ray (lane per query)
template <int D, bool Q>
__device__ __forceinline__ void ray_lane(const GP &P, int base) {
    using V = typename VT<D>::T;
    GQ_PRIVATE_STACK_DECL(kRayStackMax)
    const int i = base + threadIdx.x;
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
            load_node4<D, Q>((const void *)P.nodes, cur, nd);
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
            leaf_range(cur, (const uint32_t *)P.big, first, count);
            for (uint32_t s = first; s < first + count; ++s) {
                float t;
                if constexpr (D == 3) {
                    const float4 a = ldg(P.prims + (4 * s)), b = ldg(P.prims + (4 * s + 1)), c = ldg(P.prims + (4 * s + 2));
                    float bv, bw;
                    t = ray_triangle_robust(x, v, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), bv, bw);
                } else {
                    const float4 ab = ldg(P.prims + (s));
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
    const GRecords &o = P.out;
    bool ok = best_slot >= 0;
    const int s = best_slot;
    if (ok) {
        if constexpr (D == 3) {
            // Triangle.ray_intersect (triangle.py:48-73) with r_max = the hit distance
            const float4 a4 = ldg(P.prims + (4 * s)), b4 = ldg(P.prims + (4 * s + 1)), c4 = ldg(P.prims + (4 * s + 2));
            const V3 a = v3(a4.x, a4.y, a4.z), b = v3(b4.x, b4.y, b4.z), c = v3(c4.x, c4.y, c4.z);
            const V3 v1 = b - a, v2_ = c - a;
            float vv, ww;
            const float t = ray_triangle_robust(x, v, a, b, c, vv, ww);   // as in the traversal
            ok = t <= best;
            if (ok) {
                if (o.p) st(o.p + 3 * (size_t)qi, a + v1 * vv + v2_ * ww);
                if (o.n) st(o.n + 3 * (size_t)qi, normalize(cross(v1, v2_)));
                if (o.uv) st(o.uv + 2 * (size_t)qi, v2(1.0f - vv - ww, vv));
                if (o.d) o.d[qi] = t;
            }
        } else {
            // record of bvh.py:267-274
            const float4 ab = ldg(P.prims + (s));
            const V2 tt = v2(ab.z - ab.x, ab.w - ab.y);
            if (o.p) st(o.p + 2 * (size_t)qi, x + v * best);
            if (o.n) st(o.n + 2 * (size_t)qi, normalize(v2(tt.y, -tt.x)));
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
// One child of an SNCH pair.  3D: exact-cull cone in axis/cos_t/sin_t, reference cone (if it can
// cull) in the table at ref_cone.  2D: the reference cone inline (axis[0..1], half, radius) and
// cos_t = 2 for a subtree without silhouettes.
struct PairChild {
    float lo[3], hi[3];
    float axis[3], cos_t, sin_t;
    uint32_t child;
    int32_t ref_cone;
    float half, radius;
    float sin_h, cos_h;   // 2D: sin and cos of the reference half-angle
    float ref_axis[2];    // 2D: the reference cone axis (`axis` holds the exact-cull cone's)
    float sil_lo, sil_hi, sil_perp;   // 3D float32 records: extent of the silhouette edges below
                                      // along / across the cone axis about the box centre
};

template <int D, bool Q>
__device__ __forceinline__ void load_pair(const void *base, uint32_t idx, PairChild (&c)[2]) {
    if constexpr (Q) {
        // SPairQ<D> (64 B): origin @0, exp @4D, flags, child boxes (u8 in 3D, u16 in 2D), exact-cull
        // cone axis (int16) @28, its sine (int16), child refs, ref_base, silhouette extents (slab
        // middle int16, half-width u8, perpendicular u8)
        constexpr int kFlags = D == 3 ? 15 : 10, kLo = D == 3 ? 16 : 12, kHi = D == 3 ? 22 : 20;
        constexpr int kAx = 28, kSin = D == 3 ? 40 : 36, kChild = D == 3 ? 44 : 40, kRef = D == 3 ? 52 : 48;
        constexpr int kExt = D == 3 ? 56 : 52;
        Words<16> r;
        r.load(reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 64);
        const uint32_t flags = r.u8(kFlags);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float org = r.f32(4 * k);
            const float sc = ldexpf(1.0f, r.s8(4 * D + k));
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if constexpr (D == 3) {
                    c[j].lo[k] = fmaf((float)r.u8(kLo + 3 * j + k), sc, org);
                    c[j].hi[k] = fmaf((float)r.u8(kHi + 3 * j + k), sc, org);
                } else {
                    c[j].lo[k] = fmaf((float)r.u16(kLo + 2 * (2 * j + k)), sc, org);
                    c[j].hi[k] = fmaf((float)r.u16(kHi + 2 * (2 * j + k)), sc, org);
                }
            }
        }
        constexpr float kS = 1.0f / 32767.0f;
        const uint32_t rb = r.u32(kRef);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            float a[3] = {0.0f, 0.0f, 0.0f}, l = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                a[k] = r.s16(kAx + 2 * (D * j + k)) * kS;
                l = fmaf(a[k], a[k], l);
            }
            const float rn = l > 0.0f ? 1.0f / sqrtf(l) : 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) c[j].axis[k] = a[k] * rn;
            if constexpr (D == 2) c[j].lo[2] = c[j].hi[2] = 0.0f;
            // the cosine is formed from the rounded-up sine, so it is no larger than the cosine of
            // the widened angle (its float32 rounding sits inside the exact cull's margin)
            const int qs = r.s16(kSin + 2 * j);
            const float sn = qs * kS;
            c[j].sin_t = sn;
            c[j].cos_t = (flags & (kPairNone0 << j)) ? 2.0f
                       : qs >= 32767             ? 0.0f   // never culls
                                                 : sqrtf(fmaxf(fmaf(-sn, sn, 1.0f), 0.0f));
            c[j].child = r.u32(kChild + 4 * j);
            c[j].ref_cone = (flags & (kPairRef0 << j)) ? (int32_t)(rb + (j == 1 ? (flags & kPairRef0) : 0u)) : -1;
            // extents in units of the decoded box's half-diagonal (scaled in exact_cull)
            const float mid = r.s16(kExt + 2 * j) * kS, hw = r.u8(kExt + 4 + j) * (1.0f / 255.0f);
            c[j].sil_lo = mid - hw * hw;
            c[j].sil_hi = mid + hw * hw;
            c[j].sil_perp = r.u8(kExt + 6 + j) * (1.0f / 255.0f);
        }
    } else if constexpr (D == 2) {
        // SPairF<2> (192 B): lo[2][2] @0, hi[2][2] @16, cone axis @32, half @48,
        // radius @56, child @64, ref_cone @72, cull @80, sin(half) @88, cos(half) @96, exact-cull
        // cos @104, sin @112, axis @120, vertex extents along @136 / @144, across @152
        Words<40> r;
        r.load(reinterpret_cast<const uint8_t *>(base) + (size_t)idx * 192);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                c[j].lo[k] = r.f32(8 * j + 4 * k);
                c[j].hi[k] = r.f32(16 + 8 * j + 4 * k);
                c[j].axis[k] = r.f32(32 + 8 * j + 4 * k);
            }
            c[j].lo[2] = c[j].hi[2] = 0.0f;
            c[j].half = r.f32(48 + 4 * j);
            c[j].radius = r.f32(56 + 4 * j);
            c[j].child = r.u32(64 + 4 * j);
            c[j].ref_cone = (int32_t)r.u32(72 + 4 * j);
            c[j].sin_h = r.f32(88 + 4 * j);
            c[j].cos_h = r.f32(96 + 4 * j);
            // the exact cull's cone replaces the reference cone axis in `axis` (the reference test
            // reads its cone from ref_axis)
            c[j].ref_axis[0] = c[j].axis[0];
            c[j].ref_axis[1] = c[j].axis[1];
            c[j].cos_t = r.f32(104 + 4 * j);   // 2 = no silhouette below
            c[j].sin_t = r.f32(112 + 4 * j);
            c[j].axis[0] = r.f32(120 + 8 * j);
            c[j].axis[1] = r.f32(124 + 8 * j);
            c[j].axis[2] = 0.0f;
            c[j].sil_lo = r.f32(136 + 4 * j);
            c[j].sil_hi = r.f32(144 + 4 * j);
            c[j].sil_perp = r.f32(152 + 4 * j);
        }
    } else {
        // SPairF<D> (128 B): lo[2][D] @0, hi[2][D] @8D, cones (axis3, cos, sin) @16D, child @16D+40,
        // ref_cone @16D+48, sil_lo @16D+56, sil_hi @16D+64, sil_perp @16D+72
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
            c[j].sil_lo = r.f32(16 * D + 56 + 4 * j);
            c[j].sil_hi = r.f32(16 * D + 64 + 4 * j);
            c[j].sil_perp = r.f32(16 * D + 72 + 4 * j);
        }
    }
}

// True when every silhouette predicate in the child is provably false for query x: the normals
// lie within the cone and the dot product of every view direction into the box with every cone
// normal has one sign (bounded away from zero by a margin covering float32 rounding: 1e-5 of the
// terms plus mabs, the query's absolute coordinate slack).
template <int D, bool Q>
__device__ __forceinline__ bool exact_cull(const PairChild &c, const float *xc, float mabs) {
    if (c.cos_t > 1.5f) return true;     // no silhouette can ever pass in this subtree
    if (!(c.cos_t > 0.0f)) return false; // cone too wide to separate anything
    // a, b: the box centre relative to x along the cone axis u and across it
    float a, b, hb2 = 0.0f;
    if constexpr (D == 3) {
        float w[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            w[k] = 0.5f * (c.lo[k] + c.hi[k]) - xc[k];
            const float h = 0.5f * (c.hi[k] - c.lo[k]);
            hb2 = fmaf(h, h, hb2);
        }
        a = w[0] * c.axis[0] + w[1] * c.axis[1] + w[2] * c.axis[2];
        const float cx = w[1] * c.axis[2] - w[2] * c.axis[1];
        const float cy = w[2] * c.axis[0] - w[0] * c.axis[2];
        const float cz = w[0] * c.axis[1] - w[1] * c.axis[0];
        b = sqrtf(cx * cx + cy * cy + cz * cz);
    } else {
        const float w0 = 0.5f * (c.lo[0] + c.hi[0]) - xc[0], w1 = 0.5f * (c.lo[1] + c.hi[1]) - xc[1];
        const float h0 = 0.5f * (c.hi[0] - c.lo[0]), h1 = 0.5f * (c.hi[1] - c.lo[1]);
        hb2 = fmaf(h0, h0, h1 * h1);
        a = w0 * c.axis[0] + w1 * c.axis[1];
        b = fabsf(w0 * c.axis[1] - w1 * c.axis[0]);
    }
    // Every silhouette point p below lies, relative to the centre, within [lo_u, hi_u] along u and
    // within rho of the line through the centre along u.  With n = cos(phi) u + sin(phi) t
    // (phi <= theta, t _|_ u), dot(p - x, n) lies in
    // [cos (a + lo_u) - sin (b + rho), cos (a + hi_u) + sin (b + rho)] at phi = theta; a sphere of
    // radius R around the centre holding every point bounds it by
    // [a cos - b sin - R, a cos + b sin + R].  Either proof culls.
    float lo_u = c.sil_lo, hi_u = c.sil_hi, rho = c.sil_perp;
    float R;
    if constexpr (Q) {
        // quantized records: the extents are in units of the decoded box's half-diagonal, and the
        // sphere of that radius around the centre holds the whole box
        const float Rb = sqrtf(hb2);
        lo_u *= Rb;
        hi_u *= Rb;
        rho *= Rb;
        R = fminf(Rb, sqrtf(fmaxf(lo_u * lo_u, hi_u * hi_u) + rho * rho));
    } else {
        R = sqrtf(fmaxf(lo_u * lo_u, hi_u * hi_u) + rho * rho);
    }
    const float m = 1e-5f * (fabsf(a) + b + R) + mabs;
    const bool sphere = (a * c.cos_t - b * c.sin_t > R + m) || (a * c.cos_t + b * c.sin_t < -(R + m));
    const bool slab = ((a + lo_u) * c.cos_t - (b + rho) * c.sin_t > m) ||
                      ((a + hi_u) * c.cos_t + (b + rho) * c.sin_t < -m);
    return sphere || slab;
}


// The reference cone test (SNCHNode.overlaps, snch.py) of a child whose reference cone can cull,
// evaluated with the reference's float32 box and cone: inline in the 2D pair record, from the
// reference-cone table in 3D.  d_min receives the box distance.
// 2D reference cone test with a fast path that needs no transcendental.  With c the cosine between
// the cone axis and the view direction to the box centre, the reference's tests
// |acos(c) - pi/2| <= half and |acos(c) - pi/2| <= half + asin(radius / l) are, in exact
// arithmetic, |c| <= sin(half) and |c| <= sin(half) sqrt(1 - rho^2) + cos(half) rho
// (rho = radius / l < cos(half); rho > cos(half) means the widened half-angle exceeds pi/2: an
// overlap).  Where |c| or rho is further than 4e-6 from a threshold (ten times the float32
// rounding of either side, acos and asin included) the outcome equals the reference's float32
// evaluation, which runs for everything else.
__device__ __forceinline__ bool ref_cone_2d(const PairChild &c, V2 x, float r, float &d_min) {
    const V2 lo = v2(c.lo[0], c.lo[1]), hi = v2(c.hi[0], c.hi[1]);
    float d_max;
    box_distance(lo, hi, x, d_min, d_max);
    if (!(d_min <= r)) return false;
    constexpr float m = 4e-6f;
    const V2 va = (lo + hi) / 2.0f - x;
    const float l2 = dot(va, va);
    const float il = rsqrtf(l2);
    const float cs = fabsf(c.ref_axis[0] * va.x + c.ref_axis[1] * va.y) * il;
    if (cs <= c.sin_h - m) return true;
    if (cs >= c.sin_h + m) {
        const float l = l2 * il;
        if (l > c.radius * (1.0f + m)) {
            const float rho = c.radius * il;
            if (rho >= c.cos_h + m) return true;
            if (rho <= c.cos_h - m) {
                const float t2 = c.sin_h * sqrtf(fmaxf(1.0f - rho * rho, 0.0f)) + c.cos_h * rho;
                if (cs <= t2 - m) return true;
                if (cs >= t2 + m) return false;
            }
        }
    }
    return snch_overlaps(lo, hi, v2(c.ref_axis[0], c.ref_axis[1]), c.half, c.radius, x, r, d_min);
}

// 3D reference cone test with the same fast path as ref_cone_2d: with c the cosine between the cone
// axis and the view direction to the box centre and rho = radius / l, the reference's tests are,
// in exact arithmetic, |c| <= sin(half), then rho > cos(half) or
// |c| <= sin(half) sqrt(1 - rho^2) + cos(half) rho (rho < 1).  Outside a 4e-6 band around each
// threshold the outcome equals the reference's float32 evaluation (its rounding, acos and asin
// included, moves c by under 1e-6: the band is four times that), which runs for everything else
// and for a query inside the node's bounding sphere.  Every table entry has 0 <= half <= pi/2.
__device__ __forceinline__ bool ref_cone_3d(const RefCone &rc, const V3 &x, float r, float &d_min) {
    const V3 lo = v3(rc.lo[0], rc.lo[1], rc.lo[2]), hi = v3(rc.hi[0], rc.hi[1], rc.hi[2]);
    const V3 axis = v3(rc.axis[0], rc.axis[1], rc.axis[2]);
    float d_max;
    box_distance(lo, hi, x, d_min, d_max);
    if (!(d_min <= r)) return false;
    constexpr float m = 4e-6f;
    const V3 va = (lo + hi) / 2.0f - x;
    const float l2 = dot(va, va);
    const float il = rsqrtf(l2);
    const float cs = fabsf(dot(axis, va)) * il;   // NaN at l = 0: the float32 evaluation decides
    if (cs <= rc.sin_h - m) return true;
    if (cs >= rc.sin_h + m) {
        const float l = l2 * il;
        if (l > rc.radius * (1.0f + m)) {
            const float rho = rc.radius * il;
            if (rho >= rc.cos_h + m) return true;
            if (rho <= rc.cos_h - m) {
                const float t2 = rc.sin_h * sqrtf(fmaxf(1.0f - rho * rho, 0.0f)) + rc.cos_h * rho;
                if (cs <= t2 - m) return true;
                if (cs >= t2 + m) return false;
            }
        }
    }
    return snch_overlaps(lo, hi, axis, rc.half, rc.radius, x, r, d_min);
}

template <int D, bool Q>
__device__ __forceinline__ bool ref_cone_pass(const GP &P, const PairChild &c, const typename VT<D>::T &x, float r,
                                              float &d_min) {
    if constexpr (D == 2) {
        if constexpr (Q) {
            // quantized 2D records keep each child's reference box and cone in a table (RefCone2:
            // lo, hi | axis, half, radius | sin, cos of the half-angle), fetched only for a child
            // that passed the box test and the exact cull
            const GQ_G float4 *e = (const GQ_G float4 *)P.ref_cones + 3 * c.ref_cone;
            const float4 e0 = ldg(e), e1 = ldg(e + 1), e2 = ldg(e + 2);
            PairChild t = c;
            t.lo[0] = e0.x; t.lo[1] = e0.y; t.hi[0] = e0.z; t.hi[1] = e0.w;
            t.ref_axis[0] = e1.x; t.ref_axis[1] = e1.y; t.half = e1.z; t.radius = e1.w;
            t.sin_h = e2.x; t.cos_h = e2.y;
            return ref_cone_2d(t, x, r, d_min);
        } else {
            return ref_cone_2d(c, x, r, d_min);
        }
    } else {
        return ref_cone_3d(ldg(P.ref_cones + c.ref_cone), x, r, d_min);
    }
}

template <int D, bool Q>
__device__ __forceinline__ void sil_packet(const GP &P, int base) {
    using V = typename VT<D>::T;
    __shared__ uint32_t wstack[kPacketBlock / 64][kPacketDepth];
    __shared__ unsigned long long wmask[kPacketBlock / 64][kPacketDepth];
    uint32_t *wst = wstack[threadIdx.x >> 6];
    unsigned long long *wmk = wmask[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int i = base + threadIdx.x;
    const bool live = i < P.n;
    const uint32_t qi = live ? query_index<D>(P, i) : 0u;
    V x;
    if (live) x = ldv<D>(P.x + (size_t)D * qi);
    else x = ldv<D>(P.x);
    float xc[3] = {0.0f, 0.0f, 0.0f}, cen[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = comp<D>(x, k);
    packet_centroid<D>(live, xc, cen);
    const float cslack = coord_slack<D>(P, xc);
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
                load_pair<D, Q>((const void *)P.nodes, cur, c);
                bool any[2], hh[2];
                unsigned long long hm[2];
                float key[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float dm, dx;
                    box_distance(ldv<D>(c[j].lo), ldv<D>(c[j].hi), x, dm, dx);
                    bool h = act && dm <= r;
                    // the exact cull first: it needs no fetch, the reference cone test a table entry
                    if (h && exact_cull<D, Q>(c[j], xc, cslack)) h = false;
                    if (h && c[j].ref_cone >= 0) h = ref_cone_pass<D, Q>(P, c[j], x, r, dm);
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
                leaf_range(cur, (const uint32_t *)P.big, first, count);
                first = uniform(first);
                count = uniform(count);
                for (uint32_t s = first; s < first + count; ++s) {
                    if constexpr (D == 3) {
                        const float4 s0 = ldg(P.prims + (3 * s)), s1 = ldg(P.prims + (3 * s + 1)), s2 = ldg(P.prims + (3 * s + 2));
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
                        const float4 s0 = ldg(P.prims + (2 * s)), s1 = ldg(P.prims + (2 * s + 1));
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
    const GRecords &o = P.out;
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

// Output record of a closest-silhouette query from the winning silhouette slot (-1: none).  The
// point and distance are recomputed exactly as the leaf test computed them.
template <int D>
__device__ __forceinline__ void sil_record(const GP &P, uint32_t qi, const typename VT<D>::T &x,
                                           int best_slot) {
    using V = typename VT<D>::T;
    const GRecords &o = P.out;
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
        const float4 s0 = ldg(P.prims + (3 * s)), s1 = ldg(P.prims + (3 * s + 1));
        bd = closest_point_segment(x, v3(s0.x, s0.y, s0.z), v3(s0.w, s1.x, s1.y), bp);
    } else {
        const float4 s0 = ldg(P.prims + (2 * s));
        bp = v2(s0.z, s0.w);
        bd = norm(x - bp);
    }
    if (o.valid) o.valid[qi] = 1;
    if (o.p) st(o.p + (size_t)D * qi, bp);
    if (o.d) o.d[qi] = bd;
    if (o.prim_id) o.prim_id[qi] = P.ids[s];
}

// ---------------------------------------------------------------------------------------------
// closest silhouette, lane per query: the same node tests as the packet pass, one query per lane
// over the Morton order with a per-lane stack (one entry per binary level at most): in LDS in 2D,
// in private memory in 3D (measured above).  Cheaper than the packet pass when per-query work is
// small and culling is effective (2D silhouettes).
template <int D, bool Q>
__device__ __forceinline__ void sil_lane(const GP &P, int base) {
    using V = typename VT<D>::T;
    extern __shared__ uint2 lane_stack_lds[];
    uint2 lane_stack_priv[D == 3 ? kPacketDepth : 1];
    uint2 *const lane_stack = D == 3 ? lane_stack_priv : lane_stack_lds;
    const int lane_col = D == 3 ? 0 : (int)threadIdx.x;
    const int lane_stride = D == 3 ? 1 : (int)blockDim.x;
    int sp = 0;
    const int i = base + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t qi = query_index<D>(P, i);
    const V x = ldv<D>(P.x + (size_t)D * qi);
    float xc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = comp<D>(x, k);
    const float cslack = coord_slack<D>(P, xc);
    float r = P.r_max ? P.r_max[qi] : P.r_max_scalar;
    int best_slot = -1;
    float d_root, d_unused;
    box_distance(ldv<D>(P.root_lo), ldv<D>(P.root_hi), x, d_root, d_unused);
    bool go = d_root <= r;   // reference: a query that misses the root box is not traversed
    uint32_t cur = P.root;
    while (go) {
        if (!ref_is_leaf(cur)) {
            PairChild c[2];
            load_pair<D, Q>((const void *)P.nodes, cur, c);
            bool h[2];
            float key[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float dm, dx;
                box_distance(ldv<D>(c[j].lo), ldv<D>(c[j].hi), x, dm, dx);
                bool hj = dm <= r;
                if (hj && exact_cull<D, Q>(c[j], xc, cslack)) hj = false;
                if (hj && c[j].ref_cone >= 0) {
                    float dr;
                    hj = ref_cone_pass<D, Q>(P, c[j], x, r, dr);
                }
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
            leaf_range(cur, (const uint32_t *)P.big, first, count);
            for (uint32_t s = first; s < first + count; ++s) {
                if constexpr (D == 3) {
                    const float4 s0 = ldg(P.prims + (3 * s)), s1 = ldg(P.prims + (3 * s + 1)), s2 = ldg(P.prims + (3 * s + 2));
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
                    const float4 s0 = ldg(P.prims + (2 * s)), s1 = ldg(P.prims + (2 * s + 1));
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


template <int D, bool Q>
__device__ __forceinline__ void sil_wave_item(const GP &P, const uint4 &item, uint2 *front, int cap, uint4 *ovf,
                                              uint32_t *ovf_n) {
    using V = typename VT<D>::T;
    const int lane = threadIdx.x & 63;
    {
        const uint32_t qi = uniform(item.x);
        float r = uniformf(__uint_as_float(item.y));
        int best_slot = (int)uniform(item.z);
        const V x = ldv<D>(P.x + (size_t)D * qi);
        float xc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < D; ++j) xc[j] = comp<D>(x, j);
        const float cslack = coord_slack<D>(P, xc);
        int size = 0;
        bool overflow = false;
        if (r >= 0.0f) {
            if (ovf != nullptr && P.n_seed > 0 && P.n_seed <= cap) {
                size = seed_front<D, false>(P, xc, front);
            } else {
                front[0] = make_uint2(P.root, 0u);
                size = 1;
            }
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
                    load_pair<D, Q>((const void *)P.nodes, e.x, c);
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        float dm, dx;
                        box_distance(ldv<D>(c[j].lo), ldv<D>(c[j].hi), x, dm, dx);
                        bool h = dm <= r;
                        if (h && exact_cull<D, Q>(c[j], xc, cslack)) h = false;
                        if (h && c[j].ref_cone >= 0) h = ref_cone_pass<D, Q>(P, c[j], x, r, dm);
                        const uint2 ent = make_uint2(c[j].child, __float_as_uint(dm));
                        if (j == 0) { h0 = h; ch0 = ent; }
                        else { h1 = h; ch1 = ent; }
                    }
                } else {
                    uint32_t first, count;
                    leaf_range(e.x, (const uint32_t *)P.big, first, count);
                    for (uint32_t s = first; s < first + count; ++s) {
                        if constexpr (D == 3) {
                            const float4 s0 = ldg(P.prims + (3 * s)), s1 = ldg(P.prims + (3 * s + 1)), s2 = ldg(P.prims + (3 * s + 2));
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
                            const float4 s0 = ldg(P.prims + (2 * s)), s1 = ldg(P.prims + (2 * s + 1));
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
            wave_argmin(cand, cand_slot);
            if (cand < r || (cand == r && best_slot >= 0 && cand_slot >= 0 && cand_slot < best_slot)) {
                r = cand;
                best_slot = cand_slot;
            }
            // append the opened children in two groups: first those farther than kSilNear x the
            // round's smallest key, then the near ones, which the next round pops first (the stack
            // is taken from the top).  The result does not depend on this order, but the frontier
            // bound does (it needs lane-order appends), so the full-frontier pass (no hand-off
            // list) appends everything in lane order (thr = inf: one group).
            float km = fminf(h0 ? __uint_as_float(ch0.y) : kInf, h1 ? __uint_as_float(ch1.y) : kInf);
            km = wave_min(km);
            const float thr = ovf != nullptr ? fmaf(km, kSilNear, 1e-6f) : kInf;
            const bool n0 = h0 && __uint_as_float(ch0.y) <= thr, n1 = h1 && __uint_as_float(ch1.y) <= thr;
            const bool f0 = h0 && !n0, f1 = h1 && !n1;
            const unsigned long long bf0 = __ballot(f0), bf1 = __ballot(f1), bn0 = __ballot(n0), bn1 = __ballot(n1);
            const int tot_far = __popcll(bf0) + __popcll(bf1);
            const int tot = tot_far + __popcll(bn0) + __popcll(bn1);
            int pf = size + lanes_below(bf0) + lanes_below(bf1);
            int pn = size + tot_far + lanes_below(bn0) + lanes_below(bn1);
            if (ovf != nullptr && size + tot > cap) {   // wave-uniform: hand the query on
                overflow = true;
                break;
            }
            if (f0 && pf < cap) front[pf] = ch0;
            pf += (int)f0;
            if (f1 && pf < cap) front[pf] = ch1;
            if (n0 && pn < cap) front[pn] = ch0;
            pn += (int)n0;
            if (n1 && pn < cap) front[pn] = ch1;
            size = size + tot < cap ? size + tot : cap;
        }
        if (overflow) {
            wave_handoff(ovf, ovf_n, make_uint4(qi, __float_as_uint(r), (uint32_t)best_slot, item.w));
            return;
        }
        sil_record<D>(P, qi, x, best_slot);
    }
}

// ---------------------------------------------------------------------------------------------
// Query ordering: a counting sort of each job's queries into Morton-ordered cells of the job's
// query bounding box (about 8 queries per cell); inside a cell the queries keep their scatter
// order.  A 64-query packet spans about 8 cells, so ordering each cell by the full Morton code
// (an insertion sort per cell) changed only the packets' edges: it cost more than it saved (the
// 64-scene batch 90.4 ms with it, 87.9 without).  No result depends on the order: every kernel
// breaks distance ties to the lowest slot.  Every pass is a plain kernel over a job table: no
// library call, no memset, no host synchronisation, so a whole batch can be recorded in a graph.
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

// Copies one job's parameters into device memory (the single-call path; batches upload a table).
__global__ void k_store_params(EngineParams P, EngineParams *dst) {
    if (threadIdx.x == 0) *dst = P;
}

// Tile t of a launch: (job, first query).  A null table means job 0, tiles of `ts` queries.
__device__ __forceinline__ uint2 tile_at(const uint2 *tiles, int t, int ts) {
    return tiles ? tiles[t] : make_uint2(0u, (uint32_t)t * (uint32_t)ts);
}

// Workgroup b of a launch of n goes to XCD b % kXcds (the dispatcher deals workgroups round robin
// over the XCDs).  With spread on, XCD x takes one contiguous run of tiles instead of every 8th
// tile, so on Morton-sorted queries each XCD's L2 sees one region of space and its part of the
// scene.  A bijection on [0, n).
__device__ __forceinline__ int xcd_tile(int b, int n, int spread) {
    if (!spread) return b;
    const int x = b % kXcds, local = b / kXcds;
    const int per = n / kXcds, rem = n % kXcds;
    return x * per + min(x, rem) + local;
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

// (1) partial bounds of each 256-query tile, order-preserving uint encoding, into part[tile][6]
template <int D>
__global__ void __launch_bounds__(kBlock) k_sort_bounds(const EngineParams *J, const uint2 *tiles) {
    const uint2 tl = tile_at(tiles, blockIdx.x, kBlock);
    const EngineParams &P = J[tl.x];
    const int i = (int)tl.y + threadIdx.x;
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    if (i < P.n) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float f = P.x[(size_t)D * i + k];
            if (isfinite(f)) {
                mn[k] = f2ord(f);
                mx[k] = mn[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
        }
    }
    __shared__ uint32_t red[kBlock / 64][6];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        red[w][k] = mn[k];   // every lane stores the same wave value
        red[w][3 + k] = mx[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int j = 1; j < kBlock / 64; ++j) v = k < 3 ? min(v, red[j][k]) : max(v, red[j][k]);
        P.part[(tl.y / kBlock) * 6 + k] = v;
    }
    // this tile's share of the job's cell counters, zeroed for the count pass (a kernel, not a
    // memset, so batches record into graphs)
    const int nt = (P.n + kBlock - 1) / kBlock, t = (int)(tl.y / kBlock);
    const int per = (P.n_cells + nt - 1) / nt;
    const int c1 = min((t + 1) * per, P.n_cells);
    for (int c = t * per + threadIdx.x; c < c1; c += kBlock) P.counts[c] = 0u;
}

constexpr int kScanBlock = 1024;

// (2) one block per job: final bounds from the tile partials
__global__ void __launch_bounds__(kScanBlock) k_sort_final(const EngineParams *J) {
    const EngineParams &P = J[blockIdx.x];
    if (!P.perm) return;
    const int nt = (P.n + kBlock - 1) / kBlock;
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (int t = threadIdx.x; t < nt; t += kScanBlock)
        for (int k = 0; k < 6; ++k) v[k] = k < 3 ? min(v[k], P.part[t * 6 + k]) : max(v[k], P.part[t * 6 + k]);
    for (int k = 0; k < 6; ++k)
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)v[k], off);
            v[k] = k < 3 ? min(v[k], o) : max(v[k], o);
        }
    __shared__ uint32_t red[kScanBlock / 64][6];
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < 6; ++k) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t r = red[0][k];
        for (int j = 1; j < kScanBlock / 64; ++j) r = k < 3 ? min(r, red[j][k]) : max(r, red[j][k]);
        P.bounds[k] = r;
    }
}

// (3) fine Morton code of every query (10 bits per axis in 3D, 16 in 2D) and the histogram of
// its cell (the top cell_bits per axis)
template <int D>
__global__ void __launch_bounds__(kBlock) k_sort_count(const EngineParams *J, const uint2 *tiles) {
    const uint2 tl = tile_at(tiles, blockIdx.x, kBlock);
    const EngineParams &P = J[tl.x];
    const int i = (int)tl.y + threadIdx.x;
    if (i >= P.n) return;
    constexpr float kLevels = D == 3 ? 1023.0f : 65535.0f;
    uint32_t code = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float lo = ord2f(P.bounds[k]), hi = ord2f(P.bounds[3 + k]);
        const float ext = hi - lo;
        float t = ext > 0.0f ? (P.x[(size_t)D * i + k] - lo) / ext : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);  // NaN -> 0
        const uint32_t c = (uint32_t)(t * kLevels);
        code |= (D == 3 ? spread3(c) : spread2(c)) << k;
    }
    P.cell[i] = code;
    atomicAdd(&P.counts[code >> P.cell_shift], 1u);
}

// (4) exclusive prefix sum of the cell counters, in place, in three passes over chunks of
// kScanChunk cells: the chunk sums (into P.part, free again once the bounds are final: a job has
// at most n / 16,384 + 1 chunks and n / 256 * 6 part entries), their exclusive scan (one block per
// job), then each chunk's scan plus its offset.  Blocks are (chunk, job); a job with fewer chunks
// leaves the rest idle.
constexpr int kScanChunk = 4 * kScanBlock;
// Exclusive scan of one value per thread over a kScanBlock block; *total receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int j = 0; j < kScanBlock / 64; ++j) {
        before += j < w ? wsum[j] : 0u;
        all += wsum[j];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}
__global__ void __launch_bounds__(kScanBlock) k_scan_sums(const EngineParams *J) {
    const EngineParams &P = J[blockIdx.y];
    const int c0 = (int)blockIdx.x * kScanChunk;
    if (!P.perm || c0 >= P.n_cells) return;
    const int c = c0 + 4 * (int)threadIdx.x;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) v += c + k < P.n_cells ? P.counts[c + k] : 0u;
    uint32_t total;
    (void)block_exclusive_scan(v, &total);
    if (threadIdx.x == 0) P.part[blockIdx.x] = total;
}
__global__ void __launch_bounds__(kScanBlock) k_scan_top(const EngineParams *J) {
    const EngineParams &P = J[blockIdx.x];
    if (!P.perm) return;
    const int nch = (P.n_cells + kScanChunk - 1) / kScanChunk;
    uint32_t carry = 0;
    for (int b = 0; b < nch; b += kScanBlock) {
        const int i = b + (int)threadIdx.x;
        const uint32_t v = i < nch ? P.part[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, &total);
        if (i < nch) P.part[i] = carry + ex;
        carry += total;
    }
}
__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const EngineParams *J) {
    const EngineParams &P = J[blockIdx.y];
    const int c0 = (int)blockIdx.x * kScanChunk;
    if (!P.perm || c0 >= P.n_cells) return;
    const int c = c0 + 4 * (int)threadIdx.x;
    uint32_t n[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) n[k] = c + k < P.n_cells ? P.counts[c + k] : 0u;
    uint32_t total;
    uint32_t run = P.part[blockIdx.x] + block_exclusive_scan(n[0] + n[1] + n[2] + n[3], &total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (c + k < P.n_cells) P.counts[c + k] = run;
        run += n[k];
    }
}
// (5) scatter: each query takes the next slot of its cell
__global__ void __launch_bounds__(kBlock) k_sort_scatter(const EngineParams *J, const uint2 *tiles) {
    const uint2 tl = tile_at(tiles, blockIdx.x, kBlock);
    const EngineParams &P = J[tl.x];
    const int i = (int)tl.y + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pos = atomicAdd(&P.counts[P.cell[i] >> P.cell_shift], 1u);
    P.perm[pos] = (uint32_t)i;
}


// ---------------------------------------------------------------------------------------------
// Traversal launches over a job table.  Packet kernels take 256-query tiles, lane kernels 64-query
// tiles; the wave-per-query passes drain the deferred list of their launch group, whose entries
// name their job.
template <int D, bool Q>
__global__ void __launch_bounds__(kPacketBlock) k_cp_packet(const EngineParams *J, const uint2 *tiles, int spread) {
    const uint2 tl = tile_at(tiles, xcd_tile(blockIdx.x, gridDim.x, spread), kPacketBlock);
    const GP P = gview(J[tl.x]);
    cp_packet<D, Q>(P, (int)tl.y);
}

template <int D, bool Q>
__global__ void __launch_bounds__(kPacketBlock) k_sil_packet(const EngineParams *J, const uint2 *tiles, int spread) {
    const uint2 tl = tile_at(tiles, xcd_tile(blockIdx.x, gridDim.x, spread), kPacketBlock);
    const GP P = gview(J[tl.x]);
    sil_packet<D, Q>(P, (int)tl.y);
}

// Occupancy hints (measured on 1M queries against 327,680 triangles): rays at 7 wavefronts per
// SIMD (71 VGPRs) 0.48 ms vs 0.50 at the compiler's 6; the silhouette wave pass at 8 (64 VGPRs,
// 64 B of spills) 23.8 ms vs 29.0 at 5 (87 VGPRs): its rounds wait on dependent node loads, and
// more resident wavefronts hide them.
template <int D, bool Q>
__global__ void __launch_bounds__(kRayBlock) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_ray_lane(const EngineParams *J, const uint2 *tiles, int spread) {
    const uint2 tl = tile_at(tiles, xcd_tile(blockIdx.x, gridDim.x, spread), kRayBlock);
    const GP P = gview(J[tl.x]);
    ray_lane<D, Q>(P, (int)tl.y);
}

template <int D, bool Q>
__global__ void __launch_bounds__(kRayBlock) k_sil_lane(const EngineParams *J, const uint2 *tiles, int spread) {
    const uint2 tl = tile_at(tiles, xcd_tile(blockIdx.x, gridDim.x, spread), kRayBlock);
    const GP P = gview(J[tl.x]);
    sil_lane<D, Q>(P, (int)tl.y);
}

// Persistent wave-per-query passes.  All lanes take part in the fetch (lane 0 adds 1, the others
// 0) and all lanes store the same record: no lane-dependent branch inside the loop, so the
// compiler cannot split the wavefront across iterations.
template <int D, bool Q>
__global__ void __launch_bounds__(kWaveBlock) k_cp_wave(const EngineParams *J, uint32_t *dc, const uint4 *deferred,
                                                       int cap, uint4 *ovf, uint32_t *ovf_n) {
    extern __shared__ uint2 front[];   // cap entries
    const int lane = threadIdx.x & 63;
    const uint32_t total = uniform(dc[0]);
    for (;;) {
        const uint32_t k = uniform(atomicAdd(dc + 1, lane == 0 ? 1u : 0u));
        if (k >= total) break;
        const uint4 item = deferred[k];
        const GP P = gview(J[uniform(item.w)]);
        cp_wave_item<D, Q>(P, item, front, cap, ovf, ovf_n);
    }
}

template <int D, bool Q>
__global__ void __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) k_sil_wave(const EngineParams *J, uint32_t *dc, const uint4 *deferred,
                                                        int cap, uint4 *ovf, uint32_t *ovf_n) {
    extern __shared__ uint2 front[];   // cap entries
    const int lane = threadIdx.x & 63;
    const uint32_t total = uniform(dc[0]);
    for (;;) {
        const uint32_t k = uniform(atomicAdd(dc + 1, lane == 0 ? 1u : 0u));
        if (k >= total) break;
        const uint4 item = deferred[k];
        const GP P = gview(J[uniform(item.w)]);
        sil_wave_item<D, Q>(P, item, front, cap, ovf, ovf_n);
    }
}

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// ---- host side: plan of a set of jobs -------------------------------------------------------
// A launch group is one (query kind, dimension, node format, traversal) combination: its jobs
// share the traversal kernels and one deferred list.
struct Group {
    QueryKind kind;
    int dim;
    bool quantized;
    bool lane;                        // lane-per-query kernel (rays, lane silhouettes)
    bool dense;                       // 3D silhouette jobs on the long default budget (no wave pass
                                      // to speak of): their own group, so in a large batch the
                                      // other jobs' wave pass runs beside their long packet walks
    std::vector<int> jobs;
    size_t tiles_off = 0, n_tiles = 0;   // tiles of this group's traversal kernel
    size_t dc_off = 0, def_off = 0, ovf_off = 0;
    bool defer = false;
    int max_stack = 0;
    int front = 0;                    // frontier entries of the wave-per-query pass
    int front_small = 0;              // frontier entries of its first (small-frontier) pass
};

struct Plan {
    std::vector<EngineParams> P;
    std::vector<uint2> tiles;         // all tile tables, back to back
    size_t sort_tiles_off[2] = {0, 0}, n_sort_tiles[2] = {0, 0};   // by dimension (2D, 3D)
    std::vector<Group> groups;
    size_t table_bytes = 0;           // job table + tiles (the uploaded head of the workspace)
    size_t ws_bytes = 0;              // whole device workspace
};

int cell_bits_for(int64_t n, int dim) {
    // 4 to 32 queries per cell on average (uniform queries): small enough for the per-cell sort
    int b = 0;
    while (b < (dim == 3 ? 7 : 10) && std::pow(2.0, (double)dim * (b + 1)) <= (double)n / 4.0) ++b;
    return std::max(b, 2);
}

// Whether a job's queries are Morton-sorted before traversal.  Rays are not: origins alone do not
// make rays coherent (directions stay random), and on 1M uniform rays the traversal kernel is as
// fast in input order (3D 0.70 vs 0.69 ms, 2D 0.29 vs 0.33 ms) while the sort costs 0.2 ms.
// Closest-point and silhouette queries gain 1.2x (3D silhouettes) to 18x (2D closest point).
bool sort_job(int kind, int64_t n) { return n >= kSortMin && kind != kIntersect; }
bool sort_job(const EngineJob &jb) { return !jb.input_order && sort_job(jb.kind, jb.q.n); }

// Fills the plan for `jobs` with every device pointer relative to `base` (null when only the
// sizes are wanted).  tiles_implicit: the launch reads no tile table (a single job), so only the
// tile counts are kept.  A plan object reused for the same shape of job list is refilled without
// host allocation (its vectors keep their capacity).
hipError_t make_plan(const std::vector<EngineJob> &jobs, uint8_t *base, Plan &plan, bool tiles_implicit = false) {
    const size_t nj = jobs.size();
    plan.P.assign(nj, EngineParams{});
    plan.tiles.clear();
    plan.table_bytes = plan.ws_bytes = 0;
    size_t n_groups = 0;
    size_t n_tiles_all = 0;
    // groups
    for (size_t j = 0; j < nj; ++j) {
        const EngineJob &jb = jobs[j];
        const WideView &w = *jb.w;
        const bool sil_lane = jb.kind == kSilhouette &&
                              (jb.traversal == 2 || (jb.traversal == 0 && default_sil_lane(w.dim, jb.q.n, w.n_prims))) &&
                              w.bin_depth + 2 <= kRayStackMax;
        const bool lane = jb.kind == kIntersect || sil_lane;
        const bool dense = jb.kind == kSilhouette && w.dim == 3 && !lane && jb.budget == 0 && w.packet_budget == 0 &&
                           (double)jb.q.n >= kSil3DenseRho * (double)std::max<int64_t>(w.n_prims, 1);
        int g = -1;
        for (size_t k = 0; k < n_groups; ++k) {
            const Group &G = plan.groups[k];
            if (G.kind == jb.kind && G.dim == w.dim && G.quantized == jb.quantized && G.lane == lane && G.dense == dense)
                g = (int)k;
        }
        if (g < 0) {
            if (n_groups == plan.groups.size()) plan.groups.emplace_back();
            Group &G = plan.groups[n_groups];
            G.jobs.clear();
            G.tiles_off = G.n_tiles = G.dc_off = G.def_off = G.ovf_off = 0;
            G.defer = false;
            G.max_stack = G.front = G.front_small = 0;
            G.kind = jb.kind;
            G.dim = w.dim;
            G.quantized = jb.quantized;
            G.lane = lane;
            G.dense = dense;
            g = (int)n_groups++;
        }
        plan.groups[g].jobs.push_back((int)j);
    }
    plan.groups.resize(n_groups);
    // tile tables: sort tiles per dimension, then traversal tiles per group
    auto add_tiles = [&](size_t j, int ts) {
        const int64_t n = jobs[j].q.n;
        if (tiles_implicit) {
            n_tiles_all += (size_t)((n + ts - 1) / ts);
            return;
        }
        for (int64_t t = 0; t < n; t += ts) plan.tiles.push_back(make_uint2((uint32_t)j, (uint32_t)t));
        n_tiles_all = plan.tiles.size();
    };
    for (int di = 0; di < 2; ++di) {
        plan.sort_tiles_off[di] = n_tiles_all;
        for (size_t j = 0; j < nj; ++j) {
            if (jobs[j].w->dim != di + 2 || !sort_job(jobs[j])) continue;
            add_tiles(j, kBlock);
        }
        plan.n_sort_tiles[di] = n_tiles_all - plan.sort_tiles_off[di];
    }
    for (Group &G : plan.groups) {
        const int ts = G.lane ? kRayBlock : kPacketBlock;
        G.tiles_off = n_tiles_all;
        for (int j : G.jobs) add_tiles((size_t)j, ts);
        G.n_tiles = n_tiles_all - G.tiles_off;
    }
    // workspace layout: [job table][tiles][per job scratch][per group deferred lists]
    size_t off = align256(nj * sizeof(EngineParams));
    const size_t tiles_at = off;
    off = align256(off + (tiles_implicit ? 0 : n_tiles_all) * sizeof(uint2));
    plan.table_bytes = off;
    auto at = [&](size_t bytes) {
        const size_t o = off;
        off = align256(off + bytes);
        return base ? (void *)(base + o) : nullptr;
    };
    for (size_t j = 0; j < nj; ++j) {
        const EngineJob &jb = jobs[j];
        const WideView &w = *jb.w;
        EngineParams &P = plan.P[j];
        P = EngineParams{};
        P.x = jb.q.x;
        P.v = jb.q.v;
        P.r_max = jb.q.r_max;
        P.r_max_scalar = jb.q.r_max_scalar;
        P.n = (int)jb.q.n;
        P.out = jb.q.out;
        P.big = w.big_leaves;
        P.job = (uint32_t)j;
        P.stack_depth = 3 * w.wide_depth + 4;
        if (jb.kind == kSilhouette) {
            P.nodes = jb.quantized ? w.qsnodes : w.snodes;
            P.prims = w.sils;
            P.ids = w.sil_ids;
            // 2D: the quantized pair records' table (RefCone2); the float32 ones carry theirs inline
            P.ref_cones = w.dim == 2 ? (const RefCone *)w.ref_cones2 : w.ref_cones;
            P.root = w.sroot_ref;
            for (int k = 0; k < 3; ++k) { P.root_lo[k] = w.sroot_lo[k]; P.root_hi[k] = w.sroot_hi[k]; }
            P.seed = (const Seed *)w.sil_seed;
            P.n_seed = w.n_sil_seed;
        } else {
            if (jb.kind == kIntersect) {
                P.nodes = jb.quantized ? w.rqnodes : w.rnodes;
                P.stack_depth = 3 * w.ray_depth + 4;
            } else {
                P.nodes = jb.quantized ? w.qnodes : w.nodes;
                P.cyl = jb.quantized ? w.cbq : w.cbf;
            }
            P.prims = w.prims;
            P.ids = w.prim_ids;
            P.root = w.root_ref;
            P.seed = (const Seed *)w.bvh_seed;
            P.n_seed = w.n_bvh_seed;
        }
        if (sort_job(jb)) {
            P.cell_bits = cell_bits_for(jb.q.n, w.dim);
            P.n_cells = 1 << (P.cell_bits * w.dim);
            P.cell_shift = w.dim * ((w.dim == 3 ? 10 : 16) - P.cell_bits);
            const size_t ncell = (size_t)P.n_cells;
            P.perm = (uint32_t *)at((size_t)jb.q.n * 4);
            P.cell = (uint32_t *)at((size_t)jb.q.n * 4);
            P.counts = (uint32_t *)at(ncell * 4);
            P.part = (uint32_t *)at((size_t)((jb.q.n + kBlock - 1) / kBlock) * 24);
            P.bounds = (uint32_t *)at(32);
        }
    }
    for (Group &G : plan.groups) {
        size_t nq = 0;
        int64_t group_q = 0;
        bool shared = false;
        for (int j : G.jobs) shared = shared || jobs[j].shared;
        // launch size for the small-launch budgets: the queries of every group of this kind,
        // dimension, node format and traversal (dense and other 3D silhouette groups run together)
        for (const Group &H : plan.groups)
            if (H.kind == G.kind && H.dim == G.dim && H.quantized == G.quantized && H.lane == G.lane)
                for (int j : H.jobs) group_q += jobs[j].q.n;
        // beside other launches a small group does not leave the GPU idle: budget it as a full one
        if (shared) group_q = std::max<int64_t>(group_q, kFullLaunch);
        for (int j : G.jobs) {
            EngineParams &P = plan.P[j];
            const WideView &w = *jobs[j].w;
            if (G.lane) {
                if (G.kind == kSilhouette) P.stack_depth = w.bin_depth + 2;
                P.sil_lane = G.kind == kSilhouette;
                G.max_stack = std::max(G.max_stack, P.stack_depth);
                continue;
            }
            // frontier bound: 256 entries per BVH4 level, 128 per binary SNCH level
            const int need = G.kind == kClosestPoint ? 256 * (w.wide_depth + 1) : 128 * (w.bin_depth + 2);
            const bool fits = need <= kFrontMax;
            const int budget = jobs[j].budget != 0 ? jobs[j].budget : w.packet_budget;
            if (fits && budget >= 0) {
                P.budget = budget > 0 ? budget : default_budget((int)G.kind, G.dim, group_q, jobs[j].q.n, w.n_prims);
                G.defer = true;
                G.front = std::max(G.front, need);
                const int small = w.wave_front > 0 ? w.wave_front
                                : G.kind == kSilhouette ? kFrontSmallSil : kFrontSmallCp;
                G.front_small = G.front_small > 0 ? std::min(G.front_small, small) : small;
                nq += (size_t)jobs[j].q.n;
            }
        }
        if (G.defer) {
            uint32_t *dc = (uint32_t *)at(256);
            uint4 *def = (uint4 *)at(nq * sizeof(uint4));
            G.dc_off = base ? (size_t)((uint8_t *)dc - base) : 0;
            G.def_off = base ? (size_t)((uint8_t *)def - base) : 0;
            if (G.front > G.front_small) {
                uint4 *ovf = (uint4 *)at(nq * sizeof(uint4));
                G.ovf_off = base ? (size_t)((uint8_t *)ovf - base) : 0;
            }
            for (int j : G.jobs) {
                if (plan.P[j].budget > 0) {
                    plan.P[j].defer_count = dc;
                    plan.P[j].deferred = def;
                }
            }
        }
    }
    plan.ws_bytes = off;
    (void)tiles_at;
    return hipSuccess;
}

// XCD-contiguous tile order for the traversal launches (env GQ_XCD_SPREAD=0/1 overrides).
int xcd_spread() {
    static const int v = [] {
        const char *e = getenv("GQ_XCD_SPREAD");
        return e ? atoi(e) : 0;
    }();
    return v;
}

// Workgroups of one wavefront that the device holds at once for a kernel with `lds` bytes of
// dynamic LDS (the persistent grid of a wave-per-query pass), per device.
int resident_waves(const void *fn, size_t lds) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    static std::mutex mu;
    static std::map<std::tuple<int, const void *, size_t>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(dev, fn, lds);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWaveBlock, lds) != hipSuccess || per_cu <= 0)
        per_cu = 8;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cache[key] = per_cu * cus;
}

// The wave-per-query pass of a group: a small-frontier pass over the deferred list, then, for the
// queries it handed on, a pass with the group's full frontier.  Counters: dc[0..1] deferred list
// (total, next), dc[2..3] hand-off list.
template <typename K>
void launch_wave_pass(K kernel, const Group &G, const EngineParams *J, uint32_t *dc, const uint4 *def, uint4 *ovf,
                      hipStream_t st) {
    if (G.front > G.front_small) {
        const size_t lds = (size_t)G.front_small * sizeof(uint2);
        hipLaunchKernelGGL(kernel, dim3(resident_waves((const void *)kernel, lds)), dim3(kWaveBlock), lds, st, J, dc,
                           def, G.front_small, ovf, dc + 2);
        hipLaunchKernelGGL(kernel, dim3(kWavePassBlocks), dim3(kWaveBlock), (size_t)G.front * sizeof(uint2), st, J,
                           dc + 2, (const uint4 *)ovf, G.front, (uint4 *)nullptr, (uint32_t *)nullptr);
    } else {
        const size_t lds = (size_t)G.front * sizeof(uint2);
        hipLaunchKernelGGL(kernel, dim3(resident_waves((const void *)kernel, lds)), dim3(kWaveBlock), lds, st, J, dc,
                           def, G.front, (uint4 *)nullptr, (uint32_t *)nullptr);
    }
}

template <int D, bool Q>
void launch_group_t(const Group &G, const EngineParams *J, const uint2 *tiles, uint8_t *base, hipStream_t st) {
    const uint2 *gt = tiles ? tiles + G.tiles_off : nullptr;   // null: implicit tiles of job 0
    const unsigned nt = (unsigned)G.n_tiles;
    if (nt == 0) return;
    uint32_t *dc = G.defer ? (uint32_t *)(base + G.dc_off) : nullptr;
    const uint4 *def = G.defer ? (const uint4 *)(base + G.def_off) : nullptr;
    uint4 *ovf = G.defer && G.front > G.front_small ? (uint4 *)(base + G.ovf_off) : nullptr;
    if (G.kind == kIntersect) {
        hipLaunchKernelGGL((k_ray_lane<D, Q>), dim3(nt), dim3(kRayBlock), 0, st, J, gt, xcd_spread());
        prof_mark_first(st);
    } else if (G.kind == kSilhouette && G.lane) {
        const size_t lds = D == 3 ? 0 : (size_t)G.max_stack * kRayBlock * sizeof(uint2);
        hipLaunchKernelGGL((k_sil_lane<D, Q>), dim3(nt), dim3(kRayBlock), lds, st, J, gt, xcd_spread());
        prof_mark_first(st);
    } else if (G.kind == kSilhouette) {
        if (dc) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, dc, 4);
        hipLaunchKernelGGL((k_sil_packet<D, Q>), dim3(nt), dim3(kPacketBlock), 0, st, J, gt, xcd_spread());
        prof_mark_first(st);
        if (dc) launch_wave_pass(k_sil_wave<D, Q>, G, J, dc, def, ovf, st);
    } else {
        if (dc) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, dc, 4);
        hipLaunchKernelGGL((k_cp_packet<D, Q>), dim3(nt), dim3(kPacketBlock), 0, st, J, gt, xcd_spread());
        prof_mark_first(st);
        if (dc) launch_wave_pass(k_cp_wave<D, Q>, G, J, dc, def, ovf, st);
    }
}

void launch_group(const Group &G, const EngineParams *J, const uint2 *tiles, uint8_t *base, hipStream_t st) {
    if (G.dim == 3) {
        if (G.quantized) launch_group_t<3, true>(G, J, tiles, base, st);
        else launch_group_t<3, false>(G, J, tiles, base, st);
    } else {
        if (G.quantized) launch_group_t<2, true>(G, J, tiles, base, st);
        else launch_group_t<2, false>(G, J, tiles, base, st);
    }
}

// Auxiliary streams of the calling thread on the current device (created once, default priority,
// kept until the thread ends): with the caller's stream, the four hardware queues a process gets.
// Per thread, so a stream that joins one thread's graph capture is never used by another thread.
constexpr int kAuxStreams = 3;
hipError_t aux_streams(hipStream_t (&out)[kAuxStreams]) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    thread_local std::map<int, std::array<hipStream_t, kAuxStreams>> pool;
    auto it = pool.find(dev);
    if (it == pool.end()) {
        std::array<hipStream_t, kAuxStreams> a{};
        for (auto &x : a)
            if ((e = hipStreamCreateWithFlags(&x, hipStreamNonBlocking)) != hipSuccess) return e;
        it = pool.emplace(dev, a).first;
    }
    for (int k = 0; k < kAuxStreams; ++k) out[k] = it->second[k];
    return hipSuccess;
}

// Relative cost of a group's tiles, for spreading groups over the streams.
double group_weight(const Group &G) {
    const double w = G.kind == kIntersect ? 0.3 : G.kind == kClosestPoint ? 1.0 : (G.dim == 3 ? 4.0 : 1.0);
    return w * (double)G.n_tiles;
}

// Enqueues every group's traversal.  Groups are independent (their own tiles, deferred lists and
// counters), so in a large batch with more than one they are spread over `st` and the auxiliary
// streams (largest first, each to the least loaded stream): the auxiliary streams wait for
// everything enqueued on `st` so far and `st` waits for them at the end, so one group's tail -- its
// last waves and the draining wave-per-query pass -- overlaps the other groups' work.  Under a graph
// capture the fork and join record parallel branches.  Batches under kForkMinQueries stay on `st`:
// their callers overlap launches themselves (the serving loop's stream slots), and the four
// hardware queues are shared by all of a process's streams (served at 90% load: 78.8 Mq/s
// unforked vs 54.6 forked; the 40M-query batch: 154.6 ms forked vs 167.7).  A null `tiles` lets a
// single job use the implicit tile table (tile_at with a null table) so the single-call path
// uploads nothing.
constexpr int64_t kForkMinQueries = int64_t(1) << 22;
hipError_t launch_groups(const Plan &plan, uint8_t *base, const uint2 *tiles, hipStream_t st, const char *prof_name) {
    const EngineParams *J = (const EngineParams *)base;
    if (prof_name) prof_begin(st);
    const size_t ng = plan.groups.size();
    int64_t nq = 0;
    for (const EngineParams &P : plan.P) nq += P.n;
    hipStream_t aux[kAuxStreams];
    hipEvent_t fork = nullptr, join[kAuxStreams] = {};
    bool forked = ng > 1 && nq >= kForkMinQueries && aux_streams(aux) == hipSuccess &&
                  hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; forked && k < kAuxStreams; ++k)
        forked = hipEventCreateWithFlags(&join[k], hipEventDisableTiming) == hipSuccess;
    if (forked) forked = hipEventRecord(fork, st) == hipSuccess;
    hipError_t err = hipSuccess;
    if (!forked) {
        for (const Group &G : plan.groups) launch_group(G, J, tiles, base, st);
    } else {
        std::vector<size_t> order(ng);
        for (size_t g = 0; g < ng; ++g) order[g] = g;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return group_weight(plan.groups[a]) > group_weight(plan.groups[b]);
        });
        double load[kAuxStreams + 1] = {};
        bool waited[kAuxStreams] = {};
        for (size_t g : order) {
            int sx = 0;
            for (int k = 1; k <= kAuxStreams; ++k)
                if (load[k] < load[sx]) sx = k;
            load[sx] += group_weight(plan.groups[g]) + 1.0;
            hipStream_t gs = st;
            if (sx > 0) {
                if (!waited[sx - 1]) waited[sx - 1] = hipStreamWaitEvent(aux[sx - 1], fork, 0) == hipSuccess;
                if (waited[sx - 1]) gs = aux[sx - 1];   // a failed wait keeps the group on st
            }
            launch_group(plan.groups[g], J, tiles, base, gs);
        }
        for (int k = 0; k < kAuxStreams; ++k) {
            if (!waited[k]) continue;
            hipError_t e = hipEventRecord(join[k], aux[k]);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, join[k], 0);
            if (e != hipSuccess && err == hipSuccess) err = e;
        }
    }
    if (fork) hipEventDestroy(fork);
    for (hipEvent_t e : join)
        if (e) hipEventDestroy(e);
    if (prof_name) prof_end(st, prof_name);
    return err;
}

// Counting sort of every job with at least kSortMin queries, then the traversal of every group.
hipError_t launch_plan(const Plan &plan, uint8_t *base, bool implicit_tiles, hipStream_t st, const char *prof_name) {
    if (implicit_tiles && (plan.P.size() != 1 || plan.groups.size() != 1)) return hipErrorInvalidValue;
    const EngineParams *J = (const EngineParams *)base;
    const uint2 *tiles = implicit_tiles ? nullptr : (const uint2 *)(base + align256(plan.P.size() * sizeof(EngineParams)));
    const int nj = (int)plan.P.size();
    const bool any_sort = plan.n_sort_tiles[0] + plan.n_sort_tiles[1] > 0;
    if (any_sort) {
        for (int di = 0; di < 2; ++di) {
            const unsigned nt = (unsigned)plan.n_sort_tiles[di];
            if (!nt) continue;
            const uint2 *t = tiles ? tiles + plan.sort_tiles_off[di] : nullptr;
            if (di == 0) hipLaunchKernelGGL(k_sort_bounds<2>, dim3(nt), dim3(kBlock), 0, st, J, t);
            else hipLaunchKernelGGL(k_sort_bounds<3>, dim3(nt), dim3(kBlock), 0, st, J, t);
        }
        hipLaunchKernelGGL(k_sort_final, dim3(nj), dim3(kScanBlock), 0, st, J);
        for (int di = 0; di < 2; ++di) {
            const unsigned nt = (unsigned)plan.n_sort_tiles[di];
            if (!nt) continue;
            const uint2 *t = tiles ? tiles + plan.sort_tiles_off[di] : nullptr;
            if (di == 0) hipLaunchKernelGGL(k_sort_count<2>, dim3(nt), dim3(kBlock), 0, st, J, t);
            else hipLaunchKernelGGL(k_sort_count<3>, dim3(nt), dim3(kBlock), 0, st, J, t);
        }
        int max_cells = 0;
        for (const EngineParams &P : plan.P) max_cells = std::max(max_cells, P.n_cells);
        const unsigned chunks = (unsigned)((max_cells + kScanChunk - 1) / kScanChunk);
        if (chunks > 0) {
            hipLaunchKernelGGL(k_scan_sums, dim3(chunks, nj), dim3(kScanBlock), 0, st, J);
            hipLaunchKernelGGL(k_scan_top, dim3(nj), dim3(kScanBlock), 0, st, J);
            hipLaunchKernelGGL(k_scan_apply, dim3(chunks, nj), dim3(kScanBlock), 0, st, J);
        }
        for (int di = 0; di < 2; ++di) {
            const unsigned nt = (unsigned)plan.n_sort_tiles[di];
            if (!nt) continue;
            const uint2 *t = tiles ? tiles + plan.sort_tiles_off[di] : nullptr;
            hipLaunchKernelGGL(k_sort_scatter, dim3(nt), dim3(kBlock), 0, st, J, t);
        }
    }
    const hipError_t e = launch_groups(plan, base, tiles, st, prof_name);
    const hipError_t l = hipGetLastError();
    return e != hipSuccess ? e : l;
}

} // namespace

bool engine_supported(QueryKind kind, const WideView &w) {
    if (kind == kSilhouette) return w.bin_depth + 2 <= kPacketDepth;
    if (kind == kClosestPoint) return 3 * w.wide_depth + 4 <= kPacketDepth;
    return 3 * w.ray_depth + 4 <= kRayStackMax;
}

hipError_t launch_engine(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q, hipStream_t st) {
    static const char *names[2][3] = {
        {"engine_closest_point_2d", "engine_intersect_2d", "engine_silhouette_2d"},
        {"engine_closest_point_3d", "engine_intersect_3d", "engine_silhouette_3d"}};
    if (q.n <= 0) return hipSuccess;
    const std::vector<EngineJob> jobs{EngineJob{&w, kind, quantized, q, 0, w.traversal}};
    Plan plan;
    hipError_t e = make_plan(jobs, nullptr, plan, true);
    if (e != hipSuccess) return e;
    // The counting sort of the batch path (no library call): 2.61 vs 2.64 ms with a rocPRIM radix
    // sort of the full Morton codes at 1M queries, 1.99 vs 2.06 ms at 512k.
    void *buf = nullptr;   // stream-ordered scratch: calls on different streams never share it
    if ((e = hipMallocAsync(&buf, plan.ws_bytes, st)) != hipSuccess) return e;
    uint8_t *base = (uint8_t *)buf;
    make_plan(jobs, base, plan, true);
    hipLaunchKernelGGL(k_store_params, dim3(1), dim3(64), 0, st, plan.P[0], (EngineParams *)base);
    e = launch_plan(plan, base, true, st, names[w.dim - 2][kind]);
    const hipError_t f = hipFreeAsync(buf, st);
    return e != hipSuccess ? e : f;
}

// ---- allocation-free single call ---------------------------------------------------------------
// The scratch of gq_closest_point_bvh_noalloc: a device array of the code object, one copy per GPU,
// held until the process ends.  Host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (the name itself is a host-side stand-in).  Every pass of the call
// clears what it needs with a kernel (the counting sort's cell counters, the deferral counters), so
// the array's contents from the previous call do not matter; calls that overlap in time on one GPU
// share it and race.
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Address of `scratch` on the current device, looked up on the first call per device.
static hipError_t scratch_on_device(uint8_t **out) {
    static std::mutex mu;
    static std::map<int, uint8_t *> ptrs;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(mu);
    auto it = ptrs.find(dev);
    if (it == ptrs.end()) {
        void *p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        it = ptrs.emplace(dev, (uint8_t *)p).first;
    }
    *out = it->second;
    return hipSuccess;
}

hipError_t launch_engine_noalloc(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q, hipStream_t st,
                                 size_t *need) {
    *need = 0;
    if (q.n <= 0) return hipSuccess;
    // host-side plan reused across calls on this thread: no host allocation once it has its size
    thread_local std::vector<EngineJob> jobs(1);
    thread_local Plan plan;
    jobs[0] = EngineJob{&w, kind, quantized, q, 0, w.traversal};
    hipError_t e = make_plan(jobs, nullptr, plan, true);
    if (e != hipSuccess) return e;
    *need = plan.ws_bytes;
    if (plan.ws_bytes > kNoAllocScratchBytes) return hipErrorInvalidValue;
    uint8_t *base = nullptr;
    if ((e = scratch_on_device(&base)) != hipSuccess) return e;
    make_plan(jobs, base, plan, true);
    hipLaunchKernelGGL(k_store_params, dim3(1), dim3(64), 0, st, plan.P[0], (EngineParams *)base);
    // the counting sort of the batch path (no library temporary storage), then the traversal
    return launch_plan(plan, base, true, st, nullptr);
}

hipError_t engine_batch_sizes(const std::vector<EngineJob> &jobs, size_t &ws_bytes, size_t &table_bytes) {
    Plan plan;
    const hipError_t e = make_plan(jobs, nullptr, plan);
    ws_bytes = plan.ws_bytes;
    table_bytes = plan.table_bytes;
    return e;
}

hipError_t launch_engine_batch(const std::vector<EngineJob> &jobs, void *ws, size_t ws_bytes, void *stage,
                               size_t stage_bytes, hipStream_t st) {
    Plan plan;
    hipError_t e = make_plan(jobs, (uint8_t *)ws, plan);
    if (e != hipSuccess) return e;
    if (plan.ws_bytes > ws_bytes || plan.table_bytes > stage_bytes) return hipErrorInvalidValue;
    uint8_t *h = (uint8_t *)stage;
    std::memcpy(h, plan.P.data(), plan.P.size() * sizeof(EngineParams));
    std::memcpy(h + align256(plan.P.size() * sizeof(EngineParams)), plan.tiles.data(), plan.tiles.size() * sizeof(uint2));
    if ((e = hipMemcpyAsync(ws, stage, plan.table_bytes, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    return launch_plan(plan, (uint8_t *)ws, false, st, nullptr);
}

} // namespace gqa
