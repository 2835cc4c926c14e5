// Engine layout, built on the host from the reference-layout BVH / SNCH:
//
//  * BVH4: the binary SAH tree collapsed to 4-wide nodes (each node's children are the up to four
//    largest-area descendants of a binary node), small subtrees folded into leaves.  Two formats,
//    both one or two 64-byte cache lines, structure-of-arrays over the 4 children:
//      QNode<D>  64 B: child boxes quantized to 8 bits (3D) / 16 bits (2D) relative to the node's
//                      box, rounded outward so every decoded child box contains the true one;
//      FNode<D> 128 B: child boxes in float32.
//  * SNCH pairs: one record per binary interior node holding both children (box, reference
//    normal-cone index, exact-cull cone, child reference) so a visit is one record fetch.
//      SPairQ<D> 64 B (boxes quantized to 8 bits in 3D / 16 bits in 2D, int16 cones, 8-bit
//      silhouette extents, reference cones by table index), SPairF<D> 128 B in 3D / 192 B in 2D
//      (float32; the 2D record carries its reference cones inline).
//  * primitives in leaf order as float4 records (3D: a, b, c with a circumscribing disk in the
//    free lanes and a fourth record; 2D: a|b), silhouettes in leaf
//    order (3D: a, b, n0, n1 with the predicate normals precomputed; 2D: a|b, c).
// Child references: bit 31 set = leaf, bits 24..30 = count, bits 0..23 = first primitive /
// silhouette; otherwise the index of the child node / pair record.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <cstring>
#include <vector>

#include "builder.h"
#include "gq_internal.h"

namespace gqa {

constexpr uint32_t kLeafBit = 0x80000000u;
// Quiet-NaN bit pattern in the n1.x slot of a 3D silhouette record: the edge has one adjacent face
// and is a silhouette for every query point.
constexpr uint32_t kSilAlwaysTag = 0x7FC5B0EDu;
inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
constexpr int kMaxLeafPrims = 127;
__host__ __device__ inline bool ref_is_leaf(uint32_t r) { return (r & kLeafBit) != 0; }
__host__ __device__ inline uint32_t ref_count(uint32_t r) { return (r >> 24) & 0x7Fu; }
__host__ __device__ inline uint32_t ref_first(uint32_t r) { return r & 0xFFFFFFu; }
__host__ __device__ inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return kLeafBit | (count << 24) | first;
}

template <int D> struct QBits;                 // quantized coordinate type per dimension
template <> struct QBits<3> { using T = uint8_t; static constexpr int levels = 255; };
template <> struct QBits<2> { using T = uint16_t; static constexpr int levels = 65535; };

// 4-wide BVH node, quantized child boxes.
template <int D> struct QNode;
template <> struct alignas(64) QNode<3> {
    float origin[3];
    int8_t exp[3];          // per-axis scale 2^exp
    uint8_t nchild;
    uint8_t qlo[3][4], qhi[3][4];
    uint32_t child[4];
    uint8_t pad[8];
};
template <> struct alignas(64) QNode<2> {
    float origin[2];
    int8_t exp[2];
    uint8_t nchild, pad0;
    uint16_t qlo[2][4], qhi[2][4];
    uint32_t child[4];
    uint8_t pad[4];
};
static_assert(sizeof(QNode<3>) == 64, "QNode<3> must be one 64-byte line");
static_assert(sizeof(QNode<2>) == 64, "QNode<2> must be one 64-byte line");

// 4-wide BVH node, float32 child boxes.
template <int D> struct alignas(64) FNode {
    float lo[D][4], hi[D][4];
    uint32_t child[4];
    uint32_t nchild;
    uint8_t pad[128 - (2 * D * 16 + 20)];
};
static_assert(sizeof(FNode<3>) == 128 && sizeof(FNode<2>) == 128, "FNode is two lines");

// Bounding cylinders of the BVH4 children (closest point), one record per node beside QNode /
// FNode.  Child j's primitives lie in the slab lo[j] <= (p - c).u <= hi[j] along an axis u and
// within distance rad[j] of the line c + t u, where u is the child's area-weighted mean normal (3D
// cross(b - a, c - a), 2D (dy, -dx): the reference's normal conventions) and c the centre of the
// child box as the kernels compute it, in float32: fma(qlo + qhi, 2^(e - 1), origin) for the
// quantized nodes, 0.5f * (lo + hi) for the float32 ones (so each node format has its own array).  For |u| <= 1, with s = (x - c).u and q = (x - c) - s u,
// sqrt(max(lo - s, s - hi, 0)^2 + max(|q| - rad, 0)^2) is a lower bound on the distance from x to
// anything below the child: near the true distance for the thin curved patches of a surface, where
// the box bound is loose (a query behind a patch is inside its box, one in front sees the box's
// corners).  Every field is the upper half of a float32 pattern, widened by the kernels with a
// shift: axis components rounded toward zero (so |u| <= 1), lo down, hi and rad up.  A zero axis
// (normals that cancel) leaves a sphere about c.
template <int D> struct alignas(64) CBound {
    uint16_t ax[D][4];
    uint16_t lo[4], hi[4], rad[4];
    uint8_t pad[64 - 8 * D - 24];
};
static_assert(sizeof(CBound<3>) == 64 && sizeof(CBound<2>) == 64, "CBound is one line");

// Exact-cull cone of an SNCH child: every predicate normal (oriented consistently) lies within
// angle acos(cos_t) of `axis`; cos_t/sin_t are rounded so the cone is conservative.  cos_t <= 0
// marks a cone that never culls.
struct ConeF {
    float axis[3];
    float cos_t, sin_t;
};

// SNCH pair record (both children of one binary interior node), float32.
template <int D> struct alignas(64) SPairF {
    float lo[2][D], hi[2][D];
    ConeF cone[2];
    uint32_t child[2];
    int32_t ref_cone[2];    // index into the reference-cone table, -1 = the reference test cannot cull
    // extent of the silhouette edges below about the child box centre c = 0.5f * (lo + hi), in the
    // frame of the child's exact-cull cone axis u: every edge endpoint p has (p - c).u in
    // [sil_lo, sil_hi] and lies within sil_perp of the line c + t u (rounded outward)
    float sil_lo[2], sil_hi[2], sil_perp[2];
};
// 2D: the reference cone test can cull at nearly every node (its float32 half-angles are small),
// and it needs the node's float32 box and cone exactly, so the 2D pair record carries both
// children's reference boxes and cones inline: one record per visit instead of a pair record plus
// two reference-cone fetches.  It also carries each child's exact-cull cone (over the vertex
// predicate's segment normals) and the extent of the silhouette vertices below along and across
// its axis.  2D silhouettes use this record for both node formats.
template <> struct alignas(64) SPairF<2> {
    float lo[2][2], hi[2][2];            // the children's boxes (SNCH node data, float32)
    float cone_axis[2][2];               // reference cone axis of each child
    float cone_half[2], cone_radius[2];  // reference half-angle and radius
    uint32_t child[2];
    int32_t ref_cone[2];                 // 0 = the reference cone test can cull, -1 = it cannot
    int32_t cull[2];                     // 1 = no silhouette below this child (always skipped)
    float cone_sin[2], cone_cos[2];      // sin and cos of the half-angle (the kernels' fast path)
    float ecos[2], esin[2];              // exact-cull cone: cos / sin of its angle (cos 2 = no
                                         // silhouette below, cos <= 0 = never culls)
    float eaxis[2][2];                   // exact-cull cone axis
    float elo[2], ehi[2], eperp[2];      // vertices below about the box centre: along / across the axis
    uint8_t pad[32];
};
static_assert(sizeof(SPairF<3>) == 128 && sizeof(SPairF<2>) == 192, "SPairF: two lines in 3D, three in 2D");

// SNCH pair record, quantized: child boxes in 8 bits (3D) / 16 bits (2D) relative to the pair's
// frame, rounded outward; exact-cull cones as an int16 axis (its rounding folded into the angle)
// and the int16 sine of the widened angle, rounded up (32767: never culls; the kernels form the
// cosine from it, which keeps narrow cones narrow); the silhouette points' extent about the decoded
// box centre, in units of R, the decoded box's half-diagonal: along the decoded axis a slab of
// centre ext_mid * R / 32767 (rounded to nearest) and half-width (ext_half / 255)^2 * R (rounded
// up, fine near zero, where thin slabs sit), across the axis ext_perp * R / 255 (rounded up);
// reference cones in a table (3D: RefCone, 2D: RefCone2), child j's at
// ref_base + (j == 1 && child 0 has one).
constexpr uint8_t kPairRef0 = 1, kPairRef1 = 2, kPairNone0 = 4, kPairNone1 = 8;   // SPairQ flags
template <int D> struct SPairQ;
template <> struct alignas(64) SPairQ<3> {
    float origin[3];
    int8_t exp[3];
    uint8_t flags;             // kPairRef<j>: child j has a reference cone; kPairNone<j>: no silhouette below
    uint8_t qlo[2][3], qhi[2][3];
    int16_t cone_axis[2][3];
    int16_t cone_sin[2];
    uint32_t child[2];
    int32_t ref_base;
    int16_t ext_mid[2];
    uint8_t ext_half[2], ext_perp[2];
};
template <> struct alignas(64) SPairQ<2> {
    float origin[2];
    int8_t exp[2];
    uint8_t flags, pad0;
    uint16_t qlo[2][2], qhi[2][2];
    int16_t cone_axis[2][2];
    int16_t cone_sin[2];
    uint32_t child[2];
    int32_t ref_base;
    int16_t ext_mid[2];
    uint8_t ext_half[2], ext_perp[2];
    uint8_t pad[4];
};
static_assert(sizeof(SPairQ<3>) == 64 && sizeof(SPairQ<2>) == 64, "SPairQ is one line");
// byte offsets the kernels decode (engine.hip load_pair / load_node4)
static_assert(offsetof(SPairQ<3>, flags) == 15 && offsetof(SPairQ<3>, qlo) == 16 && offsetof(SPairQ<3>, qhi) == 22 &&
              offsetof(SPairQ<3>, cone_axis) == 28 && offsetof(SPairQ<3>, cone_sin) == 40 &&
              offsetof(SPairQ<3>, child) == 44 && offsetof(SPairQ<3>, ref_base) == 52 &&
              offsetof(SPairQ<3>, ext_mid) == 56 && offsetof(SPairQ<3>, ext_half) == 60 &&
              offsetof(SPairQ<3>, ext_perp) == 62, "SPairQ<3>");
static_assert(offsetof(SPairQ<2>, flags) == 10 && offsetof(SPairQ<2>, qlo) == 12 && offsetof(SPairQ<2>, qhi) == 20 &&
              offsetof(SPairQ<2>, cone_axis) == 28 && offsetof(SPairQ<2>, cone_sin) == 36 &&
              offsetof(SPairQ<2>, child) == 40 && offsetof(SPairQ<2>, ref_base) == 48 &&
              offsetof(SPairQ<2>, ext_mid) == 52 && offsetof(SPairQ<2>, ext_half) == 56 &&
              offsetof(SPairQ<2>, ext_perp) == 58, "SPairQ<2>");
static_assert(offsetof(SPairF<3>, cone) == 48 && offsetof(SPairF<3>, child) == 88 && offsetof(SPairF<3>, ref_cone) == 96 &&
              offsetof(SPairF<3>, sil_lo) == 104 && offsetof(SPairF<3>, sil_perp) == 120, "SPairF<3>");
static_assert(offsetof(SPairF<2>, cone_axis) == 32 && offsetof(SPairF<2>, cone_half) == 48 &&
              offsetof(SPairF<2>, cone_radius) == 56 && offsetof(SPairF<2>, child) == 64 &&
              offsetof(SPairF<2>, ref_cone) == 72 && offsetof(SPairF<2>, cull) == 80 &&
              offsetof(SPairF<2>, cone_sin) == 88 && offsetof(SPairF<2>, cone_cos) == 96 &&
              offsetof(SPairF<2>, ecos) == 104 && offsetof(SPairF<2>, esin) == 112 &&
              offsetof(SPairF<2>, eaxis) == 120 && offsetof(SPairF<2>, elo) == 136 &&
              offsetof(SPairF<2>, ehi) == 144 && offsetof(SPairF<2>, eperp) == 152, "SPairF<2>");
static_assert(offsetof(QNode<3>, qlo) == 16 && offsetof(QNode<3>, qhi) == 28 && offsetof(QNode<3>, child) == 40, "QNode<3>");
static_assert(offsetof(QNode<2>, nchild) == 10 && offsetof(QNode<2>, qlo) == 12 && offsetof(QNode<2>, qhi) == 28 &&
              offsetof(QNode<2>, child) == 44, "QNode<2>");
static_assert(offsetof(FNode<3>, child) == 96 && offsetof(FNode<3>, nchild) == 112, "FNode<3>");
static_assert(offsetof(FNode<2>, child) == 64 && offsetof(FNode<2>, nchild) == 80, "FNode<2>");

// Reference normal cone (snch.cpp:102-154 values, float32) of a node whose cone test can cull
// (0 <= half_angle <= pi/2); tested exactly as SNCHNode.overlaps does.
struct alignas(16) RefCone {
    float axis[3];
    float half, radius;
    float lo[3], hi[3];   // the node's box (the reference test uses its centroid and p_max)
    float sin_h, cos_h;   // sin and cos of the half-angle (the kernels' fast path)
    float pad[3];
};
static_assert(sizeof(RefCone) == 64, "RefCone is four float4");
// 2D reference cone of a child of a quantized pair record: the float32 box and cone the 2D pair
// record carries inline (three float4).
struct RefCone2 {
    float lo[2], hi[2];
    float axis[2];
    float half, radius;
    float sin_h, cos_h;   // sin and cos of the half-angle
    float pad[2];
};
static_assert(sizeof(RefCone2) == 48, "RefCone2 is three float4");

// One entry of a seed cut: a node (child ref) below the root and its float32 box.  The
// wave-per-query passes start from the cut (the root's descendants at a fixed depth, as far as the
// reference cone tests allow) instead of from the root.
constexpr int kSeedMax = 64;
struct alignas(32) Seed {
    float lo[3];
    uint32_t ref;
    float hi[3];
    float pad;
};
static_assert(sizeof(Seed) == 32, "Seed is two float4");

struct WideHost {
    int width = 4;
    int wide_depth = 0;
    size_t n_nodes = 0, n_snch_nodes = 0;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    uint32_t root_ref = 0;       // BVH4 root (node 0, or a leaf ref for a one-leaf tree)
    uint32_t sroot_ref = 0;      // SNCH root pair (0) or leaf ref
    float sroot_lo[3] = {0, 0, 0}, sroot_hi[3] = {0, 0, 0};
    std::vector<uint8_t> nodes, qnodes, snodes, qsnodes;  // raw records
    std::vector<uint8_t> cbq, cbf;                        // CBound<D> per BVH4 node, quantized / float32 centres
    // the rays' BVH4: the same binary tree collapsed with its own leaf size (leaf references index
    // the same primitive slots)
    int ray_depth = 0;
    size_t n_ray_nodes = 0;
    std::vector<uint8_t> rnodes, rqnodes;
    std::vector<float4> prims;
    std::vector<int32_t> prim_ids;
    std::vector<float4> sils;
    std::vector<int32_t> sil_ids;
    std::vector<RefCone> ref_cones;
    std::vector<RefCone2> ref_cones2;
    std::vector<uint32_t> big_leaves;  // (first, count) pairs of leaves with more than 126 entries
    std::vector<Seed> bvh_seed, sil_seed;   // seed cuts of the BVH4 and of the SNCH pairs
};

// max_leaf_prims: leaf size of the closest-point BVH4; ray_leaf_prims: of the rays' BVH4.
void build_wide(const HostStructure &h, WideHost &out, int max_leaf_prims, int ray_leaf_prims);

// One query batch for the engine: a scene's device layout, the query kind and node format, the
// caller's arrays, and per-batch overrides (0 = the scene's setting).
struct EngineJob {
    const WideView *w;
    QueryKind kind;
    bool quantized;
    QueryArgs q;
    int budget;
    int traversal;
    bool input_order = false;   // traverse in the caller's order (no Morton sort)
    bool shared = false;        // other launches share the device: full-launch default budgets
};

// One batch on `st`: scratch comes from the stream-ordered allocator and is released after use.
hipError_t launch_engine(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q, hipStream_t st);
// As launch_engine without allocating: scratch comes from a fixed device array of
// kNoAllocScratchBytes (one per GPU).  *need receives the bytes the call needs; more than the array
// holds gives hipErrorInvalidValue and launches nothing.
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
hipError_t launch_engine_noalloc(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q, hipStream_t st,
                                 size_t *need);
// Many batches (any scenes, kinds and sizes on one device) in one enqueue: sort passes over all of
// them, then one traversal launch per (kind, dimension, node format, traversal) group.  The job
// table and tile lists are staged in `stage` (pinned host memory, `table_bytes` long, left
// untouched until the stream has passed this call) and copied to the head of `ws`.  Nothing is
// allocated or cleared with memset, so the enqueue can be captured into a graph.
hipError_t engine_batch_sizes(const std::vector<EngineJob> &jobs, size_t &ws_bytes, size_t &table_bytes);
hipError_t launch_engine_batch(const std::vector<EngineJob> &jobs, void *ws, size_t ws_bytes, void *stage,
                               size_t stage_bytes, hipStream_t st);
// Whether the engine's fixed-capacity traversal stack covers this scene's tree depth.
bool engine_supported(QueryKind kind, const WideView &w);

} // namespace gqa
