// Internal definitions shared by the C-ABI layer and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/gquery_amd.h"

namespace gqa {

// Device view of a scene in the reference layout (flat float32 AoS records).
struct PlainView {
    int dim;
    int64_t n_nodes, n_prims, n_sils;
    const float *bvh_nodes;   // stride 7 / 9
    const float *prims;       // stride 5 / 10
    const float *snch_nodes;  // stride 13 / 16 (nullptr for a BVH-only scene)
    const float *sils;        // stride 10 / 19
};

// Engine layout (built on the host by wide.cpp, uploaded by the API layer).
struct RefCone;
// Engine BVH4 leaf size when the scene options give none (the autotuner's choice on the 3D and 2D
// closest-point workloads; rays are within 3% of their best at any size from 2 to 6).
// Default BVH4 leaf sizes: the closest-point tree (its bounding cylinders and queued leaf drains
// make large leaves pay; 1M uniform queries, 3D 327,680 triangles: leaf 6 3.47 ms, 16 2.91, 24
// 2.76, 32 2.79, 48 3.07; 2D 65,536 segments: 6 2.30, 16 2.12, 24 2.20) and the ray tree (a lane
// tests its leaf serially: 3D 0.49 ms at 6, 0.63 at 16, 0.75 at 24).
constexpr int kDefaultCpLeaf3 = 24, kDefaultCpLeaf2 = 16;
inline int default_cp_leaf(int dim) { return dim == 3 ? kDefaultCpLeaf3 : kDefaultCpLeaf2; }
// Rays test a leaf's primitives one after another in their lane, so their tree wants small leaves
// (1M uniform rays, ms: 327,680 triangles 0.463 at 2, 0.487 at 4, 0.498 at 6, 0.522 at 8; 20,480
// triangles 0.356 at 2, 0.366 at 6; 524,288-segment star 0.435 at 2 or 3, 0.501 at 4, 0.535 at 6),
// except a mid-size polyline, whose segment tests are cheaper than a node step (65,536-segment
// star 0.348 at 6, 0.350 at 4, 0.401 at 3, 0.442 at 2).
constexpr int kDefaultRayLeaf3 = 2, kDefaultRayLeaf2 = 6, kDefaultRayLeaf2Large = 3;
constexpr int64_t kRayLeaf2LargeScene = int64_t(1) << 18;
inline int default_ray_leaf(int dim, int64_t n_prims) {
    if (dim == 3) return kDefaultRayLeaf3;
    return n_prims >= kRayLeaf2LargeScene ? kDefaultRayLeaf2Large : kDefaultRayLeaf2;
}
constexpr int kMaxLeafOption = 64;   // largest gq_scene_options.leaf_prims / ray_leaf_prims

struct WideView {
    int dim;
    int width;                // children per node (4)
    int wide_depth;           // deepest node of the closest-point BVH4 (root = 0)
    int ray_depth;            // deepest node of the ray BVH4
    int bin_depth;            // deepest binary node
    int64_t n_nodes, n_snch_nodes, n_prims, n_sils;
    const void *nodes;        // FNode<D>[n_nodes]  float32 BVH4
    const void *qnodes;       // QNode<D>[n_nodes]  quantized BVH4
    const void *cbq, *cbf;    // CBound<D>[n_nodes]  child bounding cylinders about the QNode / FNode box centres
    const void *rnodes;       // FNode<D>  float32 BVH4 of the rays (its own leaf size)
    const void *rqnodes;      // QNode<D>  quantized BVH4 of the rays
    const void *snodes;       // SPairF<D>[n_snch_nodes]
    const void *qsnodes;      // SPairQ<D>[n_snch_nodes]
    const float4 *prims;      // 3D: 3 float4 per triangle (a, b, c); 2D: 1 float4 per segment (a|b)
    const int32_t *prim_ids;  // original id per ordered slot
    const float4 *sils;       // 3D: 3 float4 per silhouette (a, b, n0, n1); 2D: 2 float4 (a|b, c)
    const int32_t *sil_ids;
    const RefCone *ref_cones;
    const void *ref_cones2;   // RefCone2[]: 2D reference cones of the quantized pair records
    const uint32_t *big_leaves;
    const void *bvh_seed, *sil_seed;   // Seed[n_bvh_seed], Seed[n_sil_seed]
    int n_bvh_seed, n_sil_seed;
    uint32_t root_ref, sroot_ref;
    float sroot_lo[3], sroot_hi[3];
    float scene_lo[3], scene_hi[3];
    int packet_budget;        // 0 = default, < 0 = no deferral to the wave-per-query pass
    int traversal;            // silhouettes: 0 auto, 1 packet + wave pass, 2 lane per query
    int wave_front;           // first wave-pass frontier entries (0 = default)
    int sil_pairs;            // silhouettes under the quantized method: 0 float32 pairs, 1 quantized
};

struct QueryArgs {
    const float *x;
    const float *v;           // rays only
    const float *r_max;       // optional per-query
    float r_max_scalar;
    int64_t n;
    gq_records out;
};

enum QueryKind { kClosestPoint = 0, kIntersect = 1, kSilhouette = 2 };

// Boundary sampling: query points, ball radii, per-query PCG32 state (advanced in place) and
// stream increment.
struct SampleArgs {
    const float *x;
    const float *R;           // optional per-query
    float R_scalar;
    uint64_t *state;
    const uint64_t *inc;
    int64_t n;
    gq_sample_records out;
};
hipError_t launch_sample(const PlainView &s, const SampleArgs &q, hipStream_t st);

hipError_t launch_plain(QueryKind kind, const PlainView &s, const QueryArgs &q, hipStream_t st);

// Device view of a polyline / mesh for the brute-force (baseline) queries and the normals.
struct ShapeView {
    int dim;
    int64_t n_vertices, n_prims, n_sils;
    const float *V;              // n_vertices x dim
    const int32_t *F;            // n_prims x dim
    const float *face_normals;   // 3D: n_prims x 3
    const float *sils;           // 2D: configure_silhouette records (10 floats); 3D: flat edges (19)
    const int32_t *vf_start;     // 3D: faces of each vertex (CSR, configure_normal order)
    const int32_t *vf_list;
};

hipError_t launch_baseline(QueryKind kind, const ShapeView &s, const QueryArgs &q, const float *normal,
                           const uint8_t *on_boundary, hipStream_t st);
hipError_t launch_normals(const ShapeView &s, float *face_normals, float *vertex_normals, hipStream_t st);
// Silhouette setups on the device (silprep.hip).  *bad receives 1 (index out of range), 2 (non-finite
// coordinate) or 4 (out holds fewer than *n_floats floats); they wait for `st` before returning.
hipError_t configure_silhouette_device(const float *V, int64_t nv, const int32_t *E, int64_t ne, float *out,
                                       int64_t capacity, int64_t *n_floats, uint32_t *bad, hipStream_t st);
hipError_t flat_silhouette_edges_device(const float *V, int64_t nv, const int32_t *F, int64_t nf, float *out,
                                        int64_t capacity, int64_t *n_floats, uint32_t *bad, hipStream_t st);

void set_error(const std::string &msg);

// Traversal-kernel timing (gq_profile_begin / gq_profile_end); no-ops unless enabled.
void prof_begin(hipStream_t st);
void prof_end(hipStream_t st, const char *kernel);
// Marks the end of the first traversal kernel of the launch being profiled.
void prof_mark_first(hipStream_t st);

} // namespace gqa
