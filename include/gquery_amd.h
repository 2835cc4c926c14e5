/* gquery_amd -- C ABI of the MI355X traversal engine for gQuery's BVH / SNCH queries.
 *
 * The library replaces two layers of gQuery:
 *   1. the nanobind extension gquery_ext (src/gquery.cpp:180-250) that builds the BVH / SNCH on
 *      the CPU and hands flat float32 arrays (node_data / primitive_data / silhouette_data) to
 *      Python -> gq_build_* / gq_structure_* / gq_build_flat_silhouette_edges;
 *   2. the DrJit-traced traversals BVH.closest_point / BVH.intersect / SNCH.closest_silhouette and
 *      their 3D twins (gquery/shapes/{bvh,bvh3d,snch,snch3d}.py) -> gq_closest_point_bvh /
 *      gq_intersect_bvh / gq_closest_silhouette_snch, run as HIP kernels on a device scene.
 *
 * Conventions: plain pointers and sizes only.  Query inputs and record outputs are device
 * pointers (any hipMalloc'd / torch memory of the current device) laid out row-major:
 * a point batch is n x dim float32, records are structure-of-arrays (one pointer per field).
 * `stream` is a hipStream_t passed as void* (NULL = default stream); every query call is
 * stream-ordered and returns as soon as its work is enqueued.  Every function returns GQ_OK (0) or
 * a negative status; gq_last_error() gives the message of the last failure on the calling thread.
 * Scene buffers are read-only after gq_scene_create, so concurrent queries on distinct streams
 * against one scene are safe.
 *
 * Environment: GQ_XCD_SPREAD=1 hands each of the 8 XCDs one contiguous run of the engine's query
 * tiles instead of dealing tiles round robin (read once per process; default 0).
 */
#ifndef GQUERY_AMD_H
#define GQUERY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GQ_OK 0
#define GQ_ERR_INVALID (-1)     /* bad argument (null pointer, wrong dim, index out of range) */
#define GQ_ERR_OOM (-2)         /* device or host allocation failed */
#define GQ_ERR_DEVICE (-3)      /* HIP runtime error */
#define GQ_ERR_UNSUPPORTED (-4) /* query not available on this scene (e.g. silhouettes of a BVH) */

/* Traversal method of a query call. */
#define GQ_METHOD_PLAIN 0       /* reference layout: binary AoS float32 nodes, lane per query */
#define GQ_METHOD_ENGINE_FP32 1 /* wide SoA nodes with float32 bounds, sorted queries */
#define GQ_METHOD_ENGINE 2      /* wide nodes with 8-bit (3D) / 16-bit (2D) quantized bounds;
                                   silhouettes on float32 pairs unless sil_pairs = 1 (below) */

typedef struct gq_structure gq_structure; /* host-side BVH or SNCH (CPU builder output) */
typedef struct gq_scene gq_scene;         /* device-resident scene */

/* Output records, one field per pointer (a field may be NULL to skip it).  Replaces the DrJit
 * dataclasses of gquery/shapes/primitive.py:9-84.  Misses are all-zero (valid = 0, d = 0,
 * prim_id = 0), as dr.zeros(record) is in the reference (bvh.py:194, bvh3d.py:170). */
typedef struct gq_records {
    uint8_t *valid;       /* n                                                         */
    float *p;             /* n x dim                                                   */
    float *n;             /* n x dim   (closest point / intersection only)             */
    float *uv;            /* 3D: n x 2 uv; 2D: n  t   (closest point / intersection)   */
    float *d;             /* n                                                         */
    int32_t *prim_id;     /* n         original primitive / silhouette id               */
    uint8_t *on_boundary; /* n         (intersection only)                             */
} gq_records;

/* Boundary sample records: BoundarySamplingRecord / BoundarySamplingRecord3D
 * (gquery/shapes/primitive.py:51-67) and the hit flag BVH3D.sample_boundary returns beside them
 * (bvh3d.py:247).  A query without a sample is all zero. */
typedef struct gq_sample_records {
    uint8_t *hit;     /* n                                                                  */
    float *p;         /* n x dim   sampled point                                            */
    float *n;         /* n x dim   unit normal of the sampled primitive                     */
    float *uv;        /* 3D: n x 2 (u, v) weights of a and b; 2D: n  t along the segment   */
    float *pdf;       /* n         pmf x density of the point on the primitive             */
    float *pmf;       /* n         probability of the sampled primitive                    */
    int32_t *prim_id; /* n         original primitive id                                   */
} gq_sample_records;

int gq_version(void);
const char *gq_last_error(void);

/* ---- CPU builders (replace gquery_ext.BVH / BVH3D / SNCH / SNCH3D, src/gquery.cpp:181-247) ---- */

/* vertices: n_vertices x dim float32; indices: n_prims x dim int32 (segments in 2D, triangles in
 * 3D).  max_prims_in_node defaults to 10 in the reference (include/shapes/bvh.h:158-168). */
int gq_build_bvh(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                 int64_t n_prims, int32_t max_prims_in_node, gq_structure **out);
int gq_build_snch(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                  int64_t n_prims, int32_t max_prims_in_node, gq_structure **out);
void gq_structure_free(gq_structure *s);

/* info[0..7] = dim, is_snch, n_nodes, n_prims, n_silhouettes, n_raw_silhouettes, max_depth, 0 */
int gq_structure_info(const gq_structure *s, int64_t info[8]);

/* Copies of the flat float32 arrays in the reference wire format (bvh.cpp:329-353,
 * snch.cpp:248-285).  `out` must hold the number of floats reported by passing out = NULL
 * (the count is written to *n_floats). */
int gq_structure_node_data(const gq_structure *s, float *out, int64_t *n_floats);      /* BVH.node_data  */
int gq_structure_snch_node_data(const gq_structure *s, float *out, int64_t *n_floats); /* SNCH.node_data */
int gq_structure_primitive_data(const gq_structure *s, float *out, int64_t *n_floats);
int gq_structure_silhouette_data(const gq_structure *s, float *out, int64_t *n_floats);
int gq_structure_raw_silhouette_data(const gq_structure *s, float *out, int64_t *n_floats);

/* Scene files (no counterpart in the reference, which rebuilds the trees in every process): the
 * whole structure -- flat arrays, counts and the input geometry -- in one binary file with a
 * checksum, so a serving process loads a scene without running the builder.  gq_structure_load
 * rejects a file whose magic, sizes or checksum do not match (GQ_ERR_INVALID). */
int gq_structure_save(const gq_structure *s, const char *path);
int gq_structure_load(const char *path, gq_structure **out);
/* The input geometry of a structure: n_vertices x dim floats and n_prims x dim indices.  Pass
 * NULL outputs to get the counts in *n_floats and *n_ints. */
int gq_structure_geometry(const gq_structure *s, float *vertices, int64_t *n_floats, int32_t *indices,
                          int64_t *n_ints);

/* gquery_ext.build_flat_silhouette_edges (src/shapes/silhouette_edge.cpp:62-71). */
int gq_build_flat_silhouette_edges(const float *vertices, int64_t n_vertices, const int32_t *faces,
                                   int64_t n_faces, float *out, int64_t *n_floats);

/* ---- Device scenes (replace the Float(...) uploads of bvh.py:124-125, snch.py:100-104) ---- */

/* Uploads every layout of `s` to the current HIP device.  A BVH structure supports the closest
 * point and ray queries; an SNCH structure supports all three. */
int gq_scene_create(const gq_structure *s, gq_scene **out);
int gq_scene_destroy(gq_scene *scene);

/* Engine layout options (no counterpart in the reference; chosen by the autotuner). */
typedef struct gq_scene_options {
    int32_t leaf_prims;    /* max primitives per leaf of the closest-point BVH4, 1..64; 0 = default
                              (24 in 3D, 16 in 2D) */
    int32_t packet_budget; /* traversal steps a wavefront of 64 queries may take before its queries
                              move to the wave-per-query pass; 0 = default, by scene size and
                              queries per primitive (closest point 128-1000, 300 for 65k-330k
                              primitives; silhouettes 128 or 3000 in 3D, 750 in 2D; smaller for
                              launches too small to fill the GPU), < 0 = never */
    int32_t traversal;     /* silhouette queries: 0 = auto (lane per query in 2D below 393,216
                              segments, and in 3D at 24-128 queries per triangle; otherwise
                              packet + wave pass), 1 = packet + wave pass, 2 = lane per query */
    int32_t wave_front;    /* LDS frontier entries of the first wave-per-query pass, 1..8192;
                              0 = default (closest point 1024, silhouettes 640).  A query whose
                              frontier outgrows it is finished by a second pass with the full
                              frontier: results do not change */
    int32_t sil_pairs;     /* silhouette pair records under GQ_METHOD_ENGINE: 0 = float32 (default;
                              128 B per pair in 3D, 192 B in 2D with the reference cones inline),
                              1 = quantized (64 B per pair: 8-bit boxes in 3D, 16-bit in 2D, int16
                              cones, 2D reference cones from a table).  Results do not change */
    int32_t ray_leaf_prims; /* max primitives per leaf of the rays' BVH4 (the same binary tree
                              collapsed separately), 1..64; 0 = default (2 in 3D; in 2D 6,
                              or 3 from 262,144 segments) */
    int32_t reserved[2];
} gq_scene_options;
int gq_scene_create_ex(const gq_structure *s, const gq_scene_options *opt, gq_scene **out);
/* info[0..7] = dim, has_snch, device bytes, n_nodes, n_wide_nodes, n_prims, n_silhouettes, 0 */
int gq_scene_info(const gq_scene *scene, int64_t info[8]);

/* ---- Queries ---- */

/* Polyline/Mesh.closest_point_bvh (polyline.py:171-173, mesh.py:119-121) ->
 * BVH.closest_point (bvh.py:127-198) / BVH3D.closest_point (bvh3d.py:108-174).
 * Fills valid, p, n, uv (3D) / t (2D), d, prim_id. */
int gq_closest_point_bvh(gq_scene *scene, int method, const float *x, int64_t n,
                         const gq_records *out, void *stream);

/* No counterpart in the reference: gq_closest_point_bvh (engine methods) without any device
 * allocation, for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its scratch (sort keys, permutation, deferred-query lists) is a
 * fixed 48 MiB device array held by each GPU the library runs on until the process ends: enough for
 * 1,048,576 closest-point queries; larger calls fail with GQ_ERR_INVALID and launch nothing.  Calls
 * that overlap in time on one GPU share that array and race: order them on one stream.  Results
 * equal gq_closest_point_bvh's bit for bit. */
int gq_closest_point_bvh_noalloc(gq_scene *scene, int method, const float *x, int64_t n,
                                 const gq_records *out, void *stream);

/* Polyline/Mesh.intersect_bvh (polyline.py:132-135, mesh.py:86-89) -> BVH.intersect
 * (bvh.py:200-276) / BVH3D.intersect (bvh3d.py:176-239).  r_max: per-query device array, or NULL
 * to use r_max_scalar for every ray.  Fills valid, p, n, uv/t, d, prim_id, on_boundary. */
int gq_intersect_bvh(gq_scene *scene, int method, const float *x, const float *v,
                     const float *r_max, float r_max_scalar, int64_t n, const gq_records *out,
                     void *stream);

/* Polyline/Mesh.closest_silhouette_snch (polyline.py:225-227, mesh.py:137-139) ->
 * SNCH.closest_silhouette (snch.py:106-164) / SNCH3D.closest_silhouette (snch3d.py:143-203).
 * Fills valid, p, d, prim_id. */
int gq_closest_silhouette_snch(gq_scene *scene, int method, const float *x, const float *r_max,
                               float r_max_scalar, int64_t n, const gq_records *out, void *stream);

/* BVH3D.sample_boundary (bvh3d.py:246-266) -> BVH3D._sample_boundary (bvh3d.py:268-346): per
 * query one BVH-guided walk from the root to a leaf through children whose boxes meet the ball
 * (x, R), an area-weighted choice among the leaf's primitives that meet the ball, and a uniform
 * point on the chosen one (Triangle.sample_point, triangle.py:27-42).  Runs on the reference
 * binary BVH of the scene.  R: per-query device array, or NULL to use R_scalar.  rng_state /
 * rng_inc: per-query PCG32 state and odd stream increment (the two fields of DrJit's PCG32, the
 * reference's `sampler`), device arrays; rng_state is advanced in place by the numbers drawn.
 * Polylines run the same walk over segments (LineSegment.surface_area / sphere_intersect /
 * sample_point, line_segment.py:23-36; the reference has these pieces but no 2D walk).  Where
 * the reference cannot run as written (bvh3d.py:265,303 read fields that do not exist), the chosen
 * primitive is its leaf-order slot and the record has no `type`; a leaf whose visible primitives
 * all have zero area gives no sample. */
int gq_sample_boundary(gq_scene *scene, const float *x, const float *R, float R_scalar, uint64_t *rng_state,
                       const uint64_t *rng_inc, int64_t n, const gq_sample_records *out, void *stream);

/* ---- Batches (no counterpart in the reference, which runs one DrJit-traced query per call; this
 * is the batched path for many scenes and query batches in one enqueue) ---- */

#define GQ_QUERY_CLOSEST_POINT 0
#define GQ_QUERY_INTERSECT 1
#define GQ_QUERY_SILHOUETTE 2

typedef struct gq_job {
    gq_scene *scene;         /* any scene on the stream's device                               */
    int32_t kind;            /* GQ_QUERY_*                                                      */
    int32_t method;          /* GQ_METHOD_ENGINE_FP32 or GQ_METHOD_ENGINE                       */
    const float *x;          /* n x dim query points / ray origins                              */
    const float *v;          /* rays: n x dim directions; else NULL                             */
    const float *r_max;      /* rays and silhouettes: n limits, or NULL for r_max_scalar        */
    float r_max_scalar;
    int32_t packet_budget;   /* 0 = the scene's setting (gq_scene_options)                      */
    int64_t n;
    gq_records out;          /* as for the single-batch calls of the same kind                  */
    int32_t traversal;       /* 0 = the scene's setting                                         */
    int32_t order;           /* 0 = Morton-sort closest-point and silhouette batches of 8,192 or
                                more queries (the default); 1 = traverse in the caller's order
                                (queries already coherent, or to measure the unsorted case)     */
    int32_t shared;          /* 0 = this enqueue has the device to itself: a launch group too
                                small to fill the GPU moves its work to the wave-per-query pass
                                early, which shortens its latency; 1 = other launches run beside
                                it (a serving loop's stream slots): default packet budgets are
                                those of a full launch, the cheaper total work.  Results do not
                                change                                                           */
    int32_t reserved[1];
} gq_job;

/* Bytes of device workspace and of pinned host staging that gq_batch_enqueue needs for `jobs`. */
int gq_batch_sizes(int n_jobs, const gq_job *jobs, int64_t *device_bytes, int64_t *host_bytes);

/* Enqueues every job on `stream`: the job table is written into `host_stage` (pinned host memory,
 * left untouched until the stream has passed this call) and copied into `device_ws`, then one set
 * of sort passes covers all jobs and one traversal launch runs per (kind, dimension, node format,
 * traversal) group.  Allocates nothing, clears nothing with memset and never synchronises, so the
 * call can be recorded in a HIP graph and replayed (replays re-read host_stage and the job arrays).
 * Results equal those of the single-batch calls bit for bit. */
int gq_batch_enqueue(int n_jobs, const gq_job *jobs, void *device_ws, int64_t device_bytes,
                     void *host_stage, int64_t host_bytes, void *stream);

/* ---- Brute-force queries, normals and silhouette configuration ----
 * A gq_shape is a polyline (dim 2, segments) or triangle mesh (dim 3) on the current device, for the
 * reference's brute-force methods, which test every primitive or silhouette for every query.  Its
 * silhouettes are those of Polyline.configure_silhouette (2D, polyline.py:44-122) or of
 * build_flat_silhouette_edges (3D, the Mesh constructor, mesh.py:26-29); a mesh also gets its face
 * and vertex normals (Mesh.configure_normal, mesh.py:35-50). */
typedef struct gq_shape gq_shape;
int gq_shape_create(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                    int64_t n_prims, gq_shape **out);
int gq_shape_destroy(gq_shape *shape);
/* info[0..3] = dim, n_vertices, n_prims, n_silhouettes */
int gq_shape_info(const gq_shape *shape, int64_t info[4]);
/* Mesh.configure_normal (mesh.py:35-50): copies the face normals (n_prims x 3) and the vertex
 * normals (n_vertices x 3) into device buffers (either may be NULL).  3D only. */
int gq_shape_normals(const gq_shape *shape, float *face_normals, float *vertex_normals, void *stream);

/* Polyline/Mesh.closest_point_baseline (polyline.py:176-209, mesh.py:92-118). */
int gq_closest_point_baseline(gq_shape *shape, const float *x, int64_t n, const gq_records *out,
                              void *stream);
/* Polyline/Mesh.intersect_baseline (polyline.py:138-170, mesh.py:61-85).  With normal and
 * on_boundary (n x dim floats, n bytes) it is Polyline/Mesh.intersect (polyline.py:125-131,
 * mesh.py:53-59): origins on the boundary move by -1e-5 n first.  Either may be NULL only together. */
int gq_intersect_baseline(gq_shape *shape, const float *x, const float *v, const float *normal,
                          const uint8_t *on_boundary, const float *r_max, float r_max_scalar, int64_t n,
                          const gq_records *out, void *stream);
/* Polyline/Mesh.closest_silhouette_baseline (polyline.py:212-223, mesh.py:124-136). */
int gq_closest_silhouette_baseline(gq_shape *shape, const float *x, const float *r_max,
                                   float r_max_scalar, int64_t n, const gq_records *out, void *stream);

/* Polyline.configure_silhouette (polyline.py:44-122) on the host: records of 10 floats (a, b, c,
 * prev, vertex, next, first-appearance index) of the polyline's silhouette vertices.  Pass
 * out = NULL to get the float count in *n_floats. */
int gq_configure_silhouette(const float *vertices, int64_t n_vertices, const int32_t *edges,
                            int64_t n_edges, float *out, int64_t *n_floats);

/* The two silhouette setups on the device, bit-identical to gq_configure_silhouette and
 * gq_build_flat_silhouette_edges (Polyline.configure_silhouette, polyline.py:44-122;
 * build_flat_silhouette_edges, silhouette_edge.cpp:5-71).  vertices, edges / faces and out are
 * device pointers; out holds *n_floats floats on entry (at most 10 * n_vertices are written in 2D,
 * 19 * 3 * n_faces in 3D) or is NULL to get the count.  On return *n_floats is the float count
 * (also when out is too small, which fails with GQ_ERR_INVALID).  Inputs are validated on the
 * device as the host versions validate them.  Runs on `stream` with stream-ordered scratch and
 * waits for it before returning. */
int gq_configure_silhouette_device(const float *vertices, int64_t n_vertices, const int32_t *edges,
                                   int64_t n_edges, float *out, int64_t *n_floats, void *stream);
int gq_build_flat_silhouette_edges_device(const float *vertices, int64_t n_vertices, const int32_t *faces,
                                          int64_t n_faces, float *out, int64_t *n_floats, void *stream);

/* ---- Engine layout inspection (no counterpart in the reference: exposes the host-built engine
 * arrays so tests can check them on machines without a GPU) ---- */

/* which: 0 quantized BVH4 nodes, 1 float32 BVH4 nodes, 2 quantized SNCH pairs (8-bit boxes in 3D,
 * 16-bit in 2D), 3 float32 SNCH pairs (2D: with the reference cones inline), 4 primitives
 * (float4), 5 primitive ids, 6 silhouettes (float4), 7 silhouette ids, 8 reference cones (3D),
 * 9 big-leaf table, 10 scalars (root_ref, sroot_ref, n_nodes, n_pairs, wide_depth as uint32),
 * 11 reference cones of the 2D quantized pairs, 12 seed cut of the BVH4, 13 seed cut of the SNCH
 * pairs (32-byte entries: box low corner, child reference, box high corner, pad), 14 / 15
 * bounding cylinders of the BVH4 children about the centres of their quantized / float32 boxes
 * (64 bytes per node, every field the upper 16 bits of a float32: axis [dim][4], lo[4], hi[4],
 * rad[4]), 16 / 17 float32 / quantized BVH4 nodes of the rays (ray leaf size).  Arrays 0, 1,
 * 4, 5, 9, 12, 14, 15 are built with the default closest-point leaf size.  Pass out = NULL to get
 * the byte count in *n_bytes. */
int gq_structure_engine_layout(const gq_structure *s, int which, void *out, int64_t *n_bytes);

/* ---- Kernel timing (replaces the time.time() brackets around dr.eval in the reference demos,
 * demo/closest_point_demo_3d.py:103-107) ---- */

/* Starts recording a pair of hipEvents around every traversal kernel launched from now on (on the
 * stream of its query call).  Clears earlier records. */
int gq_profile_begin(void);
/* Stops recording, waits for the recorded events and returns the summed traversal-kernel time,
 * the number of launches and the kernel's name (truncated to name_len bytes). */
int gq_profile_end(double *total_ms, int64_t *launches, char *name, int32_t name_len);
/* As gq_profile_end, and *first_kernel_ms receives the summed time of each launch's first traversal
 * kernel (the packet kernel, or the lane kernel of rays and lane-per-query silhouettes): the
 * dominant kernel, timed on its own stream.  No counterpart in the reference (its demos time whole
 * dr.eval calls, demo/closest_point_demo_3d.py:103-107). */
int gq_profile_end_split(double *total_ms, double *first_kernel_ms, int64_t *launches, char *name,
                         int32_t name_len);

#ifdef __cplusplus
}
#endif
#endif /* GQUERY_AMD_H */
