// CPU construction of the binned-SAH BVH and the SNCH (BVH + normal cones + per-leaf
// silhouette lists) for 2D polylines and 3D triangle meshes.
//
// The algorithm and the flat float32 output are the ones gQuery's builder produces
// (src/shapes/bvh.cpp:95-353, src/shapes/snch.cpp:11-285, src/shapes/silhouette_edge.cpp:5-71):
// the arrays returned by node_data()/primitive_data()/silhouette_data() are byte-identical
// records in depth-first order.  Construction is serial; the reference's two-way fork above
// 131,072 primitives (bvh.cpp:281-292) only changes which thread builds a subtree, not the tree.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gqa {

// Record strides (floats) of the flat wire format.
constexpr int kBvhNodeStride[4]   = {0, 0, 7, 9};
constexpr int kSnchNodeStride[4]  = {0, 0, 13, 16};
constexpr int kPrimStride[4]      = {0, 0, 5, 10};
constexpr int kSilStride[4]       = {0, 0, 10, 19};

struct HostStructure {
    int dim = 0;          // 2 or 3
    bool is_snch = false; // SNCH carries cones + silhouettes on top of the BVH
    int64_t n_nodes = 0;
    int64_t n_prims = 0;
    int64_t n_silhouettes = 0;      // ordered (per-leaf, duplicated across leaves)
    int64_t n_raw_silhouettes = 0;  // one per vertex (2D) / edge (3D)
    int32_t max_depth = 0;          // deepest node, root = 0
    std::vector<float> bvh_node_data;   // stride kBvhNodeStride[dim]
    std::vector<float> snch_node_data;  // stride kSnchNodeStride[dim] (SNCH only)
    std::vector<float> primitive_data;  // stride kPrimStride[dim], leaf (depth-first) order
    std::vector<float> silhouette_data; // stride kSilStride[dim], leaf order (SNCH only)
    std::vector<float> raw_silhouette_data;
    std::vector<int32_t> ordered_prim_index;  // original primitive id of each ordered slot
    std::vector<float> vertices;              // the input geometry (n_vertices x dim)
    std::vector<int32_t> indices;             // n_prims x dim
};

// Scene file: the whole HostStructure (flat arrays, counts, input geometry) in one binary file, so
// a serving process loads a scene without running the builder.  Little-endian records: magic
// "GQSCENE1", a header of counts, then each array as (int64 count, data); an FNV-1a 64-bit
// checksum of everything before it closes the file.  load_structure checks the magic, the counts
// against the record strides, and the checksum.
bool save_structure(const HostStructure &h, const char *path, std::string &err);
bool load_structure(const char *path, HostStructure &h, std::string &err);

// Builds the BVH (and, if with_snch, the SNCH) of a polyline (dim 2: indices are segments) or a
// triangle mesh (dim 3).  Returns false and fills err on invalid input.
bool build_structure(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                     int64_t n_prims, int max_prims_in_node, bool with_snch, HostStructure &out,
                     std::string &err);

// Unordered silhouette edges of a mesh in edge-id order (src/shapes/silhouette_edge.cpp:5-71).
bool build_flat_silhouette_edges(const float *vertices, int64_t n_vertices, const int32_t *faces,
                                 int64_t n_faces, std::vector<float> &out, std::string &err);

// Polyline.configure_silhouette (gquery/shapes/polyline.py:44-122): silhouette vertices of a
// polyline as records of 10 floats (a, b, c, prev, vertex, next, index), a vertex with both
// neighbours in first-appearance order, nearly collinear ones (|dot| of the unit edge directions
// >= 1 - 1e-5) dropped.  `index` is the vertex's position in first-appearance order.
bool configure_silhouette_2d(const float *vertices, int64_t n_vertices, const int32_t *edges, int64_t n_edges,
                             std::vector<float> &out, std::string &err);

// Faces incident to each vertex in the order Mesh.configure_normal (gquery/shapes/mesh.py:35-50)
// accumulates them: the faces naming the vertex first, then second, then third, each in face
// order.  CSR: faces of vertex v are faces[start[v] .. start[v + 1]).
void vertex_faces(const int32_t *faces, int64_t n_faces, int64_t n_vertices, std::vector<int32_t> &start,
                  std::vector<int32_t> &list);

bool validate_shape(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices, int64_t n_prims,
                    std::string &err);

} // namespace gqa
