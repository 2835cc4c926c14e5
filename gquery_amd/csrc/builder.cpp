// CPU BVH / SNCH construction.  See builder.h for the contract.
//
// Float semantics follow the reference builder exactly (it is compiled without FMA contraction):
//  * every point expansion pads by +-FLT_EPSILON (include/shapes/bounding_box.h:12-26);
//  * min/max are std::min/std::max ((y < x) ? y : x), the forms Eigen's cwiseMin/cwiseMax use;
//  * fixed-size reductions (dot, squaredNorm) of 3 terms associate as c0 + (c1 + c2), Eigen's
//    divide-and-conquer unrolled redux; 2-term sums are c0 + c1;
//  * normalized() divides by sqrt(squaredNorm) and returns the input when the norm is zero.
// This file must be compiled with -ffp-contract=off.
#include "builder.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <unordered_set>

namespace gqa {
namespace {

constexpr float kEps = FLT_EPSILON;            // include/core/fwd.h:10
constexpr float kPi = (float)3.14159265358979323846;  // M_PI narrowed to float (fwd.h:13)
constexpr int kBuckets = 12;                   // bvh.cpp:91

inline float smin(float x, float y) { return (y < x) ? y : x; }
inline float smax(float x, float y) { return (x < y) ? y : x; }

template <int D> inline float rsum(const float *c) {
    if constexpr (D == 2) return c[0] + c[1];
    else return c[0] + (c[1] + c[2]);
}
template <int D> inline float dotv(const float *a, const float *b) {
    float p[D];
    for (int i = 0; i < D; ++i) p[i] = a[i] * b[i];
    return rsum<D>(p);
}
template <int D> inline float normv(const float *a) { return std::sqrt(dotv<D>(a, a)); }
template <int D> inline void normalize_inplace(float *a) {
    float z = dotv<D>(a, a);
    if (z > 0.0f) {
        float s = std::sqrt(z);
        for (int i = 0; i < D; ++i) a[i] = a[i] / s;
    }
}
inline void cross3(const float *a, const float *b, float *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

template <int D> struct Box {
    float lo[D], hi[D];
    Box() {
        for (int i = 0; i < D; ++i) { lo[i] = FLT_MAX; hi[i] = -FLT_MAX; }
    }
    void expand_point(const float *p) {
        for (int i = 0; i < D; ++i) {
            float a = p[i] - kEps, b = p[i] + kEps;
            lo[i] = smin(lo[i], a);
            hi[i] = smax(hi[i], b);
        }
    }
    void expand_box(const Box &o) {
        for (int i = 0; i < D; ++i) {
            lo[i] = smin(lo[i], o.lo[i]);
            hi[i] = smax(hi[i], o.hi[i]);
        }
    }
    float centroid(int i) const { return (lo[i] + hi[i]) * 0.5f; }
    float surface_area() const {
        float e[D];
        for (int i = 0; i < D; ++i) e[i] = hi[i] - lo[i];
        if constexpr (D == 2) return 2.0f * (e[0] + e[1]);
        else return 2.0f * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2]);
    }
    int max_dimension() const {  // first index of the largest extent
        int idx = 0;
        float m = hi[0] - lo[0];
        for (int i = 1; i < D; ++i) {
            float e = hi[i] - lo[i];
            if (e > m) { m = e; idx = i; }
        }
        return idx;
    }
    float offset(float p, int i) const {
        float ext = smax(hi[i] - lo[i], kEps);
        return (p - lo[i]) / ext;
    }
};

template <int D> struct BuildPrim {
    size_t index;  // original primitive id
    Box<D> box;
};

template <int D> struct FlatNode {
    Box<D> box;
    uint32_t prim_offset = 0;   // leaf
    uint32_t second_child = 0;  // interior: right child index - own index
    uint32_t n_prims = 0;       // 0 for interior nodes
};

template <int D> inline int bucket_of(const Box<D> &cbox, const BuildPrim<D> &p, int dim) {
    int b = (int)((float)kBuckets * cbox.offset(p.box.centroid(dim), dim));
    if (b == kBuckets) b = kBuckets - 1;
    return b;
}

// bvh.cpp:29-88 -- binned SAH over 12 buckets of the centroid box's widest axis.
template <int D>
void find_best_split(const BuildPrim<D> *first, const BuildPrim<D> *last, const Box<D> &cbox,
                     int dim, int &split, float &cost) {
    Box<D> bbox[kBuckets];
    int count[kBuckets] = {};
    for (const BuildPrim<D> *p = first; p != last; ++p) {
        int b = bucket_of(cbox, *p, dim);
        count[b]++;
        bbox[b].expand_box(p->box);
    }
    constexpr int n_splits = kBuckets - 1;
    float costs[n_splits] = {};
    int below = 0;
    Box<D> box_below;
    for (int i = 0; i < n_splits; ++i) {
        box_below.expand_box(bbox[i]);
        below += count[i];
        costs[i] = (float)below * box_below.surface_area();
    }
    int above = 0;
    Box<D> box_above;
    for (int i = n_splits - 1; i >= 0; --i) {
        box_above.expand_box(bbox[i + 1]);
        above += count[i + 1];
        costs[i] = costs[i] + (float)above * box_above.surface_area();
    }
    split = 0;
    cost = costs[0];
    for (int i = 1; i < n_splits; ++i) {
        if (costs[i] < cost) { cost = costs[i]; split = i; }
    }
}

template <int D> struct BvhBuilder {
    int max_prims_in_node;
    std::vector<FlatNode<D>> nodes;
    std::vector<size_t> ordered;  // original ids in leaf order
    int max_depth = 0;

    void make_leaf(size_t node, BuildPrim<D> *first, BuildPrim<D> *last, const Box<D> &box) {
        nodes[node].box = box;
        nodes[node].prim_offset = (uint32_t)ordered.size();
        nodes[node].n_prims = (uint32_t)(last - first);
        for (BuildPrim<D> *p = first; p != last; ++p) ordered.push_back(p->index);
    }

    // bvh.cpp:181-304 (serial); nodes are numbered in creation order == depth-first order.
    size_t build(BuildPrim<D> *first, BuildPrim<D> *last, int depth) {
        size_t node = nodes.size();
        nodes.emplace_back();
        if (depth > max_depth) max_depth = depth;
        const size_t n = (size_t)(last - first);

        Box<D> box;
        for (BuildPrim<D> *p = first; p != last; ++p) box.expand_box(p->box);
        if (box.surface_area() == 0 || n == 1) {
            make_leaf(node, first, last, box);
            return node;
        }
        Box<D> cbox;
        for (BuildPrim<D> *p = first; p != last; ++p) {
            float c[D];
            for (int i = 0; i < D; ++i) c[i] = p->box.centroid(i);
            cbox.expand_point(c);
        }
        const int dim = cbox.max_dimension();
        if (cbox.hi[dim] == cbox.lo[dim]) {
            make_leaf(node, first, last, box);
            return node;
        }
        size_t mid = n / 2;
        if (n <= 2) {
            std::nth_element(first, first + mid, last,
                             [dim](const BuildPrim<D> &a, const BuildPrim<D> &b) {
                                 return a.box.centroid(dim) < b.box.centroid(dim);
                             });
        } else {
            int split;
            float min_cost;
            find_best_split(first, last, cbox, dim, split, min_cost);
            const float leaf_cost = (float)n;
            min_cost = 0.5f + min_cost / box.surface_area();
            if (n > (size_t)max_prims_in_node || min_cost < leaf_cost) {
                BuildPrim<D> *m = std::partition(first, last, [&](const BuildPrim<D> &p) {
                    return bucket_of(cbox, p, dim) <= split;
                });
                mid = (size_t)(m - first);
                // The reference recurses forever on an empty side; that cannot occur because
                // the lowest and highest centroids always fall in buckets 0 and 11.  Guard anyway.
                if (mid == 0 || mid == n) mid = n / 2;
            } else {
                make_leaf(node, first, last, box);
                return node;
            }
        }
        size_t left = build(first, first + mid, depth + 1);
        size_t right = build(first + mid, last, depth + 1);
        Box<D> ib;  // BVHBuildNode::init_interior: empty box grown by both children
        ib.expand_box(nodes[left].box);
        ib.expand_box(nodes[right].box);
        nodes[node].box = ib;
        nodes[node].n_prims = 0;
        nodes[node].second_child = (uint32_t)(right - node);
        return node;
    }
};

// 2D silhouette vertex / 3D silhouette edge records (include/shapes/silhouette_{vertex,edge}.h).
// Unset fields are zero, as value-initialised std::vector elements are in the reference.
struct SilV {
    float m[3][2] = {};
    uint64_t idx[3] = {};
    uint64_t prim = 0;
};
struct SilE {
    float m[4][3] = {};
    uint64_t idx[4] = {};
    uint64_t face[2] = {};
    uint64_t prim = 0;
};

// Per-silhouette quantities the cone builder needs, computed once (they are pure functions of
// the record, so caching them does not change any value).
template <int D> struct SilGeom {
    float normal[D];   // normal(): normalised sum of the unnormalised face normals
    float n0[D], n1[D];
    float centroid[D];
    bool all_faces;    // has_face_0 && has_face_1
};

SilGeom<2> sil_geom(const SilV &s) {
    SilGeom<2> g;
    float v0[2] = {s.m[1][0] - s.m[0][0], s.m[1][1] - s.m[0][1]};
    float v1[2] = {s.m[2][0] - s.m[1][0], s.m[2][1] - s.m[1][1]};
    float a[2] = {v0[1], -v0[0]}, b[2] = {v1[1], -v1[0]};
    const bool f0 = s.idx[0] != (uint64_t)-1, f1 = s.idx[2] != (uint64_t)-1;
    float n[2] = {0.0f, 0.0f};
    if (f0) { n[0] = n[0] + a[0]; n[1] = n[1] + a[1]; }
    if (f1) { n[0] = n[0] + b[0]; n[1] = n[1] + b[1]; }
    normalize_inplace<2>(n);
    std::memcpy(g.normal, n, sizeof n);
    std::memcpy(g.n0, a, sizeof a);
    normalize_inplace<2>(g.n0);
    std::memcpy(g.n1, b, sizeof b);
    normalize_inplace<2>(g.n1);
    g.centroid[0] = s.m[1][0];
    g.centroid[1] = s.m[1][1];
    g.all_faces = f0 && f1;
    return g;
}

SilGeom<3> sil_geom(const SilE &s) {
    SilGeom<3> g;
    float a0[3], b0[3], a1[3], b1[3], n0[3], n1[3];
    for (int i = 0; i < 3; ++i) {
        a0[i] = s.m[2][i] - s.m[1][i];  // normal_0: (m2 - m1) x (m0 - m1)
        b0[i] = s.m[0][i] - s.m[1][i];
        a1[i] = s.m[0][i] - s.m[2][i];  // normal_1: (m0 - m2) x (m1 - m2)
        b1[i] = s.m[1][i] - s.m[2][i];
    }
    cross3(a0, b0, n0);
    cross3(a1, b1, n1);
    const bool f0 = s.face[0] != (uint64_t)-1, f1 = s.face[1] != (uint64_t)-1;
    float n[3] = {0.0f, 0.0f, 0.0f};
    if (f0) for (int i = 0; i < 3; ++i) n[i] = n[i] + n0[i];
    if (f1) for (int i = 0; i < 3; ++i) n[i] = n[i] + n1[i];
    normalize_inplace<3>(n);
    std::memcpy(g.normal, n, sizeof n);
    std::memcpy(g.n0, n0, sizeof n0);
    normalize_inplace<3>(g.n0);
    std::memcpy(g.n1, n1, sizeof n1);
    normalize_inplace<3>(g.n1);
    for (int i = 0; i < 3; ++i) g.centroid[i] = 0.5f * (s.m[1][i] + s.m[2][i]);
    g.all_faces = f0 && f1;
    return g;
}

void flatten_sil(const SilV &s, float *o) {
    for (int k = 0; k < 3; ++k) { o[2 * k] = s.m[k][0]; o[2 * k + 1] = s.m[k][1]; }
    for (int k = 0; k < 3; ++k) o[6 + k] = (float)s.idx[k];
    o[9] = (float)s.prim;
}
void flatten_sil(const SilE &s, float *o) {
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 3; ++i) o[3 * k + i] = s.m[k][i];
    for (int k = 0; k < 4; ++k) o[12 + k] = (float)s.idx[k];
    o[16] = (float)s.face[0];
    o[17] = (float)s.face[1];
    o[18] = (float)s.prim;
}

struct PairKeyHash {
    size_t operator()(uint64_t k) const { return (size_t)(k * 0x9E3779B97F4A7C15ull >> 17) ^ (size_t)k; }
};

// snch.cpp:186-241 / silhouette_edge.cpp:5-60: edge ids in first-seen order over (face, corner);
// each face (in order) overwrites m[1], m[2] with its own traversal order of the edge.
void build_edges(const float *V, const int32_t *F, int64_t nf, std::vector<SilE> &sil,
                 std::unordered_map<uint64_t, uint32_t, PairKeyHash> &emap) {
    emap.clear();
    emap.reserve((size_t)(nf * 3 / 2 + 16));
    uint32_t n_edges = 0;
    for (int64_t i = 0; i < nf; ++i)
        for (int j = 0; j < 3; ++j) {
            uint64_t i0 = (uint64_t)F[3 * i + j], i1 = (uint64_t)F[3 * i + (j + 1) % 3];
            if (i0 > i1) std::swap(i0, i1);
            uint64_t key = (i0 << 32) | i1;
            if (emap.find(key) == emap.end()) emap.emplace(key, n_edges++);
        }
    sil.assign(n_edges, SilE());
    for (int64_t i = 0; i < nf; ++i)
        for (int j = 0; j < 3; ++j) {
            uint64_t i0 = (uint64_t)F[3 * i + j], i1 = (uint64_t)F[3 * i + (j + 1) % 3],
                     i2 = (uint64_t)F[3 * i + (j + 2) % 3];
            const float *v0 = V + 3 * i0, *v1 = V + 3 * i1, *v2 = V + 3 * i2;
            bool swapped = false;
            if (i0 > i1) { std::swap(i0, i1); swapped = true; }
            uint32_t e = emap[(i0 << 32) | i1];
            SilE &s = sil[e];
            std::memcpy(s.m[1], v0, 3 * sizeof(float));
            std::memcpy(s.m[2], v1, 3 * sizeof(float));
            std::memcpy(s.m[swapped ? 3 : 0], v2, 3 * sizeof(float));
            s.idx[1] = i0;
            s.idx[2] = i1;
            s.idx[swapped ? 3 : 0] = i2;
            s.face[swapped ? 1 : 0] = (uint64_t)i;
            s.prim = e;
        }
}

// snch.cpp:156-184: one silhouette vertex per mesh vertex.
void build_vertices(const float *V, int64_t nv, const int32_t *E, int64_t ne, std::vector<SilV> &sil) {
    sil.assign((size_t)nv, SilV());
    for (int64_t k = 0; k < ne; ++k) {
        uint64_t i0 = (uint64_t)E[2 * k], i1 = (uint64_t)E[2 * k + 1];
        const float *v0 = V + 2 * i0, *v1 = V + 2 * i1;
        SilV &s0 = sil[i0];
        std::memcpy(s0.m[1], v0, 2 * sizeof(float));
        std::memcpy(s0.m[2], v1, 2 * sizeof(float));
        s0.idx[1] = i0;
        s0.idx[2] = i1;
        s0.prim = i0;
        SilV &s1 = sil[i1];
        std::memcpy(s1.m[0], v0, 2 * sizeof(float));
        std::memcpy(s1.m[1], v1, 2 * sizeof(float));
        s1.idx[0] = i0;
        s1.idx[1] = i1;
        s1.prim = i1;
    }
}

struct Cone {
    float axis[3] = {0.0f, 0.0f, 0.0f};
    float half = kPi;    // BoundingCone() default (include/shapes/bounding_cone.h:16)
    float radius = 0.0f;
};

// snch.cpp:102-154: the cone of node `start` covers every silhouette of the leaves in the
// depth-first range [start, end).
template <int D>
void build_cones(const std::vector<FlatNode<D>> &nodes, const std::vector<uint32_t> &sil_off,
                 const std::vector<uint32_t> &sil_cnt, const std::vector<SilGeom<D>> &geom,
                 const std::vector<uint32_t> &ordered_sil, std::vector<Cone> &cones) {
    cones.assign(nodes.size(), Cone());
    struct Range { size_t start, end; };
    std::vector<Range> work;
    work.push_back({0, nodes.size()});
    while (!work.empty()) {
        Range r = work.back();
        work.pop_back();
        Cone &c = cones[r.start];
        const FlatNode<D> &root = nodes[r.start];
        float cen[D];
        for (int i = 0; i < D; ++i) cen[i] = root.box.centroid(i);
        bool any = false, all_adjacent = true;
        for (size_t i = r.start; i < r.end; ++i)
            for (uint32_t j = 0; j < sil_cnt[i]; ++j) {
                const SilGeom<D> &g = geom[ordered_sil[sil_off[i] + j]];
                for (int k = 0; k < D; ++k) c.axis[k] = c.axis[k] + g.normal[k];
                float d[D];
                for (int k = 0; k < D; ++k) d[k] = g.centroid[k] - cen[k];
                c.radius = smax(c.radius, normv<D>(d));
                all_adjacent = all_adjacent && g.all_faces;
                any = true;
            }
        if (!any) {
            c.half = -kPi;
        } else if (!all_adjacent) {
            c.half = kPi;
        } else {
            float an = normv<D>(c.axis);
            if (an > kEps) {
                for (int k = 0; k < D; ++k) c.axis[k] = c.axis[k] / an;
                c.half = 0.0f;
                for (size_t i = r.start; i < r.end; ++i)
                    for (uint32_t j = 0; j < sil_cnt[i]; ++j) {
                        const SilGeom<D> &g = geom[ordered_sil[sil_off[i] + j]];
                        float x0 = std::clamp(dotv<D>(c.axis, g.n0), -1.0f, 1.0f);
                        float x1 = std::clamp(dotv<D>(c.axis, g.n1), -1.0f, 1.0f);
                        float a0 = std::acos(x0), a1 = std::acos(x1);
                        c.half = smax(c.half, smax(a0, a1));
                    }
            }
        }
        if (root.n_prims == 0) {
            // right subtree pushed first so the left one is finished first (order is irrelevant
            // to the values; kept for a depth-first visit).
            work.push_back({r.start + root.second_child, r.end});
            work.push_back({r.start + 1, r.start + root.second_child});
        }
    }
}

template <int D>
bool build_impl(const float *V, int64_t nv, const int32_t *I, int64_t np, int max_prims,
                bool with_snch, HostStructure &out, std::string &err) {
    // Primitives with their boxes (bvh.cpp:100-130, 146-149).
    std::vector<BuildPrim<D>> prims((size_t)np);
    for (int64_t i = 0; i < np; ++i) {
        BuildPrim<D> &p = prims[(size_t)i];
        p.index = (size_t)i;
        for (int k = 0; k < D; ++k) p.box.expand_point(V + (size_t)D * (size_t)I[D * i + k]);
    }
    BvhBuilder<D> b;
    b.max_prims_in_node = max_prims;
    b.nodes.reserve((size_t)(2 * np));
    b.ordered.reserve((size_t)np);
    b.build(prims.data(), prims.data() + prims.size(), 0);
    prims.clear();
    prims.shrink_to_fit();

    out = HostStructure();
    out.dim = D;
    out.is_snch = with_snch;
    out.vertices.assign(V, V + (size_t)nv * D);
    out.indices.assign(I, I + (size_t)np * D);
    out.n_nodes = (int64_t)b.nodes.size();
    out.n_prims = np;
    out.max_depth = b.max_depth;

    const int ns = kBvhNodeStride[D];
    out.bvh_node_data.resize((size_t)out.n_nodes * ns);
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const FlatNode<D> &n = b.nodes[i];
        float *o = out.bvh_node_data.data() + i * ns;
        for (int k = 0; k < D; ++k) { o[k] = n.box.lo[k]; o[D + k] = n.box.hi[k]; }
        o[2 * D] = (float)n.prim_offset;  // 0 for interior nodes
        o[2 * D + 1] = (float)n.second_child;
        o[2 * D + 2] = (float)n.n_prims;
    }
    const int ps = kPrimStride[D];
    out.primitive_data.resize((size_t)np * ps);
    out.ordered_prim_index.resize((size_t)np);
    for (size_t s = 0; s < b.ordered.size(); ++s) {
        const size_t id = b.ordered[s];
        float *o = out.primitive_data.data() + s * ps;
        for (int k = 0; k < D; ++k)
            for (int c = 0; c < D; ++c) o[D * k + c] = V[(size_t)D * (size_t)I[(size_t)D * id + k] + c];
        o[D * D] = (float)id;
        out.ordered_prim_index[s] = (int32_t)id;
    }
    if (!with_snch) return true;

    // SNCH: silhouettes, per-leaf silhouette lists, cones (snch.cpp:11-154).
    std::vector<SilV> sv;
    std::vector<SilE> se;
    std::unordered_map<uint64_t, uint32_t, PairKeyHash> emap;
    if constexpr (D == 2) build_vertices(V, nv, I, np, sv);
    else build_edges(V, I, np, se, emap);
    const size_t n_raw = (D == 2) ? sv.size() : se.size();

    std::vector<uint32_t> sil_off(b.nodes.size(), 0), sil_cnt(b.nodes.size(), 0);
    std::vector<uint32_t> ordered_sil;
    ordered_sil.reserve((size_t)np * (D == 2 ? 2 : 3));
    std::unordered_set<uint32_t> seen;
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const FlatNode<D> &n = b.nodes[i];
        if (n.n_prims == 0) continue;
        const size_t start = ordered_sil.size();
        seen.clear();
        for (uint32_t j = 0; j < n.n_prims; ++j) {
            const size_t id = b.ordered[n.prim_offset + j];
            const int32_t *f = I + (size_t)D * id;
            if constexpr (D == 2) {
                for (int k = 0; k < 2; ++k) {
                    uint32_t vi = (uint32_t)f[k];
                    if (seen.insert(vi).second) ordered_sil.push_back(vi);
                }
            } else {
                for (int k = 0; k < 3; ++k) {
                    uint64_t i0 = (uint64_t)f[k], i1 = (uint64_t)f[(k + 1) % 3];
                    if (i0 > i1) std::swap(i0, i1);
                    uint32_t e = emap.find((i0 << 32) | i1)->second;
                    if (seen.insert(e).second) ordered_sil.push_back(e);
                }
            }
        }
        sil_off[i] = (uint32_t)start;
        sil_cnt[i] = (uint32_t)(ordered_sil.size() - start);
    }

    std::vector<SilGeom<D>> geom(n_raw);
    for (size_t k = 0; k < n_raw; ++k) {
        if constexpr (D == 2) geom[k] = sil_geom(sv[k]);
        else geom[k] = sil_geom(se[k]);
    }
    std::vector<Cone> cones;
    build_cones<D>(b.nodes, sil_off, sil_cnt, geom, ordered_sil, cones);

    const int ss = kSilStride[D];
    out.n_silhouettes = (int64_t)ordered_sil.size();
    out.n_raw_silhouettes = (int64_t)n_raw;
    out.silhouette_data.resize(ordered_sil.size() * ss);
    for (size_t k = 0; k < ordered_sil.size(); ++k) {
        if constexpr (D == 2) flatten_sil(sv[ordered_sil[k]], out.silhouette_data.data() + k * ss);
        else flatten_sil(se[ordered_sil[k]], out.silhouette_data.data() + k * ss);
    }
    out.raw_silhouette_data.resize(n_raw * ss);
    for (size_t k = 0; k < n_raw; ++k) {
        if constexpr (D == 2) flatten_sil(sv[k], out.raw_silhouette_data.data() + k * ss);
        else flatten_sil(se[k], out.raw_silhouette_data.data() + k * ss);
    }
    const int nss = kSnchNodeStride[D];
    out.snch_node_data.resize(b.nodes.size() * nss);
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const FlatNode<D> &n = b.nodes[i];
        const Cone &c = cones[i];
        float *o = out.snch_node_data.data() + i * nss;
        for (int k = 0; k < D; ++k) { o[k] = n.box.lo[k]; o[D + k] = n.box.hi[k]; }
        for (int k = 0; k < D; ++k) o[2 * D + k] = c.axis[k];
        o[3 * D] = c.half;
        o[3 * D + 1] = c.radius;
        o[3 * D + 2] = n.n_prims > 0 ? (float)n.prim_offset : 0.0f;
        o[3 * D + 3] = (float)n.second_child;
        o[3 * D + 4] = (float)n.n_prims;
        o[3 * D + 5] = (float)sil_off[i];
        o[3 * D + 6] = (float)sil_cnt[i];
    }
    return true;
}

bool validate(int dim, const float *V, int64_t nv, const int32_t *I, int64_t np, std::string &err) {
    if (dim != 2 && dim != 3) { err = "dim must be 2 or 3"; return false; }
    if (nv <= 0 || np <= 0 || !V || !I) { err = "empty vertex or index array"; return false; }
    if (np >= (1 << 24) || nv >= (1 << 24)) {
        err = "more than 2^24 primitives or vertices: indices are stored as float32";
        return false;
    }
    for (int64_t k = 0; k < np * dim; ++k)
        if (I[k] < 0 || (int64_t)I[k] >= nv) { err = "index out of range"; return false; }
    for (int64_t k = 0; k < nv * dim; ++k)
        if (!std::isfinite(V[k])) { err = "non-finite vertex coordinate"; return false; }
    return true;
}

} // namespace

bool build_structure(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                     int64_t n_prims, int max_prims_in_node, bool with_snch, HostStructure &out,
                     std::string &err) {
    if (!validate(dim, vertices, n_vertices, indices, n_prims, err)) return false;
    if (max_prims_in_node < 1) { err = "max_prims_in_node must be >= 1"; return false; }
    if (dim == 2)
        return build_impl<2>(vertices, n_vertices, indices, n_prims, max_prims_in_node, with_snch, out, err);
    return build_impl<3>(vertices, n_vertices, indices, n_prims, max_prims_in_node, with_snch, out, err);
}

bool validate_shape(int dim, const float *V, int64_t nv, const int32_t *I, int64_t np, std::string &err) {
    return validate(dim, V, nv, I, np, err);
}

bool configure_silhouette_2d(const float *V, int64_t nv, const int32_t *E, int64_t ne, std::vector<float> &out,
                             std::string &err) {
    if (!validate(2, V, nv, E, ne, err)) return false;
    // neighbours of each vertex: the last edge leaving it sets `next`, the last entering it `prev`
    std::vector<int32_t> prev((size_t)nv, -2), next((size_t)nv, -2), order;
    order.reserve((size_t)nv);
    for (int64_t k = 0; k < ne; ++k) {
        const int32_t a = E[2 * k], b = E[2 * k + 1];
        for (int32_t v : {a, b}) {
            if (prev[(size_t)v] == -2) {
                prev[(size_t)v] = next[(size_t)v] = -1;
                order.push_back(v);
            }
        }
        next[(size_t)a] = b;
        prev[(size_t)b] = a;
    }
    out.clear();
    // float32 unit direction, as the query kernels normalise: v * (1 / sqrt(dot(v, v)))
    auto unit = [](float x, float y, float &ux, float &uy) {
        const float inv = 1.0f / std::sqrt(std::fma(y, y, x * x));
        ux = x * inv;
        uy = y * inv;
    };
    for (size_t i = 0; i < order.size(); ++i) {
        const int32_t v = order[i], p = prev[(size_t)v], n = next[(size_t)v];
        if (p == -1 || n == -1) continue;
        const float *a = V + 2 * (size_t)p, *b = V + 2 * (size_t)v, *c = V + 2 * (size_t)n;
        float e1x, e1y, e2x, e2y;
        unit(b[0] - a[0], b[1] - a[1], e1x, e1y);
        unit(c[0] - b[0], c[1] - b[1], e2x, e2y);
        const float col = std::fabs(std::fma(e1y, e2y, e1x * e2x));
        if (!(col < 1.0f - 1e-5f)) continue;   // nearly collinear (or degenerate: NaN)
        const float r[10] = {a[0], a[1], b[0], b[1], c[0], c[1], (float)p, (float)v, (float)n, (float)i};
        out.insert(out.end(), r, r + 10);
    }
    return true;
}

void vertex_faces(const int32_t *F, int64_t nf, int64_t nv, std::vector<int32_t> &start, std::vector<int32_t> &list) {
    start.assign((size_t)nv + 1, 0);
    for (int64_t k = 0; k < 3 * nf; ++k) ++start[(size_t)F[k] + 1];
    for (int64_t v = 0; v < nv; ++v) start[(size_t)v + 1] += start[(size_t)v];
    list.assign((size_t)(3 * nf), 0);
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int slot = 0; slot < 3; ++slot)
        for (int64_t f = 0; f < nf; ++f) list[(size_t)fill[(size_t)F[3 * f + slot]]++] = (int32_t)f;
}

namespace {
struct Fnv {
    uint64_t h = 14695981039346656037ull;   // FNV-1a 64-bit offset basis
    void add(const void *p, size_t n) {
        const uint8_t *b = (const uint8_t *)p;
        for (size_t k = 0; k < n; ++k) { h ^= b[k]; h *= 1099511628211ull; }
    }
};
const char kSceneMagic[8] = {'G', 'Q', 'S', 'C', 'E', 'N', 'E', '1'};
} // namespace

bool save_structure(const HostStructure &h, const char *path, std::string &err) {
    std::FILE *f = path ? std::fopen(path, "wb") : nullptr;
    if (!f) { err = std::string("cannot open for writing: ") + (path ? path : "(null)"); return false; }
    Fnv sum;
    bool ok = true;
    auto put = [&](const void *p, size_t n) {
        if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
        sum.add(p, n);
    };
    auto put_vec = [&](const auto &v) {
        const int64_t n = (int64_t)v.size();
        put(&n, 8);
        put(v.data(), v.size() * sizeof(v[0]));
    };
    const int64_t head[8] = {h.dim, h.is_snch ? 1 : 0, h.n_nodes, h.n_prims, h.n_silhouettes, h.n_raw_silhouettes,
                             h.max_depth, 0};
    put(kSceneMagic, 8);
    put(head, sizeof head);
    put_vec(h.bvh_node_data);
    put_vec(h.snch_node_data);
    put_vec(h.primitive_data);
    put_vec(h.silhouette_data);
    put_vec(h.raw_silhouette_data);
    put_vec(h.ordered_prim_index);
    put_vec(h.vertices);
    put_vec(h.indices);
    const uint64_t c = sum.h;
    if (ok && std::fwrite(&c, 1, 8, f) != 8) ok = false;
    if (std::fclose(f) != 0) ok = false;
    if (!ok) err = std::string("write failed: ") + path;
    return ok;
}

// A float field that holds a non-negative integer (the wire format stores offsets and counts as
// float32): exact, finite, below 2^24.
static bool int_field(float v, int64_t &out) {
    if (!(v >= 0.0f) || !(v < 16777216.0f) || v != std::floor(v)) return false;
    out = (int64_t)v;
    return true;
}

// The checks a loaded structure needs before any kernel indexes with it: every node of a tree is
// reached exactly once from the root, child offsets and leaf ranges (primitives; silhouettes for
// an SNCH) lie inside their arrays, integer fields are exact, and the deepest node matches the
// stored max_depth (the plain kernels' stack bound and the engine's depth checks read it).
static bool check_tree(const std::vector<float> &nd, int stride, int64_t n_nodes, int off_f, int second_f,
                       int count_f, int64_t n_items, int sil_off_f, int sil_cnt_f, int64_t n_sils, int32_t &depth,
                       std::string &err) {
    std::vector<uint8_t> seen((size_t)n_nodes, 0);
    std::vector<std::pair<int64_t, int32_t>> stack{{0, 0}};
    int64_t visited = 0;
    depth = 0;
    while (!stack.empty()) {
        const auto [i, d] = stack.back();
        stack.pop_back();
        if (i < 0 || i >= n_nodes || seen[(size_t)i]) { err = "node reached twice or out of range"; return false; }
        seen[(size_t)i] = 1;
        ++visited;
        depth = std::max(depth, d);
        const float *n = nd.data() + (size_t)i * stride;
        int64_t off, second, count;
        if (!int_field(n[off_f], off) || !int_field(n[second_f], second) || !int_field(n[count_f], count)) {
            err = "non-integer offset or count in node " + std::to_string(i);
            return false;
        }
        if (sil_off_f >= 0) {
            int64_t so, sc;
            if (!int_field(n[sil_off_f], so) || !int_field(n[sil_cnt_f], sc) || so + sc > n_sils) {
                err = "silhouette range out of bounds in node " + std::to_string(i);
                return false;
            }
        }
        if (count > 0) {
            if (off + count > n_items) { err = "leaf range out of bounds in node " + std::to_string(i); return false; }
        } else {
            if (second < 2 || i + second >= n_nodes) { err = "child offset out of bounds in node " + std::to_string(i); return false; }
            stack.push_back({i + second, d + 1});
            stack.push_back({i + 1, d + 1});
        }
    }
    if (visited != n_nodes) { err = "nodes unreachable from the root"; return false; }
    return true;
}

static bool check_loaded(const HostStructure &s, std::string &err) {
    const int D = s.dim;
    int32_t depth = 0;
    if (!check_tree(s.bvh_node_data, kBvhNodeStride[D], s.n_nodes, 2 * D, 2 * D + 1, 2 * D + 2, s.n_prims, -1, -1, 0,
                    depth, err))
        return false;
    if (depth != s.max_depth) { err = "stored max_depth does not match the tree"; return false; }
    if (s.is_snch) {
        // SNCH node: box, cone axis, half-angle, radius, then prim_off, second_child, n_prims,
        // sil_off, n_sils (snch.h:44-50)
        const int b = 2 * D + D + 2;
        int32_t sd = 0;
        if (!check_tree(s.snch_node_data, kSnchNodeStride[D], s.n_nodes, b, b + 1, b + 2, s.n_prims, b + 3, b + 4,
                        s.n_silhouettes, sd, err))
            return false;
    }
    const int64_t nv = (int64_t)(s.vertices.size() / (size_t)D);
    for (int32_t v : s.indices)
        if (v < 0 || v >= nv) { err = "vertex index out of range"; return false; }
    for (int32_t v : s.ordered_prim_index)
        if (v < 0 || v >= s.n_prims) { err = "primitive index out of range"; return false; }
    for (int64_t k = 0; k < s.n_prims; ++k) {
        int64_t id;
        if (!int_field(s.primitive_data[(size_t)(k * kPrimStride[D] + kPrimStride[D] - 1)], id) || id >= s.n_prims) {
            err = "primitive record id out of range";
            return false;
        }
    }
    return true;
}

bool load_structure(const char *path, HostStructure &h, std::string &err) {
    std::FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) { err = std::string("cannot open: ") + (path ? path : "(null)"); return false; }
    Fnv sum;
    bool ok = true;
    auto get = [&](void *p, size_t n) {
        if (ok && n && std::fread(p, 1, n, f) != n) ok = false;
        if (ok) sum.add(p, n);
    };
    auto get_vec = [&](auto &v, int64_t want) {
        int64_t n = -1;
        get(&n, 8);
        if (!ok || n < 0 || (want >= 0 && n != want) || n > ((int64_t)1 << 40)) { ok = false; return; }
        v.resize((size_t)n);
        get(v.data(), v.size() * sizeof(v[0]));
    };
    char magic[8] = {0};
    int64_t head[8] = {0};
    get(magic, 8);
    get(head, sizeof head);
    HostStructure s;
    if (ok && std::memcmp(magic, kSceneMagic, 8) != 0) { ok = false; }
    const int D = (int)head[0];
    if (ok && (D != 2 && D != 3)) ok = false;
    if (ok) {
        s.dim = D;
        s.is_snch = head[1] != 0;
        s.n_nodes = head[2];
        s.n_prims = head[3];
        s.n_silhouettes = head[4];
        s.n_raw_silhouettes = head[5];
        s.max_depth = (int32_t)head[6];
        if (s.n_nodes < 1 || s.n_prims < 1 || s.n_silhouettes < 0 || s.n_raw_silhouettes < 0) ok = false;
    }
    if (ok) {
        get_vec(s.bvh_node_data, s.n_nodes * kBvhNodeStride[D]);
        get_vec(s.snch_node_data, s.is_snch ? s.n_nodes * kSnchNodeStride[D] : 0);
        get_vec(s.primitive_data, s.n_prims * kPrimStride[D]);
        get_vec(s.silhouette_data, s.is_snch ? s.n_silhouettes * kSilStride[D] : 0);
        get_vec(s.raw_silhouette_data, s.is_snch ? s.n_raw_silhouettes * kSilStride[D] : 0);
        get_vec(s.ordered_prim_index, s.n_prims);
        get_vec(s.vertices, -1);
        get_vec(s.indices, s.n_prims * D);
    }
    uint64_t want = sum.h, c = 0;
    if (ok && std::fread(&c, 1, 8, f) != 8) ok = false;
    std::fclose(f);
    if (ok && c != want) { err = std::string("checksum mismatch: ") + path; return false; }
    if (ok && s.vertices.size() % (size_t)D != 0) ok = false;
    if (!ok) { err = std::string("not a valid scene file: ") + path; return false; }
    std::string why;
    if (!check_loaded(s, why)) { err = std::string("inconsistent scene file (") + why + "): " + path; return false; }
    h = std::move(s);
    return true;
}

bool build_flat_silhouette_edges(const float *V, int64_t nv, const int32_t *F, int64_t nf,
                                 std::vector<float> &out, std::string &err) {
    if (!validate(3, V, nv, F, nf, err)) return false;
    std::vector<SilE> se;
    std::unordered_map<uint64_t, uint32_t, PairKeyHash> emap;
    build_edges(V, F, nf, se, emap);
    out.resize(se.size() * 19);
    for (size_t k = 0; k < se.size(); ++k) flatten_sil(se[k], out.data() + k * 19);
    return true;
}

} // namespace gqa
