// C-ABI layer: builders, device scenes and query dispatch (declarations in include/gquery_amd.h).
#include <cstring>
#include <mutex>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "builder.h"
#include "gq_internal.h"
#include "wide.h"

struct gq_structure {
    gqa::HostStructure h;
};

struct gq_shape {
    int dim = 0;
    int device = 0;
    size_t bytes = 0;
    std::vector<void *> allocs;
    gqa::ShapeView view{};
    float *face_normals = nullptr, *vertex_normals = nullptr;
};

struct gq_scene {
    int dim = 0;
    bool has_snch = false;
    int device = 0;
    int max_depth = 0;
    size_t bytes = 0;
    std::vector<void *> allocs;
    gqa::PlainView plain{};
    gqa::WideView wide{};
};

namespace gqa {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// ---- traversal-kernel timing ----
namespace {
struct Prof {
    std::mutex mu;
    bool on = false;
    // per traversal launch: events before it, after its first kernel (null when none was marked)
    // and after it
    struct Rec { hipEvent_t begin, first, end; };
    std::vector<Rec> ev;
    std::string name;
    hipEvent_t pending = nullptr, pending_first = nullptr;
};
Prof &prof() {
    static Prof p;
    return p;
}
} // namespace

void prof_begin(hipStream_t st) {
    Prof &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    if (!p.on) return;
    if (hipEventCreate(&p.pending) != hipSuccess) { p.pending = nullptr; return; }
    hipEventRecord(p.pending, st);
}

void prof_mark_first(hipStream_t st) {
    Prof &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    if (!p.on || !p.pending || p.pending_first) return;
    if (hipEventCreate(&p.pending_first) != hipSuccess) { p.pending_first = nullptr; return; }
    hipEventRecord(p.pending_first, st);
}

void prof_end(hipStream_t st, const char *kernel) {
    Prof &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    if (!p.on || !p.pending) return;
    hipEvent_t e1;
    if (hipEventCreate(&e1) != hipSuccess) return;
    hipEventRecord(e1, st);
    p.ev.push_back({p.pending, p.pending_first, e1});
    p.pending = nullptr;
    p.pending_first = nullptr;
    p.name = kernel;
}

static int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

static int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? GQ_ERR_OOM : GQ_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

template <typename O, typename T>
static int upload(O *s, const T *host, size_t count, const T **dev) {
    *dev = nullptr;
    if (count == 0) return GQ_OK;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    s->allocs.push_back(p);
    s->bytes += count * sizeof(T);
    e = hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
    *dev = (const T *)p;
    return GQ_OK;
}

} // namespace gqa

using namespace gqa;

extern "C" {

int gq_version(void) { return 100; }

const char *gq_last_error(void) { return g_last_error.c_str(); }

static int build_common(int dim, const float *V, int64_t nv, const int32_t *I, int64_t np,
                        int32_t max_prims, bool snch, gq_structure **out) {
    if (!out) return fail(GQ_ERR_INVALID, "out is null");
    *out = nullptr;
    std::unique_ptr<gq_structure> s(new (std::nothrow) gq_structure());
    if (!s) return fail(GQ_ERR_OOM, "host allocation failed");
    std::string err;
    try {
        if (!build_structure(dim, V, nv, I, np, max_prims, snch, s->h, err))
            return fail(GQ_ERR_INVALID, err);
    } catch (const std::bad_alloc &) {
        return fail(GQ_ERR_OOM, "host allocation failed during build");
    }
    *out = s.release();
    return GQ_OK;
}

int gq_build_bvh(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                 int64_t n_prims, int32_t max_prims_in_node, gq_structure **out) {
    return build_common(dim, vertices, n_vertices, indices, n_prims, max_prims_in_node, false, out);
}

int gq_build_snch(int dim, const float *vertices, int64_t n_vertices, const int32_t *indices,
                  int64_t n_prims, int32_t max_prims_in_node, gq_structure **out) {
    return build_common(dim, vertices, n_vertices, indices, n_prims, max_prims_in_node, true, out);
}

void gq_structure_free(gq_structure *s) { delete s; }

int gq_structure_info(const gq_structure *s, int64_t info[8]) {
    if (!s || !info) return fail(GQ_ERR_INVALID, "null argument");
    const HostStructure &h = s->h;
    info[0] = h.dim;
    info[1] = h.is_snch ? 1 : 0;
    info[2] = h.n_nodes;
    info[3] = h.n_prims;
    info[4] = h.n_silhouettes;
    info[5] = h.n_raw_silhouettes;
    info[6] = h.max_depth;
    info[7] = 0;
    return GQ_OK;
}

static int copy_out(const std::vector<float> &v, float *out, int64_t *n_floats) {
    if (!n_floats) return fail(GQ_ERR_INVALID, "n_floats is null");
    if (out) {
        if (*n_floats < (int64_t)v.size()) return fail(GQ_ERR_INVALID, "output buffer too small");
        if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(float));
    }
    *n_floats = (int64_t)v.size();
    return GQ_OK;
}

int gq_structure_node_data(const gq_structure *s, float *out, int64_t *n) {
    if (!s) return fail(GQ_ERR_INVALID, "null structure");
    return copy_out(s->h.bvh_node_data, out, n);
}
int gq_structure_snch_node_data(const gq_structure *s, float *out, int64_t *n) {
    if (!s) return fail(GQ_ERR_INVALID, "null structure");
    if (!s->h.is_snch) return fail(GQ_ERR_UNSUPPORTED, "structure is not an SNCH");
    return copy_out(s->h.snch_node_data, out, n);
}
int gq_structure_primitive_data(const gq_structure *s, float *out, int64_t *n) {
    if (!s) return fail(GQ_ERR_INVALID, "null structure");
    return copy_out(s->h.primitive_data, out, n);
}
int gq_structure_silhouette_data(const gq_structure *s, float *out, int64_t *n) {
    if (!s) return fail(GQ_ERR_INVALID, "null structure");
    if (!s->h.is_snch) return fail(GQ_ERR_UNSUPPORTED, "structure is not an SNCH");
    return copy_out(s->h.silhouette_data, out, n);
}
int gq_structure_raw_silhouette_data(const gq_structure *s, float *out, int64_t *n) {
    if (!s) return fail(GQ_ERR_INVALID, "null structure");
    if (!s->h.is_snch) return fail(GQ_ERR_UNSUPPORTED, "structure is not an SNCH");
    return copy_out(s->h.raw_silhouette_data, out, n);
}

int gq_structure_save(const gq_structure *s, const char *path) {
    if (!s || !path) return fail(GQ_ERR_INVALID, "null argument");
    std::string err;
    if (!save_structure(s->h, path, err)) return fail(GQ_ERR_INVALID, err);
    return GQ_OK;
}

int gq_structure_load(const char *path, gq_structure **out) {
    if (!path || !out) return fail(GQ_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<gq_structure> s(new (std::nothrow) gq_structure());
    if (!s) return fail(GQ_ERR_OOM, "host allocation failed");
    std::string err;
    try {
        if (!load_structure(path, s->h, err)) return fail(GQ_ERR_INVALID, err);
    } catch (const std::bad_alloc &) {
        return fail(GQ_ERR_OOM, "host allocation failed while loading");
    }
    *out = s.release();
    return GQ_OK;
}

int gq_structure_geometry(const gq_structure *s, float *vertices, int64_t *n_floats, int32_t *indices,
                          int64_t *n_ints) {
    if (!s || !n_floats || !n_ints) return fail(GQ_ERR_INVALID, "null argument");
    const HostStructure &h = s->h;
    if ((vertices && *n_floats < (int64_t)h.vertices.size()) || (indices && *n_ints < (int64_t)h.indices.size()))
        return fail(GQ_ERR_INVALID, "output buffer too small");
    if (vertices) std::memcpy(vertices, h.vertices.data(), h.vertices.size() * sizeof(float));
    if (indices) std::memcpy(indices, h.indices.data(), h.indices.size() * sizeof(int32_t));
    *n_floats = (int64_t)h.vertices.size();
    *n_ints = (int64_t)h.indices.size();
    return GQ_OK;
}

int gq_build_flat_silhouette_edges(const float *vertices, int64_t n_vertices, const int32_t *faces,
                                   int64_t n_faces, float *out, int64_t *n_floats) {
    std::vector<float> v;
    std::string err;
    try {
        if (!build_flat_silhouette_edges(vertices, n_vertices, faces, n_faces, v, err))
            return fail(GQ_ERR_INVALID, err);
    } catch (const std::bad_alloc &) {
        return fail(GQ_ERR_OOM, "host allocation failed");
    }
    return copy_out(v, out, n_floats);
}

int gq_scene_destroy(gq_scene *scene) {
    if (!scene) return GQ_OK;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(scene->device);
    for (void *p : scene->allocs) hipFree(p);
    hipSetDevice(dev);
    delete scene;
    return GQ_OK;
}

int gq_scene_create(const gq_structure *st, gq_scene **out) { return gq_scene_create_ex(st, nullptr, out); }

int gq_scene_create_ex(const gq_structure *st, const gq_scene_options *opt, gq_scene **out) {
    if (!st || !out) return fail(GQ_ERR_INVALID, "null argument");
    const int leaf_prims = (opt && opt->leaf_prims > 0) ? opt->leaf_prims : default_cp_leaf(st->h.dim);
    const int ray_leaf = (opt && opt->ray_leaf_prims > 0) ? opt->ray_leaf_prims : default_ray_leaf(st->h.dim, st->h.n_prims);
    if (leaf_prims > kMaxLeafOption || ray_leaf > kMaxLeafOption)
        return fail(GQ_ERR_INVALID, "leaf_prims and ray_leaf_prims must be in 1..64");
    if (opt && (opt->traversal < 0 || opt->traversal > 2)) return fail(GQ_ERR_INVALID, "traversal must be 0, 1 or 2");
    if (opt && (opt->wave_front < 0 || opt->wave_front > 8192)) return fail(GQ_ERR_INVALID, "wave_front must be 0..8192");
    if (opt && (opt->sil_pairs < 0 || opt->sil_pairs > 1)) return fail(GQ_ERR_INVALID, "sil_pairs must be 0 or 1");
    *out = nullptr;
    const HostStructure &h = st->h;
    std::unique_ptr<gq_scene> s(new (std::nothrow) gq_scene());
    if (!s) return fail(GQ_ERR_OOM, "host allocation failed");
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    s->dim = h.dim;
    s->has_snch = h.is_snch;
    s->max_depth = h.max_depth;
    PlainView &pv = s->plain;
    pv.dim = h.dim;
    pv.n_nodes = h.n_nodes;
    pv.n_prims = h.n_prims;
    pv.n_sils = h.n_silhouettes;
    int rc;
    auto cleanup = [&](int code) {
        gq_scene_destroy(s.release());
        return code;
    };
    if ((rc = upload(s.get(), h.bvh_node_data.data(), h.bvh_node_data.size(), &pv.bvh_nodes))) return cleanup(rc);
    if ((rc = upload(s.get(), h.primitive_data.data(), h.primitive_data.size(), &pv.prims))) return cleanup(rc);
    if (h.is_snch) {
        if ((rc = upload(s.get(), h.snch_node_data.data(), h.snch_node_data.size(), &pv.snch_nodes))) return cleanup(rc);
        if ((rc = upload(s.get(), h.silhouette_data.data(), h.silhouette_data.size(), &pv.sils))) return cleanup(rc);
    }
    // Engine layouts.
    WideHost wh;
    try {
        build_wide(h, wh, leaf_prims, ray_leaf);
    } catch (const std::bad_alloc &) {
        return cleanup(fail(GQ_ERR_OOM, "host allocation failed while building wide nodes"));
    }
    WideView &wv = s->wide;
    wv.dim = h.dim;
    wv.width = wh.width;
    wv.wide_depth = wh.wide_depth;
    wv.ray_depth = wh.ray_depth;
    wv.packet_budget = opt ? opt->packet_budget : 0;
    wv.traversal = opt ? opt->traversal : 0;
    wv.wave_front = opt ? opt->wave_front : 0;
    wv.sil_pairs = opt ? opt->sil_pairs : 0;
    wv.bin_depth = h.max_depth;
    wv.n_nodes = (int64_t)wh.n_nodes;
    wv.n_snch_nodes = (int64_t)wh.n_snch_nodes;
    wv.n_prims = h.n_prims;
    wv.n_sils = h.n_silhouettes;
    wv.root_ref = wh.root_ref;
    wv.sroot_ref = wh.sroot_ref;
    for (int k = 0; k < 3; ++k) {
        wv.scene_lo[k] = wh.scene_lo[k]; wv.scene_hi[k] = wh.scene_hi[k];
        wv.sroot_lo[k] = wh.sroot_lo[k]; wv.sroot_hi[k] = wh.sroot_hi[k];
    }
    if ((rc = upload(s.get(), wh.nodes.data(), wh.nodes.size(), (const uint8_t **)&wv.nodes))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.qnodes.data(), wh.qnodes.size(), (const uint8_t **)&wv.qnodes))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.cbq.data(), wh.cbq.size(), (const uint8_t **)&wv.cbq))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.cbf.data(), wh.cbf.size(), (const uint8_t **)&wv.cbf))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.rnodes.data(), wh.rnodes.size(), (const uint8_t **)&wv.rnodes))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.rqnodes.data(), wh.rqnodes.size(), (const uint8_t **)&wv.rqnodes))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.prims.data(), wh.prims.size(), &wv.prims))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.prim_ids.data(), wh.prim_ids.size(), &wv.prim_ids))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.big_leaves.data(), wh.big_leaves.size(), &wv.big_leaves))) return cleanup(rc);
    wv.n_bvh_seed = (int)wh.bvh_seed.size();
    wv.n_sil_seed = (int)wh.sil_seed.size();
    if ((rc = upload(s.get(), wh.bvh_seed.data(), wh.bvh_seed.size(), (const Seed **)&wv.bvh_seed))) return cleanup(rc);
    if ((rc = upload(s.get(), wh.sil_seed.data(), wh.sil_seed.size(), (const Seed **)&wv.sil_seed))) return cleanup(rc);
    if (h.is_snch) {
        if ((rc = upload(s.get(), wh.snodes.data(), wh.snodes.size(), (const uint8_t **)&wv.snodes))) return cleanup(rc);
        if ((rc = upload(s.get(), wh.qsnodes.data(), wh.qsnodes.size(), (const uint8_t **)&wv.qsnodes))) return cleanup(rc);
        if ((rc = upload(s.get(), wh.sils.data(), wh.sils.size(), &wv.sils))) return cleanup(rc);
        if ((rc = upload(s.get(), wh.sil_ids.data(), wh.sil_ids.size(), &wv.sil_ids))) return cleanup(rc);
        if ((rc = upload(s.get(), wh.ref_cones.data(), wh.ref_cones.size(), &wv.ref_cones))) return cleanup(rc);
        if ((rc = upload(s.get(), wh.ref_cones2.data(), wh.ref_cones2.size(), (const RefCone2 **)&wv.ref_cones2)))
            return cleanup(rc);
    }
    *out = s.release();
    return GQ_OK;
}

int gq_scene_info(const gq_scene *s, int64_t info[8]) {
    if (!s || !info) return fail(GQ_ERR_INVALID, "null argument");
    info[0] = s->dim;
    info[1] = s->has_snch ? 1 : 0;
    info[2] = (int64_t)s->bytes;
    info[3] = s->plain.n_nodes;
    info[4] = s->wide.n_nodes;
    info[5] = s->plain.n_prims;
    info[6] = s->plain.n_sils;
    info[7] = 0;
    return GQ_OK;
}

// Whether a query runs on the quantized records.  Silhouettes use the float32 pair records unless
// the scene asks for the quantized ones: exact boxes open fewer nodes, and the 2D record carries
// the reference cones inline where the quantized one fetches them from a table (1M queries: 14.9
// vs 18.4 ms against 327,680 triangles, 0.92 vs 1.54 ms against 65,536 segments); that pass is
// bound by node tests and fetch latency, not by bytes.
static bool quantized_layout(int method, QueryKind kind, const WideView &w) {
    if (method != GQ_METHOD_ENGINE) return false;
    return !(kind == kSilhouette && w.sil_pairs == 0);
}

static int run_query(gq_scene *s, QueryKind kind, int method, const QueryArgs &q, void *stream) {
    if (!s) return fail(GQ_ERR_INVALID, "null scene");
    if (q.n < 0) return fail(GQ_ERR_INVALID, "negative query count");
    if (q.n == 0) return GQ_OK;
    if (q.n > (int64_t)INT32_MAX) return fail(GQ_ERR_INVALID, "more than 2^31-1 queries in one call");
    if (!q.x || (kind == kIntersect && !q.v)) return fail(GQ_ERR_INVALID, "null query pointer");
    if (kind == kSilhouette && !s->has_snch)
        return fail(GQ_ERR_UNSUPPORTED, "closest_silhouette_snch needs a scene built from an SNCH");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (method == GQ_METHOD_PLAIN) {
        if (s->max_depth + 2 > 64)
            return fail(GQ_ERR_UNSUPPORTED, "tree deeper than the plain kernels' 64-entry stack");
        e = launch_plain(kind, s->plain, q, st);
    } else if (method == GQ_METHOD_ENGINE || method == GQ_METHOD_ENGINE_FP32) {
        if (!engine_supported(kind, s->wide))
            return fail(GQ_ERR_UNSUPPORTED, "tree deeper than the engine kernels' traversal stack");
        e = launch_engine(kind, s->wide, quantized_layout(method, kind, s->wide), q, st);
    } else {
        return fail(GQ_ERR_INVALID, "unknown method");
    }
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return GQ_OK;
}

int gq_closest_point_bvh(gq_scene *scene, int method, const float *x, int64_t n,
                         const gq_records *out, void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    QueryArgs q{x, nullptr, nullptr, __builtin_inff(), n, *out};
    return run_query(scene, kClosestPoint, method, q, stream);
}

int gq_closest_point_bvh_noalloc(gq_scene *s, int method, const float *x, int64_t n, const gq_records *out,
                                 void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    if (!s) return fail(GQ_ERR_INVALID, "null scene");
    if (n < 0) return fail(GQ_ERR_INVALID, "negative query count");
    if (n == 0) return GQ_OK;
    if (n > (int64_t)INT32_MAX) return fail(GQ_ERR_INVALID, "more than 2^31-1 queries in one call");
    if (!x) return fail(GQ_ERR_INVALID, "null query pointer");
    if (method != GQ_METHOD_ENGINE && method != GQ_METHOD_ENGINE_FP32)
        return fail(GQ_ERR_UNSUPPORTED, "the allocation-free call runs the engine methods only");
    if (!engine_supported(kClosestPoint, s->wide))
        return fail(GQ_ERR_UNSUPPORTED, "tree deeper than the engine kernels' traversal stack");
    QueryArgs q{x, nullptr, nullptr, __builtin_inff(), n, *out};
    size_t need = 0;
    const hipError_t e = launch_engine_noalloc(kClosestPoint, s->wide, quantized_layout(method, kClosestPoint, s->wide), q,
                                               (hipStream_t)stream, &need);
    if (need > kNoAllocScratchBytes)
        return fail(GQ_ERR_INVALID, "gq_closest_point_bvh_noalloc: " + std::to_string(n) + " queries need " +
                                        std::to_string(need) + " bytes of scratch, more than the fixed " +
                                        std::to_string(kNoAllocScratchBytes) + " (split the batch)");
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return GQ_OK;
}

int gq_intersect_bvh(gq_scene *scene, int method, const float *x, const float *v,
                     const float *r_max, float r_max_scalar, int64_t n, const gq_records *out,
                     void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    QueryArgs q{x, v, r_max, r_max_scalar, n, *out};
    return run_query(scene, kIntersect, method, q, stream);
}

int gq_closest_silhouette_snch(gq_scene *scene, int method, const float *x, const float *r_max,
                               float r_max_scalar, int64_t n, const gq_records *out, void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    QueryArgs q{x, nullptr, r_max, r_max_scalar, n, *out};
    return run_query(scene, kSilhouette, method, q, stream);
}

int gq_sample_boundary(gq_scene *scene, const float *x, const float *R, float R_scalar, uint64_t *rng_state,
                       const uint64_t *rng_inc, int64_t n, const gq_sample_records *out, void *stream) {
    if (!scene || !out) return fail(GQ_ERR_INVALID, "null argument");
    if (n < 0) return fail(GQ_ERR_INVALID, "negative query count");
    if (n == 0) return GQ_OK;
    if (n > (int64_t)INT32_MAX) return fail(GQ_ERR_INVALID, "more than 2^31-1 queries in one call");
    if (!x || !rng_state || !rng_inc) return fail(GQ_ERR_INVALID, "null query or generator pointer");
    SampleArgs q{x, R, R_scalar, rng_state, rng_inc, n, *out};
    const hipError_t e = launch_sample(scene->plain, q, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return GQ_OK;
}

static int batch_jobs(int n_jobs, const gq_job *jobs, std::vector<EngineJob> &out) {
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) return fail(GQ_ERR_INVALID, "bad job list");
    out.clear();
    out.reserve((size_t)n_jobs);
    int device = -1;
    for (int j = 0; j < n_jobs; ++j) {
        const gq_job &b = jobs[j];
        const std::string at = "job " + std::to_string(j) + ": ";
        if (!b.scene) return fail(GQ_ERR_INVALID, at + "null scene");
        if (b.kind < GQ_QUERY_CLOSEST_POINT || b.kind > GQ_QUERY_SILHOUETTE) return fail(GQ_ERR_INVALID, at + "unknown query kind");
        if (b.method != GQ_METHOD_ENGINE && b.method != GQ_METHOD_ENGINE_FP32)
            return fail(GQ_ERR_UNSUPPORTED, at + "batches run the engine methods only");
        if (b.n < 0 || b.n > (int64_t)INT32_MAX) return fail(GQ_ERR_INVALID, at + "bad query count");
        const QueryKind kind = (QueryKind)b.kind;
        if (b.n > 0 && (!b.x || (kind == kIntersect && !b.v))) return fail(GQ_ERR_INVALID, at + "null query pointer");
        if (kind == kSilhouette && !b.scene->has_snch) return fail(GQ_ERR_UNSUPPORTED, at + "silhouettes need an SNCH scene");
        if (!engine_supported(kind, b.scene->wide)) return fail(GQ_ERR_UNSUPPORTED, at + "tree deeper than the engine's stack");
        if (b.traversal < 0 || b.traversal > 2) return fail(GQ_ERR_INVALID, at + "traversal must be 0, 1 or 2");
        if (b.order < 0 || b.order > 1) return fail(GQ_ERR_INVALID, at + "order must be 0 or 1");
        if (b.shared < 0 || b.shared > 1) return fail(GQ_ERR_INVALID, at + "shared must be 0 or 1");
        if (device >= 0 && b.scene->device != device) return fail(GQ_ERR_INVALID, at + "scenes on different devices");
        device = b.scene->device;
        if (b.n == 0) continue;
        QueryArgs q{b.x, b.v, b.r_max, b.r_max_scalar, b.n, b.out};
        out.push_back(EngineJob{&b.scene->wide, kind, quantized_layout(b.method, kind, b.scene->wide), q, b.packet_budget,
                                b.traversal != 0 ? b.traversal : b.scene->wide.traversal, b.order == 1,
                                b.shared == 1});
    }
    return GQ_OK;
}

int gq_batch_sizes(int n_jobs, const gq_job *jobs, int64_t *device_bytes, int64_t *host_bytes) {
    if (!device_bytes || !host_bytes) return fail(GQ_ERR_INVALID, "null argument");
    std::vector<EngineJob> ej;
    const int r = batch_jobs(n_jobs, jobs, ej);
    if (r != GQ_OK) return r;
    size_t ws = 0, table = 0;
    if (!ej.empty()) engine_batch_sizes(ej, ws, table);
    *device_bytes = (int64_t)ws;
    *host_bytes = (int64_t)table;
    return GQ_OK;
}

int gq_batch_enqueue(int n_jobs, const gq_job *jobs, void *device_ws, int64_t device_bytes, void *host_stage,
                     int64_t host_bytes, void *stream) {
    std::vector<EngineJob> ej;
    const int r = batch_jobs(n_jobs, jobs, ej);
    if (r != GQ_OK) return r;
    if (ej.empty()) return GQ_OK;
    size_t ws = 0, table = 0;
    engine_batch_sizes(ej, ws, table);
    if (!device_ws || !host_stage || device_bytes < (int64_t)ws || host_bytes < (int64_t)table)
        return fail(GQ_ERR_INVALID, "workspace or staging buffer smaller than gq_batch_sizes reports");
    const hipError_t e = launch_engine_batch(ej, device_ws, (size_t)device_bytes, host_stage, (size_t)host_bytes,
                                             (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "batch launch");
    return GQ_OK;
}

int gq_structure_engine_layout(const gq_structure *st, int which, void *out, int64_t *n_bytes) {
    if (!st || !n_bytes) return fail(GQ_ERR_INVALID, "null argument");
    WideHost wh;
    try {
        build_wide(st->h, wh, default_cp_leaf(st->h.dim), default_ray_leaf(st->h.dim, st->h.n_prims));
    } catch (const std::bad_alloc &) {
        return fail(GQ_ERR_OOM, "host allocation failed");
    }
    const void *src = nullptr;
    size_t bytes = 0;
    uint32_t scal[5] = {wh.root_ref, wh.sroot_ref, (uint32_t)wh.n_nodes, (uint32_t)wh.n_snch_nodes, (uint32_t)wh.wide_depth};
    switch (which) {
    case 0: src = wh.qnodes.data(); bytes = wh.qnodes.size(); break;
    case 1: src = wh.nodes.data(); bytes = wh.nodes.size(); break;
    case 2: src = wh.qsnodes.data(); bytes = wh.qsnodes.size(); break;
    case 3: src = wh.snodes.data(); bytes = wh.snodes.size(); break;
    case 4: src = wh.prims.data(); bytes = wh.prims.size() * sizeof(float4); break;
    case 5: src = wh.prim_ids.data(); bytes = wh.prim_ids.size() * 4; break;
    case 6: src = wh.sils.data(); bytes = wh.sils.size() * sizeof(float4); break;
    case 7: src = wh.sil_ids.data(); bytes = wh.sil_ids.size() * 4; break;
    case 8: src = wh.ref_cones.data(); bytes = wh.ref_cones.size() * sizeof(RefCone); break;
    case 9: src = wh.big_leaves.data(); bytes = wh.big_leaves.size() * 4; break;
    case 10: src = scal; bytes = sizeof scal; break;
    case 11: src = wh.ref_cones2.data(); bytes = wh.ref_cones2.size() * sizeof(RefCone2); break;
    case 12: src = wh.bvh_seed.data(); bytes = wh.bvh_seed.size() * sizeof(Seed); break;
    case 13: src = wh.sil_seed.data(); bytes = wh.sil_seed.size() * sizeof(Seed); break;
    case 14: src = wh.cbq.data(); bytes = wh.cbq.size(); break;
    case 15: src = wh.cbf.data(); bytes = wh.cbf.size(); break;
    case 16: src = wh.rnodes.data(); bytes = wh.rnodes.size(); break;
    case 17: src = wh.rqnodes.data(); bytes = wh.rqnodes.size(); break;
    default: return fail(GQ_ERR_INVALID, "unknown layout array");
    }
    if (out) {
        if (*n_bytes < (int64_t)bytes) return fail(GQ_ERR_INVALID, "output buffer too small");
        if (bytes) std::memcpy(out, src, bytes);
    }
    *n_bytes = (int64_t)bytes;
    return GQ_OK;
}

static void prof_destroy(Prof::Rec &e) {
    hipEventDestroy(e.begin);
    if (e.first) hipEventDestroy(e.first);
    hipEventDestroy(e.end);
}

int gq_profile_begin(void) {
    Prof &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    for (auto &e : p.ev) prof_destroy(e);
    p.ev.clear();
    p.name.clear();
    p.on = true;
    return GQ_OK;
}

int gq_profile_end_split(double *total_ms, double *first_kernel_ms, int64_t *launches, char *name,
                         int32_t name_len) {
    Prof &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    p.on = false;
    double tot = 0.0, first = 0.0;
    for (auto &e : p.ev) {
        hipError_t r = hipEventSynchronize(e.end);
        if (r != hipSuccess) return hip_fail(r, "hipEventSynchronize");
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e.begin, e.end);
        tot += ms;
        if (e.first) {
            hipEventElapsedTime(&ms, e.begin, e.first);
            first += ms;
        }
        prof_destroy(e);
    }
    if (total_ms) *total_ms = tot;
    if (first_kernel_ms) *first_kernel_ms = first;
    if (launches) *launches = (int64_t)p.ev.size();
    if (name && name_len > 0) {
        std::strncpy(name, p.name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    p.ev.clear();
    return GQ_OK;
}

int gq_profile_end(double *total_ms, int64_t *launches, char *name, int32_t name_len) {
    return gq_profile_end_split(total_ms, nullptr, launches, name, name_len);
}

// ---- Brute-force queries and normals (Polyline / Mesh *_baseline, configure_normal,
// configure_silhouette) ----

int gq_configure_silhouette(const float *V, int64_t nv, const int32_t *E, int64_t ne, float *out, int64_t *n_floats) {
    if (!n_floats) return fail(GQ_ERR_INVALID, "n_floats is null");
    std::vector<float> rec;
    std::string err;
    try {
        if (!configure_silhouette_2d(V, nv, E, ne, rec, err)) return fail(GQ_ERR_INVALID, err);
    } catch (const std::bad_alloc &) {
        return fail(GQ_ERR_OOM, "host allocation failed");
    }
    if (out) {
        if (*n_floats < (int64_t)rec.size()) return fail(GQ_ERR_INVALID, "output buffer too small");
        std::memcpy(out, rec.data(), rec.size() * sizeof(float));
    }
    *n_floats = (int64_t)rec.size();
    return GQ_OK;
}

static int silprep_status(hipError_t e, uint32_t bad, const char *what) {
    if (e != hipSuccess) return hip_fail(e, what);
    if (bad & 1u) return fail(GQ_ERR_INVALID, "index out of range");
    if (bad & 2u) return fail(GQ_ERR_INVALID, "non-finite vertex coordinate");
    if (bad & 4u) return fail(GQ_ERR_INVALID, "output buffer too small");
    return GQ_OK;
}

static int silprep_args(const float *V, int64_t nv, const int32_t *I, int64_t np, const int64_t *n_floats) {
    if (!n_floats) return fail(GQ_ERR_INVALID, "n_floats is null");
    if (nv <= 0 || np <= 0 || !V || !I) return fail(GQ_ERR_INVALID, "empty vertex or index array");
    if (np >= (1 << 24) || nv >= (1 << 24))
        return fail(GQ_ERR_INVALID, "more than 2^24 primitives or vertices: indices are stored as float32");
    return GQ_OK;
}

int gq_configure_silhouette_device(const float *V, int64_t nv, const int32_t *E, int64_t ne, float *out,
                                   int64_t *n_floats, void *stream) {
    int rc = silprep_args(V, nv, E, ne, n_floats);
    if (rc) return rc;
    uint32_t bad = 0;
    const hipError_t e = configure_silhouette_device(V, nv, E, ne, out, out ? *n_floats : 0, n_floats, &bad,
                                                     (hipStream_t)stream);
    return silprep_status(e, bad, "configure_silhouette");
}

int gq_build_flat_silhouette_edges_device(const float *V, int64_t nv, const int32_t *F, int64_t nf, float *out,
                                          int64_t *n_floats, void *stream) {
    int rc = silprep_args(V, nv, F, nf, n_floats);
    if (rc) return rc;
    uint32_t bad = 0;
    const hipError_t e = flat_silhouette_edges_device(V, nv, F, nf, out, out ? *n_floats : 0, n_floats, &bad,
                                                      (hipStream_t)stream);
    return silprep_status(e, bad, "build_flat_silhouette_edges");
}

int gq_shape_destroy(gq_shape *shape) {
    if (!shape) return GQ_OK;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(shape->device);
    for (void *p : shape->allocs) hipFree(p);
    hipSetDevice(dev);
    delete shape;
    return GQ_OK;
}

int gq_shape_create(int dim, const float *V, int64_t nv, const int32_t *I, int64_t np, gq_shape **out) {
    if (!out) return fail(GQ_ERR_INVALID, "out is null");
    *out = nullptr;
    std::string err;
    if (!validate_shape(dim, V, nv, I, np, err)) return fail(GQ_ERR_INVALID, err);
    std::unique_ptr<gq_shape, int (*)(gq_shape *)> s(new (std::nothrow) gq_shape(), gq_shape_destroy);
    if (!s) return fail(GQ_ERR_OOM, "host allocation failed");
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::vector<int32_t> vf_start, vf_list;
    try {
        if (dim == 3) vertex_faces(I, np, nv, vf_start, vf_list);
    } catch (const std::bad_alloc &) {
        return fail(GQ_ERR_OOM, "host allocation failed");
    }
    ShapeView &v = s->view;
    v.dim = dim;
    v.n_vertices = nv;
    v.n_prims = np;
    int rc;
    if ((rc = upload(s.get(), V, (size_t)(nv * dim), &v.V))) return rc;
    if ((rc = upload(s.get(), I, (size_t)(np * dim), &v.F))) return rc;
    {
        // the silhouette records, built on the device from the uploaded arrays: a counting pass,
        // then the records into an allocation of exactly that size
        auto setup = dim == 2 ? configure_silhouette_device : flat_silhouette_edges_device;
        int64_t nf = 0;
        uint32_t bad = 0;
        if ((rc = silprep_status(setup(v.V, nv, v.F, np, nullptr, 0, &nf, &bad, nullptr), bad, "silhouette setup")))
            return rc;
        v.n_sils = nf / (dim == 2 ? 10 : 19);
        void *sp = nullptr;
        if (nf > 0) {
            if ((e = hipMalloc(&sp, (size_t)nf * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc");
            s->allocs.push_back(sp);
            s->bytes += (size_t)nf * sizeof(float);
            if ((rc = silprep_status(setup(v.V, nv, v.F, np, (float *)sp, nf, &nf, &bad, nullptr), bad,
                                     "silhouette setup")))
                return rc;
        }
        v.sils = (const float *)sp;
    }
    if (dim == 3) {
        if ((rc = upload(s.get(), vf_start.data(), vf_start.size(), &v.vf_start))) return rc;
        if ((rc = upload(s.get(), vf_list.data(), vf_list.size(), &v.vf_list))) return rc;
        void *fn = nullptr, *vn = nullptr;
        if ((e = hipMalloc(&fn, (size_t)np * 3 * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc");
        s->allocs.push_back(fn);
        if ((e = hipMalloc(&vn, (size_t)nv * 3 * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc");
        s->allocs.push_back(vn);
        s->bytes += (size_t)(np + nv) * 3 * sizeof(float);
        s->face_normals = (float *)fn;
        s->vertex_normals = (float *)vn;
        v.face_normals = s->face_normals;
        if ((e = launch_normals(v, s->face_normals, s->vertex_normals, nullptr)) != hipSuccess)
            return hip_fail(e, "normals");
        if ((e = hipStreamSynchronize(nullptr)) != hipSuccess) return hip_fail(e, "normals");
    }
    *out = s.release();
    return GQ_OK;
}

int gq_shape_info(const gq_shape *shape, int64_t info[4]) {
    if (!shape || !info) return fail(GQ_ERR_INVALID, "null argument");
    info[0] = shape->view.dim;
    info[1] = shape->view.n_vertices;
    info[2] = shape->view.n_prims;
    info[3] = shape->view.n_sils;
    return GQ_OK;
}

int gq_shape_normals(const gq_shape *shape, float *face_normals, float *vertex_normals, void *stream) {
    if (!shape) return fail(GQ_ERR_INVALID, "null shape");
    if (shape->view.dim != 3) return fail(GQ_ERR_UNSUPPORTED, "normals are computed for meshes (3D)");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (face_normals &&
        (e = hipMemcpyAsync(face_normals, shape->face_normals, (size_t)shape->view.n_prims * 3 * sizeof(float),
                            hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync");
    if (vertex_normals &&
        (e = hipMemcpyAsync(vertex_normals, shape->vertex_normals, (size_t)shape->view.n_vertices * 3 * sizeof(float),
                            hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync");
    return GQ_OK;
}

static int run_baseline(gq_shape *s, QueryKind kind, const QueryArgs &q, const float *normal,
                        const uint8_t *on_boundary, void *stream) {
    if (!s) return fail(GQ_ERR_INVALID, "null shape");
    if (q.n < 0) return fail(GQ_ERR_INVALID, "negative query count");
    if (q.n == 0) return GQ_OK;
    if (q.n > (int64_t)INT32_MAX) return fail(GQ_ERR_INVALID, "more than 2^31-1 queries in one call");
    if (!q.x || (kind == kIntersect && !q.v)) return fail(GQ_ERR_INVALID, "null query pointer");
    if (on_boundary && !normal) return fail(GQ_ERR_INVALID, "on_boundary needs normals");
    const hipError_t e = launch_baseline(kind, s->view, q, normal, on_boundary, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return GQ_OK;
}

int gq_closest_point_baseline(gq_shape *shape, const float *x, int64_t n, const gq_records *out, void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    QueryArgs q{x, nullptr, nullptr, __builtin_inff(), n, *out};
    return run_baseline(shape, kClosestPoint, q, nullptr, nullptr, stream);
}

int gq_intersect_baseline(gq_shape *shape, const float *x, const float *v, const float *normal,
                          const uint8_t *on_boundary, const float *r_max, float r_max_scalar, int64_t n,
                          const gq_records *out, void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    QueryArgs q{x, v, r_max, r_max_scalar, n, *out};
    return run_baseline(shape, kIntersect, q, normal, on_boundary, stream);
}

int gq_closest_silhouette_baseline(gq_shape *shape, const float *x, const float *r_max, float r_max_scalar,
                                   int64_t n, const gq_records *out, void *stream) {
    if (!out) return fail(GQ_ERR_INVALID, "null records");
    QueryArgs q{x, nullptr, r_max, r_max_scalar, n, *out};
    return run_baseline(shape, kSilhouette, q, nullptr, nullptr, stream);
}

} // extern "C"
