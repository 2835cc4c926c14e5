// Scene preprocessing on the device: the silhouette records of a polyline (Polyline.configure_silhouette,
// gquery/shapes/polyline.py:44-122) and of a mesh (build_flat_silhouette_edges,
// src/shapes/silhouette_edge.cpp:5-71), bit-identical to the host restatements in builder.cpp
// (configure_silhouette_2d; build_edges + flatten_sil).
//
// Both are orderings of first appearances, which map onto sorts and scans:
//  * 2D: each vertex's first position in the edge list (atomicMin over 2k / 2k + 1) places it in
//    a position array; the last edge leaving / entering it (atomicMax over k) gives next / prev.
//    One scan over the positions yields each vertex's first-appearance index and its output slot.
//  * 3D: the half-edges (face i, corner j) are keyed by their sorted vertex pair and stably radix
//    sorted, so each edge's half-edges sit together in (i, j) order; the edge id is the rank of its
//    first half-edge among the first half-edges (a scan over a mark per half-edge), and the record's
//    fields come from the last half-edge writing each slot, as in the host loop.
// Every kernel is a plain grid-stride loop over its items; inputs are validated by a first kernel
// (indices in range, finite coordinates) before any kernel indexes with them.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <cstdint>

#include "gq_internal.h"

namespace gqa {
namespace {

constexpr int kBlk = 256;

inline unsigned grid_for(int64_t n) {
    const int64_t b = (n + kBlk - 1) / kBlk;
    return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

#define GQ_FOR(t, n) for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (n); \
                          t += (int64_t)gridDim.x * blockDim.x)

// bit 0: an index outside [0, nv); bit 1: a non-finite coordinate
__global__ void __launch_bounds__(kBlk) k_check(const float *V, int64_t nvc, const int32_t *I, int64_t ni, int64_t nv,
                                                uint32_t *flags) {
    uint32_t f = 0;
    GQ_FOR(t, ni) {
        if (I[t] < 0 || (int64_t)I[t] >= nv) f |= 1u;
    }
    GQ_FOR(t, nvc) {
        if (!isfinite(V[t])) f |= 2u;
    }
    if (f) atomicOr(flags, f);
}

// ---- 2D: silhouette vertices ----------------------------------------------------------------
__global__ void __launch_bounds__(kBlk) k_sv_init(int64_t nv, int64_t npos, uint32_t *first, int32_t *lout,
                                                  int32_t *lin, int32_t *slot_vertex) {
    GQ_FOR(t, nv) {
        first[t] = 0xFFFFFFFFu;
        lout[t] = -1;
        lin[t] = -1;
    }
    GQ_FOR(t, npos) slot_vertex[t] = -1;
}

__global__ void __launch_bounds__(kBlk) k_sv_edges(const int32_t *E, int64_t ne, uint32_t *first, int32_t *lout,
                                                   int32_t *lin) {
    GQ_FOR(k, ne) {
        const int32_t a = E[2 * k], b = E[2 * k + 1];
        atomicMin(first + a, (uint32_t)(2 * k));
        atomicMin(first + b, (uint32_t)(2 * k + 1));
        atomicMax(lout + a, (int32_t)k);   // next[a] = b of the last edge leaving a
        atomicMax(lin + b, (int32_t)k);    // prev[b] = a of the last edge entering b
    }
}

__global__ void __launch_bounds__(kBlk) k_sv_place(int64_t nv, const uint32_t *first, int32_t *slot_vertex) {
    GQ_FOR(v, nv) {
        if (first[v] != 0xFFFFFFFFu) slot_vertex[first[v]] = (int32_t)v;
    }
}

// float32 unit direction, as the query kernels normalise: v * (1 / sqrt(dot(v, v)))
__device__ __forceinline__ void unit2(float x, float y, float &ux, float &uy) {
    const float inv = 1.0f / sqrtf(fmaf(y, y, x * x));
    ux = x * inv;
    uy = y * inv;
}

// Whether the vertex at position p is emitted (it has both neighbours and they are not nearly
// collinear); prev / next receive them.
__device__ __forceinline__ bool sv_emit(const float *V, const int32_t *E, int32_t v, const int32_t *lout,
                                        const int32_t *lin, int32_t &p, int32_t &n) {
    const int32_t lo = lout[v], li = lin[v];
    if (lo < 0 || li < 0) return false;
    p = E[2 * (int64_t)li];
    n = E[2 * (int64_t)lo + 1];
    const float *a = V + 2 * (int64_t)p, *b = V + 2 * (int64_t)v, *c = V + 2 * (int64_t)n;
    float e1x, e1y, e2x, e2y;
    unit2(b[0] - a[0], b[1] - a[1], e1x, e1y);
    unit2(c[0] - b[0], c[1] - b[1], e2x, e2y);
    const float col = fabsf(fmaf(e1y, e2y, e1x * e2x));
    return col < 1.0f - 1e-5f;   // nearly collinear (or degenerate: NaN) is dropped
}

// (occupied << 32) | emitted per position: one scan gives both the first-appearance index and
// the output slot
__global__ void __launch_bounds__(kBlk) k_sv_flags(const float *V, const int32_t *E, int64_t npos,
                                                   const int32_t *slot_vertex, const int32_t *lout,
                                                   const int32_t *lin, uint64_t *flag) {
    GQ_FOR(q, npos) {
        const int32_t v = slot_vertex[q];
        uint64_t f = 0;
        if (v >= 0) {
            int32_t p, n;
            f = (1ull << 32) | (sv_emit(V, E, v, lout, lin, p, n) ? 1ull : 0ull);
        }
        flag[q] = f;
    }
}

__global__ void __launch_bounds__(kBlk) k_sv_write(const float *V, const int32_t *E, int64_t npos,
                                                   const int32_t *slot_vertex, const int32_t *lout,
                                                   const int32_t *lin, const uint64_t *flag, const uint64_t *pre,
                                                   float *out) {
    GQ_FOR(q, npos) {
        if (!(flag[q] & 1ull)) continue;
        const int32_t v = slot_vertex[q];
        int32_t p = 0, n = 0;
        sv_emit(V, E, v, lout, lin, p, n);
        const uint64_t s = pre[q];
        const uint32_t order = (uint32_t)(s >> 32), slot = (uint32_t)(s & 0xFFFFFFFFull);
        const float *a = V + 2 * (int64_t)p, *b = V + 2 * (int64_t)v, *c = V + 2 * (int64_t)n;
        float *o = out + 10 * (int64_t)slot;
        o[0] = a[0]; o[1] = a[1];
        o[2] = b[0]; o[3] = b[1];
        o[4] = c[0]; o[5] = c[1];
        o[6] = (float)p; o[7] = (float)v; o[8] = (float)n; o[9] = (float)order;
    }
}

// ---- 3D: silhouette edges -------------------------------------------------------------------
__global__ void __launch_bounds__(kBlk) k_se_keys(const int32_t *F, int64_t nh, int bits, uint64_t *key,
                                                  uint32_t *val) {
    GQ_FOR(h, nh) {
        const int64_t i = h / 3, j = h % 3;
        const uint32_t i0 = (uint32_t)F[3 * i + j], i1 = (uint32_t)F[3 * i + (j + 1) % 3];
        const uint32_t lo = i0 < i1 ? i0 : i1, hi = i0 < i1 ? i1 : i0;
        key[h] = ((uint64_t)lo << bits) | hi;
        val[h] = (uint32_t)h;
    }
}

// mark[h] = 1 for the first half-edge (lowest h: the sort is stable) of every edge
__global__ void __launch_bounds__(kBlk) k_se_heads(int64_t nh, const uint64_t *key, const uint32_t *val,
                                                   uint32_t *mark) {
    GQ_FOR(t, nh) {
        mark[val[t]] = (t == 0 || key[t] != key[t - 1]) ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(kBlk) k_se_write(const float *V, const int32_t *F, int64_t nh, const uint64_t *key,
                                                   const uint32_t *val, const uint32_t *eid, float *out) {
    GQ_FOR(t, nh) {
        if (t != 0 && key[t] == key[t - 1]) continue;
        // the edge's half-edges in (face, corner) order; each one overwrites m1, m2 (its own
        // traversal of the edge) and the opposite-vertex slot of its side
        float m[4][3] = {};
        float idx[4] = {0.0f, 0.0f, 0.0f, 0.0f}, face[2] = {0.0f, 0.0f};
        for (int64_t u = t; u < nh && key[u] == key[t]; ++u) {
            const uint32_t h = val[u];
            const int64_t i = h / 3, j = h % 3;
            const int32_t i0 = F[3 * i + j], i1 = F[3 * i + (j + 1) % 3], i2 = F[3 * i + (j + 2) % 3];
            const bool swapped = i0 > i1;
            const float *v0 = V + 3 * (int64_t)i0, *v1 = V + 3 * (int64_t)i1, *v2 = V + 3 * (int64_t)i2;
            const int side = swapped ? 3 : 0;
            for (int k = 0; k < 3; ++k) {
                m[1][k] = v0[k];
                m[2][k] = v1[k];
                m[side][k] = v2[k];
            }
            idx[1] = (float)(swapped ? i1 : i0);
            idx[2] = (float)(swapped ? i0 : i1);
            idx[side] = (float)i2;
            face[swapped ? 1 : 0] = (float)i;
        }
        const uint32_t e = eid[val[t]];
        float *o = out + 19 * (int64_t)e;
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) o[3 * k + c] = m[k][c];
        for (int k = 0; k < 4; ++k) o[12 + k] = idx[k];
        o[16] = face[0];
        o[17] = face[1];
        o[18] = (float)e;
    }
}

// Stream-ordered scratch carved into aligned pieces.
struct Carve {
    uint8_t *base = nullptr;
    size_t off = 0;
    template <typename T> T *take(size_t count) {
        T *p = (T *)(base ? base + off : nullptr);
        off += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

// Validates the inputs on the device: 0, or 1 / 2 as k_check's flags.
hipError_t check_inputs(const float *V, int64_t nv, int dim, const int32_t *I, int64_t ni, hipStream_t st,
                        uint32_t &flags) {
    uint32_t *d = nullptr;
    hipError_t e = hipMallocAsync((void **)&d, sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(d, 0, sizeof(uint32_t), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check, dim3(grid_for(ni > nv * dim ? ni : nv * dim)), dim3(kBlk), 0, st, V, nv * dim, I,
                           ni, nv, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&flags, d, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    const hipError_t f = hipFreeAsync(d, st);
    if (e == hipSuccess) e = f;
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

} // namespace

hipError_t configure_silhouette_device(const float *V, int64_t nv, const int32_t *E, int64_t ne, float *out,
                                       int64_t capacity, int64_t *n_floats, uint32_t *bad, hipStream_t st) {
    *bad = 0;
    hipError_t e = check_inputs(V, nv, 2, E, 2 * ne, st, *bad);
    if (e != hipSuccess || *bad) return e;
    const int64_t npos = 2 * ne;
    size_t scan_bytes = 0;
    e = rocprim::exclusive_scan(nullptr, scan_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr, (uint64_t)0,
                                (size_t)npos, rocprim::plus<uint64_t>(), st);
    if (e != hipSuccess) return e;
    Carve c;
    auto lay = [&](Carve &k, uint32_t *&first, int32_t *&lout, int32_t *&lin, int32_t *&sv, uint64_t *&flag,
                   uint64_t *&pre, void *&tmp) {
        first = k.take<uint32_t>((size_t)nv);
        lout = k.take<int32_t>((size_t)nv);
        lin = k.take<int32_t>((size_t)nv);
        sv = k.take<int32_t>((size_t)npos);
        flag = k.take<uint64_t>((size_t)npos);
        pre = k.take<uint64_t>((size_t)npos);
        tmp = k.take<uint8_t>(scan_bytes);
    };
    uint32_t *first; int32_t *lout, *lin, *sv; uint64_t *flag, *pre; void *tmp;
    lay(c, first, lout, lin, sv, flag, pre, tmp);
    void *buf = nullptr;
    if ((e = hipMallocAsync(&buf, c.off, st)) != hipSuccess) return e;
    Carve k;
    k.base = (uint8_t *)buf;
    lay(k, first, lout, lin, sv, flag, pre, tmp);
    hipLaunchKernelGGL(k_sv_init, dim3(grid_for(nv > npos ? nv : npos)), dim3(kBlk), 0, st, nv, npos, first, lout, lin,
                       sv);
    hipLaunchKernelGGL(k_sv_edges, dim3(grid_for(ne)), dim3(kBlk), 0, st, E, ne, first, lout, lin);
    hipLaunchKernelGGL(k_sv_place, dim3(grid_for(nv)), dim3(kBlk), 0, st, nv, first, sv);
    hipLaunchKernelGGL(k_sv_flags, dim3(grid_for(npos)), dim3(kBlk), 0, st, V, E, npos, sv, lout, lin, flag);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(tmp, scan_bytes, flag, pre, (uint64_t)0, (size_t)npos, rocprim::plus<uint64_t>(),
                                    st);
    uint64_t last[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&last[0], pre + (npos - 1), 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], flag + (npos - 1), 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    int64_t count = 0;
    if (e == hipSuccess) {
        count = (int64_t)((last[0] & 0xFFFFFFFFull) + (last[1] & 0xFFFFFFFFull));
        *n_floats = 10 * count;
        if (out && 10 * count > capacity) *bad = 4;
    }
    if (e == hipSuccess && out && !*bad && count > 0) {
        hipLaunchKernelGGL(k_sv_write, dim3(grid_for(npos)), dim3(kBlk), 0, st, V, E, npos, sv, lout, lin, flag, pre,
                           out);
        e = hipGetLastError();
    }
    const hipError_t f = hipFreeAsync(buf, st);
    if (e == hipSuccess) e = f;
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

hipError_t flat_silhouette_edges_device(const float *V, int64_t nv, const int32_t *F, int64_t nf, float *out,
                                        int64_t capacity, int64_t *n_floats, uint32_t *bad, hipStream_t st) {
    *bad = 0;
    hipError_t e = check_inputs(V, nv, 3, F, 3 * nf, st, *bad);
    if (e != hipSuccess || *bad) return e;
    const int64_t nh = 3 * nf;
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < nv) ++bits;
    size_t sort_bytes = 0, scan_bytes = 0;
    e = rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint32_t *)nullptr,
                                  (uint32_t *)nullptr, (size_t)nh, 0, 2 * bits, st);
    if (e != hipSuccess) return e;
    e = rocprim::exclusive_scan(nullptr, scan_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u, (size_t)nh,
                                rocprim::plus<uint32_t>(), st);
    if (e != hipSuccess) return e;
    uint64_t *key, *key_s; uint32_t *val, *val_s, *mark, *eid; void *tsort, *tscan;
    auto lay = [&](Carve &k) {
        key = k.take<uint64_t>((size_t)nh);
        key_s = k.take<uint64_t>((size_t)nh);
        val = k.take<uint32_t>((size_t)nh);
        val_s = k.take<uint32_t>((size_t)nh);
        mark = k.take<uint32_t>((size_t)nh);
        eid = k.take<uint32_t>((size_t)nh);
        tsort = k.take<uint8_t>(sort_bytes);
        tscan = k.take<uint8_t>(scan_bytes);
    };
    Carve c;
    lay(c);
    void *buf = nullptr;
    if ((e = hipMallocAsync(&buf, c.off, st)) != hipSuccess) return e;
    Carve k;
    k.base = (uint8_t *)buf;
    lay(k);
    hipLaunchKernelGGL(k_se_keys, dim3(grid_for(nh)), dim3(kBlk), 0, st, F, nh, bits, key, val);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(tsort, sort_bytes, key, key_s, val, val_s, (size_t)nh, 0, 2 * bits, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_se_heads, dim3(grid_for(nh)), dim3(kBlk), 0, st, nh, key_s, val_s, mark);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(tscan, scan_bytes, mark, eid, 0u, (size_t)nh, rocprim::plus<uint32_t>(), st);
    uint32_t last[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&last[0], eid + (nh - 1), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], mark + (nh - 1), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    int64_t count = 0;
    if (e == hipSuccess) {
        count = (int64_t)last[0] + (int64_t)last[1];
        *n_floats = 19 * count;
        if (out && 19 * count > capacity) *bad = 4;
    }
    if (e == hipSuccess && out && !*bad) {
        hipLaunchKernelGGL(k_se_write, dim3(grid_for(nh)), dim3(kBlk), 0, st, V, F, nh, key_s, val_s, eid, out);
        e = hipGetLastError();
    }
    const hipError_t f = hipFreeAsync(buf, st);
    if (e == hipSuccess) e = f;
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

} // namespace gqa
