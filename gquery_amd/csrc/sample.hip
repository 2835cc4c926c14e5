// BVH-guided boundary sampling: BVH3D.sample_boundary / _sample_boundary
// (gquery/shapes/bvh3d.py:246-346), and the same walk over 2D segments.
//
// One lane per query walks from the root to one leaf of the reference binary BVH (the stride-9 /
// stride-7 node array and the stride-10 / stride-5 primitive array the plain kernels read): at an
// interior node each child whose box meets the ball (x, R) has weight 1, one uniform number picks
// one of them (the pmf takes weight / total, r_max becomes that child's farthest-corner distance);
// at the leaf the visible primitives (r_max <= R, or the primitive meets the ball) form a one-pass
// reservoir weighted by surface area, one number per visible primitive; then the point is drawn
// on the chosen primitive (triangle.py:27-42: u1 = sqrt(r0), u = 1 - u1, v = r1 u1, w = 1 - u - v,
// p = a u + b v + c w, pdf = 2 / |cross|; line_segment.py:27-32: p = a + (b - a) t, pdf = 1 / |b-a|).
// Where the reference cannot run as written (bvh3d.py:265,303 name fields that do not exist), the
// chosen primitive is its leaf-order slot and the record has no type; a leaf whose visible
// primitives all have zero area gives no sample.
//
// Random numbers: PCG32 (O'Neill's XSH RR 64/32, as DrJit's PCG32): state' = state * 6364136223846793005
// + inc, output rotr32(((old >> 18) ^ old) >> 27, old >> 59), next_float32 = float bits
// ((out >> 9) | 0x3f800000) - 1.  Each query carries its own state, advanced in place.
// Float32 arithmetic in the reference's order, compiled with -ffp-contract=off: the records and the
// final states equal the CPU restatement's bit for bit.
#include "gq_device.h"
#include "gq_internal.h"

namespace gqa {
namespace {

using namespace gqd;

constexpr int kBlockS = 256;

__device__ __forceinline__ uint32_t pcg32_next(uint64_t &state, uint64_t inc) {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

__device__ __forceinline__ float pcg32_float(uint64_t &state, uint64_t inc) {
    return __uint_as_float((pcg32_next(state, inc) >> 9) | 0x3f800000u) - 1.0f;
}

template <int D> struct Prim;
template <> struct Prim<3> {
    static constexpr int kStride = 10;
    V3 a, b, c;
    int32_t index;
    __device__ __forceinline__ void load(const float *pr) {
        a = v3(pr[0], pr[1], pr[2]);
        b = v3(pr[3], pr[4], pr[5]);
        c = v3(pr[6], pr[7], pr[8]);
        index = (int32_t)pr[9];
    }
    // Triangle.surface_area (triangle.py:23-25)
    __device__ __forceinline__ float area() const { return norm(cross(b - a, c - a)) / 2.0f; }
    // Triangle.sphere_intersect (triangle.py:44-47)
    __device__ __forceinline__ bool in_ball(V3 x, float R) const {
        V3 pt;
        V2 uv;
        return closest_point_triangle(x, a, b, c, pt, uv) <= R;
    }
};
template <> struct Prim<2> {
    static constexpr int kStride = 5;
    V2 a, b;
    int32_t index;
    __device__ __forceinline__ void load(const float *pr) {
        a = v2(pr[0], pr[1]);
        b = v2(pr[2], pr[3]);
        index = (int32_t)pr[4];
    }
    // LineSegment.surface_area (line_segment.py:23-25)
    __device__ __forceinline__ float area() const { return norm(b - a); }
    // LineSegment.sphere_intersect (line_segment.py:34-36)
    __device__ __forceinline__ bool in_ball(V2 x, float R) const {
        V2 pt;
        return closest_point_segment(x, a, b, pt) <= R;
    }
};

template <int D>
__global__ void __launch_bounds__(kBlockS) k_sample_boundary(PlainView s, SampleArgs q) {
    using V = typename VT<D>::T;
    constexpr int NS = 2 * D + 3;          // node stride: box, reference_offset, second child, count
    constexpr int PS = Prim<D>::kStride;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const V x = ldv<D>(q.x + D * i);
    const float R = q.R ? q.R[i] : q.R_scalar;
    uint64_t rng = q.state[i];
    const uint64_t inc = q.inc[i];
    const float *nodes = s.bvh_nodes;
    int slot = -1;
    float pdf = 1.0f, dmin, dmax;
    box_distance(ldv<D>(nodes), ldv<D>(nodes + D), x, dmin, dmax);
    if (dmin <= R) {
        float r_max = __builtin_inff();
        int64_t node = 0;
        for (;;) {
            const float *nd = nodes + NS * node;
            const int n_refs = (int)nd[2 * D + 2];
            if (n_refs > 0) {
                const int off = (int)nd[2 * D];
                float total = 0.0f, chosen_area = 0.0f;
                for (int j = 0; j < n_refs; ++j) {
                    Prim<D> pr;
                    pr.load(s.prims + (int64_t)PS * (off + j));
                    const float area = pr.area();
                    if ((r_max <= R) || pr.in_ball(x, R)) {
                        total += area;
                        const float sel = area / total;
                        const float u = pcg32_float(rng, inc);
                        if (u < sel) {
                            slot = off + j;
                            chosen_area = area;
                        }
                    }
                }
                if (total > 0.0f) pdf *= chosen_area / total;
                break;
            }
            const int64_t left = node + 1, right = node + (int64_t)(int)nd[2 * D + 1];
            const float *b0 = nodes + NS * left, *b1 = nodes + NS * right;
            float dmin0, dmax0, dmin1, dmax1;
            box_distance(ldv<D>(b0), ldv<D>(b0 + D), x, dmin0, dmax0);
            box_distance(ldv<D>(b1), ldv<D>(b1 + D), x, dmin1, dmax1);
            const float w0 = dmin0 <= R ? 1.0f : 0.0f, w1 = dmin1 <= R ? 1.0f : 0.0f;
            const float tw = w0 + w1;
            if (!(tw > 0.0f)) break;
            const float p0 = w0 / tw, p1 = 1.0f - p0;
            const float u = pcg32_float(rng, inc);
            if (u < p0) {
                node = left;
                pdf *= p0;
                r_max = dmax0;
            } else {
                node = right;
                pdf *= p1;
                r_max = dmax1;
            }
        }
    }
    const gq_sample_records &o = q.out;
    V P{}, N{};
    float uv0 = 0.0f, uv1 = 0.0f, pdf_rec = 0.0f, pmf_rec = 0.0f;
    int32_t id = 0;
    if (slot >= 0) {
        Prim<D> pr;
        pr.load(s.prims + (int64_t)PS * slot);
        if constexpr (D == 3) {
            const float u1 = sqrtf(pcg32_float(rng, inc));
            const float u2 = pcg32_float(rng, inc);
            const float u = 1.0f - u1, v = u2 * u1, w = 1.0f - u - v;
            P = pr.a * u + pr.b * v + pr.c * w;
            const V3 c = cross(pr.b - pr.a, pr.c - pr.a);
            pdf_rec = pdf * (2.0f / norm(c));
            N = normalize(c);
            uv0 = u;
            uv1 = v;
        } else {
            const float t = pcg32_float(rng, inc);
            const V2 ba = pr.b - pr.a;
            P = pr.a + ba * t;
            pdf_rec = pdf * (1.0f / norm(ba));
            N = normalize(v2(ba.y, -ba.x));
            uv0 = t;
        }
        pmf_rec = pdf;
        id = pr.index;
    }
    if (o.hit) o.hit[i] = slot >= 0 ? 1 : 0;
    if (o.p) st(o.p + D * i, P);
    if (o.n) st(o.n + D * i, N);
    if (o.uv) {
        o.uv[(D - 1) * i] = uv0;
        if (D == 3) o.uv[(D - 1) * i + 1] = uv1;
    }
    if (o.pdf) o.pdf[i] = pdf_rec;
    if (o.pmf) o.pmf[i] = pmf_rec;
    if (o.prim_id) o.prim_id[i] = id;
    q.state[i] = rng;
}

} // namespace

hipError_t launch_sample(const PlainView &s, const SampleArgs &q, hipStream_t st) {
    if (q.n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((q.n + kBlockS - 1) / kBlockS);
    prof_begin(st);
    if (s.dim == 3) hipLaunchKernelGGL(k_sample_boundary<3>, dim3(grid), dim3(kBlockS), 0, st, s, q);
    else hipLaunchKernelGGL(k_sample_boundary<2>, dim3(grid), dim3(kBlockS), 0, st, s, q);
    prof_end(st, s.dim == 3 ? "sample_boundary_3d" : "sample_boundary_2d");
    return hipGetLastError();
}

} // namespace gqa
