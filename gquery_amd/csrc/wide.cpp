// Host construction of the engine layout (see wide.h).
//
// Quantization: a child box coordinate is stored as q with decode(q) = fma(q, 2^e, origin), the
// exact operation the kernels perform; q is chosen (and verified with that same decode) so the
// decoded box contains the true child box.  Cones of the exact cull are computed in double from
// the float32 silhouette records and widened by a fixed slack before rounding.
#include "wide.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>

namespace gqa {
namespace {

constexpr double kConeSlack = 1e-5;   // radians added to every exact-cull cone
constexpr double kCullMaxAngle = 1.45; // cones wider than this never cull (cos_t <= 0)

struct BinNode {
    float lo[3], hi[3];
    bool leaf;
    uint32_t off, cnt, second;  // leaf: prim offset/count; interior: right child offset
    uint32_t pbegin, pcount;    // subtree primitive range (contiguous in leaf order)
};

template <int D> std::vector<BinNode> read_bvh(const HostStructure &h) {
    const int ns = 2 * D + 3;
    std::vector<BinNode> b((size_t)h.n_nodes);
    for (size_t i = 0; i < b.size(); ++i) {
        const float *n = h.bvh_node_data.data() + i * ns;
        BinNode &x = b[i];
        for (int k = 0; k < 3; ++k) { x.lo[k] = k < D ? n[k] : 0.0f; x.hi[k] = k < D ? n[D + k] : 0.0f; }
        x.cnt = (uint32_t)n[2 * D + 2];
        x.leaf = x.cnt > 0;
        x.off = (uint32_t)n[2 * D];
        x.second = (uint32_t)n[2 * D + 1];
    }
    for (size_t i = b.size(); i-- > 0;) {
        BinNode &x = b[i];
        if (x.leaf) { x.pbegin = x.off; x.pcount = x.cnt; }
        else {
            const BinNode &l = b[i + 1], &r = b[i + x.second];
            x.pbegin = l.pbegin;
            x.pcount = l.pcount + r.pcount;
        }
    }
    return b;
}

inline float decode(int q, int e, float origin) { return std::fmaf((float)q, std::ldexp(1.0f, e), origin); }

// Per-axis quantization frame of a box: exponent so that decode(levels) >= hi.
inline int frame_exp(float lo, float hi, int levels) {
    double ext = (double)hi - (double)lo;
    if (!(ext > 0)) return -100;
    int e = (int)std::ceil(std::log2(ext / levels));
    e = std::max(e, -100);
    while (decode(levels, e, lo) < hi) ++e;
    while (e > -100 && decode(levels, e - 1, lo) >= hi) --e;
    return e;
}

inline void quantize(float clo, float chi, float origin, int e, int levels, int &qlo, int &qhi) {
    double s = std::ldexp(1.0, e);
    qlo = (int)std::floor(((double)clo - origin) / s);
    qhi = (int)std::ceil(((double)chi - origin) / s);
    qlo = std::clamp(qlo, 0, levels);
    qhi = std::clamp(qhi, 0, levels);
    while (qlo > 0 && decode(qlo, e, origin) > clo) --qlo;
    while (qlo < levels && decode(qlo + 1, e, origin) <= clo) ++qlo;
    while (qhi < levels && decode(qhi, e, origin) < chi) ++qhi;
    while (qhi > qlo && decode(qhi - 1, e, origin) >= chi) --qhi;
}

// Leaf references: counts above 126 go through a side table (count field 127).
struct LeafTable {
    std::vector<uint32_t> big;  // pairs (first, count)
    uint32_t ref(uint32_t first, uint32_t count) {
        if (count <= 126) return make_leaf_ref(first, count);
        uint32_t idx = (uint32_t)(big.size() / 2);
        big.push_back(first);
        big.push_back(count);
        return make_leaf_ref(idx, 127);
    }
};

// Upper halves of float32 patterns (the kernels widen them with a 16-bit shift), rounded so the
// widened value bounds x: dir < 0 below, dir > 0 above, dir == 0 toward zero.
inline double half_value(uint16_t h) { return (double)u2f((uint32_t)h << 16); }
inline uint16_t half_step(uint16_t h, bool up) {
    const bool neg = (h & 0x8000u) != 0;
    const uint16_t mag = h & 0x7FFFu;
    if (up) return neg ? (mag == 0 ? (uint16_t)0x0001u : (uint16_t)(h - 1)) : (uint16_t)(h + 1);
    return neg ? (uint16_t)(h + 1) : (mag == 0 ? (uint16_t)0x8001u : (uint16_t)(h - 1));
}
inline uint16_t half_bound(double x, int dir) {
    const float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    uint16_t h = (uint16_t)(u >> 16);
    if (dir < 0) {
        while (half_value(h) > x) h = half_step(h, false);
    } else if (dir > 0) {
        while (half_value(h) < x) h = half_step(h, true);
    } else {
        while (std::fabs(half_value(h)) > std::fabs(x)) h = (uint16_t)((h & 0x8000u) | ((h & 0x7FFFu) - 1u));
    }
    return h;
}

// Area-weighted normal sum of the reference primitives [pb, pb + pc) (3D cross(b - a, c - a), 2D
// (dy, -dx)).
template <int D> void normal_sum(const HostStructure &h, uint32_t pb, uint32_t pc, double *n) {
    const int ps = kPrimStride[D];
    n[0] = n[1] = n[2] = 0.0;
    for (uint32_t s = pb; s < pb + pc; ++s) {
        const float *r = h.primitive_data.data() + (size_t)s * ps;
        if constexpr (D == 3) {
            double e1[3], e2[3];
            for (int k = 0; k < 3; ++k) { e1[k] = (double)r[3 + k] - r[k]; e2[k] = (double)r[6 + k] - r[k]; }
            n[0] += e1[1] * e2[2] - e1[2] * e2[1];
            n[1] += e1[2] * e2[0] - e1[0] * e2[2];
            n[2] += e1[0] * e2[1] - e1[1] * e2[0];
        } else {
            n[0] += (double)r[3] - r[1];
            n[1] -= (double)r[2] - r[0];
        }
    }
}

// Bounding cylinder of child j (primitives [pb, pb + pc)) about the centre c the kernels compute
// from its box, in both record arrays' sense: the axis halves, then lo / hi / rad over every vertex
// in double with the widened axis, rounded outward.
template <int D>
void child_cylinder(const HostStructure &h, uint32_t pb, uint32_t pc, const uint16_t *ax, const float *c,
                    CBound<D> &cb, int j) {
    const int ps = kPrimStride[D];
    double u[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < D; ++k) {
        cb.ax[k][j] = ax[k];
        u[k] = half_value(ax[k]);
    }
    double lo = std::numeric_limits<double>::infinity(), hi = -lo, rad2 = 0.0;
    for (uint32_t s = pb; s < pb + pc; ++s) {
        const float *r = h.primitive_data.data() + (size_t)s * ps;
        for (int v = 0; v < D; ++v) {
            double dv[3] = {0.0, 0.0, 0.0}, t = 0.0;
            for (int k = 0; k < D; ++k) {
                dv[k] = (double)r[v * D + k] - (double)c[k];
                t += dv[k] * u[k];
            }
            double q2 = 0.0;
            for (int k = 0; k < D; ++k) q2 += (dv[k] - t * u[k]) * (dv[k] - t * u[k]);
            lo = std::min(lo, t);
            hi = std::max(hi, t);
            rad2 = std::max(rad2, q2);
        }
    }
    cb.lo[j] = half_bound(lo, -1);
    cb.hi[j] = half_bound(hi, 1);
    cb.rad[j] = half_bound(std::sqrt(rad2), 1);
}

// Collapses the binary tree into 4-wide nodes with leaves of at most max_leaf primitives.  The
// closest-point tree (ray = false) fills nodes / qnodes / wide_depth and the bounding cylinders;
// the rays' tree fills rnodes / rqnodes / ray_depth.
template <int D>
void build_bvh4(const HostStructure &h, const std::vector<BinNode> &b, int max_leaf, WideHost &out, LeafTable &lt,
                bool ray) {
    int &depth_out = ray ? out.ray_depth : out.wide_depth;
    struct Tmp {
        int n = 0;
        int bin[4];
        uint32_t ref[4];
    };
    std::vector<Tmp> tmp;
    auto area = [&](int i) {
        const BinNode &x = b[i];
        double e[3];
        for (int k = 0; k < 3; ++k) e[k] = k < D ? (double)x.hi[k] - x.lo[k] : 0.0;
        return D == 2 ? e[0] + e[1] : e[0] * e[1] + e[0] * e[2] + e[1] * e[2];
    };
    auto is_leafish = [&](int i) { return b[i].leaf || (int)b[i].pcount <= max_leaf; };
    std::function<uint32_t(int, int)> make = [&](int bi, int depth) -> uint32_t {
        depth_out = std::max(depth_out, depth);
        uint32_t idx = (uint32_t)tmp.size();
        tmp.emplace_back();
        int list[4], n = 0;
        if (is_leafish(bi)) {
            list[n++] = bi;  // only for the root of a tiny tree
        } else {
            list[n++] = bi + 1;
            list[n++] = bi + (int)b[bi].second;
            while (n < 4) {
                int best = -1;
                double ba = -1;
                for (int j = 0; j < n; ++j)
                    if (!is_leafish(list[j]) && area(list[j]) > ba) { ba = area(list[j]); best = j; }
                if (best < 0) break;
                int c = list[best];
                list[best] = c + 1;
                list[n++] = c + (int)b[c].second;
            }
        }
        Tmp t;
        t.n = n;
        for (int j = 0; j < n; ++j) {
            t.bin[j] = list[j];
            if (is_leafish(list[j])) t.ref[j] = lt.ref(b[list[j]].pbegin, b[list[j]].pcount);
            else t.ref[j] = make(list[j], depth + 1);
        }
        tmp[idx] = t;
        return idx;
    };
    make(0, 0);
    const size_t n = tmp.size();
    std::vector<uint8_t> &fnodes = ray ? out.rnodes : out.nodes, &qnodes = ray ? out.rqnodes : out.qnodes;
    (ray ? out.n_ray_nodes : out.n_nodes) = n;
    fnodes.assign(n * sizeof(FNode<D>), 0);
    qnodes.assign(n * sizeof(QNode<D>), 0);
    if (!ray) {
        out.cbq.assign(n * sizeof(CBound<D>), 0);
        out.cbf.assign(n * sizeof(CBound<D>), 0);
    }
    constexpr int L = QBits<D>::levels;
    for (size_t i = 0; i < n; ++i) {
        const Tmp &t = tmp[i];
        FNode<D> fn;
        std::memset(&fn, 0, sizeof fn);
        QNode<D> qn;
        std::memset(&qn, 0, sizeof qn);
        CBound<D> cq, cf;
        std::memset(&cq, 0, sizeof cq);
        std::memset(&cf, 0, sizeof cf);
        fn.nchild = (uint32_t)t.n;
        qn.nchild = (uint8_t)t.n;
        // node frame = union of the children
        float nlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, nhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int j = 0; j < t.n; ++j)
            for (int k = 0; k < D; ++k) {
                nlo[k] = std::min(nlo[k], b[t.bin[j]].lo[k]);
                nhi[k] = std::max(nhi[k], b[t.bin[j]].hi[k]);
            }
        int e[3];
        for (int k = 0; k < D; ++k) {
            e[k] = frame_exp(nlo[k], nhi[k], L);
            qn.origin[k] = nlo[k];
            qn.exp[k] = (int8_t)std::clamp(e[k], -127, 127);
            e[k] = qn.exp[k];
        }
        for (int j = 0; j < 4; ++j) {
            if (j < t.n) {
                const BinNode &c = b[t.bin[j]];
                fn.child[j] = qn.child[j] = t.ref[j];
                float cfc[3] = {0, 0, 0}, cqc[3] = {0, 0, 0};
                for (int k = 0; k < D; ++k) {
                    fn.lo[k][j] = c.lo[k];
                    fn.hi[k][j] = c.hi[k];
                    int ql, qh;
                    quantize(c.lo[k], c.hi[k], nlo[k], e[k], L, ql, qh);
                    qn.qlo[k][j] = (typename QBits<D>::T)ql;
                    qn.qhi[k][j] = (typename QBits<D>::T)qh;
                    // box centres exactly as the kernels compute them (no contraction here):
                    // quantized from the integer corner sum, fma(qlo + qhi, 2^(e - 1), origin)
                    const float sf = c.lo[k] + c.hi[k];
                    cqc[k] = std::fmaf((float)(ql + qh), std::ldexp(1.0f, e[k] - 1), nlo[k]);
                    cfc[k] = 0.5f * sf;
                }
                if (ray) continue;
                double nsum[3];
                normal_sum<D>(h, c.pbegin, c.pcount, nsum);
                const double len = std::sqrt(nsum[0] * nsum[0] + nsum[1] * nsum[1] + nsum[2] * nsum[2]);
                uint16_t ax[3] = {0, 0, 0};
                if (len > 0.0 && std::isfinite(len))
                    for (int k = 0; k < D; ++k) ax[k] = half_bound(nsum[k] / len, 0);
                child_cylinder<D>(h, c.pbegin, c.pcount, ax, cqc, cq, j);
                child_cylinder<D>(h, c.pbegin, c.pcount, ax, cfc, cf, j);
            } else {
                // empty slot: an inverted box that no query overlaps
                for (int k = 0; k < D; ++k) {
                    fn.lo[k][j] = FLT_MAX;
                    fn.hi[k][j] = -FLT_MAX;
                    qn.qlo[k][j] = (typename QBits<D>::T)L;
                    qn.qhi[k][j] = 0;
                }
                fn.child[j] = qn.child[j] = 0;
            }
        }
        std::memcpy(fnodes.data() + i * sizeof(FNode<D>), &fn, sizeof fn);
        std::memcpy(qnodes.data() + i * sizeof(QNode<D>), &qn, sizeof qn);
        if (!ray) {
            std::memcpy(out.cbq.data() + i * sizeof(CBound<D>), &cq, sizeof cq);
            std::memcpy(out.cbf.data() + i * sizeof(CBound<D>), &cf, sizeof cf);
        }
    }
}

// float32 predicate normals exactly as the traversal computes them: normalize(cross(u, v)) with
// the fma cross product and v * (1 / sqrt(dot(v, v))).
inline void fcross(const float *a, const float *b, float *o) {
    o[0] = std::fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = std::fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = std::fmaf(a[0], b[1], -(a[1] * b[0]));
}
inline void fnormalize(float *v) {
    float d = std::fmaf(v[2], v[2], std::fmaf(v[1], v[1], v[0] * v[0]));
    float s = 1.0f / std::sqrt(d);
    for (int k = 0; k < 3; ++k) v[k] = v[k] * s;
}

struct SilNormals {
    double n0[3], n1[3];
    bool ok;      // the predicate can hold for some query point
    bool always;  // boundary edge: a silhouette for every query point
};

// Quantized pair record from the float32 children (boxes, exact-cull cones, child refs).
// `points(j, f)` calls f(x, y, z) for every silhouette point below child j (3D: both edge
// endpoints; 2D: the vertex, z = 0).  Reference-cone flags and ref_base are the caller's.
template <int D, typename Points>
void make_qpair(const float (&lo)[2][D], const float (&hi)[2][D], const ConeF (&cone)[2], const uint32_t (&child)[2],
                Points points, SPairQ<D> &pq) {
    constexpr int L = QBits<D>::levels;
    int e[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        const float flo = std::min(lo[0][k], lo[1][k]), fhi = std::max(hi[0][k], hi[1][k]);
        e[k] = std::clamp(frame_exp(flo, fhi, L), -127, 127);
        pq.origin[k] = flo;
        pq.exp[k] = (int8_t)e[k];
    }
    for (int j = 0; j < 2; ++j) {
        float dlo[3] = {0, 0, 0}, dhi[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            int ql, qh;
            quantize(lo[j][k], hi[j][k], pq.origin[k], e[k], L, ql, qh);
            pq.qlo[j][k] = (typename QBits<D>::T)ql;
            pq.qhi[j][k] = (typename QBits<D>::T)qh;
            dlo[k] = decode(ql, e[k], pq.origin[k]);
            dhi[k] = decode(qh, e[k], pq.origin[k]);
        }
        pq.child[j] = child[j];
        // int16 cone: the axis rounding error is folded into the angle
        const ConeF &cf = cone[j];
        int16_t qa[3] = {0, 0, 0};
        double dq[3] = {0, 0, 0}, fa[3] = {0, 0, 0}, lq2 = 0.0, lf2 = 0.0;
        for (int k = 0; k < D; ++k) {
            qa[k] = (int16_t)std::lround(std::clamp((double)cf.axis[k], -1.0, 1.0) * 32767.0);
            pq.cone_axis[j][k] = qa[k];
            dq[k] = qa[k] / 32767.0;
            fa[k] = cf.axis[k];
            lq2 += dq[k] * dq[k];
            lf2 += fa[k] * fa[k];
        }
        const double lq = std::sqrt(lq2), lf = std::sqrt(lf2);
        // angle between the float axis and the int16 one, as atan2(|a x b|, a . b) (acos of a
        // dot product near 1 would turn the float axis's 1e-7 length error into 3e-4 rad)
        double dot = 0.0, cr2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            dot += fa[k] * dq[k];
            const double ck = fa[(k + 1) % 3] * dq[(k + 2) % 3] - fa[(k + 2) % 3] * dq[(k + 1) % 3];
            cr2 += ck * ck;
        }
        pq.cone_sin[j] = 32767;   // never culls
        if (cf.cos_t > 1.5f) {
            pq.flags |= (uint8_t)(kPairNone0 << j);
        } else if (cf.cos_t > 0.0f && lq > 0 && lf > 0) {
            const double err = std::atan2(std::sqrt(cr2), dot) + 2.0 / 32767.0;
            const double th = std::acos(std::clamp((double)cf.cos_t, -1.0, 1.0)) + err;
            if (th < kCullMaxAngle) pq.cone_sin[j] = (int16_t)std::min(32767.0, std::ceil(std::sin(th) * 32767.0));
        }
        // extent of the points about the decoded box centre (as the kernels form it) along and
        // across the decoded axis, in units of the decoded box's half-diagonal R
        float cen[3] = {0, 0, 0};
        double R2 = 0.0, u[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            cen[k] = 0.5f * (dlo[k] + dhi[k]);
            const double hk = (double)(0.5f * (dhi[k] - dlo[k]));
            R2 += hk * hk;
            u[k] = lq > 0 ? dq[k] / lq : 0.0;
        }
        const double R = std::sqrt(R2);
        double amin = 0.0, amax = 0.0, perp = 0.0;
        bool any = false;
        points(j, [&](double px, double py, double pz) {
            const double d[3] = {px - cen[0], py - cen[1], D == 3 ? pz - cen[2] : 0.0};
            const double al = d[0] * u[0] + d[1] * u[1] + d[2] * u[2];
            double q2 = 0.0;
            for (int k = 0; k < 3; ++k) q2 += (d[k] - al * u[k]) * (d[k] - al * u[k]);
            amin = any ? std::min(amin, al) : al;
            amax = any ? std::max(amax, al) : al;
            perp = std::max(perp, std::sqrt(q2));
            any = true;
        });
        int mid = 0, half = 255, sp = 255;
        if (any && R > 0) {
            mid = std::clamp((int)std::lround(0.5 * (amin + amax) / R * 32767.0), -32767, 32767);
            const double m = mid / 32767.0 * R;
            const double need = std::max(m - amin, amax - m) / R;   // half-width, in units of R
            half = (int)std::ceil(std::sqrt(std::max(need, 0.0)) * 255.0 + 1e-3);
            sp = (int)std::ceil(perp / R * 255.0 + 1e-3);
        } else if (any) {
            half = sp = 0;   // a point box: every point is the centre
        }
        pq.ext_mid[j] = (int16_t)mid;
        pq.ext_half[j] = (uint8_t)std::clamp(half, 0, 255);
        pq.ext_perp[j] = (uint8_t)std::clamp(sp, 0, 255);
    }
}

template <int D>
void build_snch_pairs(const HostStructure &h, const std::vector<BinNode> &b, WideHost &out, LeafTable &lt) {
    const int nss = 3 * D + 7, ss = kSilStride[D];
    const size_t nn = b.size();
    const float *snd = h.snch_node_data.data();
    // silhouette records in leaf order
    const size_t ns = (size_t)h.n_silhouettes;
    std::vector<SilNormals> nrm(ns);
    if (D == 3) {
        out.sils.resize(ns * 3);
        for (size_t s = 0; s < ns; ++s) {
            const float *r = h.silhouette_data.data() + s * ss;
            const float *c = r, *a = r + 3, *bb = r + 6, *e = r + 9;
            float ba[3], ca[3], ea[3], n0[3], n1[3];
            for (int k = 0; k < 3; ++k) { ba[k] = bb[k] - a[k]; ca[k] = c[k] - a[k]; ea[k] = e[k] - a[k]; }
            fcross(ba, ca, n0);
            fnormalize(n0);
            fcross(ea, ba, n1);
            fnormalize(n1);
            // an edge with one adjacent face is a silhouette for every query point
            // (silhouette_edge.py:54): tagged in the n1.x slot, which the kernels test first
            const bool boundary = (int)r[12] == -1 || (int)r[15] == -1;
            out.sils[3 * s + 0] = make_float4(a[0], a[1], a[2], bb[0]);
            out.sils[3 * s + 1] = make_float4(bb[1], bb[2], n0[0], n0[1]);
            out.sils[3 * s + 2] = make_float4(n0[2], boundary ? u2f(kSilAlwaysTag) : n1[0], n1[1], n1[2]);
            // exact-cull cones bound the float32 normals the predicate actually uses; a non-finite
            // normal makes the predicate NaN, i.e. never a silhouette
            SilNormals &sn = nrm[s];
            sn.ok = true;
            sn.always = boundary;
            for (int k = 0; k < 3; ++k) {
                sn.n0[k] = n0[k];
                sn.n1[k] = n1[k];
                sn.ok = sn.ok && std::isfinite(n0[k]) && std::isfinite(n1[k]);
            }
            sn.ok = sn.ok || boundary;
            out.sil_ids.push_back((int32_t)r[18]);
        }
    } else {
        out.sils.resize(ns * 2);
        for (size_t s = 0; s < ns; ++s) {
            const float *r = h.silhouette_data.data() + s * ss;
            // a vertex missing a neighbour is never a silhouette (silhouette_vertex.py:46): its
            // first point becomes NaN, so the kernels' cross-product predicate is false
            const bool open_end = (int)r[6] == -1 || (int)r[8] == -1;
            const float qn = std::numeric_limits<float>::quiet_NaN();
            out.sils[2 * s + 0] = make_float4(open_end ? qn : r[0], open_end ? qn : r[1], r[2], r[3]);
            out.sils[2 * s + 1] = make_float4(r[4], r[5], 0.0f, 0.0f);
            out.sil_ids.push_back((int32_t)r[9]);
            nrm[s].ok = !open_end;
            nrm[s].always = false;
            // the vertex predicate cross(b - a, x - a) * cross(c - b, x - b) < 0 is, with
            // n = perp(segment) = (-dy, dx), dot(n1, x - b) * dot(n2, x - b): the same form as the 3D
            // edge predicate, so the same cones bound it (z = 0)
            float n1[3] = {-(r[3] - r[1]), r[2] - r[0], 0.0f}, n2[3] = {-(r[5] - r[3]), r[4] - r[2], 0.0f};
            fnormalize(n1);
            fnormalize(n2);
            for (int k = 0; k < 3; ++k) {
                nrm[s].n0[k] = n1[k];
                nrm[s].n1[k] = n2[k];
                nrm[s].ok = nrm[s].ok && std::isfinite(n1[k]) && std::isfinite(n2[k]);
            }
        }
    }
    // subtree silhouette ranges (contiguous in leaf order)
    std::vector<uint32_t> sb(nn), sc(nn);
    for (size_t i = nn; i-- > 0;) {
        const float *n = snd + i * nss;
        if (b[i].leaf) { sb[i] = (uint32_t)n[3 * D + 5]; sc[i] = (uint32_t)n[3 * D + 6]; }
        else {
            size_t l = i + 1, r = i + b[i].second;
            sb[i] = sb[l];
            sc[i] = (sb[r] + sc[r]) - sb[l];
        }
    }
    // exact-cull cone of a subtree
    auto cone_of = [&](size_t i, ConeF &cf) {
        cf.axis[0] = 0; cf.axis[1] = 0; cf.axis[2] = 1;
        cf.cos_t = 0.0f; cf.sin_t = 1.0f;  // never culls
        bool any = false, always = false;
        for (uint32_t s = sb[i]; s < sb[i] + sc[i]; ++s) {
            any = any || nrm[s].ok;
            always = always || nrm[s].always;
        }
        if (!any) { cf.cos_t = 2.0f; cf.sin_t = 0.0f; return; }  // no possible silhouette below: always cull
        if (always) return;                                        // a boundary edge below: never cull
        // orient each (n0, n1) pair along a common direction (the predicate is invariant under
        // flipping both), then bound all of them by one cone
        bool have_ref = false;
        double r[3] = {0, 0, 0}, sum[3] = {0, 0, 0};
        for (uint32_t s = sb[i]; s < sb[i] + sc[i]; ++s) {
            const SilNormals &q = nrm[s];
            if (!q.ok) continue;
            double m[3] = {q.n0[0] + q.n1[0], q.n0[1] + q.n1[1], q.n0[2] + q.n1[2]};
            if (!have_ref) { std::memcpy(r, m, sizeof r); have_ref = true; }
            double sg = (m[0] * r[0] + m[1] * r[1] + m[2] * r[2]) >= 0 ? 1.0 : -1.0;
            for (int k = 0; k < 3; ++k) sum[k] += sg * m[k];
        }
        double l = std::sqrt(sum[0] * sum[0] + sum[1] * sum[1] + sum[2] * sum[2]);
        if (!(l > 0)) return;
        double ax[3] = {sum[0] / l, sum[1] / l, sum[2] / l};
        double cmin = 1.0;
        for (uint32_t s = sb[i]; s < sb[i] + sc[i]; ++s) {
            const SilNormals &q = nrm[s];
            if (!q.ok) continue;
            double l0 = std::sqrt(q.n0[0] * q.n0[0] + q.n0[1] * q.n0[1] + q.n0[2] * q.n0[2]);
            double l1 = std::sqrt(q.n1[0] * q.n1[0] + q.n1[1] * q.n1[1] + q.n1[2] * q.n1[2]);
            if (!(l0 > 0) || !(l1 > 0)) return;  // a zero normal: the dot is 0, never provably signed
            double m = (q.n0[0] + q.n1[0]) * ax[0] + (q.n0[1] + q.n1[1]) * ax[1] + (q.n0[2] + q.n1[2]) * ax[2];
            double sg = m >= 0 ? 1.0 : -1.0;
            double c0 = sg * (q.n0[0] * ax[0] + q.n0[1] * ax[1] + q.n0[2] * ax[2]) / l0;
            double c1 = sg * (q.n1[0] * ax[0] + q.n1[1] * ax[1] + q.n1[2] * ax[2]) / l1;
            cmin = std::min(cmin, std::min(c0, c1));
        }
        double th = std::acos(std::clamp(cmin, -1.0, 1.0)) + kConeSlack;
        if (th >= kCullMaxAngle) return;
        for (int k = 0; k < 3; ++k) cf.axis[k] = (float)ax[k];
        // the float axis differs from ax by < 1e-7 rad, far inside kConeSlack
        cf.cos_t = std::nextafter((float)std::cos(th), -1.0f);
        cf.sin_t = std::nextafter((float)std::sin(th), 2.0f);
    };
    auto child_ref = [&](size_t c, const std::vector<uint32_t> &pidx) -> uint32_t {
        if (b[c].leaf) return lt.ref(sb[c], sc[c]);
        return pidx[c];
    };
    // pair index of every interior node, depth-first
    std::vector<uint32_t> pidx(nn, 0xFFFFFFFFu);
    uint32_t np = 0;
    for (size_t i = 0; i < nn; ++i)
        if (!b[i].leaf) pidx[i] = np++;
    out.n_snch_nodes = np;
    out.snodes.assign((size_t)std::max<uint32_t>(np, 1) * sizeof(SPairF<D>), 0);
    const float kHalfPi = 1.57079637050628662109375f;
    if constexpr (D == 2) {
        // a float32 record per pair with the reference boxes and cones inline, and its quantized
        // twin whose reference cones sit in a table
        out.qsnodes.assign((size_t)std::max<uint32_t>(np, 1) * sizeof(SPairQ<2>), 0);
        for (size_t i = 0; i < nn; ++i) {
            if (b[i].leaf) continue;
            const size_t ch[2] = {i + 1, i + b[i].second};
            SPairF<2> pf;
            std::memset(&pf, 0, sizeof pf);
            ConeF cfs[2];
            for (int j = 0; j < 2; ++j) {
                const size_t c = ch[j];
                const float *n = snd + c * nss;   // SNCH node: box 0..3, axis 4..5, half 6, radius 7
                for (int k = 0; k < 2; ++k) {
                    pf.lo[j][k] = n[k];
                    pf.hi[j][k] = n[2 + k];
                    pf.cone_axis[j][k] = n[4 + k];
                }
                pf.cone_half[j] = n[6];
                pf.cone_radius[j] = n[7];
                pf.cone_sin[j] = (float)std::sin((double)n[6]);
                pf.cone_cos[j] = (float)std::cos((double)n[6]);
                pf.child[j] = child_ref(c, pidx);
                // the reference cone test can only cull when 0 <= half <= pi/2 (a wider cone always
                // contains a direction perpendicular to every view direction; a negative one is unset)
                pf.ref_cone[j] = (n[6] >= 0.0f && !(n[6] > kHalfPi)) ? 0 : -1;
                ConeF &cf = cfs[j];
                cone_of(c, cf);
                pf.cull[j] = cf.cos_t > 1.5f ? 1 : 0;
                pf.ecos[j] = cf.cos_t;
                pf.esin[j] = cf.sin_t;
                pf.eaxis[j][0] = cf.axis[0];
                pf.eaxis[j][1] = cf.axis[1];
                // extent of the silhouette vertices below about the centre the kernel forms
                const float cen[2] = {0.5f * (pf.lo[j][0] + pf.hi[j][0]), 0.5f * (pf.lo[j][1] + pf.hi[j][1])};
                const double u[2] = {cf.axis[0], cf.axis[1]};
                double amin = 0.0, amax = 0.0, perp = 0.0;
                bool first = true;
                for (uint32_t s = sb[c]; s < sb[c] + sc[c]; ++s) {
                    const float4 v = out.sils[2 * s];   // (a.x, a.y, b.x, b.y): the vertex is b
                    const double d[2] = {(double)v.z - cen[0], (double)v.w - cen[1]};
                    const double al = d[0] * u[0] + d[1] * u[1];
                    const double q = std::fabs(d[0] * u[1] - d[1] * u[0]);
                    amin = first ? al : std::min(amin, al);
                    amax = first ? al : std::max(amax, al);
                    perp = std::max(perp, q);
                    first = false;
                }
                constexpr float kInfF = std::numeric_limits<float>::infinity();
                pf.elo[j] = std::nextafter((float)amin, -kInfF);
                pf.ehi[j] = std::nextafter((float)amax, kInfF);
                pf.eperp[j] = std::nextafter((float)perp, kInfF);
            }
            std::memcpy(out.snodes.data() + (size_t)pidx[i] * sizeof(SPairF<2>), &pf, sizeof pf);
            SPairQ<2> pq;
            std::memset(&pq, 0, sizeof pq);
            make_qpair<2>(pf.lo, pf.hi, cfs, pf.child, [&](int j, auto f) {
                const size_t c = ch[j];
                for (uint32_t t = sb[c]; t < sb[c] + sc[c]; ++t) f(out.sils[2 * t].z, out.sils[2 * t].w, 0.0);
            }, pq);
            pq.ref_base = (int32_t)out.ref_cones2.size();
            for (int j = 0; j < 2; ++j) {
                if (pf.ref_cone[j] < 0) continue;
                RefCone2 rc;
                std::memset(&rc, 0, sizeof rc);
                for (int k = 0; k < 2; ++k) {
                    rc.lo[k] = pf.lo[j][k];
                    rc.hi[k] = pf.hi[j][k];
                    rc.axis[k] = pf.cone_axis[j][k];
                }
                rc.half = pf.cone_half[j];
                rc.radius = pf.cone_radius[j];
                rc.sin_h = pf.cone_sin[j];
                rc.cos_h = pf.cone_cos[j];
                pq.flags |= (uint8_t)(kPairRef0 << j);
                out.ref_cones2.push_back(rc);
            }
            std::memcpy(out.qsnodes.data() + (size_t)pidx[i] * sizeof(SPairQ<2>), &pq, sizeof pq);
        }
    } else {
    out.qsnodes.assign((size_t)std::max<uint32_t>(np, 1) * sizeof(SPairQ<D>), 0);
    for (size_t i = 0; i < nn; ++i) {
        if (b[i].leaf) continue;
        const size_t ch[2] = {i + 1, i + b[i].second};
        SPairF<D> pf;
        std::memset(&pf, 0, sizeof pf);
        SPairQ<D> pq;
        std::memset(&pq, 0, sizeof pq);
        pq.ref_base = (int32_t)out.ref_cones.size();
        for (int j = 0; j < 2; ++j) {
            const size_t c = ch[j];
            const BinNode &x = b[c];
            for (int k = 0; k < D; ++k) {
                pf.lo[j][k] = x.lo[k];
                pf.hi[j][k] = x.hi[k];
            }
            pf.child[j] = child_ref(c, pidx);
            // the reference cone test can only cull when 0 <= half <= pi/2 (a wider cone always
            // contains a direction perpendicular to every view direction; a negative one is unset)
            const float *n = snd + c * nss;
            const float half = n[3 * D];
            if (half >= 0.0f && !(half > kHalfPi)) {
                RefCone rc;
                std::memset(&rc, 0, sizeof rc);
                for (int k = 0; k < D; ++k) { rc.axis[k] = n[2 * D + k]; rc.lo[k] = n[k]; rc.hi[k] = n[D + k]; }
                rc.half = half;
                rc.radius = n[3 * D + 1];
                rc.sin_h = (float)std::sin((double)half);
                rc.cos_h = (float)std::cos((double)half);
                pf.ref_cone[j] = (int32_t)out.ref_cones.size();
                pq.flags |= (uint8_t)(kPairRef0 << j);
                out.ref_cones.push_back(rc);
            } else {
                pf.ref_cone[j] = -1;
            }
            ConeF cf;
            cone_of(c, cf);
            pf.cone[j] = cf;
            {
                // extent of the silhouette edges below along and across the cone axis, about the
                // centre the kernel forms (float32, 0.5f * (lo + hi)): the exact cull bounds every
                // edge point by it (a slab across a nearly flat patch is far thinner than the sphere
                // around it)
                float cen[3];
                for (int k = 0; k < 3; ++k) cen[k] = 0.5f * (pf.lo[j][k] + pf.hi[j][k]);
                const double u[3] = {cf.axis[0], cf.axis[1], cf.axis[2]};
                double amin = 0.0, amax = 0.0, perp2 = 0.0;
                bool first = true;
                for (uint32_t s = sb[c]; s < sb[c] + sc[c]; ++s) {
                    const float4 s0 = out.sils[3 * s], s1 = out.sils[3 * s + 1];
                    const double pts[2][3] = {{s0.x, s0.y, s0.z}, {s0.w, s1.x, s1.y}};
                    for (const auto &pt : pts) {
                        double d[3], al = 0.0;
                        for (int k = 0; k < 3; ++k) { d[k] = pt[k] - cen[k]; al += d[k] * u[k]; }
                        double q2 = 0.0;
                        for (int k = 0; k < 3; ++k) q2 += (d[k] - al * u[k]) * (d[k] - al * u[k]);
                        amin = first ? al : std::min(amin, al);
                        amax = first ? al : std::max(amax, al);
                        perp2 = std::max(perp2, q2);
                        first = false;
                    }
                }
                constexpr float kInfF = std::numeric_limits<float>::infinity();
                pf.sil_lo[j] = std::nextafter((float)amin, -kInfF);
                pf.sil_hi[j] = std::nextafter((float)amax, kInfF);
                pf.sil_perp[j] = std::nextafter((float)std::sqrt(perp2), kInfF);
            }
        }
        make_qpair<D>(pf.lo, pf.hi, pf.cone, pf.child, [&](int j, auto f) {
            const size_t c = ch[j];
            for (uint32_t t = sb[c]; t < sb[c] + sc[c]; ++t) {
                const float4 s0 = out.sils[3 * t], s1 = out.sils[3 * t + 1];
                f(s0.x, s0.y, s0.z);
                f(s0.w, s1.x, s1.y);
            }
        }, pq);
        std::memcpy(out.snodes.data() + (size_t)pidx[i] * sizeof(SPairF<D>), &pf, sizeof pf);
        std::memcpy(out.qsnodes.data() + (size_t)pidx[i] * sizeof(SPairQ<D>), &pq, sizeof pq);
    }
    }
    out.sroot_ref = b[0].leaf ? lt.ref(sb[0], sc[0]) : 0u;
    for (int k = 0; k < D; ++k) { out.sroot_lo[k] = b[0].lo[k]; out.sroot_hi[k] = b[0].hi[k]; }
}

// Seed cut of a tree of wide records: breadth first from the root's children, each round opens
// every interior entry (in list order) while the cut stays within kSeedMax entries.  `children`
// gives a record's (ref, box) children and whether the cut may pass below that record.
template <int D, typename Children>
std::vector<Seed> seed_cut(uint32_t root, Children children) {
    std::vector<Seed> cut, next;
    if (ref_is_leaf(root)) return cut;
    std::vector<Seed> kids;
    if (!children(root, kids)) return cut;
    cut = kids;
    for (bool opened = true; opened;) {
        opened = false;
        next.clear();
        for (size_t k = 0; k < cut.size(); ++k) {
            const Seed &e = cut[k];
            kids.clear();
            if (!ref_is_leaf(e.ref) && children(e.ref, kids) &&
                next.size() + kids.size() + (cut.size() - k - 1) <= (size_t)kSeedMax) {
                next.insert(next.end(), kids.begin(), kids.end());
                opened = true;
            } else {
                next.push_back(e);
            }
        }
        cut.swap(next);
    }
    return cut;
}

inline Seed make_seed(uint32_t ref, const float *lo, const float *hi, int d) {
    Seed e;
    std::memset(&e, 0, sizeof e);
    e.ref = ref;
    for (int k = 0; k < d; ++k) { e.lo[k] = lo[k]; e.hi[k] = hi[k]; }
    return e;
}

template <int D> void build_seeds(WideHost &out) {
    out.bvh_seed = seed_cut<D>(out.root_ref, [&](uint32_t ref, std::vector<Seed> &kids) {
        FNode<D> nd;
        std::memcpy(&nd, out.nodes.data() + (size_t)ref * sizeof nd, sizeof nd);
        for (uint32_t j = 0; j < nd.nchild; ++j) {
            float lo[3], hi[3];
            for (int k = 0; k < D; ++k) { lo[k] = nd.lo[k][j]; hi[k] = nd.hi[k][j]; }
            kids.push_back(make_seed(nd.child[j], lo, hi, D));
        }
        return true;
    });
    if (out.n_snch_nodes == 0) return;
    // below a child whose reference cone test can cull, every query must still take that test,
    // so the cut stops at such a pair (the box tests and exact culls it skips are conservative)
    out.sil_seed = seed_cut<D>(out.sroot_ref, [&](uint32_t ref, std::vector<Seed> &kids) {
        SPairF<D> pf;
        std::memcpy(&pf, out.snodes.data() + (size_t)ref * sizeof pf, sizeof pf);
        if (pf.ref_cone[0] >= 0 || pf.ref_cone[1] >= 0) return false;
        for (int j = 0; j < 2; ++j) kids.push_back(make_seed(pf.child[j], pf.lo[j], pf.hi[j], D));
        return true;
    });
}

template <int D> void build_impl(const HostStructure &h, WideHost &out, int max_leaf, int ray_leaf) {
    std::vector<BinNode> b = read_bvh<D>(h);
    LeafTable lt;
    build_bvh4<D>(h, b, max_leaf, out, lt, false);
    build_bvh4<D>(h, b, ray_leaf, out, lt, true);
    out.root_ref = 0;
    for (int k = 0; k < D; ++k) { out.scene_lo[k] = b[0].lo[k]; out.scene_hi[k] = b[0].hi[k]; }
    // primitives in leaf order
    const int ps = kPrimStride[D];
    out.prims.resize((size_t)h.n_prims * (D == 3 ? 4 : 1));
    out.prim_ids.resize((size_t)h.n_prims);
    for (int64_t s = 0; s < h.n_prims; ++s) {
        const float *r = h.primitive_data.data() + s * ps;
        if (D == 3) {
            // circumscribing disk for the closest-point prefilter: centre m (the centroid in float32,
            // in the free lanes of a, b, c), unit normal n (components rounded toward zero, so
            // |n| <= 1) and the vertices' largest distance from the line m + t n, rounded up.  With
            // h = (x - m).n and q = (x - m) - h n, sqrt(h^2 + max(|q| - rad, 0)^2) is a lower bound
            // on the distance from x to the triangle, up to the triangle's float32 thickness along
            // n, a few ulps of the coordinates; a degenerate triangle gets n = 0 and a sphere.
            float m[3];
            for (int k = 0; k < 3; ++k) m[k] = (float)(((double)r[k] + r[3 + k] + r[6 + k]) / 3.0);
            double e1[3], e2[3], n[3];
            for (int k = 0; k < 3; ++k) { e1[k] = (double)r[3 + k] - r[k]; e2[k] = (double)r[6 + k] - r[k]; }
            n[0] = e1[1] * e2[2] - e1[2] * e2[1];
            n[1] = e1[2] * e2[0] - e1[0] * e2[2];
            n[2] = e1[0] * e2[1] - e1[1] * e2[0];
            const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            float nf[3] = {0.0f, 0.0f, 0.0f};
            if (l > 0 && std::isfinite(l)) {
                for (int k = 0; k < 3; ++k) {
                    const double v = n[k] / l;
                    nf[k] = (float)v;
                    if (std::fabs((double)nf[k]) > std::fabs(v)) nf[k] = std::nextafter(nf[k], 0.0f);
                }
            }
            double rad2 = 0.0;
            for (int v = 0; v < 3; ++v) {
                double dv[3], t = 0.0;
                for (int k = 0; k < 3; ++k) { dv[k] = (double)r[3 * v + k] - m[k]; t += dv[k] * nf[k]; }
                double q2 = 0.0;
                for (int k = 0; k < 3; ++k) q2 += (dv[k] - t * nf[k]) * (dv[k] - t * nf[k]);
                rad2 = std::max(rad2, q2);
            }
            float rad = (float)std::sqrt(rad2);
            if ((double)rad < std::sqrt(rad2)) rad = std::nextafter(rad, std::numeric_limits<float>::infinity());
            out.prims[4 * s + 0] = make_float4(r[0], r[1], r[2], m[0]);
            out.prims[4 * s + 1] = make_float4(r[3], r[4], r[5], m[1]);
            out.prims[4 * s + 2] = make_float4(r[6], r[7], r[8], m[2]);
            out.prims[4 * s + 3] = make_float4(nf[0], nf[1], nf[2], rad);
        } else {
            out.prims[s] = make_float4(r[0], r[1], r[2], r[3]);
        }
        out.prim_ids[s] = (int32_t)r[D * D];
    }
    if (h.is_snch) build_snch_pairs<D>(h, b, out, lt);
    out.big_leaves = lt.big;
    build_seeds<D>(out);
}

} // namespace

void build_wide(const HostStructure &h, WideHost &out, int max_leaf_prims, int ray_leaf_prims) {
    out = WideHost();
    max_leaf_prims = std::clamp(max_leaf_prims, 1, kMaxLeafOption);
    ray_leaf_prims = std::clamp(ray_leaf_prims, 1, kMaxLeafOption);
    if (h.dim == 2) build_impl<2>(h, out, max_leaf_prims, ray_leaf_prims);
    else build_impl<3>(h, out, max_leaf_prims, ray_leaf_prims);
}

} // namespace gqa
