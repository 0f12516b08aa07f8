// Matrix-free J^T J apply (applyJtJDotWithIndexKernel,
// ApplyJtJMatrixFreeIndexed.cu:29-349) and the block-Jacobi PCG around it.
//
// y = sum_t J_t^T (J_t v) + damping v without the J^T J blocks, in the
// node-centred coordinates of the rest of the solve (rows J'_n = w_n J T_n):
//   terms  : one thread per term, u_t = lambda^2 J_t (sum_k w_k T_k v_k) over
//            the term's knn slots (graph edges: both ends' own rows) -- the
//            Jacobians are recomputed from the term records every apply;
//   gather : a row group per node over its Node2Term list,
//            y_i = sum_(i,t) J'_i(t)^T u_t + 1e-2 T_i^T T_i v_i.
// A node named by m > 1 knn slots of one term (the reference's pair
// multiplicity) appears m times in its Node2Term list; each entry subtracts
// (m - 1) lambda^2 J'^T J' v_i, so the apply equals the materialized J^T J of
// the index-driven assembly exactly, damping and cross-view term included
// (the reference's matrix-free path leaves both out).
#include "common.h"
#include "device_util.h"
#include "eval.h"
#include "pcg_util.h"
#include "terms.h"

namespace mvdr {

enum : int { kSrcDir = 0, kSrcX = 1, kSrcZ = 2 };        // the vector v applied to
enum : int { kEpiSpmv = 0, kEpiResidual = 1, kEpiCurvature = 2 };

// Node-centred twist -> reference twist: (w, v + g x w).
__device__ inline void node_to_ref(const float4 g, const float x[6], float y[6]) {
    y[0] = x[0];
    y[1] = x[1];
    y[2] = x[2];
    y[3] = x[3] + (g.y * x[2] - g.z * x[1]);
    y[4] = x[4] + (g.z * x[0] - g.x * x[2]);
    y[5] = x[5] + (g.x * x[1] - g.y * x[0]);
}

__device__ inline float dot6(const float a[6], const float b[6]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
}

struct MfLambda2 {
    float smooth, density, foreground, feature, cross;
};

__device__ inline MfLambda2 mf_lambda2(const Penalty& p) {
    return MfLambda2{p.smooth * p.smooth, p.density * p.density, p.foreground * p.foreground, p.feature * p.feature,
                     p.cross * p.cross};
}

// Term pass: u_t for every term.  kSrcDir forms the PCG direction
// d = s + beta d_prev (beta from the previous launches' partials) on the fly
// and stores it for the gather and the update.
template <typename Acc, int kSrc>
__global__ void __launch_bounds__(256) mf_terms_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t cur, Penalty pen) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    Acc beta = 0;
    if (kSrc == kSrcDir && k > 0) {
        Acc num, den;
        reduce_partials2<Acc>(slot_ptr(b, kSlotDelta + k), kVecBlocks, slot_ptr(b, kSlotDelta + k - 1), kVecBlocks, num, den);
        beta = den != (Acc)0 ? num / den : (Acc)0;
    }
    if (c.error) return;
    const float* __restrict__ prev = cur ? b.d1 : b.d0;
    float* __restrict__ dout = cur ? b.d0 : b.d1;
    const float* __restrict__ vec = kSrc == kSrcX ? b.x : b.d1;
    auto source = [&](uint32_t n, float v[6]) {
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            if (kSrc == kSrcDir) {
                const float p = prev[6 * n + a];
                v[a] = k > 0 ? (float)((Acc)b.s[6 * n + a] + beta * (Acc)p) : p;
            } else {
                v[a] = vec[6 * n + a];
            }
        }
    };
    const TermBounds tb(c);
    const MfLambda2 l2 = mf_lambda2(pen);
    const bool density = b.density && pen.density != 0.f;
    const uint32_t total = c.off[4];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        uint32_t kk;
        const int type = type_of(tb, t, kk);
        float u[3] = {0.f, 0.f, 0.f};
        if (type == kSmooth) {
            const SmoothRec sm = b.smooth[kk];
            const uint32_t n1 = sm.ij & 0xFFFF, n2 = sm.ij >> 16;
            float J1[3][6], J2[3][6], res[3], v[6], w1[6], w2[6];
            smooth_rows(sm, true, J1, res);
            smooth_rows(sm, n2 == n1, J2, res);
            source(n1, v);
            node_to_ref(f.nodes[n1], v, w1);
            source(n2, v);
            node_to_ref(f.nodes[n2], v, w2);
#pragma unroll
            for (int r = 0; r < 3; ++r) u[r] = l2.smooth * (dot6(J1[r], w1) + dot6(J2[r], w2));
        } else {
            ushort4 knn;
            float4 wk;
            float J[3][6], res[3];
            int rows = 3;
            float lam[3];
            if (type == kDense || type == kForeground) {
                const ScalarRec& rec = type == kDense ? b.dense[kk] : b.fg[kk];
                knn = rec.knn;
                wk = rec.w;
#pragma unroll
                for (int q = 0; q < 6; ++q) J[0][q] = rec.J[q];
                lam[0] = type == kDense ? 1.f : l2.foreground;
                rows = 1;
                if (type == kDense && density) {
                    const ScalarRec& d = b.density[kk];
#pragma unroll
                    for (int q = 0; q < 6; ++q) J[1][q] = d.J[q];
                    lam[1] = l2.density;
                    rows = 2;
                }
            } else {
                const P2PRec& rec = type == kFeature ? b.feat[kk] : b.cross[kk];
                knn = rec.knn;
                wk = rec.w;
                p2p_rows(rec, J, res);
                const uint32_t mask = type == kFeature ? 0xFFFFFFFFu : 0u;
                lam[0] = lam[1] = lam[2] =
                    __uint_as_float((__float_as_uint(l2.feature) & mask) | (__float_as_uint(l2.cross) & ~mask));
            }
            // sum_k w_k T_k v_k over the knn slots (slot-rule weights)
            float V[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            const uint32_t n[4] = {knn.x, knn.y, knn.z, knn.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float v[6], w6[6];
                source(n[q], v);
                node_to_ref(f.nodes[n[q]], v, w6);
                const float ws = slot_weight(knn, wk, n[q]);
#pragma unroll
                for (int a = 0; a < 6; ++a) V[a] += ws * w6[a];
            }
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (r < rows) u[r] = lam[r] * dot6(J[r], V);
        }
        b.mf_u[t] = make_float4(u[0], u[1], u[2], 0.f);
    }
    if (kSrc == kSrcDir) {
        for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
            float v[6];
            source(n, v);
#pragma unroll
            for (int a = 0; a < 6; ++a) dout[6 * n + a] = v[a];
        }
    }
}

// Gather pass: y_i over the Node2Term list of row i (kRowLanes lanes per row),
// then the epilogue of the PCG step it serves.
template <typename Acc, int kEpi>
__global__ void __launch_bounds__(256) mf_gather_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t cur, Penalty pen) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    const uint32_t sub = threadIdx.x % kRowLanes;
    const float* __restrict__ vec = kEpi == kEpiSpmv ? (cur ? b.d0 : b.d1) : (kEpi == kEpiResidual ? b.x : b.d1);
    Acc acc = 0;
    if (!c.error) {
        const TermBounds tb(c);
        const MfLambda2 l2 = mf_lambda2(pen);
        const bool density = b.density && pen.density != 0.f;
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            Acc y[6] = {0, 0, 0, 0, 0, 0};
            float vi[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (i < f.N) {
                const float4 g = f.nodes[i];
#pragma unroll
                for (int a = 0; a < 6; ++a) vi[a] = vec[6 * i + a];
                const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
                for (uint32_t e = beg + sub; e < end; e += kRowLanes) {
                    const uint32_t t = b.n2t_index[e];
                    uint32_t kk;
                    const int type = type_of(tb, t, kk);
                    float J[3][6], res[3], lam[3];
                    int rows = 3;
                    float wi = 1.f;
                    uint32_t mult = 1;
                    if (type == kSmooth) {
                        const SmoothRec sm = b.smooth[kk];
                        const uint32_t n1 = sm.ij & 0xFFFF, n2 = sm.ij >> 16;
                        smooth_rows(sm, i == n1, J, res);
                        lam[0] = lam[1] = lam[2] = l2.smooth;
                        mult = (n1 == i ? 1u : 0u) + (n2 == i ? 1u : 0u);
                    } else {
                        ushort4 knn;
                        float4 wk;
                        if (type == kDense || type == kForeground) {
                            const ScalarRec& rec = type == kDense ? b.dense[kk] : b.fg[kk];
                            knn = rec.knn;
                            wk = rec.w;
#pragma unroll
                            for (int q = 0; q < 6; ++q) J[0][q] = rec.J[q];
                            lam[0] = type == kDense ? 1.f : l2.foreground;
                            rows = 1;
                            if (type == kDense && density) {
                                const ScalarRec& d = b.density[kk];
#pragma unroll
                                for (int q = 0; q < 6; ++q) J[1][q] = d.J[q];
                                lam[1] = l2.density;
                                rows = 2;
                            }
                        } else {
                            const P2PRec& rec = type == kFeature ? b.feat[kk] : b.cross[kk];
                            knn = rec.knn;
                            wk = rec.w;
                            p2p_rows(rec, J, res);
                            const uint32_t mask = type == kFeature ? 0xFFFFFFFFu : 0u;
                            lam[0] = lam[1] = lam[2] = __uint_as_float((__float_as_uint(l2.feature) & mask) |
                                                                       (__float_as_uint(l2.cross) & ~mask));
                        }
                        wi = slot_weight(knn, wk, i);
                        mult = (knn.x == i) + (knn.y == i) + (knn.z == i) + (knn.w == i);
                    }
                    const float4 ut = b.mf_u[t];
                    const float u[3] = {ut.x, ut.y, ut.z};
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        if (r < rows) {
                            to_node_frame(J[r], g);
                            float s = wi * u[r];
                            if (mult > 1) s -= (float)(mult - 1) * lam[r] * wi * wi * dot6(J[r], vi);
#pragma unroll
                            for (int a = 0; a < 6; ++a) y[a] += (Acc)(s * J[r][a]);
                        }
                    }
                }
                if (sub == 0) {
                    // damping 1e-2 T^T T v: T v = (w, v + g x w), T^T (a, b) = (a + b x g, b)
                    float tv[6];
                    node_to_ref(g, vi, tv);
                    const float bx = tv[4] * g.z - tv[5] * g.y, by = tv[5] * g.x - tv[3] * g.z,
                                bz = tv[3] * g.y - tv[4] * g.x;
                    const float damp[6] = {tv[0] + bx, tv[1] + by, tv[2] + bz, tv[3], tv[4], tv[5]};
#pragma unroll
                    for (int a = 0; a < 6; ++a) y[a] += (Acc)(1e-2f * damp[a]);
                }
            }
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int o = kRowLanes / 2; o > 0; o >>= 1) y[a] += __shfl_xor(y[a], o, 64);
            if (i < f.N && sub == 0) {
                if (kEpi == kEpiSpmv) {
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        b.q[6 * i + a] = (float)y[a];
                        acc += (Acc)vi[a] * y[a];
                    }
                } else if (kEpi == kEpiResidual) {
                    float rr[6], z[6];
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        rr[a] = (float)((Acc)b.jtr[6 * i + a] - y[a]);
                        b.r[6 * i + a] = rr[a];
                    }
                    apply_minv<Acc>(b, i, rr, z);
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        b.d1[6 * i + a] = z[a];
                        acc += (Acc)rr[a] * (Acc)z[a];
                    }
                } else {
#pragma unroll
                    for (int a = 0; a < 6; ++a) acc += (Acc)vi[a] * y[a];
                }
            }
        }
    }
    write_partial<Acc>(b, kEpi == kEpiSpmv ? kSlotDQ + k : (kEpi == kEpiResidual ? kSlotRefNum : kSlotRefDen), acc);
}

// ------------------------------------------ matrix-free apply, J recomputed
// MVDR_JTJ_RECOMPUTE: the operator above with every Jacobian re-evaluated
// inside each apply from the solver / observation maps and the current warp
// se3[which] (eval.h), the way the evaluation pass computes it; from the term
// records only the index's knn and weights and the feature targets are read.
//   terms  : one thread per term: J_t, u_t = J_t V_t with V_t the knn blend
//            of the source vector, and g_t = sum_r lambda_r^2 u_tr J_tr
//            (6 floats; an edge term keeps lambda^2 u_t, 3 floats);
//   gather : y_i = T_i^T sum_(i,t) w_i(t) g_t + damping, an edge term's rows
//            re-evaluated from its two node transforms, and a knn term's J
//            re-evaluated when the term names node i in m > 1 slots (the
//            multiplicity correction of the stored-J apply).
// Debug forms (unit term weights): kSrcRows takes u_t from mf_u, kOutRows
// writes u_t = J_t V_t to mf_u (J x), and kEpiAdjoint writes the gather's sum
// without damping or correction to q (J^T y, the exact adjoint of J x: the
// slot weight of node i depends on i alone, so each of its m entries carries
// it once).
enum : int { kSrcRows = 3 };
enum : int { kOutG = 0, kOutRows = 1 };
enum : int { kEpiAdjoint = 3 };

// Knn and weights of a knn term (fixed by the index build).
__device__ inline void term_knn(const Buffers& b, int type, uint32_t kk, ushort4& knn, float4& w) {
    if (type == kDense || type == kForeground) {
        const ScalarRec& r = type == kDense ? b.dense[kk] : b.fg[kk];
        knn = r.knn;
        w = r.w;
    } else {
        const P2PRec& r = type == kFeature ? b.feat[kk] : b.cross[kk];
        knn = r.knn;
        w = r.w;
    }
}

// J rows and lambda^2 of a knn term (dense depth + density, foreground,
// feature, cross view) evaluated at the warp se3.
template <typename R>
__device__ inline int knn_term_rows(const Buffers& b, const FrameDev& f, const DQ* se3, int type, uint32_t kk,
                                    const MfLambda2& l2, bool density, float J[3][6], float lam[3]) {
    if (type == kDense || type == kForeground) {
        ScalarRec rec;
        term_knn(b, type, kk, rec.knn, rec.w);
        const ushort4 px = type == kDense ? b.dense_px[kk] : b.fg_px[kk];
        if (type == kDense) eval_dense<R>(f, se3, px, rec);
        else eval_image<R>(f, se3, px, false, rec);
#pragma unroll
        for (int q = 0; q < 6; ++q) J[0][q] = rec.J[q];
        lam[0] = type == kDense ? 1.f : l2.foreground;
        if (type == kDense && density) {
            eval_image<R>(f, se3, px, true, rec);
#pragma unroll
            for (int q = 0; q < 6; ++q) J[1][q] = rec.J[q];
            lam[1] = l2.density;
            return 2;
        }
        return 1;
    }
    const bool cross = type == kCross;
    const P2PRec& src = cross ? b.cross[kk] : b.feat[kk];
    P2PRec rec;
    rec.knn = src.knn;
    rec.w = src.w;
    rec.target = src.target;
    rec.warped = eval_warped<R>(se3, rec.knn, rec.w, cross ? b.cross_canon[kk] : b.feat_canon[kk]);
    float res[3];
    p2p_rows(rec, J, res);
    lam[0] = lam[1] = lam[2] = cross ? l2.cross : l2.feature;
    return 3;
}

// The weight of each Node2Term entry's node, fixed by the index build: the
// slot weight of node i in the term's knn, its sign bit set when the term
// names i in more than one slot (the rare entries whose apply needs the
// multiplicity correction).  Graph-edge entries are not read through it.
__global__ void __launch_bounds__(256) mf_entry_weights_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const TermBounds tb(c);
    const uint32_t sub = threadIdx.x % kRowLanes;
    const uint32_t groups = gridDim.x * kRowsPerBlock;
    for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
        const uint32_t i = base + threadIdx.x / kRowLanes;
        if (i >= f.N) continue;
        const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
        for (uint32_t e = beg + sub; e < end; e += kRowLanes) {
            uint32_t kk;
            const int type = type_of(tb, b.n2t_index[e], kk);
            if (type == kSmooth) continue;
            ushort4 knn;
            float4 wk;
            term_knn(b, type, kk, knn, wk);
            const uint32_t mult = (knn.x == i) + (knn.y == i) + (knn.z == i) + (knn.w == i);
            const float wi = slot_weight(knn, wk, i);
            b.n2t_w[e] = mult > 1 ? copysignf(fabsf(wi), -1.0f) : fabsf(wi);
        }
    }
}

template <typename Acc, typename R, int kSrc, int kOut>
__global__ void __launch_bounds__(256) mfr_terms_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t cur, Penalty pen,
                                                        uint32_t which) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    const DQ* __restrict__ se3 = which ? b.se3[1] : b.se3[0];
    Acc beta = 0;
    if (kSrc == kSrcDir && k > 0) {
        Acc num, den;
        reduce_partials2<Acc>(slot_ptr(b, kSlotDelta + k), kVecBlocks, slot_ptr(b, kSlotDelta + k - 1), kVecBlocks, num, den);
        beta = den != (Acc)0 ? num / den : (Acc)0;
    }
    if (c.error) return;
    const float* __restrict__ prev = cur ? b.d1 : b.d0;
    float* __restrict__ dout = cur ? b.d0 : b.d1;
    const float* __restrict__ vec = kSrc == kSrcX ? b.x : b.d1;
    auto source = [&](uint32_t n, float v[6]) {
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            if (kSrc == kSrcDir) {
                const float p = prev[6 * n + a];
                v[a] = k > 0 ? (float)((Acc)b.s[6 * n + a] + beta * (Acc)p) : p;
            } else {
                v[a] = vec[6 * n + a];
            }
        }
    };
    const TermBounds tb(c);
    const MfLambda2 l2 = mf_lambda2(pen);
    const bool density = b.density && pen.density != 0.f;
    const uint32_t total = c.off[4];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        uint32_t kk;
        const int type = type_of(tb, t, kk);
        float u[3] = {0.f, 0.f, 0.f};
        float g[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (kSrc == kSrcRows) {
            const float4 ut = b.mf_u[t];
            u[0] = ut.x;
            u[1] = ut.y;
            u[2] = ut.z;
        }
        if (type == kSmooth) {
            if (kSrc != kSrcRows) {
                const SmoothRec sm = eval_smooth<R>(f, se3, kk);
                const uint32_t n1 = sm.ij & 0xFFFF, n2 = sm.ij >> 16;
                float J1[3][6], J2[3][6], res[3], v[6], w1[6], w2[6];
                smooth_rows(sm, true, J1, res);
                smooth_rows(sm, n2 == n1, J2, res);
                source(n1, v);
                node_to_ref(f.nodes[n1], v, w1);
                source(n2, v);
                node_to_ref(f.nodes[n2], v, w2);
#pragma unroll
                for (int r = 0; r < 3; ++r) u[r] = dot6(J1[r], w1) + dot6(J2[r], w2);
            }
            const float l = kSrc == kSrcRows ? 1.f : l2.smooth;
#pragma unroll
            for (int r = 0; r < 3; ++r) g[r] = l * u[r];
        } else {
            float J[3][6], lam[3];
            const int rows = knn_term_rows<R>(b, f, se3, type, kk, l2, density, J, lam);
            if (kSrc != kSrcRows) {
                ushort4 knn;
                float4 wk;
                term_knn(b, type, kk, knn, wk);
                float V[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                const uint32_t n[4] = {knn.x, knn.y, knn.z, knn.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float v[6], w6[6];
                    source(n[q], v);
                    node_to_ref(f.nodes[n[q]], v, w6);
                    const float ws = slot_weight(knn, wk, n[q]);
#pragma unroll
                    for (int a = 0; a < 6; ++a) V[a] += ws * w6[a];
                }
#pragma unroll
                for (int r = 0; r < 3; ++r) u[r] = r < rows ? dot6(J[r], V) : 0.f;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (r < rows) {
                    const float s = (kSrc == kSrcRows ? 1.f : lam[r]) * u[r];
#pragma unroll
                    for (int a = 0; a < 6; ++a) g[a] += s * J[r][a];
                }
            }
        }
        if (kOut == kOutRows) {
            b.mf_u[t] = make_float4(u[0], u[1], u[2], 0.f);
        } else {
            b.mf_g[2 * (size_t)t] = make_float4(g[0], g[1], g[2], g[3]);
            b.mf_g[2 * (size_t)t + 1] = make_float4(g[4], g[5], 0.f, 0.f);
        }
    }
    if (kSrc == kSrcDir) {
        for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
            float v[6];
            source(n, v);
#pragma unroll
            for (int a = 0; a < 6; ++a) dout[6 * n + a] = v[a];
        }
    }
}

template <typename Acc, typename R, int kEpi>
__global__ void __launch_bounds__(256) mfr_gather_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t cur, Penalty pen,
                                                         uint32_t which) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    const DQ* __restrict__ se3 = which ? b.se3[1] : b.se3[0];
    const uint32_t sub = threadIdx.x % kRowLanes;
    const float* __restrict__ vec = kEpi == kEpiSpmv ? (cur ? b.d0 : b.d1) : (kEpi == kEpiResidual ? b.x : b.d1);
    constexpr bool correct = kEpi != kEpiAdjoint;
    Acc acc = 0;
    if (!c.error) {
        const TermBounds tb(c);
        const MfLambda2 l2 = mf_lambda2(pen);
        const bool density = b.density && pen.density != 0.f;
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            Acc y[6] = {0, 0, 0, 0, 0, 0};   // reference coordinates until the row is complete
            float vi[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            float4 gi = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < f.N) {
                gi = f.nodes[i];
                float tvi[6];
                if (correct) {
#pragma unroll
                    for (int a = 0; a < 6; ++a) vi[a] = vec[6 * i + a];
                }
                node_to_ref(gi, vi, tvi);
                const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
                for (uint32_t e = beg + sub; e < end; e += kRowLanes) {
                    const uint32_t t = b.n2t_index[e];
                    uint32_t kk;
                    const int type = type_of(tb, t, kk);
                    const float4 g0 = b.mf_g[2 * (size_t)t];
                    if (type == kSmooth) {
                        const SmoothRec sm = eval_smooth<R>(f, se3, kk);
                        const uint32_t n1 = sm.ij & 0xFFFF, n2 = sm.ij >> 16;
                        float J[3][6], res[3];
                        smooth_rows(sm, i == n1, J, res);
                        const uint32_t mult = (n1 == i ? 1u : 0u) + (n2 == i ? 1u : 0u);
                        const float u[3] = {g0.x, g0.y, g0.z};
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            float s = u[r];
                            if (correct && mult > 1) s -= (float)(mult - 1) * l2.smooth * dot6(J[r], tvi);
#pragma unroll
                            for (int a = 0; a < 6; ++a) y[a] += (Acc)(s * J[r][a]);
                        }
                    } else {
                        const float4 g1 = b.mf_g[2 * (size_t)t + 1];
                        const float ws = b.n2t_w[e];
                        const float wi = fabsf(ws);
                        const float gv[6] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y};
#pragma unroll
                        for (int a = 0; a < 6; ++a) y[a] += (Acc)(wi * gv[a]);
                        if (correct && signbit(ws)) {
                            ushort4 knn;
                            float4 wk;
                            term_knn(b, type, kk, knn, wk);
                            const uint32_t mult = (knn.x == i) + (knn.y == i) + (knn.z == i) + (knn.w == i);
                            float J[3][6], lam[3];
                            const int rows = knn_term_rows<R>(b, f, se3, type, kk, l2, density, J, lam);
                            for (int r = 0; r < rows; ++r) {
                                const float s = -(float)(mult - 1) * lam[r] * wi * wi * dot6(J[r], tvi);
#pragma unroll
                                for (int a = 0; a < 6; ++a) y[a] += (Acc)(s * J[r][a]);
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int o = kRowLanes / 2; o > 0; o >>= 1) y[a] += __shfl_xor(y[a], o, 64);
            if (i < f.N && sub == 0) {
                // to node-centred coordinates: T^T (a, b) = (a + b x g, b)
                y[0] += y[4] * (Acc)gi.z - y[5] * (Acc)gi.y;
                y[1] += y[5] * (Acc)gi.x - y[3] * (Acc)gi.z;
                y[2] += y[3] * (Acc)gi.y - y[4] * (Acc)gi.x;
                if (kEpi == kEpiAdjoint) {
#pragma unroll
                    for (int a = 0; a < 6; ++a) b.q[6 * i + a] = (float)y[a];
                    continue;
                }
                // damping 1e-2 T^T T v
                float tv[6];
                node_to_ref(gi, vi, tv);
                const float bx = tv[4] * gi.z - tv[5] * gi.y, by = tv[5] * gi.x - tv[3] * gi.z,
                            bz = tv[3] * gi.y - tv[4] * gi.x;
                const float damp[6] = {tv[0] + bx, tv[1] + by, tv[2] + bz, tv[3], tv[4], tv[5]};
#pragma unroll
                for (int a = 0; a < 6; ++a) y[a] += (Acc)(1e-2f * damp[a]);
                if (kEpi == kEpiSpmv) {
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        b.q[6 * i + a] = (float)y[a];
                        acc += (Acc)vi[a] * y[a];
                    }
                } else if (kEpi == kEpiResidual) {
                    float rr[6], z[6];
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        rr[a] = (float)((Acc)b.jtr[6 * i + a] - y[a]);
                        b.r[6 * i + a] = rr[a];
                    }
                    apply_minv<Acc>(b, i, rr, z);
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        b.d1[6 * i + a] = z[a];
                        acc += (Acc)rr[a] * (Acc)z[a];
                    }
                } else {
#pragma unroll
                    for (int a = 0; a < 6; ++a) acc += (Acc)vi[a] * y[a];
                }
            }
        }
    }
    if (kEpi != kEpiAdjoint)
        write_partial<Acc>(b, kEpi == kEpiSpmv ? kSlotDQ + k : (kEpi == kEpiResidual ? kSlotRefNum : kSlotRefDen), acc);
}

// Node-centred <-> reference coordinates of the debug products, in place on
// q: to_ref 1: T^-T y (a J^T result), 0 on d1: T^-1 x (a twist).
__global__ void __launch_bounds__(256) mfr_coords_kernel(const Buffers* __restrict__ bs, int jt_result) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    float* v = jt_result ? b.q : b.d1;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += gridDim.x * blockDim.x) {
        const float4 g = f.nodes[n];
        float a[3] = {v[6 * n], v[6 * n + 1], v[6 * n + 2]};
        const float c[3] = {v[6 * n + 3], v[6 * n + 4], v[6 * n + 5]};
        if (jt_result) {
            // (a, c) -> (a - c x g, c)
            a[0] -= c[1] * g.z - c[2] * g.y;
            a[1] -= c[2] * g.x - c[0] * g.z;
            a[2] -= c[0] * g.y - c[1] * g.x;
#pragma unroll
            for (int q = 0; q < 3; ++q) v[6 * n + q] = a[q];
        } else {
            // (a, c) -> (a, c - g x a)
            v[6 * n + 3] = c[0] - (g.y * a[2] - g.z * a[1]);
            v[6 * n + 4] = c[1] - (g.z * a[0] - g.x * a[2]);
            v[6 * n + 5] = c[2] - (g.x * a[1] - g.y * a[0]);
        }
    }
}

// ------------------------------------------------------------- host launch
void launch_pcg_init(const Batch& B, bool fp64, hipStream_t stream);
void launch_pcg_update(const Batch& B, bool fp64, uint32_t k, uint32_t spmv_blocks, uint32_t dsel, hipStream_t stream);
void launch_refine_step(const Batch& B, bool fp64, uint32_t spmv_blocks, hipStream_t stream);
uint32_t row_group_grid(uint32_t nodes);

static uint32_t term_grid(const Batch& B) { return std::min<uint32_t>(std::max<uint32_t>(64 * B.nodes / 256, 64), 2048); }

template <typename Acc>
static void mf_apply(const Batch& B, int src_epi, uint32_t k, uint32_t cur, const Penalty& pen, hipStream_t stream) {
    const uint32_t tg = term_grid(B), sg = row_group_grid(B.nodes);
    if (src_epi == 0) {
        hipLaunchKernelGGL((mf_terms_kernel<Acc, kSrcDir>), dim3(tg, B.frames), dim3(256), 0, stream, B.dev, k, cur, pen);
        hipLaunchKernelGGL((mf_gather_kernel<Acc, kEpiSpmv>), dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k, cur, pen);
    } else if (src_epi == 1) {
        hipLaunchKernelGGL((mf_terms_kernel<Acc, kSrcX>), dim3(tg, B.frames), dim3(256), 0, stream, B.dev, k, cur, pen);
        hipLaunchKernelGGL((mf_gather_kernel<Acc, kEpiResidual>), dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k, cur,
                           pen);
    } else {
        hipLaunchKernelGGL((mf_terms_kernel<Acc, kSrcZ>), dim3(tg, B.frames), dim3(256), 0, stream, B.dev, k, cur, pen);
        hipLaunchKernelGGL((mf_gather_kernel<Acc, kEpiCurvature>), dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k,
                           cur, pen);
    }
}

template <typename Acc, typename R>
static void mfr_apply(const Batch& B, int src_epi, uint32_t k, uint32_t cur, const Penalty& pen, uint32_t which,
                      hipStream_t stream) {
    const uint32_t tg = term_grid(B), sg = row_group_grid(B.nodes);
    if (src_epi == 0) {
        hipLaunchKernelGGL((mfr_terms_kernel<Acc, R, kSrcDir, kOutG>), dim3(tg, B.frames), dim3(256), 0, stream, B.dev, k,
                           cur, pen, which);
        hipLaunchKernelGGL((mfr_gather_kernel<Acc, R, kEpiSpmv>), dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k, cur,
                           pen, which);
    } else if (src_epi == 1) {
        hipLaunchKernelGGL((mfr_terms_kernel<Acc, R, kSrcX, kOutG>), dim3(tg, B.frames), dim3(256), 0, stream, B.dev, k,
                           cur, pen, which);
        hipLaunchKernelGGL((mfr_gather_kernel<Acc, R, kEpiResidual>), dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k,
                           cur, pen, which);
    } else {
        hipLaunchKernelGGL((mfr_terms_kernel<Acc, R, kSrcZ, kOutG>), dim3(tg, B.frames), dim3(256), 0, stream, B.dev, k,
                           cur, pen, which);
        hipLaunchKernelGGL((mfr_gather_kernel<Acc, R, kEpiCurvature>), dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k,
                           cur, pen, which);
    }
}

// One apply of either form: fp64 selects FP64 sums and FP64 term geometry
// (the mixed precision mode), as the evaluation pass.
static void apply_any(const Batch& B, bool recompute, bool fp64, int src_epi, uint32_t k, uint32_t cur,
                      const Penalty& pen, uint32_t which, hipStream_t stream) {
    if (recompute) {
        if (fp64) mfr_apply<double, double>(B, src_epi, k, cur, pen, which, stream);
        else mfr_apply<float, float>(B, src_epi, k, cur, pen, which, stream);
    } else {
        if (fp64) mf_apply<double>(B, src_epi, k, cur, pen, stream);
        else mf_apply<float>(B, src_epi, k, cur, pen, stream);
    }
}

// Block-Jacobi PCG with the matrix-free apply: per iteration the term pass,
// the gather (q = A d, dot(d, q)) and the update of the materialized path.
void launch_pcg_matrix_free(const Batch& B, uint32_t iters, uint32_t refine, bool fp64, const Penalty& pen,
                            bool recompute, uint32_t which, hipStream_t stream) {
    const uint32_t sg = row_group_grid(B.nodes);
    iters = std::min(iters, kMaxIters);
    launch_pcg_init(B, fp64, stream);
    uint32_t cur = 0;
    for (uint32_t k = 0; k < iters; ++k) {
        apply_any(B, recompute, fp64, 0, k, cur, pen, which, stream);
        cur ^= 1;
        launch_pcg_update(B, fp64, k, sg, cur, stream);
    }
    for (uint32_t r = 0; r < refine; ++r) {
        apply_any(B, recompute, fp64, 1, 0, 0, pen, which, stream);
        apply_any(B, recompute, fp64, 2, 0, 0, pen, which, stream);
        launch_refine_step(B, fp64, sg, stream);
    }
}

// One matrix-free apply q = A d (stage timing).
void launch_mf_apply_once(const Batch& B, bool fp64, const Penalty& pen, bool recompute, uint32_t which,
                          hipStream_t stream) {
    apply_any(B, recompute, fp64, 0, 0, 0, pen, which, stream);
}

void launch_mf_entry_weights(const Batch& B, hipStream_t stream) {
    hipLaunchKernelGGL(mf_entry_weights_kernel, dim3(row_group_grid(B.nodes), B.frames), dim3(256), 0, stream, B.dev);
}

// The term Jacobian at se3[which] with unit weights, recomputed from the maps:
// transpose 0: mf_u = J x for x (reference twists) in d1 (d1 is overwritten);
// transpose 1: q = J^T y (reference twists) for y in mf_u.
void launch_debug_jacobian(const Batch& B, bool fp64, bool transpose, uint32_t which, hipStream_t stream) {
    const Penalty unit{1.f, 1.f, 1.f, 1.f, 1.f};
    const uint32_t tg = term_grid(B), sg = row_group_grid(B.nodes), vg = (B.nodes + 255) / 256;
    if (!transpose) {
        hipLaunchKernelGGL(mfr_coords_kernel, dim3(vg, B.frames), dim3(256), 0, stream, B.dev, 0);
        if (fp64)
            hipLaunchKernelGGL((mfr_terms_kernel<double, double, kSrcZ, kOutRows>), dim3(tg, B.frames), dim3(256), 0, stream,
                               B.dev, 0u, 0u, unit, which);
        else
            hipLaunchKernelGGL((mfr_terms_kernel<float, float, kSrcZ, kOutRows>), dim3(tg, B.frames), dim3(256), 0, stream,
                               B.dev, 0u, 0u, unit, which);
        return;
    }
    if (fp64) {
        hipLaunchKernelGGL((mfr_terms_kernel<double, double, kSrcRows, kOutG>), dim3(tg, B.frames), dim3(256), 0, stream,
                           B.dev, 0u, 0u, unit, which);
        hipLaunchKernelGGL((mfr_gather_kernel<double, double, kEpiAdjoint>), dim3(sg, B.frames), dim3(256), 0, stream,
                           B.dev, 0u, 0u, unit, which);
    } else {
        hipLaunchKernelGGL((mfr_terms_kernel<float, float, kSrcRows, kOutG>), dim3(tg, B.frames), dim3(256), 0, stream,
                           B.dev, 0u, 0u, unit, which);
        hipLaunchKernelGGL((mfr_gather_kernel<float, float, kEpiAdjoint>), dim3(sg, B.frames), dim3(256), 0, stream,
                           B.dev, 0u, 0u, unit, which);
    }
    hipLaunchKernelGGL(mfr_coords_kernel, dim3(vg, B.frames), dim3(256), 0, stream, B.dev, 1);
}

}  // namespace mvdr
