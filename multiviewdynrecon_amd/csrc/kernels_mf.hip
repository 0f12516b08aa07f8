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

// Block-Jacobi PCG with the matrix-free apply: per iteration the term pass,
// the gather (q = A d, dot(d, q)) and the update of the materialized path.
void launch_pcg_matrix_free(const Batch& B, uint32_t iters, uint32_t refine, bool fp64, const Penalty& pen,
                            hipStream_t stream) {
    const uint32_t sg = row_group_grid(B.nodes);
    iters = std::min(iters, kMaxIters);
    launch_pcg_init(B, fp64, stream);
    uint32_t cur = 0;
    for (uint32_t k = 0; k < iters; ++k) {
        if (fp64) mf_apply<double>(B, 0, k, cur, pen, stream);
        else mf_apply<float>(B, 0, k, cur, pen, stream);
        cur ^= 1;
        launch_pcg_update(B, fp64, k, sg, cur, stream);
    }
    for (uint32_t r = 0; r < refine; ++r) {
        if (fp64) {
            mf_apply<double>(B, 1, 0, 0, pen, stream);
            mf_apply<double>(B, 2, 0, 0, pen, stream);
        } else {
            mf_apply<float>(B, 1, 0, 0, pen, stream);
            mf_apply<float>(B, 2, 0, 0, pen, stream);
        }
        launch_refine_step(B, fp64, sg, stream);
    }
}

// One matrix-free apply q = A d (stage timing).
void launch_mf_apply_once(const Batch& B, bool fp64, const Penalty& pen, hipStream_t stream) {
    if (fp64) mf_apply<double>(B, 0, 0, 0, pen, stream);
    else mf_apply<float>(B, 0, 0, 0, pen, stream);
}

}  // namespace mvdr
