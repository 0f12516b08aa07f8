// Block-sparse J^T J apply, block-Jacobi PCG and the warp-field update.
//
// The PCG is BlockPCG<6>::SolveNoConvergeCheck (BlockPCG.cuh:65-181,312-458)
// restructured into two launches per iteration:
//   spmv   : d = s + beta * d (beta = delta_k / delta_{k-1}) folded into the
//            gather of q = A d, plus dot(d, q) -- the reference's third update,
//            SpMV and first cublasSdot;
//   update : alpha = delta_k / dot(d, q); x += alpha d; r -= alpha q; s = M^-1 r;
//            delta_{k+1} = dot(r, s) -- its second update and second cublasSdot.
// No kernel talks across its own blocks: every dot product is left as one
// partial per block, and the next launch reduces those partials (in a fixed
// order, identically in every block), so there is no grid-wide fence and the
// results are run-to-run deterministic.  Acc is float in the plain-FP32 mode
// and double in the mixed mode, where M^-1 is also kept in FP64.
#include "common.h"
#include "device_util.h"
#include "pcg_util.h"
#include "tuning.h"
#include "terms.h"

namespace mvdr {

// Input vector entry of node j: prev_j, or s_j + beta * prev_j.
template <typename Acc>
__device__ inline void load_dir(const float* __restrict__ prev, const float* __restrict__ s, bool use_beta, Acc beta, uint32_t j,
                                float v[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const float p = prev[6 * j + a];
        v[a] = use_beta ? (float)((Acc)s[6 * j + a] + beta * (Acc)p) : p;
    }
}

// y = (A v)_i over block row i: diagonal, lower entries (B_ki^T, k < i), upper
// entries (B_ij, j > i), split over the kRowLanes lanes of the row group and
// reduced across them.
template <typename Acc>
__device__ inline void row_apply(const Buffers& b, uint32_t i, uint32_t sub, const float* prev, const float* s, bool use_beta,
                                 Acc beta, Acc y[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) y[a] = 0;
    const uint32_t lo_beg = b.low_offset[i], lo_n = b.low_offset[i + 1] - lo_beg;
    const uint32_t up_beg = b.pair_offset[i], up_n = b.pair_offset[i + 1] - up_beg;
    const uint32_t total = 1 + lo_n + up_n;
    for (uint32_t e = sub; e < total; e += kRowLanes) {
        const float* B;
        bool tr;
        uint32_t j;
        if (e == 0) {
            B = b.diag + 36 * (size_t)i;
            tr = false;
            j = i;
        } else if (e <= lo_n) {
            const uint2 le = b.low_ent[lo_beg + e - 1];
            B = b.offdiag + 36 * (size_t)le.y;
            tr = true;
            j = le.x;
        } else {
            const uint32_t p = up_beg + (e - 1 - lo_n);
            B = b.offdiag + 36 * (size_t)p;
            tr = false;
            j = b.pair_key[p] & 0xFFFF;
        }
        float v[6];
        load_dir<Acc>(prev, s, use_beta, beta, j, v);
        block_mv<Acc>(B, tr, v, y);
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int o = kRowLanes / 2; o > 0; o >>= 1) y[a] += __shfl_xor(y[a], o, 64);
}

// r = b, d = M^-1 b, x = 0, partials of delta_0 = dot(r, d)
// (blockPCGZeroInitKernel, BlockPCG.cuh:65-111, then cublasSdot).
template <typename Acc>
__global__ void __launch_bounds__(256) pcg_init_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
            float rb[6], dv[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) rb[a] = b.jtr[6 * n + a];
            apply_minv<Acc>(b, n, rb, dv);
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                b.r[6 * n + a] = rb[a];
                b.d0[6 * n + a] = dv[a];
                b.x[6 * n + a] = 0.f;
                acc += (Acc)rb[a] * (Acc)dv[a];
            }
        }
    }
    write_partial<Acc>(b, kSlotDelta + 0, acc);
}

// Iteration k: q = A d with d = s + beta d_prev (k > 0); partials of dot(d, q).
template <typename Acc>
__global__ void __launch_bounds__(256) pcg_spmv_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t cur) {
    const Buffers b = frame_buffers(bs);
    const float* __restrict__ d_prev = cur ? b.d1 : b.d0;
    float* __restrict__ d_out = cur ? b.d0 : b.d1;
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kRowLanes;
    Acc beta = 0;
    if (k > 0) {
        Acc num, den;
        reduce_partials2<Acc>(slot_ptr(b, kSlotDelta + k), kVecBlocks, slot_ptr(b, kSlotDelta + k - 1), kVecBlocks, num, den);
        beta = den != (Acc)0 ? num / den : (Acc)0;
    }
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            Acc y[6];
            row_apply<Acc>(b, i < f.N ? i : 0, sub, d_prev, b.s, k > 0, beta, y);
            if (i < f.N && sub == 0) {
                float v[6];
                load_dir<Acc>(d_prev, b.s, k > 0, beta, i, v);
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    d_out[6 * i + a] = v[a];
                    b.q[6 * i + a] = (float)y[a];
                    acc += (Acc)v[a] * y[a];
                }
            }
        }
    }
    write_partial<Acc>(b, kSlotDQ + k, acc);
}

// Iteration k: alpha = delta_k / dot(d, q); x += alpha d; r -= alpha q;
// s = M^-1 r; partials of delta_{k+1} = dot(r, s) (blockPCGSecondUpdateKernel,
// BlockPCG.cuh:120-161).  A zero denominator gives a zero step.
template <typename Acc>
__global__ void __launch_bounds__(256) pcg_update_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t spmv_blocks, uint32_t dsel) {
    const Buffers b = frame_buffers(bs);
    const float* __restrict__ d = dsel ? b.d1 : b.d0;
    const FrameDev& f = *b.frame;
    Acc delta, dq;
    reduce_partials2<Acc>(slot_ptr(b, kSlotDelta + k), kVecBlocks, slot_ptr(b, kSlotDQ + k), spmv_blocks, delta, dq);
    const Acc alpha = dq != (Acc)0 ? delta / dq : (Acc)0;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
            float t[6], sv[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const uint32_t i = 6 * n + a;
                b.x[i] = (float)((Acc)b.x[i] + alpha * (Acc)d[i]);
                t[a] = (float)((Acc)b.r[i] - alpha * (Acc)b.q[i]);
                b.r[i] = t[a];
            }
            apply_minv<Acc>(b, n, t, sv);
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                b.s[6 * n + a] = sv[a];
                acc += (Acc)t[a] * (Acc)sv[a];
            }
        }
    }
    write_partial<Acc>(b, kSlotDelta + k + 1, acc);
}

// Residual refinement (mixed mode), one preconditioned steepest-descent step
// on the true residual: r = b - A x, z = M^-1 r, num = dot(r, z) ...
template <typename Acc>
__global__ void __launch_bounds__(256) refine_residual_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kRowLanes;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            Acc y[6];
            row_apply<Acc>(b, i < f.N ? i : 0, sub, b.x, b.s, false, (Acc)0, y);
            if (i < f.N && sub == 0) {
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
            }
        }
    }
    write_partial<Acc>(b, kSlotRefNum, acc);
}

// ... den = dot(z, A z) ...
template <typename Acc>
__global__ void __launch_bounds__(256) refine_curvature_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kRowLanes;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            Acc y[6];
            row_apply<Acc>(b, i < f.N ? i : 0, sub, b.d1, b.s, false, (Acc)0, y);
            if (i < f.N && sub == 0) {
#pragma unroll
                for (int a = 0; a < 6; ++a) acc += (Acc)b.d1[6 * i + a] * y[a];
            }
        }
    }
    write_partial<Acc>(b, kSlotRefDen, acc);
}

// ... x += (num / den) z
template <typename Acc>
__global__ void __launch_bounds__(256) refine_step_kernel(const Buffers* __restrict__ bs, uint32_t spmv_blocks) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    Acc num, den;
    reduce_partials2<Acc>(slot_ptr(b, kSlotRefNum), spmv_blocks, slot_ptr(b, kSlotRefDen), spmv_blocks, num, den);
    if (b.counts->error) return;
    const Acc a = den > (Acc)0 ? num / den : (Acc)0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 6 * f.N; i += stride)
        b.x[i] = (float)((Acc)b.x[i] + a * (Acc)b.d1[i]);
}

// ------------------------------------------- fused CG (mixed precision mode)
// Block-Jacobi PCG on A x = b is CG on A^ x^ = b^ with A^ = L^-1 A L^-T,
// b^ = L^-1 b, x = L^-T x^, where M = L L^T holds the Cholesky factors of the
// diagonal blocks: the same iterates in exact arithmetic, and A^ has identity
// diagonal blocks.  The Chronopoulos-Gear form of CG forms both inner
// products of an iteration in one pass:
//   x_{i+1} = x_i + alpha_i p_i,   r_{i+1} = r_i - alpha_i s_i,
//   w_{i+1} = A^ r_{i+1},          gamma_{i+1} = (r, r), delta_{i+1} = (w, r),
//   beta_{i+1} = gamma_{i+1} / gamma_i,
//   alpha_{i+1} = gamma_{i+1} / (delta_{i+1} - beta_{i+1} gamma_{i+1} / alpha_i),
//   p_{i+1} = r_{i+1} + beta_{i+1} p_i,   s_{i+1} = w_{i+1} + beta_{i+1} s_i,
// so one launch per iteration suffices: launch i reduces gamma_i, delta_i from
// the previous launch's per-block partials (in a fixed order, in every block),
// forms s_i = w_i + beta_i s_{i-1} and r_{i+1} = r_i - alpha_i s_i for its rows
// and, redundantly, for each column its rows touch, and applies A^.  Vectors
// are FP32 in HBM, updated with fused FP32 multiply-adds, and the row products
// of A^ accumulate in FP32; the inner products, alpha / beta and the factors
// L^-1 are FP64.
constexpr uint32_t kSlotCgGamma = 0;   // gamma_i partials: slot i
constexpr uint32_t kSlotCgDelta = 72;  // delta_i partials: slot 72 + i

// Partial sums of the fused CG are per row group: the kHatRowsPerBlock block
// rows a workgroup covers in one pass (rows base .. base + 31) sum their
// values (a butterfly per wave, the four waves in order) and store them as
// partial base / kHatRowsPerBlock.  A frame of N nodes has cg_partials(N) of
// them whatever the grid, so its inner products, and its iterates, do not
// depend on the workgroup count, which may then follow the batch size (one
// pass of rows per workgroup for a lone frame, several for a full batch).
// The four waves of a pass combine without a workgroup barrier: each stores
// its sums in the pass's LDS slot and counts itself in (workgroup-scope
// acquire-release on an LDS counter); the fourth to arrive adds the four in
// wave order and stores the partial.  Slots are not reused within a launch,
// so waves may drift apart by any number of passes.
__host__ __device__ inline uint32_t cg_partials(uint32_t nodes) {
    return (nodes + kHatRowsPerBlock - 1) / kHatRowsPerBlock;
}
constexpr uint32_t kCgMaxPasses = 64;  // passes of one workgroup: ceil(65535 / 32 / kXcds)
struct GroupSums {
    double part[kCgMaxPasses][2][4];  // [pass][value][wave]
    uint32_t count[kCgMaxPasses];
};
// Zeroes the arrival counters; the caller's next workgroup barrier publishes them.
__device__ inline void group_sums_init(GroupSums& sh) {
    for (uint32_t q = threadIdx.x; q < kCgMaxPasses; q += blockDim.x) sh.count[q] = 0;
}
__device__ inline void group_partials2(const Buffers& b, GroupSums& sh, uint32_t pass, uint32_t slot0, double v0,
                                       uint32_t slot1, double v1, uint32_t base) {
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t w = threadIdx.x >> 6;
        sh.part[pass][0][w] = v0;
        sh.part[pass][1][w] = v1;
        const uint32_t n = __hip_atomic_fetch_add(&sh.count[pass], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (n == blockDim.x / 64 - 1) {
            const double* p0 = sh.part[pass][0];
            const double* p1 = sh.part[pass][1];
            slot_ptr(b, slot0)[base / kHatRowsPerBlock] = p0[0] + p0[1] + p0[2] + p0[3];
            slot_ptr(b, slot1)[base / kHatRowsPerBlock] = p1[0] + p1[1] + p1[2] + p1[3];
        }
    }
}
__device__ inline void group_partial(const Buffers& b, GroupSums& sh, uint32_t pass, uint32_t slot, double v,
                                     uint32_t base) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        sh.part[pass][0][threadIdx.x >> 6] = v;
        const uint32_t n = __hip_atomic_fetch_add(&sh.count[pass], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (n == blockDim.x / 64 - 1) {
            const double* p0 = sh.part[pass][0];
            slot_ptr(b, slot)[base / kHatRowsPerBlock] = p0[0] + p0[1] + p0[2] + p0[3];
        }
    }
}

// The transformed system.  The row Gram assembly leaves
// S_ij = sum w_i w_j lambda^2 u_i^T u_i (21 values) per pair; here
// B_ij = S_ij M_ij (the change of centre g_i -> g_j: rows e_c x (g_j - g_i) add
// the translation part to the rotation part) plus the pair blocks of the graph
// edges joining i and j, in edge order, gives A^_ij = L_i^-1 B_ij L_j^-T (FP64
// factors), stored once and named by a slot of block row i and a transposed
// slot of block row j; the nodes get b^ = L^-1 b and their block-row starts.
// B_ij itself (J^T J off-diagonal, node frame) is not stored:
// pair_blocks_kernel forms it on demand.
// B_ij of pair p = (i, j): the row Gram assembly's S_ij (offdiag, 21 values)
// times the change of centre M_ij, plus the pair blocks of the graph edges
// joining i and j, in edge order.
__device__ inline void pair_block(const Buffers& b, const Counts& c, const FrameDev& f, const Penalty& pen, uint32_t p,
                                  uint32_t i, uint32_t j, float B[36]) {
    const float* Bp = b.offdiag + 36 * (size_t)p;
    float v[21];
#pragma unroll
    for (int q = 0; q < 21; ++q) v[q] = Bp[q];
    float S[36];
    int o = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int k2 = a; k2 < 6; ++k2) {
            S[6 * a + k2] = v[o];
            S[6 * k2 + a] = v[o];
            ++o;
        }
    const float4 gi = f.nodes[i], gj = f.nodes[j];
    const float d[3] = {gj.x - gi.x, gj.y - gi.y, gj.z - gi.z};
    // rows e_c x d of the change of centre (translation -> rotation part)
    const float K[3][3] = {{0.f, -d[2], d[1]}, {d[2], 0.f, -d[0]}, {-d[1], d[0], 0.f}};
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int k2 = 0; k2 < 6; ++k2) {
            float s = S[6 * a + k2];
            if (k2 < 3) s += S[6 * a + 3] * K[0][k2] + S[6 * a + 4] * K[1][k2] + S[6 * a + 5] * K[2][k2];
            B[6 * a + k2] = s;
        }
    auto add_edge = [&](uint32_t e) {
        float Ji[3][6], Jj[3][6], li[3];
        const int rows = pair_term_rows(b, c, pen, c.off[0] + e, i, j, Ji, Jj, li);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r < rows) {
                to_node_frame(Ji[r], gi);
                to_node_frame(Jj[r], gj);
#pragma unroll
                for (int a = 0; a < 6; ++a)
#pragma unroll
                    for (int k2 = 0; k2 < 6; ++k2) B[6 * a + k2] += li[r] * Ji[r][a] * Jj[r][k2];
            }
        }
    };
    const uint32_t ne = b.pair_edge_n[p];
    if (ne <= kPairEdges) {
        uint32_t ed[kPairEdges];
#pragma unroll
        for (int q = 0; q < (int)kPairEdges; ++q)
            ed[q] = (uint32_t)q < ne ? b.pair_edge[kPairEdges * (size_t)p + q] : 0xFFFFFFFFu;
        // edge order (the edges were appended in arbitrary order)
#pragma unroll
        for (int q = 1; q < (int)kPairEdges; ++q)
#pragma unroll
            for (int r = q; r > 0; --r)
                if (ed[r] < ed[r - 1]) {
                    const uint32_t tmp = ed[r];
                    ed[r] = ed[r - 1];
                    ed[r - 1] = tmp;
                }
        for (uint32_t q = 0; q < ne; ++q) add_edge(ed[q]);
    } else {
        // repeated graph edges: row i's Node2Term list holds each edge
        // term of i once, in edge order
        const uint32_t last = b.n2t_offset[i + 1];
        uint32_t lo = b.n2t_offset[i], hi = last;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (b.n2t_index[mid] < c.off[0]) lo = mid + 1;
            else hi = mid;
        }
        for (uint32_t e = lo; e < last; ++e) {
            const uint32_t t = b.n2t_index[e];
            if (t >= c.off[1]) break;
            const ushort2 ge = f.graph[t - c.off[0]];
            if (ge.x != ge.y && (ge.x == j || ge.y == j)) add_edge(t - c.off[0]);
        }
    }
}

// One XCD per frame (frame_affine): the frame's FP64 factors, read by every
// pair of their nodes, stay in that XCD's L2.
__global__ void __launch_bounds__(256) cg_transform_kernel(const Buffers* __restrict__ bs, Penalty pen) {
    const WorkItem wi = frame_affine();
    const Buffers b = frame_buffers_at(bs, wi.frame);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t N = f.N, P = c.pairs;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t x = wi.block * blockDim.x + threadIdx.x; x < P + N; x += stride) {
        if (x < P) {
            const uint32_t p = x;
            const uint32_t key = b.pair_key[p];
            const uint32_t i = key >> 16, j = key & 0xFFFF;
            float B[36];
            pair_block(b, c, f, pen, p, i, j, B);
            const double* Li = b.linv64 + kTri * (size_t)i;
            const double* Lj = b.linv64 + kTri * (size_t)j;
            double T[36];
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int cc = 0; cc < 6; ++cc) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k <= a; ++k) t += Li[tri_index(a, k)] * (double)B[6 * k + cc];
                    T[6 * a + cc] = t;
                }
            // A^_ij stored once; row i's upper slot and row j's lower slot name it
            const uint32_t su = b.low_offset[i + 1] + p, sl = b.pair_offset[j] + b.low_pos[p];
            float Ah[36];
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int bb = 0; bb < 6; ++bb) {
                    double t = 0.0;
#pragma unroll
                    for (int cc = 0; cc <= bb; ++cc) t += T[6 * a + cc] * Lj[tri_index(bb, cc)];
                    Ah[6 * a + bb] = (float)t;
                }
            float4* up = reinterpret_cast<float4*>(b.offhat + 36 * (size_t)p);
#pragma unroll
            for (int q = 0; q < 9; ++q) up[q] = make_float4(Ah[4 * q], Ah[4 * q + 1], Ah[4 * q + 2], Ah[4 * q + 3]);
            b.row_ent[su] = make_uint2(j, p);
            b.row_ent[sl] = make_uint2(i, p | 0x80000000u);
        } else {
            const uint32_t n = x - P;
            b.row_ptr[n] = b.low_offset[n] + b.pair_offset[n];
            if (n + 1 == N) b.row_ptr[N] = b.low_offset[N] + b.pair_offset[N];
            const double* L = b.linv64 + kTri * (size_t)n;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k <= a; ++k) t += L[tri_index(a, k)] * (double)b.jtr[6 * (size_t)n + k];
                b.cg_b[6 * (size_t)n + a] = (float)t;
                b.cg_rws[0][kRws * (size_t)n + a] = (float)t;
            }
        }
    }
}

// B_ij in place of S_ij (offdiag) for inspection and the materialized SpMV
// stage; the same arithmetic as the transform.
__global__ void __launch_bounds__(256) pair_blocks_kernel(const Buffers* __restrict__ bs, Penalty pen) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < c.pairs; p += stride) {
        const uint32_t key = b.pair_key[p];
        float B[36];
        pair_block(b, c, f, pen, p, key >> 16, key & 0xFFFF, B);
        float4* out = reinterpret_cast<float4*>(b.offdiag + 36 * (size_t)p);
#pragma unroll
        for (int q = 0; q < 9; ++q) out[q] = make_float4(B[4 * q], B[4 * q + 1], B[4 * q + 2], B[4 * q + 3]);
    }
}

// y = sum over the off-diagonal blocks of row i of A^ v_j, with v_j = V(j)
// supplied by the caller; split over the row's kHatLanes lanes and reduced.
template <typename Load, typename Acc = double>
__device__ inline void hat_row_apply(const Buffers& b, uint32_t i, uint32_t sub, const Load& load, Acc y[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) y[a] = 0;
    const uint32_t beg = b.row_ptr[i], n = b.row_ptr[i + 1] - beg;
#pragma unroll 2  // two slots of a lane in flight (a row has ~12 slots over 8 lanes)
    for (uint32_t e = sub; e < n; e += kHatLanes) {
        const uint2 s = b.row_ent[beg + e];
        float v[6];
        load(s.x, v);
        block_mv<Acc>(b.offhat + 36 * (size_t)(s.y & 0x7FFFFFFFu), (s.y >> 31) != 0, v, y);
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int o = kHatLanes / 2; o > 0; o >>= 1) y[a] += __shfl_xor(y[a], o, 64);
}

// w_0 = A^ r_0 (r_0 = b^), x^ = p_{-1} = s_{-1} = 0, partials of gamma_0, delta_0.
__global__ void __launch_bounds__(256) cg_init_kernel(const Buffers* __restrict__ bs) {
    const WorkItem wi = frame_affine();
    const Buffers b = frame_buffers_at(bs, wi.frame);
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kHatLanes;
    __shared__ GroupSums sums;
    group_sums_init(sums);
    __syncthreads();
    uint32_t pass = 0;
    if (!b.counts->error) {
        float* RWS = b.cg_rws[0];
        auto load = [&](uint32_t j, float v[6]) {
            const float4 r03 = *reinterpret_cast<const float4*>(RWS + kRws * (size_t)j);
            const float2 r45 = *reinterpret_cast<const float2*>(RWS + kRws * (size_t)j + 4);
            v[0] = r03.x; v[1] = r03.y; v[2] = r03.z; v[3] = r03.w; v[4] = r45.x; v[5] = r45.y;
        };
        const uint32_t groups = gridDim.x * kHatRowsPerBlock;
        for (uint32_t base = wi.block * kHatRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kHatLanes;
            double g = 0.0, d = 0.0;
            // component a = sub of the row (lanes 0..5), its r_0 loaded before
            // the row apply
            const bool own = i < f.N && sub < 6;
            float* rec = RWS + kRws * (size_t)(i < f.N ? i : 0);
            const double r = own ? rec[sub] : 0.0;
            double y[6];
            hat_row_apply(b, i < f.N ? i : 0, sub, load, y);
            if (own) {
                const double ya = sub == 0 ? y[0] : sub == 1 ? y[1] : sub == 2 ? y[2] : sub == 3 ? y[3] : sub == 4 ? y[4] : y[5];
                const size_t k = 6 * (size_t)i + sub;
                const double w = (double)(float)(ya + r);
                rec[6 + sub] = (float)w;  // w_0
                rec[12 + sub] = 0.f;      // s_-1
                b.cg_p[k] = 0.f;
                b.cg_x[k] = 0.f;
                g += r * r;
                d += w * r;
            }
            group_partials2(b, sums, pass++, kSlotCgGamma + 0, g, kSlotCgDelta + 0, d, base);
        }
    }
}

// alpha_i, beta_i from the partials of gamma_i, delta_i and the stored
// alpha_{i-1}, gamma_{i-1}; block 0 stores alpha_i, gamma_i for launch i + 1.
__device__ inline void cg_scalars(const Buffers& b, uint32_t i, uint32_t blocks, double& alpha, double& beta,
                                  uint32_t block) {
    double gamma, delta;
    reduce_partials2<double>(slot_ptr(b, kSlotCgGamma + i), blocks, slot_ptr(b, kSlotCgDelta + i), blocks, gamma, delta);
    beta = 0.0;
    double den = delta;
    if (i > 0) {
        const double alpha_prev = b.cg_scal[2 * (i - 1)], gamma_prev = b.cg_scal[2 * (i - 1) + 1];
        beta = gamma_prev != 0.0 ? gamma / gamma_prev : 0.0;
        den = alpha_prev != 0.0 ? delta - beta * gamma / alpha_prev : 0.0;
    }
    alpha = den != 0.0 ? gamma / den : 0.0;
    if (block == 0 && threadIdx.x == 0) {
        b.cg_scal[2 * i] = alpha;
        b.cg_scal[2 * i + 1] = gamma;
    }
}

// The row kernels of the fused CG (init, iterations, refinement) keep every
// workgroup of a frame on one XCD (frame_affine), so a frame's A^ blocks and
// vectors stay in that XCD's L2 from one launch to the next.

// Launch i of the fused CG (see above).  Reads r_i, w_i, s_{i-1}; writes
// r_{i+1}, w_{i+1}, s_i, p_i, x^_{i+1}.
__global__ void __launch_bounds__(256) cg_iter_kernel(const Buffers* __restrict__ bs, uint32_t i) {
    const WorkItem wi = frame_affine();
    const Buffers b = frame_buffers_at(bs, wi.frame);
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kHatLanes;
    __shared__ GroupSums sums;
    group_sums_init(sums);  // published by the barrier of the partial reduction
    double alpha, beta;
    cg_scalars(b, i, cg_partials(f.N), alpha, beta, wi.block);
    uint32_t pass = 0;
    if (!b.counts->error) {
        // reads (r_i, w_i, s_{i-1}), writes (r_{i+1}, w_{i+1}, s_i)
        // (selects, not an index into the by-value struct: that would move it to scratch)
        const bool odd = (i & 1) != 0;
        const float* RWS = odd ? b.cg_rws[1] : b.cg_rws[0];
        float* RWSo = odd ? b.cg_rws[0] : b.cg_rws[1];
        // r_{i+1} of node j, rounded as it is stored
        // FP32 vector updates (one rounding each, fused): every row forms its
        // columns' r_{i+1} exactly as their owners store it
        const float af = (float)alpha, bf = (float)beta;
        auto next_r = [&](uint32_t j, float v[6]) {
            const float4* q = reinterpret_cast<const float4*>(RWS + kRws * (size_t)j);
            const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
            const float2 q4 = *reinterpret_cast<const float2*>(RWS + kRws * (size_t)j + 16);
            const float r[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
            const float w[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
            const float sp[6] = {q3.x, q3.y, q3.z, q3.w, q4.x, q4.y};
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const float s_i = __builtin_fmaf(bf, sp[a], w[a]);
                v[a] = __builtin_fmaf(-af, s_i, r[a]);
            }
        };
        const uint32_t groups = gridDim.x * kHatRowsPerBlock;
        for (uint32_t base = wi.block * kHatRowsPerBlock; base < f.N; base += groups) {
            const uint32_t n = base + threadIdx.x / kHatLanes;
            double g = 0.0, d = 0.0;
            // component a = sub of the row's own vectors (lanes 0..5), loaded
            // before the row apply so that their latency overlaps it
            const bool own = n < f.N && sub < 6;
            const size_t ka = 6 * (size_t)(n < f.N ? n : 0) + (sub < 6 ? sub : 0);
            const size_t kr = kRws * (size_t)(n < f.N ? n : 0) + (sub < 6 ? sub : 0);
            float rk = 0.f, wk = 0.f, sk = 0.f, pk = 0.f, xk = 0.f;
            if (own) {
                rk = RWS[kr];
                wk = RWS[kr + 6];
                sk = RWS[kr + 12];
                pk = b.cg_p[ka];
                xk = b.cg_x[ka];
            }
            float y[6];
            hat_row_apply(b, n < f.N ? n : 0, sub, next_r, y);
            if (own) {
                const float ya = sub == 0 ? y[0] : sub == 1 ? y[1] : sub == 2 ? y[2] : sub == 3 ? y[3] : sub == 4 ? y[4] : y[5];
                const float s_i = __builtin_fmaf(bf, sk, wk);
                const float r_next = __builtin_fmaf(-af, s_i, rk);
                const float p_i = __builtin_fmaf(bf, pk, rk);
                const double w_next = (double)(ya + r_next);
                RWSo[kr] = r_next;
                RWSo[kr + 6] = (float)w_next;
                RWSo[kr + 12] = s_i;
                b.cg_p[ka] = p_i;
                b.cg_x[ka] = __builtin_fmaf(af, p_i, xk);
                g += (double)r_next * (double)r_next;
                d += w_next * (double)r_next;
            }
            group_partials2(b, sums, pass++, kSlotCgGamma + i + 1, g, kSlotCgDelta + i + 1, d, base);
        }
    }
}

// Residual refinement on the CG solution (preconditioned steepest descent on
// the true residual, as refine_* above, in the A^ coordinates):
// r^ = b^ - A^ x^, num = (r^, r^) ...
__global__ void __launch_bounds__(256) cg_refine_residual_kernel(const Buffers* __restrict__ bs) {
    const WorkItem wi = frame_affine();
    const Buffers b = frame_buffers_at(bs, wi.frame);
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kHatLanes;
    __shared__ GroupSums sums;
    group_sums_init(sums);
    __syncthreads();
    uint32_t pass = 0;
    if (!b.counts->error) {
        const float* X = b.cg_x;
        auto load = [&](uint32_t j, float v[6]) {
            const float2* p = reinterpret_cast<const float2*>(X + 6 * (size_t)j);
            const float2 p0 = p[0], p1 = p[1], p2 = p[2];
            v[0] = p0.x; v[1] = p0.y; v[2] = p1.x; v[3] = p1.y; v[4] = p2.x; v[5] = p2.y;
        };
        const uint32_t groups = gridDim.x * kHatRowsPerBlock;
        for (uint32_t base = wi.block * kHatRowsPerBlock; base < f.N; base += groups) {
            const uint32_t n = base + threadIdx.x / kHatLanes;
            double acc = 0.0;
            // component a = sub of the row (lanes 0..5), its inputs loaded
            // before the row apply
            const bool own = n < f.N && sub < 6;
            const size_t k = 6 * (size_t)(n < f.N ? n : 0) + (sub < 6 ? sub : 0);
            const double bk = own ? (double)b.cg_b[k] : 0.0, xk = own ? (double)X[k] : 0.0;
            double y[6];
            hat_row_apply(b, n < f.N ? n : 0, sub, load, y);
            if (own) {
                const double ya = sub == 0 ? y[0] : sub == 1 ? y[1] : sub == 2 ? y[2] : sub == 3 ? y[3] : sub == 4 ? y[4] : y[5];
                const float rr = (float)(bk - (ya + xk));
                b.q[k] = rr;
                acc += (double)rr * (double)rr;
            }
            group_partial(b, sums, pass++, kSlotRefNum, acc, base);
        }
    }
}

// ... den = (r^, A^ r^) ...
__global__ void __launch_bounds__(256) cg_refine_curvature_kernel(const Buffers* __restrict__ bs) {
    const WorkItem wi = frame_affine();
    const Buffers b = frame_buffers_at(bs, wi.frame);
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kHatLanes;
    __shared__ GroupSums sums;
    group_sums_init(sums);
    __syncthreads();
    uint32_t pass = 0;
    if (!b.counts->error) {
        const float* Q = b.q;
        auto load = [&](uint32_t j, float v[6]) {
            const float2* p = reinterpret_cast<const float2*>(Q + 6 * (size_t)j);
            const float2 p0 = p[0], p1 = p[1], p2 = p[2];
            v[0] = p0.x; v[1] = p0.y; v[2] = p1.x; v[3] = p1.y; v[4] = p2.x; v[5] = p2.y;
        };
        const uint32_t groups = gridDim.x * kHatRowsPerBlock;
        for (uint32_t base = wi.block * kHatRowsPerBlock; base < f.N; base += groups) {
            const uint32_t n = base + threadIdx.x / kHatLanes;
            double acc = 0.0;
            const bool own = n < f.N && sub < 6;
            const double q = own ? (double)Q[6 * (size_t)n + sub] : 0.0;
            double y[6];
            hat_row_apply(b, n < f.N ? n : 0, sub, load, y);
            if (own) {
                const double ya = sub == 0 ? y[0] : sub == 1 ? y[1] : sub == 2 ? y[2] : sub == 3 ? y[3] : sub == 4 ? y[4] : y[5];
                acc += q * (ya + q);
            }
            group_partial(b, sums, pass++, kSlotRefDen, acc, base);
        }
    }
}

// ... x^ += (num / den) r^ (all refinement steps but the last; the last is
// folded into the warp update).
__global__ void __launch_bounds__(256) cg_refine_step_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    double num, den;
    const uint32_t parts = cg_partials(f.N);
    reduce_partials2<double>(slot_ptr(b, kSlotRefNum), parts, slot_ptr(b, kSlotRefDen), parts, num, den);
    if (b.counts->error) return;
    const double a = den > 0.0 ? num / den : 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < 6 * f.N; k += stride)
        b.cg_x[k] = (float)((double)b.cg_x[k] + a * (double)b.q[k]);
}

// ---------------------------------------------------------- warp update
// DualQuaternion::operator mat34 (DualQuaternion.h:309-319): normalised, with
// Quaternion::matrix() normalising the rotation part once more.
__device__ inline Mat34 mat34_of(DQ d) {
    const float inv = rsqrtf(d.q0.w * d.q0.w + d.q0.x * d.q0.x + d.q0.y * d.q0.y + d.q0.z * d.q0.z);
    float4 a = make_float4(inv * d.q0.x, inv * d.q0.y, inv * d.q0.z, inv * d.q0.w);
    float4 bq = make_float4(inv * d.q1.x, inv * d.q1.y, inv * d.q1.z, inv * d.q1.w);
    const float dd = a.w * bq.w + a.x * bq.x + a.y * bq.y + a.z * bq.z;
    bq = make_float4(bq.x - dd * a.x, bq.y - dd * a.y, bq.z - dd * a.z, bq.w - dd * a.w);
    const float inv2 = rsqrtf(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
    const float4 r = make_float4(inv2 * a.x, inv2 * a.y, inv2 * a.z, inv2 * a.w);
    Mat34 m;
    m.c0 = mk(1 - 2 * r.y * r.y - 2 * r.z * r.z, 2 * r.x * r.y + 2 * r.z * r.w, 2 * r.x * r.z - 2 * r.y * r.w);
    m.c1 = mk(2 * r.x * r.y - 2 * r.z * r.w, 1 - 2 * r.x * r.x - 2 * r.z * r.z, 2 * r.y * r.z + 2 * r.x * r.w);
    m.c2 = mk(2 * r.x * r.z + 2 * r.y * r.w, 2 * r.y * r.z - 2 * r.x * r.w, 1 - 2 * r.x * r.x - 2 * r.y * r.y);
    const float4 tq = qmul(make_float4(2 * bq.x, 2 * bq.y, 2 * bq.z, 2 * bq.w), make_float4(-a.x, -a.y, -a.z, a.w));
    m.t = mk(tq.x, tq.y, tq.z);
    return m;
}

// applyWarpFieldUpdateKernel (SolverIterationData.cu:9-57): Rodrigues twist
// (MatUtils.cpp:10-40) left-composed with the node's SE3.
// kHat: the solution is x = L^-T (x^ + a r^) of the fused CG, a the last
// refinement step (0 without refinement, refined = 0).
template <bool kHat>
__global__ void __launch_bounds__(256) apply_update_kernel(const Buffers* __restrict__ bs, uint32_t k, uint32_t refined) {
    const Buffers b = frame_buffers(bs);
    const DQ* __restrict__ se3_in = (k & 1) ? b.se3[1] : b.se3[0];
    DQ* __restrict__ se3_out = (k & 1) ? b.se3[0] : b.se3[1];
    const FrameDev& f = *b.frame;
    const uint32_t stride = gridDim.x * blockDim.x;
    double step = 0.0;
    if (kHat && refined) {
        double num, den;
        const uint32_t parts = cg_partials(f.N);
        reduce_partials2<double>(slot_ptr(b, kSlotRefNum), parts, slot_ptr(b, kSlotRefDen), parts, num, den);
        step = den > 0.0 ? num / den : 0.0;
    }
    const bool bad = b.counts->error != 0;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
        if (bad) {
            se3_out[n] = se3_in[n];
            continue;
        }
        float* xi = b.x + 6 * (size_t)n;
        if (kHat) {
            double xh[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const size_t k = 6 * (size_t)n + a;
                xh[a] = (double)b.cg_x[k] + (refined ? step * (double)b.q[k] : 0.0);
            }
            const double* L = b.linv64 + kTri * (size_t)n;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                double t = 0.0;
#pragma unroll
                for (int k = a; k < 6; ++k) t += L[tri_index(k, a)] * xh[k];
                xi[a] = (float)t;
            }
        }
        // node-centred solution (w, u) -> reference twist (w, v = u + g x w),
        // written back so that the twist buffer holds the reference quantity
        const float4 g4 = f.nodes[n];
        const f3 w = mk(xi[0], xi[1], xi[2]);
        const f3 v = mk(xi[3], xi[4], xi[5]) + cross3(mk(g4.x, g4.y, g4.z), w);
        xi[3] = v.x;
        xi[4] = v.y;
        xi[5] = v.z;
        Mat34 R;
        R.c0 = mk(1, 0, 0); R.c1 = mk(0, 1, 0); R.c2 = mk(0, 0, 1);
        if (fabsf(w.x) + fabsf(w.y) + fabsf(w.z) >= 1e-4f) {
            const float angle = sqrtf(dot3(w, w));
            const f3 a = (1.0f / angle) * w;
            const float c = cosf(angle), s = sinf(angle), t = 1.0f - c;
            R.c0 = mk(t * a.x * a.x + c, t * a.x * a.y + a.z * s, t * a.x * a.z - a.y * s);
            R.c1 = mk(t * a.x * a.y - a.z * s, t * a.y * a.y + c, t * a.y * a.z + a.x * s);
            R.c2 = mk(t * a.x * a.z + a.y * s, t * a.y * a.z - a.x * s, t * a.z * a.z + c);
        }
        R.t = v;
        const Mat34 P = mat34_of(se3_in[n]);
        Mat34 T;
        T.c0 = R.rot(P.c0);
        T.c1 = R.rot(P.c1);
        T.c2 = R.rot(P.c2);
        T.t = R.rot(P.t) + R.t;
        se3_out[n] = dq_from_mat34(T);
    }
}

// Start of a frame: the warp field the GN loop reads first is the input one,
// and the counters the index build accumulates into start at zero.
__global__ void copy_se3_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    DQ* __restrict__ dst = b.se3[0];
    const FrameDev& f = *b.frame;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Counts* c = b.counts;
        c->dense = c->smooth = c->fg = c->feat = c->cross = 0;
        c->kv1 = c->kv2 = c->pairs = c->max_seg = c->max_row = 0;
        c->error = 0;
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) dst[n] = f.node_se3_in[n];
}

// ------------------------------------------------------------- host launch
// Row-group grid for `nodes` block rows: kRowsPerBlock rows per workgroup, a
// multiple of the XCD count, at most one partial slot row.
static uint32_t spmv_grid(uint32_t nodes) {
    const uint32_t g = (nodes + kRowsPerBlock - 1) / kRowsPerBlock;
    return std::max<uint32_t>(kXcds, std::min<uint32_t>((g + kXcds - 1) / kXcds * kXcds, kPartialStride));
}

template <typename Acc>
static void pcg_impl(const Batch& B, uint32_t iters, uint32_t refine, hipStream_t stream) {
    const uint32_t sg = spmv_grid(B.nodes);
    iters = std::min(iters, kMaxIters);
    hipLaunchKernelGGL(pcg_init_kernel<Acc>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev);
    uint32_t cur = 0;
    for (uint32_t k = 0; k < iters; ++k) {
        hipLaunchKernelGGL(pcg_spmv_kernel<Acc>, dim3(sg, B.frames), dim3(256), 0, stream, B.dev, k, cur);
        cur ^= 1;
        hipLaunchKernelGGL(pcg_update_kernel<Acc>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, k, sg, cur);
    }
    for (uint32_t r = 0; r < refine; ++r) {
        hipLaunchKernelGGL(refine_residual_kernel<Acc>, dim3(sg, B.frames), dim3(256), 0, stream, B.dev);
        hipLaunchKernelGGL(refine_curvature_kernel<Acc>, dim3(sg, B.frames), dim3(256), 0, stream, B.dev);
        hipLaunchKernelGGL(refine_step_kernel<Acc>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, sg);
    }
}

void launch_pcg(const Batch& B, uint32_t iters, uint32_t refine, bool fp64, hipStream_t stream) {
    if (fp64) pcg_impl<double>(B, iters, refine, stream);
    else pcg_impl<float>(B, iters, refine, stream);
}

// One ApplyJtJ (q = A d0, no direction update) for stage timing.
void launch_spmv_once(const Batch& B, bool fp64, hipStream_t stream) {
    const uint32_t sg = spmv_grid(B.nodes);
    if (fp64) hipLaunchKernelGGL(pcg_spmv_kernel<double>, dim3(sg, B.frames), dim3(256), 0, stream, B.dev, 0u, 0u);
    else hipLaunchKernelGGL(pcg_spmv_kernel<float>, dim3(sg, B.frames), dim3(256), 0, stream, B.dev, 0u, 0u);
}

void launch_pcg_update_once(const Batch& B, bool fp64, hipStream_t stream) {
    const uint32_t sg = spmv_grid(B.nodes);
    if (fp64)
        hipLaunchKernelGGL(pcg_update_kernel<double>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, 0u, sg, 1u);
    else
        hipLaunchKernelGGL(pcg_update_kernel<float>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, 0u, sg, 1u);
}

// Warp update of GN iteration k: se3[k & 1] -> se3[(k + 1) & 1].
void launch_apply_update(const Batch& B, uint32_t k, hipStream_t stream) {
    hipLaunchKernelGGL(apply_update_kernel<false>, dim3(64, B.frames), dim3(256), 0, stream, B.dev, k, 0u);
}

// Fused CG of the mixed precision mode: transform, init, one launch per
// iteration, refinement, warp update of GN iteration k.
// Workgroups per frame of the fused CG's row kernels.  Their partial sums are
// per wave (cg_partials), so the count changes no result and follows the
// launch: a lone frame or a small batch takes one workgroup per 32 block rows
// (one pass, the shortest chain of dependent loads); a larger batch, whose
// launches are bandwidth bound, at most 64 per frame, each looping over its
// rows (two passes at ~4k nodes), which halves the workgroups that reduce the
// previous iteration's partials.
constexpr uint32_t kCgOnePassWorkgroups = 2048;  // batch total up to which every row group has its own workgroup
// A tuned count (the launch table) replaces the rule; either is clamped to the
// frame's row groups and to the passes a workgroup's LDS slots hold.
static uint32_t cg_grid(uint32_t nodes, uint32_t frames) {
    const uint32_t g = (nodes + kHatRowsPerBlock - 1) / kHatRowsPerBlock;
    const uint32_t tuned = launch_params(nodes, frames).cg_workgroups;
    uint32_t cap = tuned ? tuned : ((uint64_t)g * frames <= kCgOnePassWorkgroups ? g : 64u);
    cap = std::max(cap, (g + kCgMaxPasses - 1) / kCgMaxPasses);  // a workgroup's passes fit its LDS slots
    return std::max<uint32_t>(kXcds, (std::min<uint32_t>(g, cap) + kXcds - 1) / kXcds * kXcds);
}

// The CG transform: a grid-stride loop over pairs and nodes.
static uint32_t transform_grid(uint32_t nodes, uint32_t frames) {
    const uint32_t tuned = launch_params(nodes, frames).transform_workgroups;
    return tuned ? tuned : std::min<uint32_t>((8 * nodes + 255) / 256, 2048u);
}

uint32_t cg_workgroups_for(uint32_t nodes, uint32_t frames) { return cg_grid(nodes, frames); }
uint32_t transform_workgroups_for(uint32_t nodes, uint32_t frames) { return transform_grid(nodes, frames); }

void launch_cg_solve(const Batch& B, uint32_t iters, uint32_t refine, uint32_t k, const Penalty& pen,
                     hipStream_t stream) {
    const uint32_t sg = cg_grid(B.nodes, B.frames);
    iters = std::min(iters, kMaxIters);
    hipLaunchKernelGGL(cg_transform_kernel, dim3(transform_grid(B.nodes, B.frames), B.frames), dim3(256), 0, stream,
                       B.dev, pen);
    hipLaunchKernelGGL(cg_init_kernel, dim3(sg, B.frames), dim3(256), 0, stream, B.dev);
    for (uint32_t i = 0; i < iters; ++i)
        hipLaunchKernelGGL(cg_iter_kernel, dim3(sg, B.frames), dim3(256), 0, stream, B.dev, i);
    for (uint32_t r = 0; r < refine; ++r) {
        hipLaunchKernelGGL(cg_refine_residual_kernel, dim3(sg, B.frames), dim3(256), 0, stream, B.dev);
        hipLaunchKernelGGL(cg_refine_curvature_kernel, dim3(sg, B.frames), dim3(256), 0, stream, B.dev);
        if (r + 1 < refine)
            hipLaunchKernelGGL(cg_refine_step_kernel, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev);
    }
    hipLaunchKernelGGL(apply_update_kernel<true>, dim3(64, B.frames), dim3(256), 0, stream, B.dev, k, refine ? 1u : 0u);
}

uint32_t row_group_grid(uint32_t nodes) { return spmv_grid(nodes); }

void launch_pair_blocks(const Batch& B, const Penalty& pen, hipStream_t stream) {
    const uint32_t grid = std::min<uint32_t>((8 * B.nodes + 255) / 256, 2048u);
    hipLaunchKernelGGL(pair_blocks_kernel, dim3(grid, B.frames), dim3(256), 0, stream, B.dev, pen);
}

void launch_pcg_init(const Batch& B, bool fp64, hipStream_t stream) {
    if (fp64) hipLaunchKernelGGL(pcg_init_kernel<double>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev);
    else hipLaunchKernelGGL(pcg_init_kernel<float>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev);
}

void launch_pcg_update(const Batch& B, bool fp64, uint32_t k, uint32_t spmv_blocks, uint32_t dsel, hipStream_t stream) {
    if (fp64)
        hipLaunchKernelGGL(pcg_update_kernel<double>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, k,
                           spmv_blocks, dsel);
    else
        hipLaunchKernelGGL(pcg_update_kernel<float>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, k,
                           spmv_blocks, dsel);
}

void launch_refine_step(const Batch& B, bool fp64, uint32_t spmv_blocks, hipStream_t stream) {
    if (fp64)
        hipLaunchKernelGGL(refine_step_kernel<double>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, spmv_blocks);
    else
        hipLaunchKernelGGL(refine_step_kernel<float>, dim3(kVecBlocks, B.frames), dim3(256), 0, stream, B.dev, spmv_blocks);
}

// The CG transform alone (launch tuning).
void launch_cg_transform_once(const Batch& B, const Penalty& pen, hipStream_t stream) {
    hipLaunchKernelGGL(cg_transform_kernel, dim3(transform_grid(B.nodes, B.frames), B.frames), dim3(256), 0, stream,
                       B.dev, pen);
}

// One fused CG iteration (stage timing).
void launch_cg_iter_once(const Batch& B, hipStream_t stream) {
    hipLaunchKernelGGL(cg_iter_kernel, dim3(cg_grid(B.nodes, B.frames), B.frames), dim3(256), 0, stream, B.dev, 1u);
}

void launch_copy_se3(const Batch& B, hipStream_t stream) {
    hipLaunchKernelGGL(copy_se3_kernel, dim3(64, B.frames), dim3(256), 0, stream, B.dev);
}

}  // namespace mvdr
