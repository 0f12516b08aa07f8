// Block-sparse J^T J apply, block-Jacobi PCG and the warp-field update.
//
// The PCG is BlockPCG<6>::SolveNoConvergeCheck (BlockPCG.cuh:65-181,312-458)
// restructured into two launches per iteration:
//   spmv   : d = s + beta * d (beta = delta / delta_old) folded into the gather of
//            q = A d, plus dot(d, q) -- the reference's third update, SpMV and
//            first cublasSdot;
//   update : alpha = delta / dot(d, q); x += alpha d; r -= alpha q; s = M^-1 r;
//            delta = dot(r, s) -- its second update and second cublasSdot.
// Dot products end in the last block to arrive (ticket), which sums the block
// partials in a fixed order, so results are run-to-run deterministic.  Acc is
// float for the plain-FP32 mode and double for the mixed mode.
#include "common.h"
#include "device_util.h"

namespace mvdr {

constexpr int kRowLanes = 8;          // lanes per block row in the SpMV
constexpr int kRowsPerBlock = 256 / kRowLanes;

// scalar slots
enum : int { kDelta = 0, kDeltaOld = 1, kDotDQ = 2, kRefineNum = 3, kRefineDen = 4 };

template <typename Acc>
__device__ inline void finish_dot(const Buffers& b, Acc block_value, int slot, int shift_delta) {
    __shared__ Acc part[4];
    Acc v = wave_sum(block_value);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const Acc s = part[0] + part[1] + part[2] + part[3];
        b.partial[blockIdx.x] = (double)s;
        __threadfence();
        const uint32_t ticket = atomicAdd(&b.ticket[0], 1u);
        if (ticket == gridDim.x - 1) {
            __threadfence();
            Acc total = 0;
            for (uint32_t q = 0; q < gridDim.x; ++q) total += (Acc)b.partial[q];
            if (shift_delta) b.scalars[kDeltaOld] = b.scalars[kDelta];
            b.scalars[slot] = (double)total;
            b.ticket[0] = 0;
        }
    }
}

// 6x6 row-major block times a 6-vector (or its transpose).
__device__ inline void block_mv(const float* __restrict__ B, bool transposed, const float v[6], float y[6]) {
    float m[36];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float4 t = reinterpret_cast<const float4*>(B)[q];
        m[4 * q] = t.x; m[4 * q + 1] = t.y; m[4 * q + 2] = t.z; m[4 * q + 3] = t.w;
    }
    if (!transposed) {
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int k = 0; k < 6; ++k) y[a] += m[6 * a + k] * v[k];
    } else {
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int k = 0; k < 6; ++k) y[a] += m[6 * k + a] * v[k];
    }
}

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

// y = (A v)_i over the block row i: diagonal, lower entries (B_ki^T, k < i),
// upper entries (B_ij, j > i), in ascending column order, split over the
// kRowLanes lanes of the row group and reduced across them.
template <typename Acc>
__device__ inline void row_apply(const Buffers& b, uint32_t i, uint32_t sub, const float* prev, const float* s, bool use_beta,
                                 Acc beta, float y[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) y[a] = 0.f;
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
        block_mv(B, tr, v, y);
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int o = kRowLanes / 2; o > 0; o >>= 1) y[a] += __shfl_xor(y[a], o, 64);
}

// r = b, d = M^-1 b, x = 0, delta = dot(r, d) (blockPCGZeroInitKernel,
// BlockPCG.cuh:65-111, then cublasSdot).
template <typename Acc>
__global__ void __launch_bounds__(256) pcg_init_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
            float rb[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) rb[a] = b.jtr[6 * n + a];
            const float* M = b.diag_inv + 36 * (size_t)n;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                float dv = 0.f;
#pragma unroll
                for (int k = 0; k < 6; ++k) dv += M[6 * a + k] * rb[k];
                b.r[6 * n + a] = rb[a];
                b.d0[6 * n + a] = dv;
                b.x[6 * n + a] = 0.f;
                acc += (Acc)rb[a] * (Acc)dv;
            }
        }
    }
    finish_dot<Acc>(b, acc, kDelta, 0);
}

// q = A d with d = s + beta * d_prev folded in (iterations after the first);
// writes d to d_out and dot(d, q) to the kDotDQ slot.
template <typename Acc>
__global__ void __launch_bounds__(256) pcg_spmv_kernel(Buffers b, const float* __restrict__ d_prev, float* __restrict__ d_out,
                                                       int use_beta) {
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kRowLanes;
    Acc beta = 0;
    if (use_beta) {
        const Acc den = (Acc)b.scalars[kDeltaOld];
        beta = den != (Acc)0 ? (Acc)b.scalars[kDelta] / den : (Acc)0;
    }
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            float y[6];
            const uint32_t row = i < f.N ? i : 0;
            row_apply<Acc>(b, row, sub, d_prev, b.s, use_beta != 0, beta, y);
            if (i < f.N && sub == 0) {
                float v[6];
                load_dir<Acc>(d_prev, b.s, use_beta != 0, beta, i, v);
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    d_out[6 * i + a] = v[a];
                    b.q[6 * i + a] = y[a];
                    acc += (Acc)v[a] * (Acc)y[a];
                }
            }
        }
    }
    finish_dot<Acc>(b, acc, kDotDQ, 0);
}

// alpha = delta / dot(d, q); x += alpha d; r -= alpha q; s = M^-1 r;
// delta_old = delta; delta = dot(r, s) (blockPCGSecondUpdateKernel,
// BlockPCG.cuh:120-161).  A zero denominator gives a zero step.
template <typename Acc>
__global__ void __launch_bounds__(256) pcg_update_kernel(Buffers b, const float* __restrict__ d) {
    const FrameDev& f = *b.frame;
    const Acc dq = (Acc)b.scalars[kDotDQ];
    const Acc alpha = dq != (Acc)0 ? (Acc)b.scalars[kDelta] / dq : (Acc)0;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
            float t[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const uint32_t k = 6 * n + a;
                b.x[k] = (float)((Acc)b.x[k] + alpha * (Acc)d[k]);
                t[a] = (float)((Acc)b.r[k] - alpha * (Acc)b.q[k]);
                b.r[k] = t[a];
            }
            const float* M = b.diag_inv + 36 * (size_t)n;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                float sv = 0.f;
#pragma unroll
                for (int k = 0; k < 6; ++k) sv += M[6 * a + k] * t[k];
                b.s[6 * n + a] = sv;
                acc += (Acc)t[a] * (Acc)sv;
            }
        }
    }
    finish_dot<Acc>(b, acc, kDelta, 1);
}

// Residual refinement (mixed mode): r = b - A x, z = M^-1 r, num = dot(r, z).
template <typename Acc>
__global__ void __launch_bounds__(256) refine_residual_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kRowLanes;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            float y[6];
            row_apply<Acc>(b, i < f.N ? i : 0, sub, b.x, b.s, false, (Acc)0, y);
            if (i < f.N && sub == 0) {
                float rr[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    rr[a] = (float)((Acc)b.jtr[6 * i + a] - (Acc)y[a]);
                    b.r[6 * i + a] = rr[a];
                }
                const float* M = b.diag_inv + 36 * (size_t)i;
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    float z = 0.f;
#pragma unroll
                    for (int k = 0; k < 6; ++k) z += M[6 * a + k] * rr[k];
                    b.d1[6 * i + a] = z;
                    acc += (Acc)rr[a] * (Acc)z;
                }
            }
        }
    }
    finish_dot<Acc>(b, acc, kRefineNum, 0);
}

// q = A z, den = dot(z, q)
template <typename Acc>
__global__ void __launch_bounds__(256) refine_curvature_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    const uint32_t sub = threadIdx.x % kRowLanes;
    Acc acc = 0;
    if (!b.counts->error) {
        const uint32_t groups = gridDim.x * kRowsPerBlock;
        for (uint32_t base = blockIdx.x * kRowsPerBlock; base < f.N; base += groups) {
            const uint32_t i = base + threadIdx.x / kRowLanes;
            float y[6];
            row_apply<Acc>(b, i < f.N ? i : 0, sub, b.d1, b.s, false, (Acc)0, y);
            if (i < f.N && sub == 0) {
#pragma unroll
                for (int a = 0; a < 6; ++a) acc += (Acc)b.d1[6 * i + a] * (Acc)y[a];
            }
        }
    }
    finish_dot<Acc>(b, acc, kRefineDen, 0);
}

// x += (num / den) z
template <typename Acc>
__global__ void __launch_bounds__(256) refine_step_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    if (b.counts->error) return;
    const Acc den = (Acc)b.scalars[kRefineDen];
    const Acc a = den > (Acc)0 ? (Acc)b.scalars[kRefineNum] / den : (Acc)0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < 6 * f.N; k += stride)
        b.x[k] = (float)((Acc)b.x[k] + a * (Acc)b.d1[k]);
}

// ---------------------------------------------------------- warp update
// applyWarpFieldUpdateKernel (SolverIterationData.cu:9-57): Rodrigues twist
// (MatUtils.cpp:10-40) left-composed with the node's SE3, back to a dual
// quaternion (DualQuaternion(const mat34&), DualQuaternion.h:123-130).
__device__ inline DQ dq_from_mat34(const Mat34& T) {
    const float m00 = T.c0.x, m10 = T.c0.y, m20 = T.c0.z;
    const float m01 = T.c1.x, m11 = T.c1.y, m21 = T.c1.z;
    const float m02 = T.c2.x, m12 = T.c2.y, m22 = T.c2.z;
    float4 q;
    const float tr = m00 + m11 + m22;
    if (tr > 0) {
        const float s = sqrtf(tr + 1.0f) * 2;
        q.w = s * 0.25f; q.x = (m21 - m12) / s; q.y = (m02 - m20) / s; q.z = (m10 - m01) / s;
    } else if ((m00 > m11) && (m00 > m22)) {
        const float s = sqrtf(1.0f + m00 - m11 - m22) * 2;
        q.w = (m21 - m12) / s; q.x = 0.25f * s; q.y = (m01 + m10) / s; q.z = (m02 + m20) / s;
    } else if (m11 > m22) {
        const float s = sqrtf(1.0f + m11 - m00 - m22) * 2;
        q.w = (m02 - m20) / s; q.x = (m01 + m10) / s; q.y = 0.25f * s; q.z = (m12 + m21) / s;
    } else {
        const float s = sqrtf(1.0f + m22 - m00 - m11) * 2;
        q.w = (m10 - m01) / s; q.x = (m02 + m20) / s; q.y = (m12 + m21) / s; q.z = 0.25f * s;
    }
    DQ o;
    o.q0 = q;
    o.q1 = qmul(make_float4(0.5f * T.t.x, 0.5f * T.t.y, 0.5f * T.t.z, 0.f), q);
    return o;
}

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

__global__ void __launch_bounds__(256) apply_update_kernel(Buffers b, const DQ* __restrict__ se3_in, DQ* __restrict__ se3_out) {
    const FrameDev& f = *b.frame;
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool bad = b.counts->error != 0;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
        if (bad) {
            se3_out[n] = se3_in[n];
            continue;
        }
        const float* xi = b.x + 6 * (size_t)n;
        const f3 w = mk(xi[0], xi[1], xi[2]), v = mk(xi[3], xi[4], xi[5]);
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

// Start of a frame: the warp field the GN loop reads first is the input one.
__global__ void copy_se3_kernel(Buffers b, DQ* __restrict__ dst) {
    const FrameDev& f = *b.frame;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) dst[n] = f.node_se3_in[n];
}

// ------------------------------------------------------------- host launch
constexpr uint32_t kVecBlocks = 64;

template <typename Acc>
static void pcg_impl(const Buffers& b, uint32_t iters, uint32_t refine, uint32_t cap_nodes, hipStream_t stream) {
    const uint32_t spmv_blocks = std::max<uint32_t>(1, std::min<uint32_t>((cap_nodes + kRowsPerBlock - 1) / kRowsPerBlock, 1024u));
    hipLaunchKernelGGL(pcg_init_kernel<Acc>, dim3(kVecBlocks), dim3(256), 0, stream, b);
    // d lives in d0 after init; the SpMV writes the folded direction to the other buffer
    float* dbuf[2] = {b.d0, b.d1};
    int cur = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(pcg_spmv_kernel<Acc>, dim3(spmv_blocks), dim3(256), 0, stream, b, dbuf[cur], dbuf[cur ^ 1],
                           it > 0 ? 1 : 0);
        cur ^= 1;
        hipLaunchKernelGGL(pcg_update_kernel<Acc>, dim3(kVecBlocks), dim3(256), 0, stream, b, dbuf[cur]);
    }
    for (uint32_t k = 0; k < refine; ++k) {
        hipLaunchKernelGGL(refine_residual_kernel<Acc>, dim3(spmv_blocks), dim3(256), 0, stream, b);
        hipLaunchKernelGGL(refine_curvature_kernel<Acc>, dim3(spmv_blocks), dim3(256), 0, stream, b);
        hipLaunchKernelGGL(refine_step_kernel<Acc>, dim3(kVecBlocks), dim3(256), 0, stream, b);
    }
}

void launch_pcg(const Buffers& b, uint32_t iters, uint32_t refine, bool fp64, uint32_t cap_nodes, hipStream_t stream) {
    if (fp64) pcg_impl<double>(b, iters, refine, cap_nodes, stream);
    else pcg_impl<float>(b, iters, refine, cap_nodes, stream);
}

void launch_apply_update(const Buffers& b, const DQ* in, DQ* out, hipStream_t stream) {
    hipLaunchKernelGGL(apply_update_kernel, dim3(64), dim3(256), 0, stream, b, in, out);
}

void launch_copy_se3(const Buffers& b, DQ* dst, hipStream_t stream) {
    hipLaunchKernelGGL(copy_se3_kernel, dim3(64), dim3(256), 0, stream, b, dst);
}

}  // namespace mvdr
