// Shared pieces of the PCG launches: row-group shape, per-block partial sums
// (one partial per workgroup, reduced in a fixed order by the next launch),
// 6x6 block products and the block-Jacobi preconditioner.
#pragma once
#include "common.h"
#include "device_util.h"

namespace mvdr {

constexpr int kRowLanes = 16;         // lanes per block row in the SpMV
constexpr int kRowsPerBlock = 256 / kRowLanes;
// lanes per block row in the fused CG's row kernels (a row has ~12 slots)
constexpr int kHatLanes = 8;
constexpr int kHatRowsPerBlock = 256 / kHatLanes;
static_assert(kHatLanes >= 6 && (kHatLanes & (kHatLanes - 1)) == 0,
              "lanes 0..5 of a row own its six components; the row sum is a butterfly");
constexpr uint32_t kVecBlocks = 64;   // grid of the per-node vector kernels

// partial-sum slots (partial_stride doubles each)
constexpr uint32_t kSlotDelta = 0;     // delta_k   : slot k, k <= kMaxIters
constexpr uint32_t kSlotDQ = 72;       // dot(d, q)_k: slot 72 + k
constexpr uint32_t kSlotRefNum = 144;  // refinement numerator / denominator
constexpr uint32_t kSlotRefDen = 145;
constexpr uint32_t kMaxIters = MVDR_MAX_PCG_ITERATIONS;
static_assert(kMaxIters < kSlotDQ, "delta_k slots 0..kMaxIters precede the dot(d, q) slots");

__device__ inline double* slot_ptr(const Buffers& b, uint32_t slot) { return b.partial + (size_t)slot * b.partial_stride; }

// Sums of the partials of two slots (n0 and n1 of them), computed identically
// by every block (fixed order: strided per-thread sums, a wave butterfly, the
// waves in order).  One workgroup barrier; at most one call per kernel (the
// LDS is not reused).
template <typename Acc>
__device__ inline void reduce_partials2(const double* p0, uint32_t n0, const double* p1, uint32_t n1, Acc& s0, Acc& s1) {
    __shared__ Acc sh[2][4];
    Acc v0 = 0, v1 = 0;
    for (uint32_t i = threadIdx.x; i < n0; i += blockDim.x) v0 += (Acc)p0[i];
    for (uint32_t i = threadIdx.x; i < n1; i += blockDim.x) v1 += (Acc)p1[i];
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = v0;
        sh[1][threadIdx.x >> 6] = v1;
    }
    __syncthreads();
    s0 = sh[0][0];
    s1 = sh[1][0];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
        s0 += sh[0][w];
        s1 += sh[1][w];
    }
}

// Block sum of per-thread values, stored as this block's partial of `slot`
// (once per kernel: the LDS values are not protected against a second call).
template <typename Acc>
__device__ inline void write_partial(const Buffers& b, uint32_t slot, Acc value, uint32_t block) {
    __shared__ Acc part[4];
    value = wave_sum(value);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = value;
    __syncthreads();
    if (threadIdx.x == 0) slot_ptr(b, slot)[block] = (double)(part[0] + part[1] + part[2] + part[3]);
}
template <typename Acc>
__device__ inline void write_partial(const Buffers& b, uint32_t slot, Acc value) {
    write_partial<Acc>(b, slot, value, blockIdx.x);
}

// Two block sums stored as this block's partials of two slots.  One call, not
// two of write_partial: a second write_partial would overwrite its LDS values
// while thread 0 may still be reading the first ones.
template <typename Acc>
__device__ inline void write_partials2(const Buffers& b, uint32_t slot0, Acc v0, uint32_t slot1, Acc v1, uint32_t block) {
    __shared__ Acc part[2][4];
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = v0;
        part[1][threadIdx.x >> 6] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        slot_ptr(b, slot0)[block] = (double)(part[0][0] + part[0][1] + part[0][2] + part[0][3]);
        slot_ptr(b, slot1)[block] = (double)(part[1][0] + part[1][1] + part[1][2] + part[1][3]);
    }
}

// 6x6 row-major block times a 6-vector (or its transpose), accumulated in Acc.
template <typename Acc>
__device__ inline void block_mv(const float* __restrict__ B, bool transposed, const float v[6], Acc y[6]) {
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
            for (int k = 0; k < 6; ++k) y[a] += (Acc)m[6 * a + k] * (Acc)v[k];
    } else {
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int k = 0; k < 6; ++k) y[a] += (Acc)m[6 * k + a] * (Acc)v[k];
    }
}

// M^-1 v for node n: FP64 inverse in the mixed mode, FP32 otherwise.
template <typename Acc>
__device__ inline void apply_minv(const Buffers& b, uint32_t n, const float v[6], float out[6]) {
    if constexpr (sizeof(Acc) == 8) {
        const double* M = b.diag_inv64 + 36 * (size_t)n;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += M[6 * a + k] * (double)v[k];
            out[a] = (float)s;
        }
    } else {
        const float* M = b.diag_inv + 36 * (size_t)n;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += M[6 * a + k] * v[k];
            out[a] = s;
        }
    }
}

}  // namespace mvdr
