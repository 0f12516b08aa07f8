// Wave64 / workgroup primitives used by the solver kernels: lane prefixes,
// block sums and scans, LDS bitonic sort, bordered map fetches.
#pragma once

#include "common.h"

namespace mvdr {

// Number of set bits of `mask` in lanes below the calling lane.
__device__ inline uint32_t lane_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Orders this wave's LDS accesses for its other lanes (single-wave blocks):
// waits for the wave's outstanding LDS operations only.  A workgroup barrier
// would also wait for every outstanding global store.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0), vmcnt / expcnt untouched
    __builtin_amdgcn_wave_barrier();
}

// XCD-aware work mapping.  Workgroups are dispatched round-robin over the 8
// XCDs (linear workgroup id mod 8), each with its own L2.  For a grid-stride
// loop over `used` work items with `grid` workgroups (a multiple of 8), the
// v-th virtual item (v = blockIdx.x + s * grid) maps to a work item such that
// XCD x handles a contiguous share of every sweep of `grid` items: neighbouring
// rows / chunks, and the data they share, stay in one L2.  Returns 0xFFFFFFFF
// for a virtual item without work.
constexpr uint32_t kXcds = 8;

__device__ inline uint32_t xcd_remap(uint32_t v, uint32_t grid, uint32_t used) {
    const uint32_t base = (v / grid) * grid, w = v % grid;
    if (base >= used) return 0xFFFFFFFFu;
    const uint32_t n = min(grid, used - base);
    const uint32_t per = (n + kXcds - 1) / kXcds;
    const uint32_t x = w % kXcds, k = w / kXcds;
    const uint32_t g = x * per + k;
    return (k < per && g < n) ? base + g : 0xFFFFFFFFu;
}

// Number of virtual items a grid-stride loop over `used` items visits.
__device__ inline uint32_t xcd_span(uint32_t grid, uint32_t used) { return ((used + grid - 1) / grid) * grid; }

// Frame-to-XCD affinity for a grid (blocks, frames) whose frame count is a
// multiple of the XCD count: every workgroup of a frame runs on one XCD
// (workgroup id mod 8), so the frame's blocks and vectors share that L2.
// Within a frame the blocks keep their dispatch order.
struct WorkItem {
    uint32_t frame, block;
};
__device__ inline WorkItem frame_affine() {
    const uint32_t F = gridDim.y, blocks = gridDim.x;
    if (F % kXcds != 0) return WorkItem{blockIdx.y, blockIdx.x};
    const uint32_t L = blockIdx.x + blocks * blockIdx.y;
    const uint32_t x = L % kXcds, r = L / kXcds;
    return WorkItem{x + kXcds * (r / blocks), r % blocks};
}
__device__ inline Buffers frame_buffers_at(const Buffers* bs, uint32_t frame) {
#if __HIP_DEVICE_COMPILE__
    return *((const __attribute__((address_space(4))) Buffers*)bs + frame);
#else
    return bs[frame];
#endif
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// Sum over a block of BLOCK threads; the result is valid in every thread.
template <int BLOCK>
__device__ inline uint32_t block_sum_u32(uint32_t v) {
    __shared__ uint32_t part[BLOCK / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += part[w];
    return s;
}

// Exclusive scan of n values with one block of 1024 threads; each thread
// scans a contiguous chunk.  Optionally copies the result to `copy` and
// reports the largest input.  Returns the total (valid in every thread).
__device__ inline uint32_t block_exclusive_scan_1024(const uint32_t* in, uint32_t* out, uint32_t n,
                                                     uint32_t* copy = nullptr, uint32_t* max_out = nullptr) {
    __shared__ uint32_t wave_tot[16];
    __shared__ uint32_t wave_max_s[16];
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t beg = min(n, threadIdx.x * per), end = min(n, beg + per);
    uint32_t local = 0, mx = 0;
    for (uint32_t i = beg; i < end; ++i) {
        const uint32_t v = in[i];
        local += v;
        mx = v > mx ? v : mx;
    }
    // inclusive scan of `local` across the block
    uint32_t incl = local;
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += u;
    }
    mx = wave_max(mx);
    if (lane == 63) {
        wave_tot[threadIdx.x >> 6] = incl;
        wave_max_s[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    uint32_t before = 0, total = 0, gmax = 0;
    for (int w = 0; w < 16; ++w) {
        const uint32_t t = wave_tot[w];
        before += (w < (int)(threadIdx.x >> 6)) ? t : 0u;
        total += t;
        gmax = wave_max_s[w] > gmax ? wave_max_s[w] : gmax;
    }
    uint32_t run = before + incl - local;
    for (uint32_t i = beg; i < end; ++i) {
        const uint32_t v = in[i];
        out[i] = run;
        if (copy) copy[i] = run;
        run += v;
    }
    if (max_out) *max_out = gmax;
    __syncthreads();
    return total;
}

__device__ inline uint32_t next_pow2(uint32_t v) {
    uint32_t m = 1;
    while (m < v) m <<= 1;
    return m;
}

// Ascending bitonic sort of m (a power of two) keys in LDS by the whole block.
template <typename K>
__device__ inline void bitonic_sort_lds(K* keys, uint32_t m) {
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (m >> 1); i += blockDim.x) {
                const uint32_t lo = ((i / j) * 2 * j) + (i % j), hi = lo + j;
                const bool asc = (lo & k) == 0;
                const K a = keys[lo], c = keys[hi];
                if ((a > c) == asc) {
                    keys[lo] = c;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
}

// Map fetches with the reference textures' border addressing: zero outside.
__device__ inline float4 fetch4(const FrameDev& f, const float4* map, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)f.W || y >= (int)f.H) return make_float4(0.f, 0.f, 0.f, 0.f);
    return map[y * (int)f.W + x];
}
__device__ inline uint32_t fetch_index(const FrameDev& f, uint32_t v, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)f.W || y >= (int)f.H) return 0u;
    return f.index[v][y * (int)f.W + x];
}
__device__ inline uint8_t fetch_edge(const FrameDev& f, uint32_t v, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)f.W || y >= (int)f.H) return 0;
    return f.edge[v][y * (int)f.W + x];
}

}  // namespace mvdr
