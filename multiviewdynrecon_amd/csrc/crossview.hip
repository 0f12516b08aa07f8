// Cross-view edge pair production (include/mvdr/crossview.h): candidates by a
// windowed nearest-vertex search into the neighbouring views, one-to-one
// pairs by a per-target minimum, the back-traced subset, and in-order
// compaction of the three lists.  Reconstructs
// core/CrossViewEdgeCorrespondence/CrossViewEdgeCorrespondence.h:21-224
// (header only in the reference).
//
// Layout: per source pixel (view, y, x) of the V x H x W grid one candidate
// word pair and one flag byte; per target pixel one 64-bit minimum
// (squared-distance bits << 32 | source pixel); the work is a few hundred
// window taps per edge pixel, edge pixels being ~1 % of the grid, so every
// pass is bound by streaming the per-pixel arrays (HBM), not by arithmetic.
//
// Selection arithmetic (transforms, projection, distances) is evaluated
// without FMA contraction so that it rounds exactly like the oracle's.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/mvdr/crossview.h"
#include "common.h"
#include "device_util.h"

namespace mvdr {
namespace {

constexpr uint32_t kCvTile = 256;        // source pixels per workgroup step
constexpr uint32_t kCvLists = 3;
constexpr uint32_t kNoCandidate = 0xFFFFFFFFu;

struct CvFrame {
    uint32_t V, W, H;
    float fx[kMaxViews], fy[kMaxViews], cx[kMaxViews], cy[kMaxViews];
    Mat34 init[kMaxViews], init_inv[kMaxViews];
    const float4* vertex[kMaxViews];
    const uint8_t* edge[kMaxViews];
    const ushort4* corr[kMaxViews];
};

struct CvBuffers {
    uint2* cand;                 // per source pixel: (x2 | y2 << 16, view2) or (kNoCandidate, 0)
    float* dist2;                // per source pixel: squared distance of its candidate
    unsigned long long* best;    // per target pixel: min (dist2 bits << 32 | source pixel)
    uint8_t* flags;              // per source pixel: 1 candidate, 2 unique, 4 back-traced
    uint32_t* tile_count;        // kCvLists x tiles
    uint32_t* tile_offset;       // kCvLists x tiles
    uint32_t* counts;            // kCvLists totals (untruncated)
    mvdr_cross_pair* out[kCvLists];
    uint32_t max_pairs;
    int radius, step, track;
    float max_d2;
};

__device__ inline bool zero_vertex(float4 v) { return fabsf(v.x) < 1e-3f && fabsf(v.y) < 1e-3f && fabsf(v.z) < 1e-3f; }

__device__ inline bool edge_pixel(const CvFrame& f, uint32_t v, int x, int y, float4& vert) {
    if (x < 0 || y < 0 || x >= (int)f.W || y >= (int)f.H) return false;
    const size_t o = (size_t)y * f.W + x;
    if (f.edge[v][o] != 1) return false;
    vert = f.vertex[v][o];
    return !zero_vertex(vert);
}

// Nearest edge vertex of view n to camera-n point q in the sampled window
// around q's projection; returns its squared distance (or +inf).
__device__ inline float nearest_in_view(const CvFrame& f, const CvBuffers& b, uint32_t n, f3 q, int& bx, int& by) {
    float best = __int_as_float(0x7f800000);
    bx = by = -1;
    if (!(q.z > 0.f)) return best;
    const int u = project_round<float>(q.x, q.z, f.fx[n], f.cx[n]);
    const int w = project_round<float>(q.y, q.z, f.fy[n], f.cy[n]);
    for (int y = w - b.radius; y <= w + b.radius; y += b.step)
        for (int x = u - b.radius; x <= u + b.radius; x += b.step) {
            float4 t;
            if (!edge_pixel(f, n, x, y, t)) continue;
            const float dx = t.x - q.x, dy = t.y - q.y, dz = t.z - q.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best) {
                best = d2;
                bx = x;
                by = y;
            }
        }
    return best;
}

__global__ void __launch_bounds__(256) cv_clear_kernel(CvBuffers b, uint32_t pixels) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += gridDim.x * blockDim.x)
        b.best[p] = ~0ull;
}

// Stage 1: one candidate per edge pixel; the per-target minimum.
__global__ void __launch_bounds__(256) cv_candidates_kernel(CvFrame f, CvBuffers b) {
    const uint32_t plane = f.W * f.H, pixels = f.V * plane;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += gridDim.x * blockDim.x) {
        const uint32_t v = p / plane, o = p - v * plane;
        const int x = (int)(o % f.W), y = (int)(o / f.W);
        uint2 cand = make_uint2(kNoCandidate, 0u);
        float cd = 0.f;
        float4 src;
        if (f.V > 1 && edge_pixel(f, v, x, y, src)) {
            const f3 world = f.init[v].apply(xyz(src));
            float best = __int_as_float(0x7f800000);
            const uint32_t next = (v + 1) % f.V, last = (v + f.V - 1) % f.V;
            for (int k = 0; k < 2; ++k) {
                const uint32_t n = k == 0 ? next : last;
                if (k == 1 && last == next) break;
                int bx, by;
                const float d2 = nearest_in_view(f, b, n, f.init_inv[n].apply(world), bx, by);
                if (d2 < best) {
                    best = d2;
                    cand = make_uint2((uint32_t)bx | ((uint32_t)by << 16), n);
                }
            }
            if (!(best < b.max_d2)) cand = make_uint2(kNoCandidate, 0u);
            cd = best;
        }
        b.cand[p] = cand;
        b.dist2[p] = cd;
        if (cand.x != kNoCandidate) {
            const uint32_t t = cand.y * plane + (cand.x >> 16) * f.W + (cand.x & 0xFFFF);
            atomicMin(&b.best[t], ((unsigned long long)__float_as_uint(cd) << 32) | p);
        }
    }
}

__device__ inline uint32_t target_of(const CvFrame& f, uint2 c) {
    return c.y * f.W * f.H + (c.x >> 16) * f.W + (c.x & 0xFFFF);
}

// An optical-flow correspondence ends within the track window around (x, y).
__device__ inline bool traceable(const CvFrame& f, const CvBuffers& b, uint32_t v, int x, int y) {
    const ushort4* m = f.corr[v];
    if (!m) return false;
    for (int yy = y - b.track; yy <= y + b.track; ++yy)
        for (int xx = x - b.track; xx <= x + b.track; ++xx) {
            if (xx < 0 || yy < 0 || xx >= (int)f.W || yy >= (int)f.H) continue;
            const ushort4 e = m[(size_t)yy * f.W + xx];
            if (e.z == (uint32_t)xx && e.w == (uint32_t)yy) return true;
        }
    return false;
}

// Block-wide exclusive rank of a predicate in thread order, and its total.
__device__ inline uint32_t block_rank(bool pred, uint32_t& total, uint32_t* lds) {
    const uint64_t mask = __ballot(pred);
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        before += w < wave ? lds[w] : 0u;
        total += lds[w];
    }
    __syncthreads();
    return before + lane_prefix(mask);
}

// Stage 2 + 3 flags per source pixel; per-tile counts of the three lists.
__global__ void __launch_bounds__(kCvTile) cv_flags_kernel(CvFrame f, CvBuffers b, uint32_t tiles) {
    __shared__ uint32_t lds[kCvTile / 64];
    const uint32_t plane = f.W * f.H, pixels = f.V * plane;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t p = tile * kCvTile + threadIdx.x;
        uint32_t flag = 0;
        if (p < pixels) {
            const uint2 c = b.cand[p];
            if (c.x != kNoCandidate) {
                flag = 1;
                const uint32_t t = target_of(f, c);
                bool unique = (uint32_t)(b.best[t] & 0xFFFFFFFFull) == p;
                if (unique && t < p) {
                    // the reverse pair t -> p kept as well: it stays from the lower source pixel
                    const uint2 r = b.cand[t];
                    if (r.x != kNoCandidate && target_of(f, r) == p && (uint32_t)(b.best[p] & 0xFFFFFFFFull) == t)
                        unique = false;
                }
                if (unique) {
                    flag |= 2;
                    const uint32_t v = p / plane, o = p - v * plane;
                    if (traceable(f, b, v, (int)(o % f.W), (int)(o / f.W)) &&
                        traceable(f, b, c.y, (int)(c.x & 0xFFFF), (int)(c.x >> 16)))
                        flag |= 4;
                }
            }
            b.flags[p] = (uint8_t)flag;
        }
        for (uint32_t l = 0; l < kCvLists; ++l) {
            uint32_t total;
            (void)block_rank((flag >> l) & 1u, total, lds);
            if (threadIdx.x == 0) b.tile_count[l * tiles + tile] = total;
        }
    }
}

__global__ void __launch_bounds__(1024) cv_scan_kernel(CvBuffers b, uint32_t tiles) {
    for (uint32_t l = 0; l < kCvLists; ++l) {
        const uint32_t total = block_exclusive_scan_1024(b.tile_count + l * tiles, b.tile_offset + l * tiles, tiles);
        if (threadIdx.x == 0) b.counts[l] = total;
    }
}

__global__ void __launch_bounds__(kCvTile) cv_scatter_kernel(CvFrame f, CvBuffers b, uint32_t tiles) {
    __shared__ uint32_t lds[kCvTile / 64];
    const uint32_t plane = f.W * f.H, pixels = f.V * plane;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t p = tile * kCvTile + threadIdx.x;
        const uint32_t flag = p < pixels ? b.flags[p] : 0u;
        mvdr_cross_pair rec;
        if (flag) {
            const uint2 c = b.cand[p];
            const uint32_t v = p / plane, o = p - v * plane;
            rec.pixel[0] = (uint16_t)(o % f.W);
            rec.pixel[1] = (uint16_t)(o / f.W);
            rec.pixel[2] = (uint16_t)(c.x & 0xFFFF);
            rec.pixel[3] = (uint16_t)(c.x >> 16);
            rec.view[0] = (uint16_t)v;
            rec.view[1] = (uint16_t)c.y;
        }
        for (uint32_t l = 0; l < kCvLists; ++l) {
            uint32_t total;
            const bool pred = (flag >> l) & 1u;
            const uint32_t dst = b.tile_offset[l * tiles + tile] + block_rank(pred, total, lds);
            if (pred && dst < b.max_pairs) b.out[l][dst] = rec;
        }
    }
}

}  // namespace
}  // namespace mvdr

using namespace mvdr;

struct mvdr_crossview {
    mvdr_crossview_config cfg{};
    CvBuffers b{};
    uint32_t tiles = 0;
    void* block = nullptr;
    hipEvent_t done = nullptr;
    bool found = false;
    ~mvdr_crossview() {
        if (done) (void)hipEventDestroy(done);
        if (block) (void)hipFree(block);
    }
};

extern "C" {

void mvdr_crossview_default_config(mvdr_crossview_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_views = MVDR_MAX_VIEWS;
    c->max_width = 640;
    c->max_height = 480;
    c->max_pairs = 16384;
    c->search_radius = 20;
    c->search_step = 2;
    c->track_radius = 1;
    c->max_distance2 = 1e-4f;  // 1 cm
}

mvdr_status mvdr_crossview_create(const mvdr_crossview_config* config, mvdr_crossview** out) {
    if (!config || !out) return MVDR_E_INVALID;
    *out = nullptr;
    const mvdr_crossview_config& c = *config;
    if (c.max_views < 1 || c.max_views > MVDR_MAX_VIEWS || c.max_width < 1 || c.max_height < 1 ||
        c.max_width > 65535 || c.max_height > 65535 || c.max_pairs < 1 || c.search_step < 1 || c.search_radius > 1024 ||
        c.track_radius > 16 || !(c.max_distance2 >= 0.f))
        return MVDR_E_SHAPE;
    const uint64_t pixels = (uint64_t)c.max_views * c.max_width * c.max_height;
    if (pixels >= 0xFFFFFFFFull) return MVDR_E_SHAPE;
    auto* h = new mvdr_crossview();
    h->cfg = c;
    h->tiles = (uint32_t)((pixels + kCvTile - 1) / kCvTile);
    size_t off = 0;
    auto carve = [&](auto*& p, size_t n, char* base) {
        using T = std::remove_reference_t<decltype(*p)>;
        off = (off + 255) & ~(size_t)255;
        p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += std::max<size_t>(n, 1) * sizeof(T);
    };
    auto lay = [&](char* base) {
        off = 0;
        carve(h->b.cand, pixels, base);
        carve(h->b.dist2, pixels, base);
        carve(h->b.best, pixels, base);
        carve(h->b.flags, pixels, base);
        carve(h->b.tile_count, (size_t)kCvLists * h->tiles, base);
        carve(h->b.tile_offset, (size_t)kCvLists * h->tiles, base);
        carve(h->b.counts, kCvLists, base);
        for (uint32_t l = 0; l < kCvLists; ++l) carve(h->b.out[l], c.max_pairs, base);
        return off;
    };
    const size_t bytes = lay(nullptr);
    hipError_t e = hipMalloc(&h->block, bytes);
    if (e == hipSuccess) {
        lay(static_cast<char*>(h->block));
        e = hipMemset(h->b.counts, 0, kCvLists * sizeof(uint32_t));
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete h;
        return MVDR_E_HIP;
    }
    h->b.max_pairs = c.max_pairs;
    h->b.radius = (int)c.search_radius;
    h->b.step = (int)c.search_step;
    h->b.track = (int)c.track_radius;
    h->b.max_d2 = c.max_distance2;
    *out = h;
    return MVDR_OK;
}

void mvdr_crossview_destroy(mvdr_crossview* h) {
    if (!h) return;
    if (h->done) (void)hipEventSynchronize(h->done);
    delete h;
}

mvdr_status mvdr_crossview_find(mvdr_crossview* h, const mvdr_crossview_input* in, void* stream) {
    if (!h || !in) return MVDR_E_INVALID;
    const mvdr_crossview_config& c = h->cfg;
    if (in->num_views < 1 || in->num_views > c.max_views || in->width < 1 || in->height < 1 ||
        in->width > c.max_width || in->height > c.max_height)
        return MVDR_E_SHAPE;
    CvFrame f;
    std::memset(&f, 0, sizeof(f));
    f.V = in->num_views;
    f.W = in->width;
    f.H = in->height;
    for (uint32_t v = 0; v < f.V; ++v) {
        if (!in->vertex[v] || !in->edge_mask[v]) return MVDR_E_INVALID;
        f.fx[v] = in->intrinsic[v].fx;
        f.fy[v] = in->intrinsic[v].fy;
        f.cx[v] = in->intrinsic[v].cx;
        f.cy[v] = in->intrinsic[v].cy;
        f.init[v] = to_mat34(in->initial_camera_se3[v]);
        f.init_inv[v] = inverse(f.init[v]);
        f.vertex[v] = static_cast<const float4*>(in->vertex[v]);
        f.edge[v] = static_cast<const uint8_t*>(in->edge_mask[v]);
        f.corr[v] = static_cast<const ushort4*>(in->corr_map[v]);
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t pixels = f.V * f.W * f.H;
    const uint32_t tiles = (pixels + kCvTile - 1) / kCvTile;
    const uint32_t grid = std::min<uint32_t>(tiles, 2048u);
    hipLaunchKernelGGL(cv_clear_kernel, dim3(grid), dim3(256), 0, st, h->b, pixels);
    hipLaunchKernelGGL(cv_candidates_kernel, dim3(grid), dim3(256), 0, st, f, h->b);
    hipLaunchKernelGGL(cv_flags_kernel, dim3(grid), dim3(kCvTile), 0, st, f, h->b, tiles);
    hipLaunchKernelGGL(cv_scan_kernel, dim3(1), dim3(1024), 0, st, h->b, tiles);
    hipLaunchKernelGGL(cv_scatter_kernel, dim3(grid), dim3(kCvTile), 0, st, f, h->b, tiles);
    if (hipGetLastError() != hipSuccess) return MVDR_E_HIP;
    if (hipEventRecord(h->done, st) != hipSuccess) return MVDR_E_HIP;
    h->found = true;
    return MVDR_OK;
}

mvdr_status mvdr_crossview_get_pairs(mvdr_crossview* h, int list, const void** pairs, uint32_t* count) {
    if (!h || !pairs || !count || list < 0 || list >= (int)kCvLists) return MVDR_E_INVALID;
    if (!h->found) return MVDR_E_STATE;
    if (hipEventSynchronize(h->done) != hipSuccess) return MVDR_E_HIP;
    uint32_t n[kCvLists];
    if (hipMemcpy(n, h->b.counts, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) return MVDR_E_HIP;
    *pairs = h->b.out[list];
    *count = std::min(n[list], h->cfg.max_pairs);
    return n[list] > h->cfg.max_pairs ? MVDR_E_CAPACITY : MVDR_OK;
}

}  // extern "C"
