// Index build of one frame: term selection, Node2Term and NodePair2Term
// indices and the block-row structure of J^T J.  Every kernel reads its sizes
// from device memory (FrameDev / Counts), so the whole chain runs without a
// host round trip and can be captured into a hipGraph once per solver.
//
// Selection decisions (window searches, distance thresholds, the camera
// transforms feeding them) are evaluated without FMA contraction so that they
// round exactly like the same expressions on the host.  The pragma precedes
// the includes: the transforms are inline functions of common.h, and a
// contraction state set after their definitions would not reach them.
#pragma clang fp contract(off)

#include "common.h"
#include "device_util.h"

namespace mvdr {

constexpr int kTileThreads = 256;
constexpr int kTileRounds = 16;
constexpr int kTilePixels = kTileThreads * kTileRounds;  // 4096 pixels per tile
constexpr uint32_t kSelectBlocks = 32;  // workgroups per frame of the pair selection's first step

// The feature / cross-view pair selection shares the dense compaction's
// launches (defined with the selection below).
__device__ void select_eval(const Buffers& b, uint32_t block, int use_cross);
__device__ void select_pairs(const Buffers& b, bool is_cross, int use_cross);

// ------------------------------------------------------------- term selection
struct PixelCoord {
    uint32_t v, x, y;
};
__device__ inline PixelCoord pixel_of(const FrameDev& f, uint32_t p) {
    const uint32_t wh = f.W * f.H;
    PixelCoord c;
    c.v = p / wh;
    const uint32_t rem = p - c.v * wh;
    c.y = rem / f.W;
    c.x = rem - c.y * f.W;
    return c;
}

// View and in-view offset of flat pixel p (compares, no division).
__device__ inline uint32_t view_of(const FrameDev& f, uint32_t p, uint32_t wh) {
    uint32_t v = 0;
#pragma unroll
    for (int k = 1; k < kMaxViews; ++k) v += ((uint32_t)k < f.V && p >= (uint32_t)k * wh) ? 1u : 0u;
    return v;
}

// The index map of view v: a select among the frame's view pointers (no
// per-lane load of the pointer table).
__device__ inline const uint32_t* index_map(const FrameDev& f, uint32_t v) {
    const uint32_t* m = f.index[0];
#pragma unroll
    for (int k = 1; k < kMaxViews; ++k)
        if ((uint32_t)k == v) m = f.index[k];
    return m;
}

__device__ inline bool knn_ok(const FrameDev& f, ushort4 k) { return k.x < f.N && k.y < f.N && k.z < f.N && k.w < f.N; }

// Foreground-term predicate (markValidColorForegroundMapsPixelKernel,
// DensityForegroundMapHandler.cu:34-118) at the input warp field.
struct ForegroundPred {
    __device__ bool operator()(const FrameDev& f, const Buffers& b, PixelCoord c) const {
        const uint32_t o = c.y * f.W + c.x;
        const uint32_t s = f.index[c.v][o];
        if (s == kInvalidIndex) return false;
        if (s >= f.num_surfels || !knn_ok(f, f.surfel_knn[s])) {
            atomicOr(&b.counts->error, kErrNodes);
            return false;
        }
        const float4 cv = f.ref_v[c.v][o], cn = f.ref_n[c.v][o];
        const Mat34& I = f.init[c.v];
        const f3 v0 = I.apply(xyz(cv)), n0 = I.rot(xyz(cn));
        const Mat34 T = se3_of(blend4(f.node_se3_in, f.surfel_knn[s], f.surfel_w[s]));
        f3 pc = f.w2c[c.v].apply(T.apply(v0)), nc = f.w2c[c.v].rot(T.rot(n0));
        pc = f.init_inv[c.v].apply(pc);
        nc = f.init_inv[c.v].rot(nc);
        const float inv = 1.0f / sqrtf(dot3(pc, pc));
        const float cosine = -dot3(inv * pc, nc);
        const int u = project_round(pc.x, pc.z, f.fx[c.v], f.cx[c.v]);
        const int w = project_round(pc.y, pc.z, f.fy[c.v], f.cy[c.v]);
        if (!(u >= 0 && u < (int)f.W && w >= 0 && w < (int)f.H && cosine > 0.5f)) return false;
        return f.ffg[c.v][w * f.W + u] > 1e-5f;
    }
};

// Pass 1: valid pixels per 4096-pixel tile.
template <typename Pred>
__global__ void __launch_bounds__(kTileThreads) compact_count_kernel(const Buffers* __restrict__ bs, Pred pred) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const uint32_t M = f.V * f.W * f.H;
    const uint32_t base = blockIdx.x * kTilePixels;
    if (base >= M) return;
    uint32_t count = 0;
    for (int r = 0; r < kTileRounds; ++r) {
        const uint32_t p = base + r * kTileThreads + threadIdx.x;
        if (p < M && pred(f, b, pixel_of(f, p))) ++count;
    }
    count = block_sum_u32<kTileThreads>(count);
    if (threadIdx.x == 0) b.tile_count[blockIdx.x] = count;
}

// Pass 2: exclusive scan of the tile counts (block 0).  select >= 0: blocks 1
// and 2 compact the selected feature / cross-view pairs (select = use_cross).
__global__ void __launch_bounds__(1024) compact_scan_kernel(const Buffers* __restrict__ bs, int which, int select) {
    const Buffers b = frame_buffers(bs);
    if (blockIdx.x > 0) {
        select_pairs(b, blockIdx.x == 2, select);
        return;
    }
    const FrameDev& f = *b.frame;
    const uint32_t M = f.V * f.W * f.H;
    const uint32_t tiles = (M + kTilePixels - 1) / kTilePixels;
    const uint32_t total = block_exclusive_scan_1024(b.tile_count, b.tile_offset, tiles);
    if (threadIdx.x == 0) {
        if (which == 0) {
            if (total > b.cap_pixels) atomicOr(&b.counts->error, kErrDense);
            b.counts->dense = min(total, b.cap_pixels);
        } else {
            if (total > b.cap_pixels) atomicOr(&b.counts->error, kErrFg);
            b.counts->fg = min(total, b.cap_pixels);
        }
    }
}

// Pass 3: scatter in ascending pixel order (view, y, x), the order of the
// reference's inclusive-scan compaction (ImageTermKNNFetcher.cu:33-68).
template <typename Pred>
__global__ void __launch_bounds__(kTileThreads) compact_scatter_kernel(const Buffers* __restrict__ bs, Pred pred, int which) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const uint32_t M = f.V * f.W * f.H;
    const uint32_t base = blockIdx.x * kTilePixels;
    if (base >= M) return;
    __shared__ uint32_t wave_cnt[2][kTileThreads / 64];
    uint32_t running = b.tile_offset[blockIdx.x];
    const uint32_t cap = b.cap_pixels;
    const uint32_t wave = threadIdx.x >> 6;
    for (int r = 0; r < kTileRounds; ++r) {
        const uint32_t p = base + r * kTileThreads + threadIdx.x;
        PixelCoord c{0, 0, 0};
        bool ok = false;
        if (p < M) {
            c = pixel_of(f, p);
            ok = pred(f, b, c);
        }
        const uint64_t mask = __ballot(ok);
        const uint32_t in_wave = lane_prefix(mask);
        if ((threadIdx.x & 63) == 0) wave_cnt[r & 1][wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t before = 0, round_total = 0;
#pragma unroll
        for (int w = 0; w < kTileThreads / 64; ++w) {
            const uint32_t cw = wave_cnt[r & 1][w];
            before += (w < (int)wave) ? cw : 0u;
            round_total += cw;
        }
        if (ok) {
            const uint32_t dst = running + before + in_wave;
            if (dst < cap) {
                const uint32_t o = c.y * f.W + c.x;
                const uint32_t s = f.index[c.v][o];
                const ushort4 px = make_ushort4((uint16_t)c.x, (uint16_t)c.y, (uint16_t)c.v, 0);
                ScalarRec rec;
                rec.knn = f.surfel_knn[s];
                if (!knn_ok(f, rec.knn)) {
                    atomicOr(&b.counts->error, kErrNodes);
                    rec.knn = make_ushort4(0, 0, 0, 0);
                }
                rec.w = f.surfel_w[s];
                rec.r = 0.f;
                rec.pad0 = 0;
#pragma unroll
                for (int k = 0; k < 6; ++k) rec.J[k] = 0.f;
                rec.pad1[0] = rec.pad1[1] = 0.f;
                if (which == 0) {
                    b.dense_px[dst] = px;
                    b.dense_knn[dst] = rec.knn;
                    b.dense[dst] = rec;
                    if (b.density) b.density[dst] = rec;
                } else {
                    b.fg_px[dst] = px;
                    b.fg[dst] = rec;
                }
            }
        }
        running += round_total;
    }
}

// Dense terms: the pixels whose index map holds a surfel
// (ImageTermKNNFetcher.cu:19-31); the count and scatter passes above,
// specialised -- a tile's 16 index-map reads per thread are issued
// together, and the scatter takes every round's positions from one exchange
// of per-wave counts (one workgroup barrier per tile).
__device__ inline void dense_tile_surfels(const FrameDev& f, uint32_t base, uint32_t sv[kTileRounds]) {
    const uint32_t wh = f.W * f.H, M = f.V * wh;
#pragma unroll
    for (int r = 0; r < kTileRounds; ++r) {
        const uint32_t p = base + r * kTileThreads + threadIdx.x;
        sv[r] = kInvalidIndex;
        if (p < M) {
            const uint32_t v = view_of(f, p, wh);
            sv[r] = index_map(f, v)[p - v * wh];
        }
    }
}

// Workgroups [0, kSelectBlocks) evaluate the pair candidates; workgroup
// kSelectBlocks + t counts tile t.
__global__ void __launch_bounds__(kTileThreads) dense_count_kernel(const Buffers* __restrict__ bs, int use_cross) {
    const Buffers b = frame_buffers(bs);
    if (blockIdx.x < kSelectBlocks) {
        select_eval(b, blockIdx.x, use_cross);
        return;
    }
    const uint32_t tile = blockIdx.x - kSelectBlocks;
    const FrameDev& f = *b.frame;
    const uint32_t M = f.V * f.W * f.H;
    // clears the counters the rest of the index build accumulates into
    const uint32_t stride = (gridDim.x - kSelectBlocks) * blockDim.x;
    for (uint32_t i = tile * blockDim.x + threadIdx.x; i <= b.cap_nodes; i += stride) {
        b.n2t_count[i] = 0;
        b.row_up[i] = 0;
        b.row_low[i] = 0;
        b.row_kv2[i] = 0;
    }
    const uint32_t base = tile * kTilePixels;
    if (base >= M) return;
    uint32_t sv[kTileRounds];
    dense_tile_surfels(f, base, sv);
    uint32_t count = 0;
    bool bad = false;
#pragma unroll
    for (int r = 0; r < kTileRounds; ++r) {
        count += (sv[r] != kInvalidIndex && sv[r] < f.num_surfels) ? 1u : 0u;
        bad |= sv[r] != kInvalidIndex && sv[r] >= f.num_surfels;
    }
    if (bad) atomicOr(&b.counts->error, kErrNodes);
    count = block_sum_u32<kTileThreads>(count);
    if (threadIdx.x == 0) b.tile_count[tile] = count;
}

__global__ void __launch_bounds__(kTileThreads) dense_scatter_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const uint32_t wh = f.W * f.H, M = f.V * wh;
    const uint32_t base = blockIdx.x * kTilePixels;
    if (base >= M) return;
    constexpr int kWaves = kTileThreads / 64;
    __shared__ uint32_t wave_cnt[kTileRounds][kWaves];
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t sv[kTileRounds];
    dense_tile_surfels(f, base, sv);
    uint64_t mask[kTileRounds];
#pragma unroll
    for (int r = 0; r < kTileRounds; ++r) {
        mask[r] = __ballot(sv[r] != kInvalidIndex && sv[r] < f.num_surfels);
        if ((threadIdx.x & 63) == 0) wave_cnt[r][wave] = (uint32_t)__popcll(mask[r]);
    }
    __syncthreads();
    uint32_t running = b.tile_offset[blockIdx.x];
    const uint32_t cap = b.cap_pixels;
    // the surfels' knn and weights of kGatherRounds rounds are gathered before
    // any of their stores (the stores could alias them for the compiler)
    constexpr int kGatherRounds = 4;
#pragma unroll
    for (int r0 = 0; r0 < kTileRounds; r0 += kGatherRounds) {
        ushort4 knn_g[kGatherRounds];
        float4 w_g[kGatherRounds];
#pragma unroll
        for (int g = 0; g < kGatherRounds; ++g) {
            const uint32_t s = sv[r0 + g];
            const bool ok = s != kInvalidIndex && s < f.num_surfels;
            knn_g[g] = ok ? f.surfel_knn[s] : make_ushort4(0, 0, 0, 0);
            w_g[g] = ok ? f.surfel_w[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    #pragma unroll
        for (int g = 0; g < kGatherRounds; ++g) {
            const int r = r0 + g;
            uint32_t before = 0, round_total = 0;
    #pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t cw = wave_cnt[r][w];
                before += (w < (int)wave) ? cw : 0u;
                round_total += cw;
            }
            const uint32_t s = sv[r];
            if (s != kInvalidIndex && s < f.num_surfels) {
                const uint32_t dst = running + before + lane_prefix(mask[r]);
                if (dst < cap) {
                    const uint32_t p = base + r * kTileThreads + threadIdx.x;
                    const uint32_t v = view_of(f, p, wh), o = p - v * wh;
                    const uint32_t y = o / f.W, x = o - y * f.W;
                    ushort4 knn = knn_g[g];
                    if (!knn_ok(f, knn)) {
                        atomicOr(&b.counts->error, kErrNodes);
                        knn = make_ushort4(0, 0, 0, 0);
                    }
                    const float4 w = w_g[g];
                    b.dense_px[dst] = make_ushort4((uint16_t)x, (uint16_t)y, (uint16_t)v, 0);
                    b.dense_knn[dst] = knn;
                    // the record's first half (knn, r = 0, weights); J is written by
                    // the term evaluation before any use
                    const uint4 h0 = make_uint4((uint32_t)knn.x | ((uint32_t)knn.y << 16),
                                                (uint32_t)knn.z | ((uint32_t)knn.w << 16), 0u, 0u);
                    const uint4 h1 = make_uint4(__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z),
                                                __float_as_uint(w.w));
                    uint4* d = reinterpret_cast<uint4*>(b.dense + dst);
                    d[0] = h0;
                    d[1] = h1;
                    if (b.density) {
                        d = reinterpret_cast<uint4*>(b.density + dst);
                        d[0] = h0;
                        d[1] = h1;
                    }
                }
            }
            running += round_total;
        }
    }
}

// validDepthPixelInWindow (SparseCorrespondenceHandler.cu:67-93): the `break`
// leaves only the x loop, so the last row with a valid pixel wins; the centre
// is preferred.  Out-of-image taps read zero (cudaAddressModeBorder).
__device__ inline void depth_pixel_in_window(const FrameDev& f, uint32_t v, int cx, int cy, int& ox, int& oy) {
    ox = 0xFFFF;
    oy = 0xFFFF;
    for (int y = cy - 1; y <= cy + 1; ++y)
        for (int x = cx - 1; x <= cx + 1; ++x) {
            const float4 vv = fetch4(f, f.vertex[v], x, y);
            if (!is_zero_vertex(vv)) {
                ox = x & 0xFFFF;
                oy = y & 0xFFFF;
                break;
            }
        }
    if (!is_zero_vertex(fetch4(f, f.vertex[v], cx, cy))) {
        ox = cx;
        oy = cy;
    }
}

// validGeometryPixelInWindow, nearest-live-vertex form
// (SparseCorrespondenceHandler.cu:46-65); taps outside the image read index 0,
// which counts as valid, and a zero warp vertex.
__device__ inline void geometry_pixel_nearest(const FrameDev& f, uint32_t v, f3 obs, int cx, int cy, int half, int step,
                                              int& ox, int& oy) {
    ox = 0xFFFF;
    oy = 0xFFFF;
    float best = 100.0f;
    for (int y = cy - half; y <= cy + half; y += step)
        for (int x = cx - half; x <= cx + half; x += step) {
            if (fetch_index(f, v, x, y) == kInvalidIndex) continue;
            const float4 lv = fetch4(f, f.warp_v[v], x, y);
            const float d = (obs.x - lv.x) * (obs.x - lv.x) + (obs.y - lv.y) * (obs.y - lv.y) + (obs.z - lv.z) * (obs.z - lv.z);
            if (d < best) {
                ox = x & 0xFFFF;
                oy = y & 0xFFFF;
                best = d;
            }
        }
}

// Feature pair candidate g (chooseValidPixelKernel, USE_NEAREST_DISTANCE_PAIRS,
// SparseCorrespondenceHandler.cu:100-131).
__device__ inline bool choose_feature(const FrameDev& f, uint32_t g, P2PRec& rec, float4& canon) {
    uint32_t v = 0;
    while (v + 1 < f.V && g >= f.feat_base[v + 1]) ++v;
    const ushort4 c = f.feat[v][g - f.feat_base[v]];
    int dx, dy;
    depth_pixel_in_window(f, v, c.z, c.w, dx, dy);
    if (fetch_edge(f, v, dx, dy) == 1) return false;
    if (!(dx < (int)f.W && dy < (int)f.H)) return false;
    const float4 obs = f.vertex[v][dy * f.W + dx];
    int gx, gy;
    geometry_pixel_nearest(f, v, xyz(obs), c.x, c.y, 1, 1, gx, gy);
    if (!(gx < (int)f.W && gy < (int)f.H)) return false;
    const float4 ref = f.ref_v[v][gy * f.W + gx];
    const float d2 = (ref.x - obs.x) * (ref.x - obs.x) + (ref.y - obs.y) * (ref.y - obs.y) + (ref.z - obs.z) * (ref.z - obs.z);
    if (!(d2 < 2.5e-3f)) return false;
    // compactQueryValidPairsKernel (SparseCorrespondenceHandler.cu:140-190)
    const f3 can = f.init[v].apply(xyz(ref));
    const f3 tgt = f.c2w[v].apply(f.init[v].apply(xyz(obs)));
    const uint32_t s = f.index[v][gy * f.W + gx];
    if (s >= f.num_surfels) return false;
    rec.knn = f.surfel_knn[s];
    if (!knn_ok(f, rec.knn)) return false;
    rec.w = f.surfel_w[s];
    rec.pad0[0] = rec.pad0[1] = 0.f;
    rec.target = make_float4(tgt.x, tgt.y, tgt.z, 1.0f);
    rec.warped = make_float4(0.f, 0.f, 0.f, 1.0f);
    canon = make_float4(can.x, can.y, can.z, ref.w);
    return true;
}

// Cross-view candidate k.  The reference handler's kernels are absent
// (CrossViewCorrespondenceHandler.h:69-79, window half size 2, step 2); this
// is the feature rule applied across views: depth pixel in view 2, nearest
// warp vertex around pixel 1 in view 1 (back-traced to its previous-frame
// pixel through corrMap[v1] when the map holds a pair ending at pixel 1),
// reference vertex within 5 cm of the observed point expressed in camera 1.
__device__ inline bool choose_cross(const FrameDev& f, uint32_t k, P2PRec& rec, float4& canon) {
    const mvdr_cross_pair c = f.cross[k];
    const uint32_t v1 = c.view[0], v2 = c.view[1];
    if (v1 >= f.V || v2 >= f.V) return false;
    uint32_t cx = c.pixel[0], cy = c.pixel[1];
    if (f.corr[v1] && cx < f.W && cy < f.H) {
        const ushort4 m = f.corr[v1][cy * f.W + cx];
        if (m.z == cx && m.w == cy) { cx = m.x; cy = m.y; }
    }
    int dx, dy;
    depth_pixel_in_window(f, v2, c.pixel[2], c.pixel[3], dx, dy);
    if (fetch_edge(f, v2, dx, dy) == 1) return false;
    if (!(dx < (int)f.W && dy < (int)f.H)) return false;
    const float4 obs2 = f.vertex[v2][dy * f.W + dx];
    const f3 p0 = f.init[v2].apply(xyz(obs2));
    const f3 p1 = f.init_inv[v1].apply(p0);
    int gx, gy;
    geometry_pixel_nearest(f, v1, p1, cx, cy, 2, 2, gx, gy);
    if (!(gx < (int)f.W && gy < (int)f.H)) return false;
    const float4 ref = f.ref_v[v1][gy * f.W + gx];
    const float d2 = (ref.x - p1.x) * (ref.x - p1.x) + (ref.y - p1.y) * (ref.y - p1.y) + (ref.z - p1.z) * (ref.z - p1.z);
    if (!(d2 < 2.5e-3f)) return false;
    const f3 can = f.init[v1].apply(xyz(ref));
    const f3 tgt = f.c2w[v2].apply(p0);
    const uint32_t s = f.index[v1][gy * f.W + gx];
    if (s >= f.num_surfels) return false;
    rec.knn = f.surfel_knn[s];
    if (!knn_ok(f, rec.knn)) return false;
    rec.w = f.surfel_w[s];
    rec.pad0[0] = rec.pad0[1] = 0.f;
    rec.target = make_float4(tgt.x, tgt.y, tgt.z, 1.0f);
    rec.warped = make_float4(0.f, 0.f, 0.f, 1.0f);
    canon = make_float4(can.x, can.y, can.z, ref.w);
    return true;
}

// Selection, step 1: every candidate (feature pairs, then cross-view pairs)
// evaluated by its own thread into the staging slot of its candidate index.
// The candidate counts were checked against cap_feat / cap_cross at set-up.
// Runs as the first kSelectBlocks workgroups of the dense count's launch: it
// reads only the frame's inputs and writes only the staging arrays.
__device__ void select_eval(const Buffers& b, uint32_t block, int use_cross) {
    const FrameDev& f = *b.frame;
    const uint32_t nf = f.feat_base[f.V], total = nf + (use_cross ? f.num_cross : 0u);
    const uint32_t stride = kSelectBlocks * kTileThreads;
    for (uint32_t g = block * kTileThreads + threadIdx.x; g < total; g += stride) {
        P2PRec rec;
        float4 canon;
        const bool ok = g < nf ? choose_feature(f, g, rec, canon) : choose_cross(f, g - nf, rec, canon);
        b.sel_ok[g] = ok ? 1u : 0u;
        if (ok) {
            b.sel_rec[g] = rec;
            b.sel_canon[g] = canon;
        }
    }
}

// Step 2: one 1024-thread block compacts the feature pairs, another the
// cross-view pairs, each in candidate order with a running block-wide prefix
// (blocks 1 and 2 of the dense scan's launch).
__device__ void select_pairs(const Buffers& b, bool is_cross, int use_cross) {
    const FrameDev& f = *b.frame;
    if (is_cross && !use_cross) {
        if (threadIdx.x == 0) b.counts->cross = 0;
        return;
    }
    const uint32_t total = is_cross ? f.num_cross : f.feat_base[f.V];
    const uint32_t first = is_cross ? f.feat_base[f.V] : 0u;  // staging slot of candidate 0
    const uint32_t cap = is_cross ? b.cap_cross : b.cap_feat;
    P2PRec* out = is_cross ? b.cross : b.feat;
    float4* canon_out = is_cross ? b.cross_canon : b.feat_canon;
    __shared__ uint32_t wave_cnt[2][16];
    uint32_t running = 0;
    const uint32_t wave = threadIdx.x >> 6;
    int parity = 0;
    for (uint32_t base = 0; base < total; base += 1024, parity ^= 1) {
        const uint32_t g = base + threadIdx.x;
        const bool ok = g < total && b.sel_ok[first + g] != 0u;
        const uint64_t mask = __ballot(ok);
        if ((threadIdx.x & 63) == 0) wave_cnt[parity][wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t before = 0, chunk = 0;
        for (int w = 0; w < 16; ++w) {
            const uint32_t cw = wave_cnt[parity][w];
            before += (w < (int)wave) ? cw : 0u;
            chunk += cw;
        }
        if (ok) {
            const uint32_t dst = running + before + lane_prefix(mask);
            if (dst < cap) {
                out[dst] = b.sel_rec[first + g];
                canon_out[dst] = b.sel_canon[first + g];
            }
        }
        running += chunk;
    }
    if (threadIdx.x == 0) {
        if (running > cap) atomicOr(&b.counts->error, is_cross ? kErrCross : kErrFeat);
        if (is_cross) b.counts->cross = min(running, cap);
        else b.counts->feat = min(running, cap);
    }
}

// --------------------------------------------------------------- Node2Term
// Global term index: [dense | smooth | foreground | feature | cross]
// (TermTypeOffset, base/term_offset_types.h:21-64).
__device__ inline void term_offsets(const Counts& c, uint32_t E, uint32_t off[5]) {
    off[0] = c.dense;
    off[1] = off[0] + E;
    off[2] = off[1] + c.fg;
    off[3] = off[2] + c.feat;
    off[4] = off[3] + c.cross;
}

// Node keys of term t (buildTermKeyValueKernel, Node2TermsIndex.cu:8-110):
// four knn nodes, or the two ends of a graph edge.  Returns the key count.
__device__ inline int term_nodes(const Buffers& b, const FrameDev& f, const uint32_t off[5], uint32_t t, ushort4& k) {
    if (t < off[0]) k = b.dense_knn[t];  // 8 B per term: a sector serves four neighbouring terms
    else if (t < off[1]) {
        const ushort2 e = f.graph[t - off[0]];
        k = make_ushort4(e.x, e.y, 0, 0);
        return 2;
    } else if (t < off[2]) k = b.fg[t - off[1]].knn;
    else if (t < off[3]) k = b.feat[t - off[2]].knn;
    else k = b.cross[t - off[3]].knn;
    return 4;
}

__device__ inline uint32_t slot(ushort4 k, int q) { return q == 0 ? k.x : (q == 1 ? k.y : (q == 2 ? k.z : k.w)); }

// The reference builds Node2Term by a stable key sort of (node, term) pairs
// generated in term order (Node2TermsIndex.cu:112-166): each node's list holds
// its terms ascending, a term once per knn slot naming the node.  Here the same
// lists come from a stable counting sort over chunks of kChunkTerms (or
// kBigChunkTerms) terms, one wave per chunk:
//   n2t_hist    per-chunk node histograms H[c][n] (LDS counters)
//   n2t_colscan H[c][n] -> exclusive prefix over chunks, per-node totals
//   node_scan   per-node totals -> list offsets
//   n2t_place   each chunk writes its entries at offset[n] + H[c][n] + rank,
//               rank = position among the chunk's earlier entries of node n
// No global atomics and no sorts: the result is deterministic by construction.
// A chunk holds at most 4 kBigChunkTerms entries, so the per-chunk node
// counters fit 16 bits: two share a 32-bit LDS word (half the LDS of u32
// counters, twice the resident workgroups).
constexpr uint32_t kChunkTerms = 512;      // chunk of a launch the chip holds at once; the capacity unit
constexpr uint32_t kBigChunkTerms = 2048;  // chunk of a larger launch (half the histogram rows)
constexpr uint32_t kHistNodes = 16384;  // largest LDS node tile (32 KB of u16 counters)
static_assert(4 * kBigChunkTerms <= 0xFFFFu && kBigChunkTerms % kChunkTerms == 0,
              "per-chunk node counts must fit 16 bits");

// LDS node tile of the counting-sort kernels (dynamic shared memory).
uint32_t hist_tile(uint32_t nodes) { return std::min<uint32_t>(nodes, kHistNodes); }
size_t hist_tile_bytes(uint32_t tile) { return (size_t)(tile + 1) / 2 * sizeof(uint32_t); }
// Chunk workgroups (over all frames of a launch) up to which the placement
// keeps u32 counters: what the chip holds at once with 16 KB tiles (9 per CU).
constexpr size_t kPlaceResident = 2304;

__device__ inline uint32_t num_chunks(const Counts& c, uint32_t chunk) { return (c.off[4] + chunk - 1) / chunk; }

__global__ void __launch_bounds__(64) n2t_hist_kernel(const Buffers* __restrict__ bs, uint32_t tile, uint32_t chunk) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    Counts& c = *b.counts;
    uint32_t off[5];
    term_offsets(c, f.E, off);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.smooth = f.E;
        for (int k = 0; k < 5; ++k) c.off[k] = off[k];
    }
    const uint32_t t0 = blockIdx.x * chunk;
    if (t0 >= off[4]) return;
    const uint32_t t1 = min(off[4], t0 + chunk);
    const uint32_t N = f.N;
    extern __shared__ uint32_t hist[];
    uint32_t* H = b.n2t_hist + (size_t)blockIdx.x * N;
    for (uint32_t n0 = 0; n0 < N; n0 += tile) {
        const uint32_t span = min(tile, N - n0);
        for (uint32_t k = threadIdx.x; k < (span + 1) / 2; k += 64) hist[k] = 0;
        wave_lds_sync();
        // four rounds of term nodes in flight, then their counts
        for (uint32_t tb = t0 + threadIdx.x; tb < t1; tb += 4 * 64) {
            ushort4 k[4];
            int m[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t t = tb + 64 * r;
                k[r] = make_ushort4(0, 0, 0, 0);
                m[r] = t < t1 ? term_nodes(b, f, off, t, k[r]) : 0;
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t n = slot(k[r], q);
                    if (q >= m[r]) continue;
                    if (n >= N) {
                        if (n0 == 0) atomicOr(&c.error, kErrNodes);
                    } else if (n >= n0 && n - n0 < span) {
                        const uint32_t l = n - n0;
                        atomicAdd(&hist[l >> 1], 1u << (16 * (l & 1)));
                    }
                }
        }
        wave_lds_sync();
        for (uint32_t k = threadIdx.x; k < span; k += 64) H[n0 + k] = (hist[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
        wave_lds_sync();
    }
}

// One lane per node, 16 waves splitting the chunk range: H[c][n] becomes the
// number of node-n entries in chunks before c; n2t_count[n] the node's total.
constexpr uint32_t kScanWaves = 16;

__global__ void __launch_bounds__(64 * kScanWaves) n2t_colscan_kernel(const Buffers* __restrict__ bs, uint32_t chunk) {
    const Buffers b = frame_buffers(bs);
    const uint32_t N = b.frame->N;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t n = blockIdx.x * 64 + lane;
    const uint32_t C = num_chunks(*b.counts, chunk);
    const uint32_t per = (C + kScanWaves - 1) / kScanWaves;
    const uint32_t c0 = min(C, w * per), c1 = min(C, c0 + per);
    __shared__ uint32_t part[kScanWaves][64];
    uint32_t sum = 0;
    if (n < N) {
#pragma unroll 8
        for (uint32_t ch = c0; ch < c1; ++ch) sum += b.n2t_hist[(size_t)ch * N + n];
    }
    part[w][lane] = sum;
    __syncthreads();
    uint32_t run = 0, total = 0;
    for (uint32_t k = 0; k < kScanWaves; ++k) {
        run += k < w ? part[k][lane] : 0u;
        total += part[k][lane];
    }
    if (n >= N) return;
#pragma unroll 8
    for (uint32_t ch = c0; ch < c1; ++ch) {
        uint32_t* h = b.n2t_hist + (size_t)ch * N + n;
        const uint32_t v = *h;
        *h = run;
        run += v;
    }
    if (w == kScanWaves - 1) b.n2t_count[n] = total;
}

// Bitonic sort of a wave's 256 keys, ascending; key e sits in lane e / 4,
// register e % 4.  Pairs 4 or more keys apart are exchanged across lanes.
__device__ inline void wave_sort256(uint32_t k[4]) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t size = 2; size <= 256; size <<= 1) {
#pragma unroll
        for (uint32_t d = size >> 1; d > 0; d >>= 1) {
            if (d >= 4) {
                // bits d and size of e = 4 lane + j come from the lane
                const bool keep_min = (((4 * lane) & size) == 0) == (((4 * lane) & d) == 0);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t o = (uint32_t)__shfl_xor((int)k[j], (int)(d >> 2), 64);
                    k[j] = keep_min ? min(k[j], o) : max(k[j], o);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j & d) continue;
                    const bool up = ((4 * lane + j) & size) == 0;
                    const uint32_t lo = min(k[j], k[j ^ d]), hi = max(k[j], k[j ^ d]);
                    k[j] = up ? lo : hi;
                    k[j ^ d] = up ? hi : lo;
                }
            }
        }
    }
}

// with_nodes: also each entry's node (n2t_node), which only the segmented
// assembly of the plain path reads.
// kLdsBase: the tile's u32 LDS counters start at each node's list offset plus
// its entries in earlier chunks (loaded once per tile: a small grid, all of it
// resident at once); otherwise u16 counters start at zero and each entry loads
// that base while its round sorts (half the LDS: twice the resident
// workgroups of a large grid).  Both place every entry at the same position.
template <bool kLdsBase>
__global__ void __launch_bounds__(64) n2t_place_kernel(const Buffers* __restrict__ bs, uint32_t tile, int with_nodes, uint32_t chunk) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t t0 = blockIdx.x * chunk;
    if (t0 >= c.off[4]) return;
    const uint32_t t1 = min(c.off[4], t0 + chunk);
    const uint32_t N = f.N;
    const uint32_t lane = threadIdx.x;
    extern __shared__ uint32_t cur_words[];
    using u16_lds = uint16_t __attribute__((may_alias));
    u16_lds* cur16 = reinterpret_cast<u16_lds*>(cur_words);
    __shared__ uint4 rank_of[64];  // rank of entry (lane, slot) among the round's entries of its node
    const uint32_t* H = b.n2t_hist + (size_t)blockIdx.x * N;
    for (uint32_t n0 = 0; n0 < N; n0 += tile) {
        const uint32_t span = min(tile, N - n0);
        if (kLdsBase) {
#pragma unroll 8
            for (uint32_t k = lane; k < span; k += 64) cur_words[k] = b.n2t_offset[n0 + k] + H[n0 + k];
        } else {
            for (uint32_t k = lane; k < (span + 1) / 2; k += 64) cur_words[k] = 0;
        }
        // the next round's term nodes load while this round sorts
        ushort4 kn_next = make_ushort4(0, 0, 0, 0);
        int m_next = t0 + lane < t1 ? term_nodes(b, f, c.off, t0 + lane, kn_next) : 0;
        for (uint32_t tb = t0; tb < t1; tb += 64) {
            const uint32_t t = tb + lane;
            const ushort4 kn = kn_next;
            const int m = m_next;
            if (t + 64 < t1) m_next = term_nodes(b, f, c.off, t + 64, kn_next);
            else m_next = 0;
            uint32_t n[4] = {kn.x, kn.y, kn.z, kn.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (!(q < m && n[q] >= n0 && n[q] - n0 < span)) n[q] = 0xFFFFFFFFu;  // not placed here
            // where this chunk's entries of each node start: the node's list
            // offset plus its entries in earlier chunks (in flight during the sort)
            uint32_t base[4] = {0u, 0u, 0u, 0u};
            if (!kLdsBase) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (n[q] != 0xFFFFFFFFu) base[q] = b.n2t_offset[n[q]] + H[n[q]];
            }
            // the round's 256 entries sorted by (node, term, slot): key = local
            // node << 8 | 4 lane + slot; entries not placed here sort last
            uint32_t k[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) k[q] = n[q] == 0xFFFFFFFFu ? 0xFFFFFFFFu : ((n[q] - n0) << 8) | (4 * lane + q);
            wave_sort256(k);
            // runs of equal nodes: rank = sorted position - start of the run
            uint32_t nd[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) nd[j] = k[j] >> 8;
            const uint32_t prev = (uint32_t)__shfl_up((int)nd[3], 1, 64);
            const uint32_t next = (uint32_t)__shfl_down((int)nd[0], 1, 64);
            uint32_t start[4], run = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool head = j == 0 ? (lane == 0 || prev != nd[0]) : nd[j - 1] != nd[j];
                if (head) run = 4 * lane + j;
                start[j] = run;  // start of the run within this lane (0 if it began earlier)
            }
            // inclusive max-scan of the lanes' last run starts, then exclusive
            uint32_t scan = run;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = (uint32_t)__shfl_up((int)scan, o, 64);
                if ((int)lane >= o) scan = max(scan, v);
            }
            uint32_t before = (uint32_t)__shfl_up((int)scan, 1, 64);
            if (lane == 0) before = 0;
            uint32_t rank[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) rank[j] = 4 * lane + j - max(start[j], before);
            uint32_t* rk = reinterpret_cast<uint32_t*>(rank_of);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (k[j] != 0xFFFFFFFFu) rk[k[j] & 255] = rank[j];
            wave_lds_sync();
            const uint4 r4 = rank_of[lane];
            const uint32_t mine[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (n[q] != 0xFFFFFFFFu) {
                    const uint32_t l = n[q] - n0;
                    const uint32_t at = kLdsBase ? cur_words[l] + mine[q] : base[q] + cur16[l] + mine[q];
                    b.n2t_index[at] = t;
                    if (with_nodes) b.n2t_node[at] = (uint16_t)n[q];
                }
            }
            wave_lds_sync();  // every lane has read cur before the updates
            // the last entry of each run advances its node's counter
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool tail = j < 3 ? nd[j + 1] != nd[j] : (lane == 63 || next != nd[3]);
                if (k[j] != 0xFFFFFFFFu && tail) {
                    if (kLdsBase) cur_words[nd[j]] += rank[j] + 1;
                    else cur16[nd[j]] = (uint16_t)(cur16[nd[j]] + rank[j] + 1);
                }
            }
            wave_lds_sync();
        }
    }
}

// Exclusive scans over the N node counters.  which = 0: Node2Term lists;
// which = 1: pairs (i, j > i) per row (block 0), lower entries per row
// (block 1), pair-term entries per row (block 2).
__global__ void __launch_bounds__(1024) node_scan_kernel(const Buffers* __restrict__ bs, int which) {
    const Buffers b = frame_buffers(bs);
    const uint32_t N = b.frame->N;
    Counts& c = *b.counts;
    const int job = which == 0 ? 0 : 1 + (int)blockIdx.x;
    const uint32_t* in = job == 0 ? b.n2t_count : (job == 1 ? b.row_up : (job == 2 ? b.row_low : b.row_kv2));
    uint32_t* out = job == 0 ? b.n2t_offset : (job == 1 ? b.pair_offset : (job == 2 ? b.low_offset : b.kv2_offset));
    uint32_t mx = 0;
    const uint32_t total = block_exclusive_scan_1024(in, out, N, nullptr, &mx);
    if (threadIdx.x == 0) {
        out[N] = total;
        if (job == 0) {
            if (total > b.cap_kv1) atomicOr(&c.error, kErrKv1);
            c.kv1 = total;
            c.max_seg = mx;
            c.kv2 = c.pairs = c.max_row = 0;  // set by the pair scans when the pair index is built
        } else if (job == 1) {
            if (total > b.cap_pairs) atomicOr(&c.error, kErrPairs);
            c.pairs = total;
            c.max_row = mx;
        } else if (job == 3) {
            if (total > b.cap_kv2) atomicOr(&c.error, kErrKv2);
            c.kv2 = total;
        }
    }
}

// ------------------------------------------------------------- node pairs
// The unique pairs (i, j), i < j, of nodes sharing a term -- the key set of the
// reference's NodePair2Term index (NodePair2TermsIndex.cu:11-185) -- and, per
// block row, the lower-triangle entries (NodePairRowColumnIndex.cu:19-179).
// One wave per row i collects the partner nodes of its Node2Term list in an LDS
// hash set, then emits them ascending (wave minimum per step).  The blocks are
// single waves, so LDS traffic between lanes needs no workgroup barrier.
constexpr uint32_t kRowHash = kRowPartners;  // partner set capacity per row

// Partners of node i in term t with the reference's pair multiplicity: every
// knn slot q with knn[q] != i, or the other end of a non-loop graph edge.
__device__ inline ushort4 partners_of(ushort4 k, int m, uint32_t i) {
    constexpr uint16_t kNone = 0xFFFF;
    if (m == 2) return make_ushort4(k.x != k.y ? (k.x == i ? k.y : k.x) : kNone, kNone, kNone, kNone);
    return make_ushort4(k.x != i ? k.x : kNone, k.y != i ? k.y : kNone, k.z != i ? k.z : kNone, k.w != i ? k.w : kNone);
}

__device__ inline ushort4 term_partners(const Buffers& b, const FrameDev& f, const uint32_t off[5], uint32_t t, uint32_t i) {
    ushort4 k;
    const int m = term_nodes(b, f, off, t, k);
    return partners_of(k, m, i);
}

__device__ inline uint32_t lower_bound_lds(const uint32_t* a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// kRowCount: pairs above / below the diagonal and pair-term entries of row i.
// kRowGram: the same counts, the row's partner lists (row_part, emitted at
// their offsets by pair_emit_kernel once the counts are scanned), and for every
// Node2Term entry the weight column of each knn slot holding an upper partner
// j (2 + rank of j among the row's upper partners; 0xFF otherwise) -- what the
// row Gram assembly needs instead of pair-term lists.
// kRowTerms: pair keys, lower entries, and the pair-term lists (NodePair2Term)
// with their ranges: for pair (i, j) every Node2Term entry of row i once per
// knn slot holding j, in Node2Term order -- the order of the reference's stable
// key sort.  kWaves waves per row (one in the counting modes, kRowWaves for
// the pair-term lists, where wave w places a contiguous w-th share of the
// row's entries).
constexpr uint32_t kRowWaves = 4;
enum : int { kRowCount = 0, kRowGram = 1, kRowTerms = 2 };

template <int kMode, uint32_t kWaves>
__global__ void __launch_bounds__(64 * kWaves) pair_rows_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    Counts& c = *b.counts;
    if (c.error) return;
    // XCD-aware row order: an XCD takes a contiguous range of rows, whose term
    // records overlap, into its L2
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x, f.N);
    if (i == 0xFFFFFFFFu) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    __shared__ uint32_t set[kRowHash];
    __shared__ uint32_t ul[kRowHash];             // ascending partners j > i
    __shared__ uint32_t ll[kRowHash];             // ascending partners k < i
    __shared__ uint32_t cnt[kWaves][kRowHash]; // entries per pair and wave, then next positions
    __shared__ uint32_t start[kRowHash + 1];      // pair-term range starts
    __shared__ uint32_t part[kWaves];
    __shared__ uint32_t nup, nlow, ovf;
    for (uint32_t k = tid; k < kRowHash; k += 64 * kWaves) set[k] = 0xFFFFFFFFu;
    if (tid == 0) ovf = 0;
    __syncthreads();
    const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
    uint32_t upper_entries = 0;
    bool overflow = false;
    // four entries per lane in flight: their indices, then their records; a
    // row of at most one batch keeps its entries' term nodes for the column pass
    constexpr uint32_t kBatch = 4;
    const bool one_batch = end - beg <= kBatch * 64 * kWaves;
    ushort4 kq[kBatch];
    int mq[kBatch];
#pragma unroll
    for (int u = 0; u < (int)kBatch; ++u) {
        kq[u] = make_ushort4(0, 0, 0, 0);
        mq[u] = 0;
    }
    for (uint32_t p0 = beg + tid; p0 < end; p0 += kBatch * 64 * kWaves) {
#pragma unroll
        for (int u = 0; u < (int)kBatch; ++u) {
            const uint32_t p = p0 + u * 64 * kWaves;
            kq[u] = make_ushort4(0, 0, 0, 0);
            mq[u] = p < end ? term_nodes(b, f, c.off, b.n2t_index[p], kq[u]) : 0;
        }
        ushort4 oq[kBatch];
#pragma unroll
        for (int u = 0; u < (int)kBatch; ++u)
            oq[u] = mq[u] ? partners_of(kq[u], mq[u], i) : make_ushort4(0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF);
#pragma unroll
        for (int uq = 0; uq < (int)(4 * kBatch); ++uq) {
            const uint32_t v = slot(oq[uq / 4], uq % 4);
            if (v == 0xFFFF) continue;
            upper_entries += v > i ? 1u : 0u;
            uint32_t h = (v * 2654435761u) >> 23;  // 9 bits
            for (uint32_t probe = 0;; ++probe) {
                if (probe == kRowHash) {
                    overflow = true;
                    break;
                }
                // most partners are already present: a plain read settles them
                const uint32_t seen = set[h];
                if (seen == v) break;
                if (seen == 0xFFFFFFFFu) {
                    const uint32_t old = atomicCAS(&set[h], 0xFFFFFFFFu, v);
                    if (old == 0xFFFFFFFFu || old == v) break;
                }
                h = (h + 1) & (kRowHash - 1);
            }
        }
    }
    if (overflow) ovf = 1;
    upper_entries = wave_sum(upper_entries);
    if (lane == 0) part[w] = upper_entries;
    __syncthreads();
    if (ovf) {
        if (tid == 0) atomicOr(&c.error, kErrRow);
        return;
    }
    if (w == 0) {
        // ascending partner lists: the set's entries are compacted as
        // (node << 16 | slot) and each is ranked by a count of the smaller ones
        // (nodes are distinct, so the packed order is the node order); the
        // lists are written at the ranks.  kRowGram also writes each upper
        // partner's weight column into its set slot (node | column << 16).
        uint32_t* cand = ll;  // ll is rewritten with the lower list once every rank is known
        uint32_t total = 0;
#pragma unroll
        for (int k = 0; k < (int)(kRowHash / 64); ++k) {
            const uint32_t v = set[lane + 64 * k];
            const uint64_t full = __ballot(v != 0xFFFFFFFFu);
            if (v != 0xFFFFFFFFu) cand[total + (uint32_t)__popcll(full & below)] = v << 16 | (lane + 64 * k);
            total += (uint32_t)__popcll(full);
        }
        wave_lds_sync();
        uint32_t low = 0;
        uint32_t key[kRowHash / 64], rank[kRowHash / 64];
#pragma unroll
        for (int r = 0; r < (int)(kRowHash / 64); ++r) {
            const uint32_t j = 64 * r + lane;
            key[r] = j < total ? cand[j] : 0xFFFFFFFFu;
            rank[r] = 0;
            if (64 * (uint32_t)r >= total) continue;  // uniform
            low += (uint32_t)__popcll(__ballot(j < total && (key[r] >> 16) < i));
            if (kMode != kRowCount)
                for (uint32_t q = 0; q < total; ++q) rank[r] += cand[q] < key[r] ? 1u : 0u;
        }
        const uint32_t up = total - low;
        if (kMode != kRowCount) {
            wave_lds_sync();  // every read of cand is done before the lists overwrite ll
#pragma unroll
            for (int r = 0; r < (int)(kRowHash / 64); ++r) {
                if (64 * (uint32_t)r + lane >= total) continue;
                const uint32_t v = key[r] >> 16;
                if (v > i) ul[rank[r] - low] = v;
                else ll[rank[r]] = v;
                if (kMode == kRowGram && v > i) set[key[r] & 0xFFFFu] = v | (2 + rank[r] - low) << 16;
            }
        }
        if (lane == 0) {
            nup = up;
            nlow = low;
        }
    }
    __syncthreads();
    const uint32_t up = nup, low = nlow;
    if (kMode == kRowCount || kMode == kRowGram) {
        if (tid == 0) {
            b.row_up[i] = up;
            b.row_low[i] = low;
            uint32_t s2 = 0;
            for (uint32_t k = 0; k < kWaves; ++k) s2 += part[k];
            b.row_kv2[i] = s2;
        }
        if (kMode == kRowCount) return;
        uint16_t* rp = b.row_part + (size_t)i * kRowHash;
        for (uint32_t u = tid; u < up; u += 64 * kWaves) rp[u] = (uint16_t)ul[u];
        for (uint32_t k = tid; k < low; k += 64 * kWaves) rp[up + k] = (uint16_t)ll[k];
        for (uint32_t p0 = beg + tid; p0 < end; p0 += kBatch * 64 * kWaves) {
            if (!one_batch) {
#pragma unroll
                for (int u = 0; u < (int)kBatch; ++u) {
                    const uint32_t p = p0 + u * 64 * kWaves;
                    kq[u] = make_ushort4(0, 0, 0, 0);
                    mq[u] = p < end ? term_nodes(b, f, c.off, b.n2t_index[p], kq[u]) : 0;
                }
            }
#pragma unroll
            for (int u = 0; u < (int)kBatch; ++u) {
                const uint32_t p = p0 + u * 64 * kWaves;
                if (p >= end) continue;
                uint32_t packed = 0xFFFFFFFFu;
                if (mq[u] == 4 && up + 2 < 0xFF) {
                    packed = 0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t v = slot(kq[u], q);
                        uint32_t col = 0xFFu;
                        if (v != 0xFFFF && v > i) {
                            // v is in the set: its slot holds its column (above the node)
                            uint32_t h = (v * 2654435761u) >> 23;
                            for (uint32_t probe = 0; probe < kRowHash && (set[h] & 0xFFFFu) != v; ++probe)
                                h = (h + 1) & (kRowHash - 1);
                            col = (set[h] & 0xFFFFu) == v ? set[h] >> 16 : 0xFFu;
                        }
                        packed |= col << (8 * q);
                    }
                }
                b.n2t_col[p] = packed;
            }
        }
        return;
    }
    const uint32_t up_base = b.pair_offset[i], low_base = b.low_offset[i];
    const uint32_t len = end - beg, per = (len + kWaves - 1) / kWaves;
    const uint32_t wb = beg + min(len, w * per), we = beg + min(len, (w + 1) * per);
    uint32_t* run = cnt[w];
    if (kMode == kRowTerms && up > 0) {
        // pass 1: this wave's entries per pair
        for (uint32_t u = lane; u < up; u += 64) run[u] = 0;
        wave_lds_sync();
        for (uint32_t p = wb + lane; p < we; p += 64) {
            const ushort4 o = term_partners(b, f, c.off, b.n2t_index[p], i);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t v = slot(o, q);
                if (v != 0xFFFF && v > i) atomicAdd(&run[lower_bound_lds(ul, up, v)], 1u);
            }
        }
        __syncthreads();
        // per pair: range start, then every wave's first position
        if (tid == 0) {
            uint32_t acc = b.kv2_offset[i];
            for (uint32_t u = 0; u < up; ++u) {
                start[u] = acc;
                for (uint32_t k = 0; k < kWaves; ++k) {
                    const uint32_t n = cnt[k][u];
                    cnt[k][u] = acc;
                    acc += n;
                }
            }
            start[up] = acc;
        }
        __syncthreads();
        for (uint32_t u = tid; u < up; u += 64 * kWaves)
            b.pair_range[up_base + u] = make_uint2(start[u], start[u + 1]);
        // pass 2: entries in Node2Term order, slot order within an entry
        for (uint32_t base = wb; base < we; base += 64) {
            const uint32_t p = base + lane;
            uint32_t t = 0;
            uint32_t su[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
            if (p < we) {
                t = b.n2t_index[p];
                const ushort4 o = term_partners(b, f, c.off, t, i);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t v = slot(o, q);
                    if (v != 0xFFFF && v > i) su[q] = lower_bound_lds(ul, up, v);
                }
            }
            uint64_t pend = __ballot(su[0] != 0xFFFFFFFFu || su[1] != 0xFFFFFFFFu || su[2] != 0xFFFFFFFFu ||
                                     su[3] != 0xFFFFFFFFu);
            while (pend) {
                const int leader = __builtin_ctzll(pend);
                const uint32_t lsu = su[0] != 0xFFFFFFFFu ? su[0]
                                   : (su[1] != 0xFFFFFFFFu ? su[1] : (su[2] != 0xFFFFFFFFu ? su[2] : su[3]));
                const uint32_t u = (uint32_t)__shfl((int)lsu, leader, 64);
                uint64_t hit[4];
                uint32_t total = 0, before = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    hit[q] = __ballot(su[q] == u);
                    total += (uint32_t)__popcll(hit[q]);
                    before += (uint32_t)__popcll(hit[q] & below);
                }
                uint32_t at = 0;
                if ((int)lane == leader) at = atomicAdd(&run[u], total);
                at = (uint32_t)__shfl((int)at, leader, 64);
                uint32_t own = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (su[q] == u) {
                        b.pair_terms[at + before + own] = t;
                        b.kv2_pair[at + before + own] = up_base + u;
                        ++own;
                        su[q] = 0xFFFFFFFFu;
                    }
                pend = __ballot(su[0] != 0xFFFFFFFFu || su[1] != 0xFFFFFFFFu || su[2] != 0xFFFFFFFFu ||
                                su[3] != 0xFFFFFFFFu);
            }
        }
    }
    for (uint32_t u = tid; u < up; u += 64 * kWaves) b.pair_key[up_base + u] = (i << 16) | ul[u];
    for (uint32_t k = tid; k < low; k += 64 * kWaves) b.low_ent[low_base + k] = make_uint2(ll[k], 0u);
}

// Pair keys and lower entries of every row at their scanned offsets, from the
// partner lists kRowGram left in row_part; graph-edge counts cleared.  A
// quarter wave per row.
__global__ void __launch_bounds__(256) pair_emit_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t sub = threadIdx.x & 15;
    const uint32_t groups = gridDim.x * 16;
    for (uint32_t i = blockIdx.x * 16 + threadIdx.x / 16; i < f.N; i += groups) {
        const uint32_t up = b.row_up[i], low = b.row_low[i];
        const uint32_t up_base = b.pair_offset[i], low_base = b.low_offset[i];
        const uint16_t* rp = b.row_part + (size_t)i * kRowHash;
        for (uint32_t u = sub; u < up; u += 16) {
            b.pair_key[up_base + u] = (i << 16) | rp[u];
            b.pair_edge_n[up_base + u] = 0;
        }
        for (uint32_t k = sub; k < low; k += 16) b.low_ent[low_base + k] = make_uint2(rp[up + k], 0u);
    }
}

// Graph edges of every pair (i, j): the row Gram assembly adds their pair
// blocks in the fix-up, in edge order.  The first kPairEdges edges found are
// kept; the count tells the fix-up when a pair has more (a node graph with
// repeated edges) and it walks the row's Node2Term list instead.
__device__ void pair_edges(const Buffers& b, uint32_t block, uint32_t blocks) {
    const FrameDev& f = *b.frame;
    Counts& c = *b.counts;
    const uint32_t stride = blocks * blockDim.x;
    for (uint32_t e = block * blockDim.x + threadIdx.x; e < f.E; e += stride) {
        const ushort2 g = f.graph[e];
        if (g.x == g.y) continue;
        const uint32_t i = min(g.x, g.y), j = max(g.x, g.y);
        const uint32_t key = (i << 16) | j;
        uint32_t lo = b.pair_offset[i], hi = b.pair_offset[i + 1];
        const uint32_t end = hi;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (b.pair_key[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        if (lo == end || b.pair_key[lo] != key) {
            atomicOr(&c.error, kErrPairs);
            continue;
        }
        const uint32_t at = atomicAdd(&b.pair_edge_n[lo], 1u);
        if (at < kPairEdges) b.pair_edge[kPairEdges * (size_t)lo + at] = e;
    }
}

// Pair index of every lower entry (k, .) of row j: the position of (k, j) in
// row k's ascending upper list (and, in the mixed mode, the inverse map).
constexpr uint32_t kLinkLanes = 8;  // lanes per row of the pair link

// Workgroups [0, link_blocks) link the lower entries; the rest (row Gram
// index only) collect the graph edges of each pair (pair_edges).
__global__ void __launch_bounds__(256) pair_link_kernel(const Buffers* __restrict__ bs, uint32_t link_blocks) {
    const Buffers b = frame_buffers(bs);
    const Counts& c = *b.counts;
    if (c.error) return;
    if (blockIdx.x >= link_blocks) {
        pair_edges(b, blockIdx.x - link_blocks, gridDim.x - link_blocks);
        return;
    }
    const uint32_t N = b.frame->N;
    const uint32_t sub = threadIdx.x % kLinkLanes;
    const uint32_t stride = link_blocks * blockDim.x / kLinkLanes;
    for (uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) / kLinkLanes; j < N; j += stride) {
        const uint32_t l1 = b.low_offset[j + 1];
        for (uint32_t l = b.low_offset[j] + sub; l < l1; l += kLinkLanes) {
            const uint32_t k = b.low_ent[l].x;
            const uint32_t key = (k << 16) | j;
            uint32_t lo = b.pair_offset[k], hi = b.pair_offset[k + 1];
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (b.pair_key[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            if (lo == b.pair_offset[k + 1] || b.pair_key[lo] != key) {
                atomicOr(&b.counts->error, kErrPairs);
                continue;
            }
            b.low_ent[l].y = lo;
            if (b.low_pos) b.low_pos[lo] = l;
        }
    }
}

// ------------------------------------------------------------- host launch
// mode kIndexLists: the NodePair2Term lists of the reference's assembly;
// kIndexGram: pair keys, knn-slot columns and graph edges per pair for the row
// Gram assembly; kIndexNodes: Node2Term alone (the matrix-free solve).
// The frame's Counts start at zero: launch_copy_se3 clears them earlier on
// the same stream.  The pair selection rides in the dense count's and the
// dense scan's launches (its own workgroups; disjoint outputs).
void launch_index_build(const Batch& B, bool use_foreground, bool use_cross, int mode, hipStream_t stream) {
    const bool gram = mode == kIndexGram;
    const uint32_t F = B.frames;
    const int cross = use_cross ? 1 : 0;
    hipLaunchKernelGGL(dense_count_kernel, dim3(kSelectBlocks + B.cap_tiles, F), dim3(kTileThreads), 0, stream, B.dev,
                       cross);
    hipLaunchKernelGGL(compact_scan_kernel, dim3(3, F), dim3(1024), 0, stream, B.dev, 0, cross);
    hipLaunchKernelGGL(dense_scatter_kernel, dim3(B.cap_tiles, F), dim3(kTileThreads), 0, stream, B.dev);
    if (use_foreground) {
        hipLaunchKernelGGL(compact_count_kernel<ForegroundPred>, dim3(B.cap_tiles, F), dim3(kTileThreads), 0, stream,
                           B.dev, ForegroundPred{});
        hipLaunchKernelGGL(compact_scan_kernel, dim3(1, F), dim3(1024), 0, stream, B.dev, 1, -1);
        hipLaunchKernelGGL(compact_scatter_kernel<ForegroundPred>, dim3(B.cap_tiles, F), dim3(kTileThreads), 0, stream,
                           B.dev, ForegroundPred{}, 1);
    }
    const uint32_t tile = hist_tile(B.nodes);
    // the capacity counts kChunkTerms chunks; a launch larger than the chip
    // holds at once runs kBigChunkTerms chunks (fewer histogram rows to scan)
    const bool resident = (size_t)B.cap_chunks * F <= kPlaceResident;
    const uint32_t chunk = resident ? kChunkTerms : kBigChunkTerms;
    const uint32_t chunks = (uint32_t)(((uint64_t)B.cap_chunks * kChunkTerms + chunk - 1) / chunk);
    hipLaunchKernelGGL(n2t_hist_kernel, dim3(chunks, F), dim3(64), hist_tile_bytes(tile), stream, B.dev, tile, chunk);
    hipLaunchKernelGGL(n2t_colscan_kernel, dim3((B.nodes + 63) / 64, F), dim3(64 * kScanWaves), 0, stream, B.dev,
                       chunk);
    hipLaunchKernelGGL(node_scan_kernel, dim3(1, F), dim3(1024), 0, stream, B.dev, 0);
    const int with_nodes = mode == kIndexLists ? 1 : 0;
    if (resident)
        hipLaunchKernelGGL(n2t_place_kernel<true>, dim3(chunks, F), dim3(64), tile * sizeof(uint32_t), stream, B.dev,
                           tile, with_nodes, chunk);
    else
        hipLaunchKernelGGL(n2t_place_kernel<false>, dim3(chunks, F), dim3(64), hist_tile_bytes(tile), stream, B.dev,
                           tile, with_nodes, chunk);
    if (mode == kIndexNodes) return;
    const uint32_t row_grid = (B.nodes + kXcds - 1) / kXcds * kXcds;
    if (gram) {
        hipLaunchKernelGGL((pair_rows_kernel<kRowGram, 1>), dim3(row_grid, F), dim3(64), 0, stream, B.dev);
        hipLaunchKernelGGL(node_scan_kernel, dim3(3, F), dim3(1024), 0, stream, B.dev, 1);
        hipLaunchKernelGGL(pair_emit_kernel, dim3((B.nodes + 15) / 16, F), dim3(256), 0, stream, B.dev);
    } else {
        hipLaunchKernelGGL((pair_rows_kernel<kRowCount, 1>), dim3(row_grid, F), dim3(64), 0, stream, B.dev);
        hipLaunchKernelGGL(node_scan_kernel, dim3(3, F), dim3(1024), 0, stream, B.dev, 1);
        hipLaunchKernelGGL((pair_rows_kernel<kRowTerms, kRowWaves>), dim3(row_grid, F), dim3(64 * kRowWaves), 0, stream, B.dev);
    }
    const uint32_t link_blocks = (B.nodes * kLinkLanes + 255) / 256;
    const uint32_t edge_blocks = gram ? (8 * B.nodes + 255) / 256 : 0u;
    hipLaunchKernelGGL(pair_link_kernel, dim3(link_blocks + edge_blocks, F), dim3(256), 0, stream, B.dev, link_blocks);
}

// The NodePair2Term lists of an index built for the row Gram assembly (which
// does not need them), for inspection: the pair keys and graph-edge lists stay
// as they are.
void launch_pair_term_lists(const Batch& B, hipStream_t stream) {
    const uint32_t row_grid = (B.nodes + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL((pair_rows_kernel<kRowTerms, kRowWaves>), dim3(row_grid, B.frames), dim3(64 * kRowWaves), 0, stream, B.dev);
    const uint32_t link_blocks = (B.nodes * kLinkLanes + 255) / 256;
    hipLaunchKernelGGL(pair_link_kernel, dim3(link_blocks, B.frames), dim3(256), 0, stream, B.dev, link_blocks);
}

uint32_t index_chunks(uint64_t max_terms) { return (uint32_t)((max_terms + kChunkTerms - 1) / kChunkTerms); }

}  // namespace mvdr
