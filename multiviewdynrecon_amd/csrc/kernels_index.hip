// Index build of one frame: term selection, Node2Term and NodePair2Term
// indices and the block-row structure of J^T J.  Every kernel reads its sizes
// from device memory (FrameDev / Counts), so the whole chain runs without a
// host round trip and can be captured into a hipGraph once per solver.
#include "common.h"
#include "device_util.h"

// Selection decisions (window searches, distance thresholds) are evaluated
// without FMA contraction so that they round exactly like the same
// expressions on the host.
#pragma clang fp contract(off)

namespace mvdr {

constexpr int kTileThreads = 256;
constexpr int kTileRounds = 16;
constexpr int kTilePixels = kTileThreads * kTileRounds;  // 4096 pixels per tile
constexpr int kSegCap = 8192;                            // longest node term list sorted in LDS
constexpr int kRowCap = 4096;                            // longest pair row sorted in LDS

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

// Dense-term predicate: the index map holds a surfel (ImageTermKNNFetcher.cu:19-31).
struct DensePred {
    __device__ bool operator()(const FrameDev& f, const Buffers& b, PixelCoord c) const {
        const uint32_t s = f.index[c.v][c.y * f.W + c.x];
        if (s == kInvalidIndex) return false;
        if (s >= f.num_surfels) {
            atomicOr(&b.counts->error, kErrNodes);
            return false;
        }
        return true;
    }
};

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

// Pass 1: valid pixels per 4096-pixel tile.  Block 0 also clears the
// counters the rest of the index build accumulates into.
template <typename Pred>
__global__ void __launch_bounds__(kTileThreads) compact_count_kernel(Buffers b, Pred pred, int clear) {
    const FrameDev& f = *b.frame;
    const uint32_t M = f.V * f.W * f.H;
    if (clear) {
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= b.cap_nodes; i += stride) {
            b.n2t_count[i] = 0;
            b.row_count[i] = 0;
            b.row_pairs[i] = 0;
            b.low_count[i] = 0;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            Counts* c = b.counts;
            c->dense = c->smooth = c->fg = c->feat = c->cross = 0;
            c->kv1 = c->kv2 = c->pairs = c->max_seg = c->max_row = 0;
            c->error = 0;
        }
    }
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

// Pass 2: exclusive scan of the tile counts (one block).
__global__ void __launch_bounds__(1024) compact_scan_kernel(Buffers b, int which) {
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
__global__ void __launch_bounds__(kTileThreads) compact_scatter_kernel(Buffers b, Pred pred, int which) {
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
// warp vertex around pixel 1 in view 1, reference vertex within 5 cm of the
// observed point expressed in camera 1.
__device__ inline bool choose_cross(const FrameDev& f, uint32_t k, P2PRec& rec, float4& canon) {
    const mvdr_cross_pair c = f.cross[k];
    const uint32_t v1 = c.view[0], v2 = c.view[1];
    if (v1 >= f.V || v2 >= f.V) return false;
    int dx, dy;
    depth_pixel_in_window(f, v2, c.pixel[2], c.pixel[3], dx, dy);
    if (fetch_edge(f, v2, dx, dy) == 1) return false;
    if (!(dx < (int)f.W && dy < (int)f.H)) return false;
    const float4 obs2 = f.vertex[v2][dy * f.W + dx];
    const f3 p0 = f.init[v2].apply(xyz(obs2));
    const f3 p1 = f.init_inv[v1].apply(p0);
    int gx, gy;
    geometry_pixel_nearest(f, v1, p1, c.pixel[0], c.pixel[1], 2, 2, gx, gy);
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

// Block 0 compacts feature pairs, block 1 cross-view pairs, each in candidate
// order with a running block-wide prefix.
__global__ void __launch_bounds__(1024) select_pairs_kernel(Buffers b, int use_cross) {
    const FrameDev& f = *b.frame;
    const bool is_cross = blockIdx.x == 1;
    if (is_cross && !use_cross) {
        if (threadIdx.x == 0) b.counts->cross = 0;
        return;
    }
    const uint32_t total = is_cross ? f.num_cross : f.feat_base[f.V];
    const uint32_t cap = is_cross ? b.cap_cross : b.cap_feat;
    P2PRec* out = is_cross ? b.cross : b.feat;
    float4* canon_out = is_cross ? b.cross_canon : b.feat_canon;
    __shared__ uint32_t wave_cnt[2][16];
    uint32_t running = 0;
    const uint32_t wave = threadIdx.x >> 6;
    int parity = 0;
    for (uint32_t base = 0; base < total; base += 1024, parity ^= 1) {
        const uint32_t g = base + threadIdx.x;
        P2PRec rec;
        float4 canon;
        bool ok = false;
        if (g < total) ok = is_cross ? choose_cross(f, g, rec, canon) : choose_feature(f, g, rec, canon);
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
                out[dst] = rec;
                canon_out[dst] = canon;
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
__device__ inline int term_nodes(const Buffers& b, const FrameDev& f, const uint32_t off[5], uint32_t t, uint16_t n[4]) {
    ushort4 k;
    if (t < off[0]) k = b.dense[t].knn;
    else if (t < off[1]) {
        const ushort2 e = f.graph[t - off[0]];
        n[0] = e.x;
        n[1] = e.y;
        return 2;
    } else if (t < off[2]) k = b.fg[t - off[1]].knn;
    else if (t < off[3]) k = b.feat[t - off[2]].knn;
    else k = b.cross[t - off[3]].knn;
    n[0] = k.x; n[1] = k.y; n[2] = k.z; n[3] = k.w;
    return 4;
}

__global__ void n2t_count_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    Counts& c = *b.counts;
    uint32_t off[5];
    term_offsets(c, f.E, off);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.smooth = f.E;
        for (int k = 0; k < 5; ++k) c.off[k] = off[k];
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < off[4]; t += stride) {
        uint16_t n[4];
        const int m = term_nodes(b, f, off, t, n);
        for (int q = 0; q < m; ++q) {
            if (n[q] < f.N) atomicAdd(&b.n2t_count[n[q]], 1u);
            else atomicOr(&c.error, kErrNodes);
        }
    }
}

// Exclusive scan over the N node counters.  which: 0 = Node2Term segments,
// 1 = pair-term rows, 2 = unique pairs per row, 3 = lower block-row entries.
__global__ void __launch_bounds__(1024) node_scan_kernel(Buffers b, int which) {
    const FrameDev& f = *b.frame;
    const uint32_t N = f.N;
    Counts& c = *b.counts;
    const uint32_t* in;
    uint32_t* out;
    uint32_t* cursor = nullptr;
    switch (which) {
        case 0: in = b.n2t_count; out = b.n2t_offset; cursor = b.n2t_cursor; break;
        case 1: in = b.row_count; out = b.row_offset; break;
        case 2: in = b.row_pairs; out = b.pair_offset; break;
        default: in = b.low_count; out = b.low_offset; cursor = b.low_cursor; break;
    }
    uint32_t mx = 0;
    const uint32_t total = block_exclusive_scan_1024(in, out, N, cursor, &mx);
    if (threadIdx.x == 0) {
        out[N] = total;
        switch (which) {
            case 0:
                if (total > b.cap_kv1) atomicOr(&c.error, kErrKv1);
                if (mx > kSegCap) atomicOr(&c.error, kErrSeg);
                c.kv1 = total;
                c.max_seg = mx;
                break;
            case 1:
                if (total > b.cap_kv2) atomicOr(&c.error, kErrKv2);
                if (mx > kRowCap) atomicOr(&c.error, kErrRow);
                c.kv2 = total;
                c.max_row = mx;
                break;
            case 2:
                if (total > b.cap_pairs) atomicOr(&c.error, kErrPairs);
                c.pairs = total;
                break;
            default:
                break;
        }
    }
}

__global__ void n2t_scatter_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < c.off[4]; t += stride) {
        uint16_t n[4];
        const int m = term_nodes(b, f, c.off, t, n);
        for (int q = 0; q < m; ++q)
            if (n[q] < f.N) {
                const uint32_t pos = atomicAdd(&b.n2t_cursor[n[q]], 1u);
                b.n2t_index[pos] = t;
            }
    }
}

// Sort each node's term list ascending: the order a stable key sort of
// (node, term) pairs generated in term order produces (Node2TermsIndex.cu:136-166).
__global__ void __launch_bounds__(256) n2t_sort_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    if (b.counts->error) return;
    __shared__ uint32_t keys[kSegCap];
    for (uint32_t n = blockIdx.x; n < f.N; n += gridDim.x) {
        const uint32_t beg = b.n2t_offset[n], len = b.n2t_offset[n + 1] - beg;
        if (len < 2) continue;
        const uint32_t m = next_pow2(len);
        for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) keys[i] = i < len ? b.n2t_index[beg + i] : 0xFFFFFFFFu;
        __syncthreads();
        bitonic_sort_lds(keys, m);
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) b.n2t_index[beg + i] = keys[i];
        __syncthreads();
    }
}

// ------------------------------------------------------------ NodePair2Term
// Pairs (i, j), i < j, that term t contributes to row i, counted with the
// multiplicity of the reference's key generation: for every knn slot p holding
// i, one entry per slot q with knn[q] > i (computeKVNodePairKNN,
// NodePair2TermsIndex.cu:11-27); the slot multiplicity is already carried by
// the Node2Term list, which holds t once per slot equal to i.  A graph edge
// gives (min, max) once (NodePair2TermsIndex.cu:63-73).
__device__ inline int row_entries(const Buffers& b, const FrameDev& f, const uint32_t off[5], uint32_t t, uint32_t i,
                                  uint16_t out[4]) {
    uint16_t n[4];
    const int m = term_nodes(b, f, off, t, n);
    int k = 0;
    if (m == 2) {
        const uint16_t other = (n[0] == i) ? n[1] : n[0];
        if (other > i && n[0] != n[1]) out[k++] = other;
        return k;
    }
    for (int q = 0; q < 4; ++q)
        if (n[q] > i) out[k++] = n[q];
    return k;
}

__global__ void __launch_bounds__(256) pair_count_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    for (uint32_t i = blockIdx.x; i < f.N; i += gridDim.x) {
        const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
        uint32_t cnt = 0;
        for (uint32_t p = beg + threadIdx.x; p < end; p += blockDim.x) {
            uint16_t o[4];
            cnt += row_entries(b, f, c.off, b.n2t_index[p], i, o);
        }
        cnt = block_sum_u32<256>(cnt);
        if (threadIdx.x == 0) b.row_count[i] = cnt;
        __syncthreads();
    }
}

// Row i: gather (j, term) entries, sort, write the term column of the sorted
// pair-term array and the unique j's with their ranges.
__global__ void __launch_bounds__(256) pair_fill_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    __shared__ uint64_t keys[kRowCap];
    __shared__ uint32_t fill;
    __shared__ uint32_t wave_cnt[2][4];
    for (uint32_t i = blockIdx.x; i < f.N; i += gridDim.x) {
        const uint32_t len = b.row_count[i];
        const uint32_t out_base = b.row_offset[i];
        if (threadIdx.x == 0) fill = 0;
        __syncthreads();
        const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
        for (uint32_t p = beg + threadIdx.x; p < end; p += blockDim.x) {
            const uint32_t t = b.n2t_index[p];
            uint16_t o[4];
            const int k = row_entries(b, f, c.off, t, i, o);
            if (k) {
                const uint32_t at = atomicAdd(&fill, (uint32_t)k);
                for (int q = 0; q < k; ++q) keys[at + q] = ((uint64_t)o[q] << 32) | t;
            }
        }
        __syncthreads();
        const uint32_t m = next_pow2(len);
        for (uint32_t q = len + threadIdx.x; q < m; q += blockDim.x) keys[q] = ~0ull;
        __syncthreads();
        if (len > 1) bitonic_sort_lds(keys, m);
        // segment heads -> unique pairs, in order
        uint32_t running = 0;
        int parity = 0;
        for (uint32_t base = 0; base < len; base += blockDim.x, parity ^= 1) {
            const uint32_t q = base + threadIdx.x;
            bool head = false;
            uint32_t j = 0;
            if (q < len) {
                j = (uint32_t)(keys[q] >> 32);
                b.pair_terms[out_base + q] = (uint32_t)keys[q];
                head = q == 0 || (uint32_t)(keys[q - 1] >> 32) != j;
            }
            const uint64_t mask = __ballot(head);
            if ((threadIdx.x & 63) == 0) wave_cnt[parity][threadIdx.x >> 6] = (uint32_t)__popcll(mask);
            __syncthreads();
            uint32_t before = 0, chunk = 0;
            for (int w = 0; w < 4; ++w) {
                before += (w < (int)(threadIdx.x >> 6)) ? wave_cnt[parity][w] : 0u;
                chunk += wave_cnt[parity][w];
            }
            if (head) {
                uint32_t e = q + 1;
                while (e < len && (uint32_t)(keys[e] >> 32) == j) ++e;
                const uint32_t u = running + before + lane_prefix(mask);
                b.uniq_j[out_base + u] = j;
                b.uniq_range[out_base + u] = make_uint2(out_base + q, out_base + e);
            }
            running += chunk;
        }
        if (threadIdx.x == 0) b.row_pairs[i] = running;
        __syncthreads();
    }
}

// Unique pairs in (i, j) order with their term ranges, and the count of
// lower-triangle entries each row j receives.
__global__ void __launch_bounds__(256) pair_emit_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    for (uint32_t i = blockIdx.x; i < f.N; i += gridDim.x) {
        const uint32_t n = b.row_pairs[i], src = b.row_offset[i], dst = b.pair_offset[i];
        for (uint32_t u = threadIdx.x; u < n; u += blockDim.x) {
            const uint32_t j = b.uniq_j[src + u];
            b.pair_key[dst + u] = (i << 16) | j;
            b.pair_range[dst + u] = b.uniq_range[src + u];
            atomicAdd(&b.low_count[j], 1u);
        }
    }
}

__global__ void low_scatter_kernel(Buffers b) {
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < c.pairs; p += stride) {
        const uint32_t key = b.pair_key[p];
        const uint32_t i = key >> 16, j = key & 0xFFFF;
        const uint32_t pos = atomicAdd(&b.low_cursor[j], 1u);
        b.low_ent[pos] = make_uint2(i, p);
    }
}

// Lower-triangle entries of each block row in ascending column order.
__global__ void low_sort_kernel(Buffers b) {
    const FrameDev& f = *b.frame;
    if (b.counts->error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < f.N; i += stride) {
        const uint32_t beg = b.low_offset[i], end = b.low_offset[i + 1];
        for (uint32_t a = beg + 1; a < end; ++a) {
            const uint2 key = b.low_ent[a];
            uint32_t z = a;
            while (z > beg && b.low_ent[z - 1].x > key.x) {
                b.low_ent[z] = b.low_ent[z - 1];
                --z;
            }
            b.low_ent[z] = key;
        }
    }
}

// ------------------------------------------------------------- host launch
void launch_index_build(const Buffers& b, uint32_t cap_tiles, uint32_t cap_nodes, bool use_foreground, bool use_cross,
                        hipStream_t stream) {
    const uint32_t node_blocks = std::min<uint32_t>(cap_nodes, 8192u);
    hipLaunchKernelGGL(compact_count_kernel<DensePred>, dim3(cap_tiles), dim3(kTileThreads), 0, stream, b, DensePred{}, 1);
    hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, stream, b, 0);
    hipLaunchKernelGGL(compact_scatter_kernel<DensePred>, dim3(cap_tiles), dim3(kTileThreads), 0, stream, b, DensePred{}, 0);
    if (use_foreground) {
        hipLaunchKernelGGL(compact_count_kernel<ForegroundPred>, dim3(cap_tiles), dim3(kTileThreads), 0, stream, b,
                           ForegroundPred{}, 0);
        hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, stream, b, 1);
        hipLaunchKernelGGL(compact_scatter_kernel<ForegroundPred>, dim3(cap_tiles), dim3(kTileThreads), 0, stream, b,
                           ForegroundPred{}, 1);
    }
    hipLaunchKernelGGL(select_pairs_kernel, dim3(2), dim3(1024), 0, stream, b, use_cross ? 1 : 0);
    const uint32_t term_blocks = 1024;
    hipLaunchKernelGGL(n2t_count_kernel, dim3(term_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(node_scan_kernel, dim3(1), dim3(1024), 0, stream, b, 0);
    hipLaunchKernelGGL(n2t_scatter_kernel, dim3(term_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(n2t_sort_kernel, dim3(node_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(pair_count_kernel, dim3(node_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(node_scan_kernel, dim3(1), dim3(1024), 0, stream, b, 1);
    hipLaunchKernelGGL(pair_fill_kernel, dim3(node_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(node_scan_kernel, dim3(1), dim3(1024), 0, stream, b, 2);
    hipLaunchKernelGGL(pair_emit_kernel, dim3(node_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(node_scan_kernel, dim3(1), dim3(1024), 0, stream, b, 3);
    hipLaunchKernelGGL(low_scatter_kernel, dim3(term_blocks), dim3(256), 0, stream, b);
    hipLaunchKernelGGL(low_sort_kernel, dim3((cap_nodes + 255) / 256), dim3(256), 0, stream, b);
}

}  // namespace mvdr
