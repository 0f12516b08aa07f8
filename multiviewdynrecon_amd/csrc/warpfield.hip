// Warp-field maintenance kernels (include/mvdr/warpfield.h): node graph,
// skinning, forward and inverse warps.
//
// The two searches are all-pairs scans over the canonical nodes: each
// workgroup stages the nodes through LDS in tiles of kNodeTile (every node is
// read from HBM once per workgroup, the tile by all 256 threads), so the
// searches run at the VALU rate of ~7 flops per (point, node) pair -- about
// 1.2 G pairs for 150k surfels x 4k nodes.  The warps are streaming passes
// bound by HBM (a 16-32 B read and write per point plus the L2-resident
// warp-field gathers).
//
// Distances are evaluated without contraction except where the reference
// fuses explicitly (the skinning distance), so that they round like the
// oracle's.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/mvdr/warpfield.h"
#include "common.h"

namespace mvdr {
namespace {

constexpr uint32_t kNodeTile = 1024;
constexpr int kGraphNeighbours = 8;   // nodesGraphNeigboursNum

// The reference's 4-slot heap update (brute_foce_knn.cuh:17-39): slot x holds
// the current worst; a closer node replaces it and sifts once.
__device__ inline void heap_update(float d[4], uint32_t id[4], uint32_t idx, float dist) {
    if (!(dist < d[0])) return;
    d[0] = dist;
    id[0] = idx;
    if (d[1] < d[2]) {
        if (d[0] < d[2]) {
            const float t = d[0]; d[0] = d[2]; d[2] = t;
            const uint32_t u = id[0]; id[0] = id[2]; id[2] = u;
        }
    } else if (d[0] < d[1]) {
        float t = d[0]; d[0] = d[1]; d[1] = t;
        uint32_t u = id[0]; id[0] = id[1]; id[1] = u;
        if (d[1] < d[3]) {
            t = d[1]; d[1] = d[3]; d[3] = t;
            u = id[1]; id[1] = id[3]; id[3] = u;
        }
    }
}

// One workgroup of 256 threads per 256 points; the nodes pass through LDS.
// Points [0, num_vertices) are vertices, [num_vertices, + num_nodes) nodes.
__global__ void __launch_bounds__(256) skin_kernel(const float4* __restrict__ vertices, uint32_t num_vertices,
                                                   const float4* __restrict__ nodes, uint32_t num_nodes,
                                                   float radius, ushort4* __restrict__ vknn, float4* __restrict__ vw,
                                                   ushort4* __restrict__ nknn, float4* __restrict__ nw) {
    __shared__ float4 tile[kNodeTile];
    const uint32_t total = num_vertices + num_nodes;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    float4 pt = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < num_vertices) pt = vertices[p];
    else if (p < total) pt = nodes[p - num_vertices];
    float d[4] = {1e6f, 1e6f, 1e6f, 1e6f};
    uint32_t id[4] = {0, 0, 0, 0};
    for (uint32_t base = 0; base < num_nodes; base += kNodeTile) {
        const uint32_t n = min(kNodeTile, num_nodes - base);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) tile[k] = nodes[base + k];
        __syncthreads();
        for (uint32_t k = 0; k < n; ++k) {
            const float4 q = tile[k];
            const float dx = pt.x - q.x, dy = pt.y - q.y, dz = pt.z - q.z;
            const float dist = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            heap_update(d, id, base + k, dist);
        }
    }
    if (p >= total) return;
    float w[4];
    const float den = (2.0f * radius) * radius;
    for (int s = 0; s < 4; ++s) {
        const float4 q = nodes[id[s]];
        const float dx = pt.x - q.x, dy = pt.y - q.y, dz = pt.z - q.z;
        w[s] = expf(-(dx * dx + dy * dy + dz * dz) / den);
    }
    const float inv = 1.0f / (fabsf(w[0]) + fabsf(w[1]) + fabsf(w[2]) + fabsf(w[3]));
    const ushort4 k4 = make_ushort4(id[0], id[1], id[2], id[3]);
    const float4 w4 = make_float4(w[0] * inv, w[1] * inv, w[2] * inv, w[3] * inv);
    if (p < num_vertices) {
        vknn[p] = k4;
        vw[p] = w4;
    } else {
        nknn[p - num_vertices] = k4;
        nw[p - num_vertices] = w4;
    }
}

// BuildNodesGraphKernel: 8 slots, the worst replaced (WarpField.cu:15-59).
__global__ void __launch_bounds__(256) node_graph_kernel(const float4* __restrict__ nodes, uint32_t num_nodes,
                                                         ushort2* __restrict__ graph) {
    __shared__ float4 tile[kNodeTile];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float4 me = i < num_nodes ? nodes[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float dist[kGraphNeighbours];
    int idx[kGraphNeighbours];
    for (int k = 0; k < kGraphNeighbours; ++k) {
        dist[k] = 1e5f;
        idx[k] = -1;
    }
    int worst = 0;
    for (uint32_t base = 0; base < num_nodes; base += kNodeTile) {
        const uint32_t n = min(kNodeTile, num_nodes - base);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) tile[k] = nodes[base + k];
        __syncthreads();
        for (uint32_t k = 0; k < n; ++k) {
            const float4 q = tile[k];
            const float d = (me.x - q.x) * (me.x - q.x) + (me.y - q.y) * (me.y - q.y) + (me.z - q.z) * (me.z - q.z);
            if (d > 1e-6f && d < dist[worst]) {
                dist[worst] = d;
                idx[worst] = (int)(base + k);
                worst = 0;
                float mx = 0.f;
                for (int j = 0; j < kGraphNeighbours; ++j)
                    if (dist[j] > mx) {
                        worst = j;
                        mx = dist[j];
                    }
            }
        }
    }
    if (i >= num_nodes) return;
    for (int k = 0; k < kGraphNeighbours; ++k)
        graph[(size_t)i * kGraphNeighbours + k] = make_ushort2((uint16_t)i, (uint16_t)idx[k]);
}

__device__ inline bool knn_in_range(ushort4 k, uint32_t n) { return k.x < n && k.y < n && k.z < n && k.w < n; }

// forwardWarpVertexAndNodeKernel: surfels (vertex, normal), then nodes.  A
// point whose knn names a node outside the field is copied unwarped.
__global__ void __launch_bounds__(256) forward_warp_kernel(const DQ* __restrict__ field, const float4* __restrict__ cv,
                                                           const float4* __restrict__ cn, const ushort4* __restrict__ vknn,
                                                           const float4* __restrict__ vw, uint32_t num_vertices,
                                                           const float4* __restrict__ nodes,
                                                           const ushort4* __restrict__ nknn,
                                                           const float4* __restrict__ nw, uint32_t num_nodes,
                                                           float4* __restrict__ lv, float4* __restrict__ ln,
                                                           float4* __restrict__ lnodes) {
    const uint32_t total = num_vertices + num_nodes;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const bool is_vertex = p < num_vertices;
        const uint32_t o = is_vertex ? p : p - num_vertices;
        const float4 v = is_vertex ? cv[o] : nodes[o];
        const ushort4 k = is_vertex ? vknn[o] : nknn[o];
        const float4 w = is_vertex ? vw[o] : nw[o];
        float4 vo = v, no = is_vertex ? cn[o] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (knn_in_range(k, num_nodes)) {
            const Mat34 T = se3_of(blend4(field, k, w));
            const f3 pv = T.apply(xyz(v));
            vo = make_float4(pv.x, pv.y, pv.z, v.w);
            if (is_vertex) {
                const f3 nv = T.rot(xyz(no));
                no = make_float4(nv.x, nv.y, nv.z, no.w);
            }
        }
        if (is_vertex) {
            lv[o] = vo;
            ln[o] = no;
        } else {
            lnodes[o] = vo;
        }
    }
}

// inverseWarpVertexNormalKernel (without the correction transform).
__global__ void __launch_bounds__(256) inverse_warp_kernel(const DQ* __restrict__ field, uint32_t field_nodes,
                                                           const float4* __restrict__ lv, const float4* __restrict__ ln,
                                                           const ushort4* __restrict__ vknn,
                                                           const float4* __restrict__ vw, uint32_t num_vertices,
                                                           float4* __restrict__ rv, float4* __restrict__ rn) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < num_vertices; p += gridDim.x * blockDim.x) {
        const float4 v = lv[p], n = ln[p];
        float4 vo = v, no = n;
        const ushort4 k = vknn[p];
        if (knn_in_range(k, field_nodes)) {
            const Mat34 T = se3_of(blend4(field, k, vw[p]));
            const f3 a = T.rot_t(mk(v.x - T.t.x, v.y - T.t.y, v.z - T.t.z));
            const f3 b = T.rot_t(xyz(n));
            vo = make_float4(a.x, a.y, a.z, v.w);
            no = make_float4(b.x, b.y, b.z, n.w);
        }
        rv[p] = vo;
        rn[p] = no;
    }
}

// --------------------------------------------- large-error node correction
constexpr uint32_t kCorrHalfWindow = 2;  // CorrSearchWindowHalfSize (NonRigidSolver.h:41)
constexpr float kNodeRadius = 0.025f;    // NODE_RADIUS (GlobalConfigs.h:85)

struct CorrViews {
    uint32_t V, W, H;
    float fx[kMaxViews], fy[kMaxViews], cx[kMaxViews], cy[kMaxViews];
    Mat34 w2c[kMaxViews], w2c_inv[kMaxViews], init[kMaxViews], init_inv[kMaxViews];
    const float4* prev[kMaxViews];
    const float* guided[kMaxViews];  // 12 floats per pixel: rotation columns, translation
    const uint8_t* valid[kMaxViews];
};

// mat34::operator* (MatUtils.h:449-455): a * b
__device__ inline Mat34 compose(const Mat34& a, const Mat34& b) {
    Mat34 o;
    o.c0 = a.rot(b.c0);
    o.c1 = a.rot(b.c1);
    o.c2 = a.rot(b.c2);
    o.t = a.rot(b.t) + a.t;
    return o;
}

__device__ inline Mat34 identity34() {
    Mat34 m;
    m.c0 = mk(1.f, 0.f, 0.f);
    m.c1 = mk(0.f, 1.f, 0.f);
    m.c2 = mk(0.f, 0.f, 1.f);
    m.t = mk(0.f, 0.f, 0.f);
    return m;
}

// decodeNodesCameraView (CommonUtils.h:110-117): v with v < w <= v + 1.
__device__ inline uint32_t node_view(float w) {
    for (uint32_t v = 0; v < (uint32_t)kMaxViews; ++v)
        if ((float)v < w && w <= (float)(v + 1)) return v;
    return 0xFFFFFFFFu;
}

// The pixel coordinate of the reference's projection: the expression is
// evaluated in double (its 1e-10 literal), converted to float and rounded to
// the nearest integer (ties to even); negative values give 0.
__device__ inline uint32_t pixel_of(float x, float z, float f, float c) {
    const float p = (float)(((double)x / ((double)z + 1e-10)) * (double)f + (double)c);
    if (!(p > 0.f)) return 0u;
    if (p >= 4294967040.f) return 0xFFFFFFFFu;
    return (uint32_t)rintf(p);
}

__device__ inline float4 qconj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }

// Quaternion::pow (Quaternion.h:235-249)
__device__ inline float4 qpow(float4 q, float t) {
    const float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
    const float nt = atan2f(n, q.w) * t;
    if (fabsf(n) < 1e-6f) return make_float4(0.f, 0.f, 0.f, cosf(nt));
    const float sc = sinf(nt) / n;
    return make_float4(q.x * sc, q.y * sc, q.z * sc, cosf(nt));
}

// DualQuaternion::operator*(float) (DualQuaternion.h:165-178): the screw
// scaled by t in [1e-10, 1].
__device__ inline DQ dq_scale(const DQ& d, float ratio) {
    float t = ratio;
    if (t > 1.0f) t = 1.0f;
    else if (t < 1e-10f) t = 1e-10f;
    const float4 r = qmul(d.q1, qconj(d.q0));
    const float4 q0s = qpow(d.q0, t);
    return DQ{q0s, qmul(make_float4(t * r.x, t * r.y, t * r.z, t * r.w), q0s)};
}

// DualQuaternion::operator*(DualQuaternion) (DualQuaternion.h:152-157)
__device__ inline DQ dq_mul(const DQ& a, const DQ& b) {
    const float4 u = qmul(a.q1, b.q0), v = qmul(a.q0, b.q1);
    return DQ{qmul(a.q0, b.q0), make_float4(u.x + v.x, u.y + v.y, u.z + v.z, u.w + v.w)};
}

// computeCorrespondenceGuidedNodesSE3 with computeSuitableCorrespondenceGuidedSe3
// (NonRigidSolver.cu:100-138, :155-167): one thread per node.
__global__ void __launch_bounds__(256) correction_kernel(CorrViews r, const float4* __restrict__ nodes,
                                                         const float* __restrict__ error, uint32_t num_nodes,
                                                         DQ* __restrict__ correction) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < num_nodes; n += gridDim.x * blockDim.x) {
        const float e = error[n];
        DQ out;
        out.q0 = make_float4(0.f, 0.f, 0.f, 1.f);
        out.q1 = make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 node = nodes[n];
        const uint32_t v = node_view(node.w);
        if (e > 0.2f && v < r.V) {
            Mat34 sel = identity34();
            const f3 c = r.init_inv[v].apply(r.w2c[v].apply(xyz(node)));
            const uint32_t u = pixel_of(c.x, c.z, r.fx[v], r.cx[v]);
            const uint32_t y = pixel_of(c.y, c.z, r.fy[v], r.cy[v]);
            const bool room = r.W > 2 * kCorrHalfWindow && r.H > 2 * kCorrHalfWindow;  // the window fits the map
            if (room && !(u >= r.W - kCorrHalfWindow || u < kCorrHalfWindow || y >= r.H - kCorrHalfWindow ||
                          y < kCorrHalfWindow)) {
                float best = 1e6f;
                for (uint32_t i = u - kCorrHalfWindow; i <= u + kCorrHalfWindow; ++i)
                    for (uint32_t j = y - kCorrHalfWindow; j <= y + kCorrHalfWindow; ++j) {
                        const size_t o = (size_t)j * r.W + i;
                        const float4 pv = r.prev[v][o];
                        if (r.valid[v][o] != 1 || is_zero_vertex(pv)) continue;
                        const float* g = r.guided[v] + 12 * o;
                        Mat34 G;
                        G.c0 = mk(g[0], g[1], g[2]);
                        G.c1 = mk(g[3], g[4], g[5]);
                        G.c2 = mk(g[6], g[7], g[8]);
                        G.t = mk(g[9], g[10], g[11]);
                        const f3 d = c - xyz(pv);
                        const float d2 = dot3(d, d);
                        const f3 m = G.apply(c) - c;
                        if (best > d2 && d2 < 1e-4f && dot3(m, m) < 4 * (kNodeRadius * kNodeRadius)) {
                            best = d2;
                            sel = G;
                        }
                    }
                sel = compose(compose(compose(compose(r.w2c_inv[v], r.init[v]), sel), r.init_inv[v]), r.w2c[v]);
            }
            out = dq_scale(dq_from_mat34(sel), e);
        }
        correction[n] = out;
    }
}

// CorrectWarpVertexAndNodeKernel (NonRigidSolver.cu:170-216): surfels
// [0, S), then nodes; in place.
__global__ void __launch_bounds__(256) apply_correction_kernel(const DQ* __restrict__ correction, uint32_t num_nodes,
                                                               float4* __restrict__ lv, float4* __restrict__ ln,
                                                               const ushort4* __restrict__ sknn,
                                                               const float4* __restrict__ sw, uint32_t num_surfels,
                                                               float4* __restrict__ lnodes, DQ* __restrict__ node_se3,
                                                               const ushort4* __restrict__ nknn,
                                                               const float4* __restrict__ nw) {
    const uint32_t total = num_surfels + num_nodes;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const bool surfel = p < num_surfels;
        const uint32_t o = surfel ? p : p - num_surfels;
        const ushort4 k = surfel ? sknn[o] : nknn[o];
        if (!knn_in_range(k, num_nodes)) continue;
        const DQ avg = blend4(correction, k, surfel ? sw[o] : nw[o]);
        const Mat34 T = se3_of(avg);
        const float4 v = surfel ? lv[o] : lnodes[o];
        const f3 pv = T.apply(xyz(v));
        if (surfel) {
            const float4 n = ln[o];
            const f3 pn = T.rot(xyz(n));
            lv[o] = make_float4(pv.x, pv.y, pv.z, v.w);
            ln[o] = make_float4(pn.x, pn.y, pn.z, n.w);
        } else {
            lnodes[o] = make_float4(pv.x, pv.y, pv.z, v.w);
            node_se3[o] = dq_mul(avg, node_se3[o]);
        }
    }
}

uint32_t stream_grid(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 4096u)); }

}  // namespace
}  // namespace mvdr

using namespace mvdr;

extern "C" {

mvdr_status mvdr_warp_build_node_graph(const void* nodes, uint32_t num_nodes, void* graph, void* stream) {
    if (num_nodes == 0) return MVDR_OK;
    if (!nodes || !graph) return MVDR_E_INVALID;
    if (num_nodes > 65535) return MVDR_E_SHAPE;
    hipLaunchKernelGGL(node_graph_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4*>(nodes), num_nodes, static_cast<ushort2*>(graph));
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}

mvdr_status mvdr_warp_skin(const void* vertices, uint32_t num_vertices, const void* nodes, uint32_t num_nodes,
                           float node_radius, void* vknn, void* vw, void* nknn, void* nw, void* stream) {
    if (num_nodes == 0 || !nodes || !(node_radius > 0.f)) return MVDR_E_INVALID;
    if (num_nodes > 65535) return MVDR_E_SHAPE;
    if (num_vertices && (!vertices || !vknn || !vw)) return MVDR_E_INVALID;
    if (!nknn || !nw) return MVDR_E_INVALID;
    const uint64_t total = (uint64_t)num_vertices + num_nodes;
    if (total > 0xFFFFFFFFull - 255) return MVDR_E_SHAPE;
    hipLaunchKernelGGL(skin_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4*>(vertices), num_vertices, static_cast<const float4*>(nodes), num_nodes,
                       node_radius, static_cast<ushort4*>(vknn), static_cast<float4*>(vw), static_cast<ushort4*>(nknn),
                       static_cast<float4*>(nw));
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}

mvdr_status mvdr_warp_forward(const void* node_se3, const void* cv, const void* cn, const void* vknn, const void* vw,
                              uint32_t num_vertices, const void* nodes, const void* nknn, const void* nw,
                              uint32_t num_nodes, void* lv, void* ln, void* lnodes, void* stream) {
    if (!node_se3 || num_nodes == 0) return MVDR_E_INVALID;
    if (num_vertices && (!cv || !cn || !vknn || !vw || !lv || !ln)) return MVDR_E_INVALID;
    if (!nodes || !nknn || !nw || !lnodes) return MVDR_E_INVALID;
    if (num_nodes > 65535) return MVDR_E_SHAPE;
    const uint64_t total = (uint64_t)num_vertices + num_nodes;
    if (total > 0xFFFFFFFFull) return MVDR_E_SHAPE;
    hipLaunchKernelGGL(forward_warp_kernel, dim3(stream_grid((uint32_t)total)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const DQ*>(node_se3),
                       static_cast<const float4*>(cv), static_cast<const float4*>(cn), static_cast<const ushort4*>(vknn),
                       static_cast<const float4*>(vw), num_vertices, static_cast<const float4*>(nodes),
                       static_cast<const ushort4*>(nknn), static_cast<const float4*>(nw), num_nodes,
                       static_cast<float4*>(lv), static_cast<float4*>(ln), static_cast<float4*>(lnodes));
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}

mvdr_status mvdr_warp_inverse(const void* node_se3, uint32_t num_nodes, const void* lv, const void* ln,
                              const void* vknn, const void* vw, uint32_t num_vertices, void* rv, void* rn,
                              void* stream) {
    if (num_vertices == 0) return MVDR_OK;
    if (!node_se3 || num_nodes == 0 || !lv || !ln || !vknn || !vw || !rv || !rn) return MVDR_E_INVALID;
    if (num_nodes > 65535) return MVDR_E_SHAPE;
    hipLaunchKernelGGL(inverse_warp_kernel, dim3(stream_grid(num_vertices)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const DQ*>(node_se3), num_nodes,
                       static_cast<const float4*>(lv), static_cast<const float4*>(ln),
                       static_cast<const ushort4*>(vknn), static_cast<const float4*>(vw), num_vertices,
                       static_cast<float4*>(rv), static_cast<float4*>(rn));
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}

mvdr_status mvdr_warp_correct_nodes(const mvdr_node_correction* in, void* stream) {
    if (!in) return MVDR_E_INVALID;
    if (in->num_views < 1 || in->num_views > MVDR_MAX_VIEWS || in->width < 1 || in->height < 1 ||
        in->width > 65535 || in->height > 65535 || in->num_nodes > 65535)
        return MVDR_E_SHAPE;
    if (in->num_nodes == 0) return MVDR_OK;
    if (!in->node_error || !in->live_nodes || !in->node_se3 || !in->node_knn || !in->node_knn_weight || !in->correction)
        return MVDR_E_INVALID;
    if (in->num_surfels && (!in->live_vertex || !in->live_normal || !in->surfel_knn || !in->surfel_knn_weight))
        return MVDR_E_INVALID;
    if ((uint64_t)in->num_surfels + in->num_nodes > 0xFFFFFFFFull) return MVDR_E_SHAPE;
    CorrViews r;
    std::memset(&r, 0, sizeof(r));
    r.V = in->num_views;
    r.W = in->width;
    r.H = in->height;
    for (uint32_t v = 0; v < r.V; ++v) {
        if (!in->previous_vertex[v] || !in->guided_se3[v] || !in->guided_valid[v]) return MVDR_E_INVALID;
        r.fx[v] = in->intrinsic[v].fx;
        r.fy[v] = in->intrinsic[v].fy;
        r.cx[v] = in->intrinsic[v].cx;
        r.cy[v] = in->intrinsic[v].cy;
        r.w2c[v] = to_mat34(in->world2camera[v]);
        r.init[v] = to_mat34(in->initial_camera_se3[v]);
        r.w2c_inv[v] = inverse(r.w2c[v]);
        r.init_inv[v] = inverse(r.init[v]);
        r.prev[v] = static_cast<const float4*>(in->previous_vertex[v]);
        r.guided[v] = static_cast<const float*>(in->guided_se3[v]);
        r.valid[v] = static_cast<const uint8_t*>(in->guided_valid[v]);
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(correction_kernel, dim3(stream_grid(in->num_nodes)), dim3(256), 0, st, r,
                       static_cast<const float4*>(in->live_nodes), static_cast<const float*>(in->node_error),
                       in->num_nodes, static_cast<DQ*>(in->correction));
    hipLaunchKernelGGL(apply_correction_kernel, dim3(stream_grid(in->num_surfels + in->num_nodes)), dim3(256), 0, st,
                       static_cast<const DQ*>(in->correction), in->num_nodes, static_cast<float4*>(in->live_vertex),
                       static_cast<float4*>(in->live_normal), static_cast<const ushort4*>(in->surfel_knn),
                       static_cast<const float4*>(in->surfel_knn_weight), in->num_surfels,
                       static_cast<float4*>(in->live_nodes), static_cast<DQ*>(in->node_se3),
                       static_cast<const ushort4*>(in->node_knn), static_cast<const float4*>(in->node_knn_weight));
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}

}  // extern "C"
