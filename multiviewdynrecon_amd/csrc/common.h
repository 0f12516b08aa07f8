// Device-side types and math shared by the solver's HIP kernels.
//
// Layout conventions (byte-compatible with the reference types, see
// include/mvdr/types.h): mat34 rotation is stored as three columns
// (math/MatUtils.h:19-20); a dual quaternion is two float4 (x, y, z, w)
// (math/DualQuaternion/DualQuaternion.h:99-102).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mvdr/solver.h"

namespace mvdr {

constexpr int kMaxViews = MVDR_MAX_VIEWS;
constexpr uint32_t kInvalidIndex = 0xFFFFFFFFu;
constexpr uint32_t kPartialStride = 1024;  // per-block partial sums per reduction slot
constexpr uint32_t kPartialSlots = 160;
constexpr uint32_t kSlotCost = 150;        // partials of the total residual

// ---------------------------------------------------------------- vectors
struct f3 {
    float x, y, z;
};
__host__ __device__ inline f3 mk(float a, float b, float c) { return {a, b, c}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline f3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// Rigid transform, rotation as columns c0, c1, c2.
struct Mat34 {
    f3 c0, c1, c2, t;
    __host__ __device__ f3 rot(f3 p) const {
        return {c0.x * p.x + c1.x * p.y + c2.x * p.z, c0.y * p.x + c1.y * p.y + c2.y * p.z,
                c0.z * p.x + c1.z * p.y + c2.z * p.z};
    }
    __host__ __device__ f3 rot_t(f3 p) const { return {dot3(c0, p), dot3(c1, p), dot3(c2, p)}; }
    __host__ __device__ f3 apply(f3 p) const { return rot(p) + t; }
};

inline Mat34 to_mat34(const mvdr_mat34& m) {
    Mat34 o;
    o.c0 = mk(m.rot[0], m.rot[1], m.rot[2]);
    o.c1 = mk(m.rot[3], m.rot[4], m.rot[5]);
    o.c2 = mk(m.rot[6], m.rot[7], m.rot[8]);
    o.t = mk(m.trans[0], m.trans[1], m.trans[2]);
    return o;
}

// mat34::inverse (MatUtils.h:585-590): R^T, -R^T t
inline Mat34 inverse(const Mat34& m) {
    Mat34 o;
    o.c0 = mk(m.c0.x, m.c1.x, m.c2.x);
    o.c1 = mk(m.c0.y, m.c1.y, m.c2.y);
    o.c2 = mk(m.c0.z, m.c1.z, m.c2.z);
    f3 rt = o.rot(m.t);
    o.t = mk(-rt.x, -rt.y, -rt.z);
    return o;
}

// --------------------------------------------------------- dual quaternion
struct DQ {
    float4 q0, q1;
};

__device__ inline float4 qmul(float4 a, float4 b) {
    float4 q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return q;
}

// Linear blend of four node dual quaternions, no sign alignment and no
// weight renormalisation (averageDualQuaternion, DualQuaternion.h:381-398).
__device__ inline DQ blend4(const DQ* __restrict__ field, ushort4 knn, float4 w) {
    const DQ a = field[knn.x], b = field[knn.y], c = field[knn.z], d = field[knn.w];
    DQ o;
    o.q0.x = w.x * a.q0.x + w.y * b.q0.x + w.z * c.q0.x + w.w * d.q0.x;
    o.q0.y = w.x * a.q0.y + w.y * b.q0.y + w.z * c.q0.y + w.w * d.q0.y;
    o.q0.z = w.x * a.q0.z + w.y * b.q0.z + w.z * c.q0.z + w.w * d.q0.z;
    o.q0.w = w.x * a.q0.w + w.y * b.q0.w + w.z * c.q0.w + w.w * d.q0.w;
    o.q1.x = w.x * a.q1.x + w.y * b.q1.x + w.z * c.q1.x + w.w * d.q1.x;
    o.q1.y = w.x * a.q1.y + w.y * b.q1.y + w.z * c.q1.y + w.w * d.q1.y;
    o.q1.z = w.x * a.q1.z + w.y * b.q1.z + w.z * c.q1.z + w.w * d.q1.z;
    o.q1.w = w.x * a.q1.w + w.y * b.q1.w + w.z * c.q1.w + w.w * d.q1.w;
    return o;
}

// normalize() then rotation_matrix(false) and t = 2 q1 conj(q0)
// (DualQuaternion::se3_matrix, DualQuaternion.h:281-287,326-335)
__device__ inline Mat34 se3_of(DQ d) {
    const float inv = rsqrtf(d.q0.w * d.q0.w + d.q0.x * d.q0.x + d.q0.y * d.q0.y + d.q0.z * d.q0.z);
    float4 a = make_float4(inv * d.q0.x, inv * d.q0.y, inv * d.q0.z, inv * d.q0.w);
    float4 b = make_float4(inv * d.q1.x, inv * d.q1.y, inv * d.q1.z, inv * d.q1.w);
    const float dd = a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z;
    b = make_float4(b.x - dd * a.x, b.y - dd * a.y, b.z - dd * a.z, b.w - dd * a.w);
    Mat34 m;
    m.c0 = mk(1 - 2 * a.y * a.y - 2 * a.z * a.z, 2 * a.x * a.y + 2 * a.z * a.w, 2 * a.x * a.z - 2 * a.y * a.w);
    m.c1 = mk(2 * a.x * a.y - 2 * a.z * a.w, 1 - 2 * a.x * a.x - 2 * a.z * a.z, 2 * a.y * a.z + 2 * a.x * a.w);
    m.c2 = mk(2 * a.x * a.z + 2 * a.y * a.w, 2 * a.y * a.z - 2 * a.x * a.w, 1 - 2 * a.x * a.x - 2 * a.y * a.y);
    const float4 tq = qmul(make_float4(2 * b.x, 2 * b.y, 2 * b.z, 2 * b.w), make_float4(-a.x, -a.y, -a.z, a.w));
    m.t = mk(tq.x, tq.y, tq.z);
    return m;
}

// __float2int_rn applied to the reference's projection, which evaluates
// x / (z + 1e-10) * f + c in double and narrows to float first.
__device__ inline int project_round(float x, float z, float f, float c) {
    const float u = (float)((double)x / ((double)z + 1e-10) * (double)f + (double)c);
    return (int)rintf(u);
}

// is_zero_vertex (VectorUtils.h:214-216)
__device__ inline bool is_zero_vertex(float4 v) {
    return fabsf(v.x) < 1e-3f && fabsf(v.y) < 1e-3f && fabsf(v.z) < 1e-3f;
}

// compute_feature_huber_weight (huber_weight.h:11-27)
__device__ inline float huber(float r) {
    const float a = fabsf(r);
    return a < 7.5e-2f ? 1.0f : 7.5e-2f / a;
}

// -------------------------------------------------------------- frame / state
// Pointer table of one frame.  Lives in device memory so that a captured
// hipGraph stays valid when the caller's buffers change between frames.
struct FrameDev {
    uint32_t V, W, H, N, E, num_surfels;
    uint32_t num_feat[kMaxViews];
    uint32_t feat_base[kMaxViews + 1];  // prefix of feature candidates per view
    uint32_t num_cross;
    float fx[kMaxViews], fy[kMaxViews], cx[kMaxViews], cy[kMaxViews];
    Mat34 init[kMaxViews], init_inv[kMaxViews], w2c[kMaxViews], c2w[kMaxViews];
    const float4* vertex[kMaxViews];
    const float4* normal[kMaxViews];
    const float* ffg[kMaxViews];
    const float2* fg_grad[kMaxViews];
    const float* gray[kMaxViews];
    const float2* gray_grad[kMaxViews];
    const uint8_t* edge[kMaxViews];
    const float4* ref_v[kMaxViews];
    const float4* ref_n[kMaxViews];
    const float4* warp_v[kMaxViews];
    const uint32_t* index[kMaxViews];
    const float4* rgb[kMaxViews];
    const ushort4* surfel_knn;
    const float4* surfel_w;
    const DQ* node_se3_in;
    const float4* nodes;
    const ushort2* graph;
    const ushort4* feat[kMaxViews];
    const mvdr_cross_pair* cross;
};

// Penalty constants of one GN iteration (PenaltyConstants.h): lambda values.
struct Penalty {
    float smooth, density, foreground, feature, cross;
};

// Device-resident sizes written by the index build.
struct Counts {
    uint32_t dense, smooth, fg, feat, cross;   // term counts
    uint32_t off[5];                           // global term offsets (TermTypeOffset)
    uint32_t kv1, kv2, pairs;                  // Node2Term entries, pair-term entries, unique pairs
    uint32_t max_seg, max_row;                 // longest node / pair-row list seen
    uint32_t error;                            // bit mask of capacity overflows
};

enum ErrorBits : uint32_t {
    kErrDense = 1u << 0, kErrFg = 1u << 1, kErrFeat = 1u << 2, kErrCross = 1u << 3, kErrKv1 = 1u << 4,
    kErrKv2 = 1u << 5, kErrPairs = 1u << 6, kErrSeg = 1u << 7, kErrRow = 1u << 8, kErrNodes = 1u << 9,
};

// Term records (64 B dense / p2p, 32 B smooth) -- the MVDR_TAP_*_TERMS layouts.
struct ScalarRec {      // dense depth, density, foreground
    ushort4 knn;
    float r;
    uint32_t pad0;
    float4 w;
    float J[6];         // rotation xyz, translation xyz
    float pad1[2];
};
struct SmoothRec {
    uint32_t ij;        // i | j << 16
    uint32_t valid;
    float Ti_xj[3];
    float Tj_xj[3];
};
struct P2PRec {         // feature, cross view
    ushort4 knn;
    float pad0[2];
    float4 w;
    float4 target;
    float4 warped;
};
static_assert(sizeof(ScalarRec) == 64, "ScalarRec layout");
static_assert(sizeof(SmoothRec) == 32, "SmoothRec layout");
static_assert(sizeof(P2PRec) == 64, "P2PRec layout");

// Everything the kernels of one solver instance touch.
struct Buffers {
    const FrameDev* frame;
    Counts* counts;
    // capacities
    uint32_t cap_pixels, cap_feat, cap_cross, cap_edges, cap_kv1, cap_kv2, cap_pairs, cap_nodes;
    // term storage
    ushort4* dense_px;       // x, y, view, 0
    ScalarRec* dense;
    ScalarRec* density;
    ushort4* fg_px;
    ScalarRec* fg;
    SmoothRec* smooth;
    P2PRec* feat;
    float4* feat_canon;
    P2PRec* cross;
    float4* cross_canon;
    // compaction scratch
    uint32_t* tile_count;
    uint32_t* tile_offset;
    uint32_t* feat_flag;
    // node -> term
    uint32_t* n2t_count;     // N + 1
    uint32_t* n2t_offset;    // N + 1
    uint32_t* n2t_cursor;    // N
    uint32_t* n2t_index;     // kv1
    // node pairs
    uint32_t* row_count;     // N + 1 : pair-term entries per row
    uint32_t* row_offset;    // N + 1
    uint32_t* row_pairs;     // N + 1 : unique pairs per row
    uint32_t* pair_offset;   // N + 1
    uint32_t* uniq_j;        // kv2 scratch: unique j per row position
    uint2* uniq_range;       // kv2 scratch
    uint32_t* pair_terms;    // kv2
    uint32_t* pair_key;      // cap_pairs
    uint2* pair_range;       // cap_pairs
    // block-sparse rows
    uint32_t* low_count;     // N + 1
    uint32_t* low_offset;    // N + 1
    uint32_t* low_cursor;    // N
    uint2* low_ent;          // cap_pairs: (k, pair)
    // linear system
    float* diag;             // 36 N
    float* diag_inv;         // 36 N
    double* diag_inv64;      // 36 N, mixed mode only
    float* jtr;              // 6 N
    float* offdiag;          // 36 P
    // PCG
    float *x, *r, *s, *q, *d0, *d1;
    double* partial;         // kPartialSlots x kPartialStride per-block partial sums
    // warp field double buffer
    DQ* se3[2];
};

}  // namespace mvdr
