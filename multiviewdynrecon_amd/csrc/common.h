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
constexpr uint32_t kPartialStride = 1024;  // partial sums per reduction slot, at least (per-block sums)
constexpr uint32_t kPartialSlots = 160;
constexpr uint32_t kSlotCost = 150;        // partials of the total residual

// ---------------------------------------------------------------- vectors
// V3<float> / M34<float> carry the reference's FP32 arithmetic; the mixed
// precision mode instantiates the same code with double.
template <typename R>
struct V3 {
    R x, y, z;
};
using f3 = V3<float>;
__host__ __device__ inline f3 mk(float a, float b, float c) { return {a, b, c}; }
template <typename R> __host__ __device__ inline V3<R> operator+(V3<R> a, V3<R> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename R> __host__ __device__ inline V3<R> operator-(V3<R> a, V3<R> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__host__ __device__ inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> __host__ __device__ inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __host__ __device__ inline V3<R> cross3(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
template <typename R> __host__ __device__ inline V3<R> widen(f3 v) { return {(R)v.x, (R)v.y, (R)v.z}; }
template <typename R> __host__ __device__ inline V3<R> xyz_as(float4 v) { return {(R)v.x, (R)v.y, (R)v.z}; }

// Rigid transform, rotation as columns c0, c1, c2.
template <typename R>
struct M34 {
    V3<R> c0, c1, c2, t;
    __host__ __device__ V3<R> rot(V3<R> p) const {
        return {c0.x * p.x + c1.x * p.y + c2.x * p.z, c0.y * p.x + c1.y * p.y + c2.y * p.z,
                c0.z * p.x + c1.z * p.y + c2.z * p.z};
    }
    __host__ __device__ V3<R> rot_t(V3<R> p) const { return {dot3(c0, p), dot3(c1, p), dot3(c2, p)}; }
    __host__ __device__ V3<R> apply(V3<R> p) const { return rot(p) + t; }
};
using Mat34 = M34<float>;
template <typename R> __host__ __device__ inline M34<R> widen(const Mat34& m) {
    return {widen<R>(m.c0), widen<R>(m.c1), widen<R>(m.c2), widen<R>(m.t)};
}

inline Mat34 to_mat34(const mvdr_mat34& m) {
    Mat34 o;
    o.c0 = mk(m.rot[0], m.rot[1], m.rot[2]);
    o.c1 = mk(m.rot[3], m.rot[4], m.rot[5]);
    o.c2 = mk(m.rot[6], m.rot[7], m.rot[8]);
    o.t = mk(m.trans[0], m.trans[1], m.trans[2]);
    return o;
}

// mat34::inverse (MatUtils.h:585-590): R^T, -R^T t
template <typename R> __host__ __device__ inline M34<R> inverse(const M34<R>& m) {
    M34<R> o;
    o.c0 = {m.c0.x, m.c1.x, m.c2.x};
    o.c1 = {m.c0.y, m.c1.y, m.c2.y};
    o.c2 = {m.c0.z, m.c1.z, m.c2.z};
    const V3<R> rt = o.rot(m.t);
    o.t = {-rt.x, -rt.y, -rt.z};
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

// DualQuaternion(const mat34&) (DualQuaternion.h:123-130) with Quaternion(mat33)
// (Quaternion.h:46-77).
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

// The same blend and normalisation in R: the mixed precision mode evaluates
// term geometry in double from the stored FP32 dual quaternions.
template <typename R>
struct QR {
    R x, y, z, w;
};
template <typename R>
struct DQR {
    QR<R> q0, q1;
};
template <typename R> __device__ inline QR<R> widen(float4 q) { return {(R)q.x, (R)q.y, (R)q.z, (R)q.w}; }
template <typename R> __device__ inline QR<R> qmul(QR<R> a, QR<R> b) {
    QR<R> q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return q;
}
template <typename R>
__device__ inline DQR<R> blend4_as(const DQ* __restrict__ field, ushort4 knn, float4 w) {
    const uint32_t id[4] = {knn.x, knn.y, knn.z, knn.w};
    const R wk[4] = {(R)w.x, (R)w.y, (R)w.z, (R)w.w};
    DQR<R> o;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const QR<R> a = widen<R>(field[id[k]].q0), b = widen<R>(field[id[k]].q1);
        if (k == 0) {
            o.q0 = {wk[0] * a.x, wk[0] * a.y, wk[0] * a.z, wk[0] * a.w};
            o.q1 = {wk[0] * b.x, wk[0] * b.y, wk[0] * b.z, wk[0] * b.w};
        } else {
            o.q0 = {o.q0.x + wk[k] * a.x, o.q0.y + wk[k] * a.y, o.q0.z + wk[k] * a.z, o.q0.w + wk[k] * a.w};
            o.q1 = {o.q1.x + wk[k] * b.x, o.q1.y + wk[k] * b.y, o.q1.z + wk[k] * b.z, o.q1.w + wk[k] * b.w};
        }
    }
    return o;
}
template <typename R> __device__ inline DQR<R> widen(const DQ& d) { return {widen<R>(d.q0), widen<R>(d.q1)}; }
template <typename R> __device__ inline M34<R> se3_as(DQR<R> d) {
    const R inv = R(1) / sqrt(d.q0.w * d.q0.w + d.q0.x * d.q0.x + d.q0.y * d.q0.y + d.q0.z * d.q0.z);
    const QR<R> a = {inv * d.q0.x, inv * d.q0.y, inv * d.q0.z, inv * d.q0.w};
    QR<R> b = {inv * d.q1.x, inv * d.q1.y, inv * d.q1.z, inv * d.q1.w};
    const R dd = a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z;
    b = {b.x - dd * a.x, b.y - dd * a.y, b.z - dd * a.z, b.w - dd * a.w};
    const R one = 1, two = 2;
    M34<R> m;
    m.c0 = {one - two * a.y * a.y - two * a.z * a.z, two * a.x * a.y + two * a.z * a.w, two * a.x * a.z - two * a.y * a.w};
    m.c1 = {two * a.x * a.y - two * a.z * a.w, one - two * a.x * a.x - two * a.z * a.z, two * a.y * a.z + two * a.x * a.w};
    m.c2 = {two * a.x * a.z + two * a.y * a.w, two * a.y * a.z - two * a.x * a.w, one - two * a.x * a.x - two * a.y * a.y};
    const QR<R> tq = qmul<R>({two * b.x, two * b.y, two * b.z, two * b.w}, {-a.x, -a.y, -a.z, a.w});
    m.t = {tq.x, tq.y, tq.z};
    return m;
}
// Warp of a knn-blended point: FP32 exactly as the reference, or double.
template <typename R> __device__ inline M34<R> warp_of(const DQ* __restrict__ field, ushort4 knn, float4 w) {
    if constexpr (sizeof(R) == sizeof(float)) return se3_of(blend4(field, knn, w));
    else return se3_as<R>(blend4_as<R>(field, knn, w));
}
template <typename R> __device__ inline M34<R> node_warp(const DQ& d) {
    if constexpr (sizeof(R) == sizeof(float)) return se3_of(d);
    else return se3_as<R>(widen<R>(d));
}

// __float2int_rn applied to the reference's projection, which evaluates
// x / (z + 1e-10) * f + c in double and narrows to float first.
template <typename R> __device__ inline int project_round(R x, R z, float f, float c) {
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
    const ushort4* corr[kMaxViews];     // corrMap per view (nullable)
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
    uint32_t max_seg, max_row;                 // longest Node2Term list, most pairs of a row
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
// The smooth residual is stored, not the second warped point: the two warped
// points are ~1 m from the origin while their difference is ~1e-4 m, so their
// FP32 difference would carry a ~1e-3 relative error.  T_j x_j = Ti_xj - r is
// exact when both came from FP32 values (Sterbenz).
struct SmoothRec {
    uint32_t ij;        // i | j << 16
    uint32_t valid;
    float Ti_xj[3];
    float r[3];         // T_i x_j - T_j x_j
};
struct P2PRec {         // feature, cross view
    ushort4 knn;
    float pad0[2];
    float4 w;
    float4 target;
    float4 warped;
};
static_assert(sizeof(ScalarRec) == 64, "ScalarRec layout");
static_assert(offsetof(ScalarRec, r) == 8 && offsetof(ScalarRec, w) == 16 && offsetof(ScalarRec, J) == 32,
              "ScalarRec layout");
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
    ushort4* dense_knn;      // knn of each dense term (a packed copy of the records' first field)
    ScalarRec* dense;
    ScalarRec* density;
    ushort4* fg_px;
    ScalarRec* fg;
    SmoothRec* smooth;
    P2PRec* feat;
    float4* feat_canon;
    P2PRec* cross;
    float4* cross_canon;
    // selection staging, one slot per feature then cross-view candidate
    // (cap_feat + cap_cross): flag, record, canonical vertex
    uint32_t* sel_ok;
    P2PRec* sel_rec;
    float4* sel_canon;
    // compaction scratch
    uint32_t* tile_count;
    uint32_t* tile_offset;
    uint32_t* feat_flag;
    // node -> term
    uint32_t* n2t_hist;      // chunks x N: per-chunk node counts, then chunk offsets
    uint32_t* n2t_count;     // N + 1
    uint32_t* n2t_offset;    // N + 1
    uint32_t* n2t_index;     // kv1
    uint16_t* n2t_node;      // kv1: the node of each Node2Term entry
    uint32_t* n2t_col;       // kv1: weight columns of the entry's knn slots, 4 x u8 (row Gram assembly)
    // node pairs (i, j), i < j, in (i, j) order
    uint32_t* row_up;        // N + 1 : pairs (i, j > i) per row i
    uint32_t* row_low;       // N + 1 : pairs (k < i, i) per row i
    uint32_t* row_kv2;       // N + 1 : pair-term entries of row i's pairs
    uint32_t* pair_offset;   // N + 1 : first pair of row i
    uint32_t* kv2_offset;    // N + 1 : first pair-term entry of row i
    uint32_t* pair_key;      // cap_pairs: i << 16 | j
    uint2* pair_range;       // cap_pairs: [begin, end) into pair_terms
    uint32_t* pair_terms;    // kv2: NodePair2Term term lists
    uint32_t* kv2_pair;      // kv2: the pair of each entry
    uint32_t* pair_edge_n;   // cap_pairs: graph edges joining the pair (row Gram assembly)
    uint32_t* pair_edge;     // kPairEdges x cap_pairs: their edge indices
    uint16_t* row_part;      // N x kRowPartners: each row's partners ascending, upper then lower
    float4* mf_u;            // terms: u_t of the matrix-free apply
    float4* mf_g;            // 2 per term: J_t^T u_t of the recomputing matrix-free apply
    float* n2t_w;            // kv1: slot weight of each Node2Term entry's node (sign: several slots)
    // assembly partials of runs crossing 64-entry chunk boundaries
    float* part_diag;        // (kv1 / 64 + 1) x 2 x 27
    float* part_pair;        // (kv2 / 64 + 1) x 2 x 36
    // block-sparse rows
    uint32_t* low_offset;    // N + 1
    uint2* low_ent;          // cap_pairs: (k, pair) of row i's lower entries, k ascending
    uint32_t* low_pos;       // cap_pairs: position in low_ent of each pair's lower entry (mixed mode)
    // linear system
    float* diag;             // 36 N
    float* diag_inv;         // 36 N
    double* diag_inv64;      // 36 N, mixed mode only
    float* jtr;              // 6 N
    float* offdiag;          // 36 P
    // PCG
    float *x, *r, *s, *q, *d0, *d1;
    // fused CG on A^ = L^-1 A L^-T (mixed precision mode)
    double* linv64;          // 21 N: L^-1 of each diagonal block, packed lower triangle (tri_index)
    // L^-1 A L^-T: the upper blocks A^_ij (pair order, row-major) and, by
    // block rows, the slots of row i -- its lower entries (low_ent order), then
    // its upper entries (pair order) -- as (column node, pair | transposed << 31);
    // row i starts at slot low_offset[i] + pair_offset[i]
    float* offhat;           // 36 x P
    uint2* row_ent;          // 2P
    uint32_t* row_ptr;       // N + 1: first slot of each block row
    // fused CG: per node one record (r 0..5, w 6..11, s 12..17, 2 pad) of the
    // set read by iteration i (cg_rws[i & 1]) and of the set it writes
    float *cg_rws[2], *cg_p, *cg_x, *cg_b;
    double* cg_scal;         // alpha_i, gamma_i per iteration
    double* partial;         // kPartialSlots x partial_stride partial sums (per block, or per 32-row CG group)
    uint32_t partial_stride; // max(kPartialStride, CG row groups of max_nodes)
    // warp field double buffer
    DQ* se3[2];
};

// The Buffers of this workgroup's frame (bs[blockIdx.y]).  Read through the
// constant address space: pointers loaded from constant memory are known to
// point to global memory, so the accesses through them compile to global
// loads / stores with scalar bases instead of flat ones.  (The host pass of
// the compiler only parses device code; it takes the plain read.)
__device__ inline Buffers frame_buffers(const Buffers* bs) {
#if __HIP_DEVICE_COMPILE__
    return *((const __attribute__((address_space(4))) Buffers*)bs + blockIdx.y);
#else
    return bs[blockIdx.y];
#endif
}

// One launch over `frames` solver instances: the device array of their
// Buffers (kernels read bs[blockIdx.y]), the largest capacities among them and
// the largest node count of their current frames (rounded up to 64), which size
// the grids; every kernel bounds its work by its own frame's sizes.
struct Batch {
    const Buffers* dev;
    uint32_t frames;
    uint32_t cap_tiles, nodes, cap_chunks;
};

// Index variants (launch_index_build).
enum : int { kIndexLists = 0, kIndexGram = 1, kIndexNodes = 2 };

// Packed row-major lower triangle of a 6x6 matrix: entry (a, k), k <= a.
__host__ __device__ constexpr uint32_t tri_index(uint32_t a, uint32_t k) { return a * (a + 1) / 2 + k; }
constexpr uint32_t kTri = 21;
constexpr uint32_t kRws = 20;  // floats per node record of the fused CG's r, w, s

// Partner nodes a row may have (upper and lower pairs together).
constexpr uint32_t kRowPartners = 512;

// Graph edges per node pair the row Gram assembly can hold.
constexpr uint32_t kPairEdges = 4;

}  // namespace mvdr
