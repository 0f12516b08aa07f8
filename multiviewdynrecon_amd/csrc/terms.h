// Per-term Jacobian rows shared by the assembly, the matrix-free apply and the
// fix-up kernels: term types and bounds, knn slot weights, point-to-point and
// node-graph rows, and the node-centred coordinates of the solve.
#pragma once
#include "common.h"

namespace mvdr {

// ------------------------------------------------------- per-term Jacobians
enum : int { kDense = 0, kSmooth = 1, kForeground = 2, kFeature = 3, kCross = 4 };

// The term-type bounds c.off[0..3] held in registers: kernels that write
// global memory would otherwise reload them (with a full memory wait) at
// every use.
struct TermBounds {
    uint32_t o0, o1, o2, o3;
    __device__ explicit TermBounds(const Counts& c) : o0(c.off[0]), o1(c.off[1]), o2(c.off[2]), o3(c.off[3]) {}
};

__device__ inline int type_of(const TermBounds& c, uint32_t t, uint32_t& local) {
    if (t < c.o0) { local = t; return kDense; }
    if (t < c.o1) { local = t - c.o0; return kSmooth; }
    if (t < c.o2) { local = t - c.o1; return kForeground; }
    if (t < c.o3) { local = t - c.o2; return kFeature; }
    local = t - c.o3;
    return kCross;
}

__device__ inline int type_of(const Counts& c, uint32_t t, uint32_t& local) {
    if (t < c.off[0]) { local = t; return kDense; }
    if (t < c.off[1]) { local = t - c.off[0]; return kSmooth; }
    if (t < c.off[2]) { local = t - c.off[1]; return kForeground; }
    if (t < c.off[3]) { local = t - c.off[2]; return kFeature; }
    local = t - c.off[3];
    return kCross;
}

// Knn slot weight of node n (PreconditionerBlocked.cu:25-30): the slot index is
// the sum 1*(n==y) + 2*(n==z) + 3*(n==w).
__device__ inline float slot_weight(ushort4 k, float4 w, uint32_t n) {
    const int s = 1 * (n == k.y) + 2 * (n == k.z) + 3 * (n == k.w);
    float r = w.x;  // selects, not branches
    r = s == 1 ? w.y : r;
    r = s == 2 ? w.z : r;
    return s >= 3 ? w.w : r;
}

// Huber-weighted point-to-point Jacobian rows and residuals
// (computePointToPointICPTermJacobian, geometry_icp_jacobian.cuh:91-117).
__device__ inline void p2p_rows(const P2PRec& t, float J[3][6], float res[3]) {
    const float4 w = t.warped;
    const float r0 = w.x - t.target.x, r1 = w.y - t.target.y, r2 = w.z - t.target.z;
    const float h0 = huber(r0), h1 = huber(r1), h2 = huber(r2);
    J[0][0] = 0.f;        J[0][1] = h0 * w.z;  J[0][2] = -h0 * w.y; J[0][3] = h0;  J[0][4] = 0.f; J[0][5] = 0.f;
    J[1][0] = -h1 * w.z;  J[1][1] = 0.f;       J[1][2] = h1 * w.x;  J[1][3] = 0.f; J[1][4] = h1;  J[1][5] = 0.f;
    J[2][0] = h2 * w.y;   J[2][1] = -h2 * w.x; J[2][2] = 0.f;       J[2][3] = 0.f; J[2][4] = 0.f; J[2][5] = h2;
    res[0] = h0 * r0; res[1] = h1 * r1; res[2] = h2 * r2;
}

// computeSmoothTermJacobian (geometry_icp_jacobian.cuh:323-343) for node i or j.
__device__ inline void smooth_rows(const SmoothRec& s, bool node_i, float J[3][6], float res[3]) {
    const float sg = node_i ? 1.f : -1.f;
    const float a[3] = {node_i ? s.Ti_xj[0] : s.Ti_xj[0] - s.r[0], node_i ? s.Ti_xj[1] : s.Ti_xj[1] - s.r[1],
                        node_i ? s.Ti_xj[2] : s.Ti_xj[2] - s.r[2]};
    const float on = s.valid ? 1.f : 0.f;
    J[0][0] = 0.f;             J[0][1] = on * sg * a[2];  J[0][2] = -on * sg * a[1];
    J[1][0] = -on * sg * a[2]; J[1][1] = 0.f;             J[1][2] = on * sg * a[0];
    J[2][0] = on * sg * a[1];  J[2][1] = -on * sg * a[0]; J[2][2] = 0.f;
    J[0][3] = on * sg; J[0][4] = 0.f;     J[0][5] = 0.f;
    J[1][3] = 0.f;     J[1][4] = on * sg; J[1][5] = 0.f;
    J[2][3] = 0.f;     J[2][4] = 0.f;     J[2][5] = on * sg;
    res[0] = on * s.r[0];
    res[1] = on * s.r[1];
    res[2] = on * s.r[2];
}

// ------------------------------------------------ node-centred coordinates
// The reference twist (w, v) moves a point p by w x p + v: its rotation is
// about the world origin, so every node block has a near-null direction
// ("rotate about the node, translate back") and FP32 storage of J^T J costs
// ~1e-3 in the solution.  The solver assembles in the node-centred twist
// (w, u), u = v + w x g_n, i.e. x_ref = T_n x with T_n = [[I, 0], [[g_n]x, I]].
// Block-Jacobi PCG is invariant under this block-diagonal congruence
// (A' = T^T A T, M' = T^T M T, b' = T^T b give x' = T^-1 x in exact arithmetic),
// so the iterates are the reference's; only their FP32 rounding improves.
// A row J = (J_w, J_v) becomes J' = (J_w + J_v x g, J_v).
__device__ inline void to_node_frame(float J[6], const float4 g) {
    const float a = J[4] * g.z - J[5] * g.y;
    const float b = J[5] * g.x - J[3] * g.z;
    const float c = J[3] * g.y - J[4] * g.x;
    J[0] += a;
    J[1] += b;
    J[2] += c;
}

// The reference's damping 1e-2 I (PreconditionerBlocked.cu:272-275) seen in
// node-centred coordinates: 1e-2 T^T T = 1e-2 [[I + G^T G, G^T], [G, I]], G = [g]x.
__device__ inline void add_node_frame_damping(float D[36], const float4 g) {
    const float l = 1e-2f;
    const float gg = g.x * g.x + g.y * g.y + g.z * g.z;
    const float gv[3] = {g.x, g.y, g.z};
    const float G[3][3] = {{0.f, -g.z, g.y}, {g.z, 0.f, -g.x}, {-g.y, g.x, 0.f}};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            D[6 * r + c] += l * ((r == c ? 1.f + gg : 0.f) - gv[r] * gv[c]);
            D[6 * r + 3 + c] += l * G[c][r];        // G^T
            D[6 * (3 + r) + c] += l * G[r][c];      // G
            D[6 * (3 + r) + 3 + c] += (r == c) ? l : 0.f;
        }
}

// --------------------------------------------------------- per-term rows
// Jacobian rows of term t with respect to node n (knn slot weight of n
// applied), residuals and lambda^2.  Returns the number of rows (1..3).
__device__ inline int term_rows(const Buffers& b, const Counts& c, const Penalty& pen, uint32_t t, uint32_t n,
                                float J[3][6], float res[3], float lam2[3]) {
    uint32_t k;
    const int type = type_of(c, t, k);
    if (type == kDense || type == kForeground) {
        const ScalarRec& rec = type == kDense ? b.dense[k] : b.fg[k];
        const float w = slot_weight(rec.knn, rec.w, n);
#pragma unroll
        for (int q = 0; q < 6; ++q) J[0][q] = w * rec.J[q];
        res[0] = rec.r;
        lam2[0] = type == kDense ? 1.0f : pen.foreground * pen.foreground;
        if (type == kDense && b.density && pen.density != 0.f) {
            const ScalarRec& d = b.density[k];
#pragma unroll
            for (int q = 0; q < 6; ++q) J[1][q] = w * d.J[q];
            res[1] = d.r;
            lam2[1] = pen.density * pen.density;
            return 2;
        }
        return 1;
    }
    if (type == kSmooth) {
        const SmoothRec s = b.smooth[k];
        smooth_rows(s, n == (s.ij & 0xFFFF), J, res);
        lam2[0] = lam2[1] = lam2[2] = pen.smooth * pen.smooth;
        return 3;
    }
    const P2PRec& rec = type == kFeature ? b.feat[k] : b.cross[k];
    p2p_rows(rec, J, res);
    const float w = slot_weight(rec.knn, rec.w, n);
    const float l = type == kFeature ? pen.feature : pen.cross;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
        for (int q = 0; q < 6; ++q) J[ch][q] *= w;
        lam2[ch] = l * l;
    }
    return 3;
}

// Rows of term t for both nodes of the pair (i, j): one record read, one
// type dispatch.  lam2 holds the i-side lambda^2 (equal on both sides).
__device__ inline int pair_term_rows(const Buffers& b, const Counts& c, const Penalty& pen, uint32_t t, uint32_t i,
                                     uint32_t j, float Ji[3][6], float Jj[3][6], float lam2[3]) {
    uint32_t k;
    const int type = type_of(c, t, k);
    float res[3];
    if (type == kSmooth) {
        const SmoothRec sm = b.smooth[k];
        const uint32_t first = sm.ij & 0xFFFF;
        smooth_rows(sm, i == first, Ji, res);
        smooth_rows(sm, j == first, Jj, res);
        lam2[0] = lam2[1] = lam2[2] = pen.smooth * pen.smooth;
        return 3;
    }
    if (type == kDense || type == kForeground) {
        const ScalarRec& rec = type == kDense ? b.dense[k] : b.fg[k];
        const ushort4 knn = rec.knn;
        const float4 w = rec.w;
        const float wi = slot_weight(knn, w, i), wj = slot_weight(knn, w, j);
        float J0[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) J0[q] = rec.J[q];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Ji[0][q] = wi * J0[q];
            Jj[0][q] = wj * J0[q];
        }
        lam2[0] = type == kDense ? 1.0f : pen.foreground * pen.foreground;
        if (type == kDense && b.density && pen.density != 0.f) {
            const ScalarRec& d = b.density[k];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                Ji[1][q] = wi * d.J[q];
                Jj[1][q] = wj * d.J[q];
            }
            lam2[1] = pen.density * pen.density;
            return 2;
        }
        return 1;
    }
    const P2PRec& rec = type == kFeature ? b.feat[k] : b.cross[k];
    float J[3][6];
    p2p_rows(rec, J, res);
    const float wi = slot_weight(rec.knn, rec.w, i), wj = slot_weight(rec.knn, rec.w, j);
    const float l = type == kFeature ? pen.feature : pen.cross;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Ji[ch][q] = wi * J[ch][q];
            Jj[ch][q] = wj * J[ch][q];
        }
        lam2[ch] = l * l;
    }
    return 3;
}

}  // namespace mvdr
