// One Gauss-Newton iteration on the device: term residuals / Jacobians,
// J^T J diagonal + J^T r + block-Jacobi inverse per node, J^T J off-diagonal
// blocks per node pair, block-sparse PCG and the warp-field update.
#include "common.h"
#include "device_util.h"

namespace mvdr {

// ---------------------------------------------------------- term evaluation
// R = float is the reference's arithmetic; R = double (mixed precision mode)
// evaluates the warp, camera chain and projection in FP64 and stores J and r in
// FP32.  In FP32, p_live carries ~1e-7 m of rounding: enough to move ~0.1% of
// the projections to a neighbouring depth pixel and to cost ~1e-3 in the twist.
template <typename R> __device__ inline M34<R> init_inverse(const FrameDev& f, uint32_t v) {
    if constexpr (sizeof(R) == sizeof(float)) return f.init_inv[v];
    else return inverse(widen<R>(f.init[v]));
}

// computeDenseDepthJacobianKernel (DenseDepthJacobianResidual.cu:24-104).  The
// camera of a term is the view stored with its compacted pixel.
template <typename R>
__device__ inline void eval_dense(const FrameDev& f, const DQ* se3, ushort4 px, ScalarRec& rec) {
    const uint32_t v = px.z;
    const uint32_t o = px.y * f.W + px.x;
    const V3<R> cv = xyz_as<R>(f.ref_v[v][o]), cn = xyz_as<R>(f.ref_n[v][o]);
    const M34<R> I = widen<R>(f.init[v]), W2C = widen<R>(f.w2c[v]), Iinv = init_inverse<R>(f, v);
    const V3<R> v0 = I.apply(cv), n0 = I.rot(cn);
    const M34<R> T = warp_of<R>(se3, rec.knn, rec.w);
    const V3<R> pl = T.apply(v0), nl = T.rot(n0);
    V3<R> pc = W2C.apply(pl), nc = W2C.rot(nl);
    pc = Iinv.apply(pc);
    nc = Iinv.rot(nc);
    float r = 0.f, J[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int u = project_round<R>(pc.x, pc.z, f.fx[v], f.cx[v]);
    const int w = project_round<R>(pc.y, pc.z, f.fy[v], f.cy[v]);
    if (u >= 0 && u < (int)f.W && w >= 0 && w < (int)f.H) {
        const float4 dv = f.vertex[v][w * f.W + u];
        const float4 dn = f.normal[v][w * f.W + u];
        const V3<R> d = xyz_as<R>(dv), n = xyz_as<R>(dn);
        bool valid = !is_zero_vertex(dv);
        if (dot3(n, nc) < R(0.7f)) valid = false;
        const V3<R> diff = d - pc;
        if (dot3(diff, diff) > R(0.075f) * R(0.075f)) valid = false;
        if (valid) {
            r = (float)dot3(n, pc - d);
            const V3<R> jt = W2C.rot_t(I.rot(n));
            const V3<R> jr = cross3(pl, jt);
            J[0] = (float)jr.x; J[1] = (float)jr.y; J[2] = (float)jr.z;
            J[3] = (float)jt.x; J[4] = (float)jt.y; J[5] = (float)jt.z;
        }
    }
    rec.r = r;
#pragma unroll
    for (int k = 0; k < 6; ++k) rec.J[k] = J[k];
}

// computeImageDensityJacobainAndResidual (density_map_jacobian.cuh:12-79) as
// called for density and foreground terms (DensityForegroundJacobianResidual.cu:41-142).
// The camera-v reference vertex is warped without the move to camera 0, as there.
template <typename R>
__device__ inline void eval_image(const FrameDev& f, const DQ* se3, ushort4 px, bool density, ScalarRec& rec) {
    const uint32_t v = px.z;
    const uint32_t o = px.y * f.W + px.x;
    const V3<R> cv = xyz_as<R>(f.ref_v[v][o]);
    R geometry = 0;
    if (density) {
        const float4 rgb = f.rgb[v][o];
        geometry = R(0.299f) * rgb.x + R(0.587f) * rgb.y + R(0.114f) * rgb.z;
    }
    const M34<R> T = warp_of<R>(se3, rec.knn, rec.w);
    const V3<R> pl = T.apply(cv);
    const V3<R> pc = init_inverse<R>(f, v).apply(widen<R>(f.w2c[v]).apply(pl));
    float r = 0.f, J[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const R fx = f.fx[v], fy = f.fy[v];
    const int u = project_round<R>(pc.x, pc.z, f.fx[v], f.cx[v]);
    const int w = project_round<R>(pc.y, pc.z, f.fy[v], f.cy[v]);
    if (u >= 0 && u < (int)f.W && w >= 0 && w < (int)f.H) {
        const uint32_t q = w * f.W + u;
        const R img = density ? f.gray[v][q] : f.ffg[v][q];
        const float2 g = density ? f.gray_grad[v][q] : f.fg_grad[v][q];
        V3<R> d;
        d.x = R(g.x) * fx / pc.z;
        d.y = R(g.y) * fy / pc.z;
        d.z = -R(1) / (pc.z * pc.z) * (fx * pc.x * R(g.x) + fy * pc.y * R(g.y));
        d = widen<R>(f.init[v]).rot(d);
        d = widen<R>(f.w2c[v]).rot_t(d);
        J[0] = (float)(-d.y * pl.z + d.z * pl.y);
        J[1] = (float)(d.x * pl.z - d.z * pl.x);
        J[2] = (float)(-d.x * pl.y + d.y * pl.x);
        J[3] = (float)d.x; J[4] = (float)d.y; J[5] = (float)d.z;
        r = (float)(img - geometry);
    }
    rec.r = r;
#pragma unroll
    for (int k = 0; k < 6; ++k) rec.J[k] = J[k];
}

// One thread per global term; smooth terms forward-warp the graph edge
// (NodeGraphSmoothHandler.cu:10-43), feature / cross terms their canonical
// vertex (SparseCorrespondenceHandler.cu:193-220).
template <typename R>
__global__ void __launch_bounds__(256) eval_terms_kernel(Buffers b, const DQ* __restrict__ se3) {
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < c.off[4]; t += stride) {
        if (t < c.off[0]) {
            const ushort4 px = b.dense_px[t];
            ScalarRec rec = b.dense[t];
            eval_dense<R>(f, se3, px, rec);
            b.dense[t].r = rec.r;
#pragma unroll
            for (int k = 0; k < 6; ++k) b.dense[t].J[k] = rec.J[k];
            if (b.density) {
                ScalarRec d = b.density[t];
                eval_image<R>(f, se3, px, true, d);
                b.density[t].r = d.r;
#pragma unroll
                for (int k = 0; k < 6; ++k) b.density[t].J[k] = d.J[k];
            }
        } else if (t < c.off[1]) {
            const uint32_t e = t - c.off[0];
            const ushort2 ij = f.graph[e];
            const float4 xi = f.nodes[ij.x], xj = f.nodes[ij.y];
            const M34<R> Ti = node_warp<R>(se3[ij.x]), Tj = node_warp<R>(se3[ij.y]);
            const V3<R> a = Ti.apply(xyz_as<R>(xj)), s = Tj.apply(xyz_as<R>(xj));
            const V3<R> r = a - s;
            const f3 d = xyz(xi) - xyz(xj);
            SmoothRec rec;
            rec.ij = (uint32_t)ij.x | ((uint32_t)ij.y << 16);
            rec.valid = dot3(d, d) > 64 * 0.025f * 0.025f ? 0u : 1u;
            rec.Ti_xj[0] = (float)a.x; rec.Ti_xj[1] = (float)a.y; rec.Ti_xj[2] = (float)a.z;
            rec.r[0] = (float)r.x; rec.r[1] = (float)r.y; rec.r[2] = (float)r.z;
            b.smooth[e] = rec;
        } else if (t < c.off[2]) {
            const uint32_t k = t - c.off[1];
            ScalarRec rec = b.fg[k];
            eval_image<R>(f, se3, b.fg_px[k], false, rec);
            b.fg[k].r = rec.r;
#pragma unroll
            for (int q = 0; q < 6; ++q) b.fg[k].J[q] = rec.J[q];
        } else {
            const bool is_cross = t >= c.off[3];
            const uint32_t k = is_cross ? t - c.off[3] : t - c.off[2];
            P2PRec* recs = is_cross ? b.cross : b.feat;
            const float4 can = is_cross ? b.cross_canon[k] : b.feat_canon[k];
            const M34<R> T = warp_of<R>(se3, recs[k].knn, recs[k].w);
            const V3<R> w = T.apply(xyz_as<R>(can));
            recs[k].warped = make_float4((float)w.x, (float)w.y, (float)w.z, 1.0f);
        }
    }
}

// ------------------------------------------------------- per-term Jacobians
enum : int { kDense = 0, kSmooth = 1, kForeground = 2, kFeature = 3, kCross = 4 };

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
    return s == 0 ? w.x : (s == 1 ? w.y : (s == 2 ? w.z : w.w));
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

// -------------------------------------------------- 6x6 SPD inverse (Cholesky)
// inv(A) = L^-T L^-1 with A = L L^T; A is symmetric positive definite (the
// reference adds 1e-2 to every diagonal, PreconditionerBlocked.cu:272-275).
template <typename T>
__device__ inline void spd_inverse6(const float* A, float* out, double* out64) {
    T L[21];  // packed lower triangle, row-major
#define LI(r, c) L[(r) * ((r) + 1) / 2 + (c)]
    for (int r = 0; r < 6; ++r) {
        for (int col = 0; col <= r; ++col) {
            T s = (T)A[6 * r + col];
            for (int k = 0; k < col; ++k) s -= LI(r, k) * LI(col, k);
            if (r == col) LI(r, r) = s > (T)0 ? sqrt(s) : (T)1e-30;
            else LI(r, col) = s / LI(col, col);
        }
    }
    // X = L^-1 (lower), in place of a dense 6x6
    T X[36];
    for (int c2 = 0; c2 < 6; ++c2)
        for (int r = 0; r < 6; ++r) {
            T s = (r == c2) ? (T)1 : (T)0;
            for (int k = 0; k < r; ++k) s -= LI(r, k) * X[6 * k + c2];
            X[6 * r + c2] = s / LI(r, r);
        }
#undef LI
    // inv = X^T X
    for (int r = 0; r < 6; ++r)
        for (int col = 0; col < 6; ++col) {
            T s = 0;
            for (int k = 0; k < 6; ++k) s += X[6 * k + r] * X[6 * k + col];
            out[6 * r + col] = (float)s;
            if (out64) out64[6 * r + col] = (double)s;
        }
}

// ------------------------------------------------------ fused row assembly
// One block per node row i walks the row's Node2Term list once and produces
//   D_i  = sum lambda^2 J_i^T J_i + 1e-2 I
//          (computeBlockDiagonalPreconditionerKernel, PreconditionerBlocked.cu:172-284),
//   b_i  = -sum lambda^2 J_i^T r
//          (computeJtResidualWithIndexKernel, JtDotResidualIndexed.cu:116-224),
//   B_ij = sum lambda^2 J_i^T J_j for every pair (i, j), j > i, of the row
//          (computeJtJNonDiagonalBlockKernel, JtJMaterializer.cu:270-365),
// then the block-Jacobi inverse (blockDiagonalInverseKernel,
// BlockDiagonalPreconditionerInverse.cu:6-41).  A term enters B_ij once per
// Node2Term entry of row i times the number of its knn slots holding j -- the
// multiplicity of the reference's NodePair2Term lists -- so no per-pair term
// list is needed.  Batches of kAsmBatch entries are staged in LDS (node-frame
// i-side rows, raw partner rows, residuals, weights, one coefficient per pair);
// lane (u, a) then accumulates row a of block u (u = 0: the diagonal) over the
// batch in term order.  Deterministic: a fixed summation order per output.
constexpr int kAsmThreads = 128;
constexpr int kAsmBatch = 64;
constexpr int kAsmPairs = 20;   // pairs per pass: (1 + 20) * 6 = 126 lanes
constexpr int kAsmStride = 63;  // floats per staged entry (odd: conflict-free LDS)
constexpr int kOffJi = 0, kOffJp = 18, kOffRes = 36, kOffLam = 39, kOffRows = 42, kOffCoef = 43;
static_assert(kOffCoef + kAsmPairs == kAsmStride, "staged entry layout");

// Rows of term t seen from node i: i-side rows (slot weight of i applied), the
// raw rows a partner node multiplies by its own weight (a graph edge: the rows
// of the other end), residuals and lambda^2.  coef(j) for a partner j follows.
struct RowTerm {
    int rows;
    float Ji[3][6], Jp[3][6], res[3], lam2[3];
    bool smooth;
    uint32_t partner;  // smooth: the other end (0xFFFFFFFF for a loop)
    ushort4 knn;
    float4 w;
};

__device__ inline void row_term(const Buffers& b, const Counts& c, const Penalty& pen, uint32_t t, uint32_t i,
                                RowTerm& o) {
    uint32_t k;
    const int type = type_of(c, t, k);
    o.smooth = false;
    o.partner = 0xFFFFFFFFu;
    if (type == kDense || type == kForeground) {
        const ScalarRec& rec = type == kDense ? b.dense[k] : b.fg[k];
        o.knn = rec.knn;
        o.w = rec.w;
        const float wi = slot_weight(rec.knn, rec.w, i);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            o.Jp[0][q] = rec.J[q];
            o.Ji[0][q] = wi * rec.J[q];
        }
        o.res[0] = rec.r;
        o.lam2[0] = type == kDense ? 1.0f : pen.foreground * pen.foreground;
        o.rows = 1;
        if (type == kDense && b.density && pen.density != 0.f) {
            const ScalarRec& d = b.density[k];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                o.Jp[1][q] = d.J[q];
                o.Ji[1][q] = wi * d.J[q];
            }
            o.res[1] = d.r;
            o.lam2[1] = pen.density * pen.density;
            o.rows = 2;
        }
        return;
    }
    if (type == kSmooth) {
        const SmoothRec s = b.smooth[k];
        const uint32_t first = s.ij & 0xFFFF, second = s.ij >> 16;
        const bool is_first = i == first;
        float r2[3];
        smooth_rows(s, is_first, o.Ji, o.res);
        smooth_rows(s, !is_first, o.Jp, r2);
        o.smooth = true;
        o.partner = first == second ? 0xFFFFFFFFu : (is_first ? second : first);
        o.lam2[0] = o.lam2[1] = o.lam2[2] = pen.smooth * pen.smooth;
        o.rows = 3;
        return;
    }
    const P2PRec& rec = type == kFeature ? b.feat[k] : b.cross[k];
    o.knn = rec.knn;
    o.w = rec.w;
    p2p_rows(rec, o.Jp, o.res);
    const float wi = slot_weight(rec.knn, rec.w, i);
    const float l = type == kFeature ? pen.feature : pen.cross;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
        for (int q = 0; q < 6; ++q) o.Ji[ch][q] = wi * o.Jp[ch][q];
        o.lam2[ch] = l * l;
    }
    o.rows = 3;
}

// Weight of partner j in term o: (knn slots holding j) x slot weight of j.
__device__ inline float partner_coef(const RowTerm& o, uint32_t j) {
    if (o.smooth) return o.partner == j ? 1.0f : 0.0f;
    const int cnt = (o.knn.x == j) + (o.knn.y == j) + (o.knn.z == j) + (o.knn.w == j);
    return cnt ? (float)cnt * slot_weight(o.knn, o.w, j) : 0.0f;
}

template <typename InvT>
__global__ void __launch_bounds__(kAsmThreads) assemble_kernel(Buffers b, Penalty pen) {
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t i = blockIdx.x;
    if (i >= f.N) return;
    __shared__ float stage[kAsmBatch * kAsmStride];
    __shared__ float sD[36];
    __shared__ float4 sG[kAsmPairs + 1];
    __shared__ uint32_t sJ[kAsmPairs];
    const uint32_t tid = threadIdx.x;
    const float4 gi = f.nodes[i];
    const uint32_t beg = b.n2t_offset[i], end = b.n2t_offset[i + 1];
    const uint32_t pbeg = b.pair_offset[i], U = b.pair_offset[i + 1] - pbeg;
    const uint32_t passes = U == 0 ? 1u : (U + kAsmPairs - 1) / kAsmPairs;
    const uint32_t u = tid / 6, a = tid % 6;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const uint32_t u0 = pass * kAsmPairs;
        const uint32_t np = min((uint32_t)kAsmPairs, U - min(U, u0));
        if (tid < np) {
            const uint32_t j = b.pair_key[pbeg + u0 + tid] & 0xFFFF;
            sJ[tid] = j;
            sG[tid + 1] = f.nodes[j];
        }
        if (tid == 0) sG[0] = make_float4(0.f, 0.f, 0.f, 0.f);
        // lanes: u = 0 the diagonal (first pass only), u = 1..np the row's pairs
        const bool active = tid < 6 * (1 + np) && (u > 0 || pass == 0);
        float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float accr = 0.f;
        __syncthreads();
        const float4 g = sG[u < np + 1 ? u : 0];
        for (uint32_t base = beg; base < end; base += kAsmBatch) {
            const uint32_t nb = min((uint32_t)kAsmBatch, end - base);
            if (tid < nb) {
                RowTerm o;
                row_term(b, c, pen, b.n2t_index[base + tid], i, o);
                float* E = stage + tid * kAsmStride;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    if (r < o.rows) {
                        to_node_frame(o.Ji[r], gi);
#pragma unroll
                        for (int q = 0; q < 6; ++q) {
                            E[kOffJi + 6 * r + q] = o.Ji[r][q];
                            E[kOffJp + 6 * r + q] = o.Jp[r][q];
                        }
                        E[kOffRes + r] = o.res[r];
                        E[kOffLam + r] = o.lam2[r];
                    }
                }
                E[kOffRows] = (float)o.rows;
                for (uint32_t q = 0; q < np; ++q) E[kOffCoef + q] = partner_coef(o, sJ[q]);
            }
            __syncthreads();
            if (active) {
                for (uint32_t e = 0; e < nb; ++e) {
                    const float* E = stage + e * kAsmStride;
                    const float coef = u == 0 ? 1.0f : E[kOffCoef + u - 1];
                    if (coef == 0.f) continue;
                    const int rows = (int)E[kOffRows];
                    for (int r = 0; r < rows; ++r) {
                        const float A = E[kOffLam + r] * E[kOffJi + 6 * r + a] * coef;
                        const float* P = E + (u == 0 ? kOffJi : kOffJp) + 6 * r;
                        float p[6] = {P[0], P[1], P[2], P[3], P[4], P[5]};
                        // partner rows into the partner's node frame (identity for u = 0: g = 0)
                        p[0] += p[4] * g.z - p[5] * g.y;
                        p[1] += p[5] * g.x - p[3] * g.z;
                        p[2] += p[3] * g.y - p[4] * g.x;
#pragma unroll
                        for (int q = 0; q < 6; ++q) acc[q] += A * p[q];
                        accr += A * E[kOffRes + r];
                    }
                }
            }
            __syncthreads();
        }
        if (active && u > 0) {
            float* out = b.offdiag + 36 * (size_t)(pbeg + u0 + u - 1) + 6 * a;
#pragma unroll
            for (int q = 0; q < 6; ++q) out[q] = acc[q];
        }
        if (active && u == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) sD[6 * a + q] = acc[q];
            b.jtr[6 * (size_t)i + a] = -accr;
        }
        __syncthreads();
    }
    if (tid == 0) {
        float D[36];
#pragma unroll
        for (int q = 0; q < 36; ++q) D[q] = sD[q];
        add_node_frame_damping(D, gi);
        float* Dout = b.diag + 36 * (size_t)i;
#pragma unroll
        for (int q = 0; q < 36; ++q) Dout[q] = D[q];
        spd_inverse6<InvT>(D, b.diag_inv + 36 * (size_t)i, sizeof(InvT) == 8 ? b.diag_inv64 + 36 * (size_t)i : nullptr);
    }
}

// ---------------------------------------------------------- total residual
// computeTermResidualKernel (ResidualEvaluator.cu:71-142): sum of squared
// lambda-weighted residuals, accumulated in double.
__global__ void __launch_bounds__(256) cost_kernel(Buffers b, Penalty pen, int density_in_jtj) {
    const Counts& c = *b.counts;
    double acc = 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < c.off[4] && !c.error; t += stride) {
        uint32_t k;
        const int type = type_of(c, t, k);
        if (type == kDense) {
            const double r = b.dense[k].r;
            acc += r * r;
            if (density_in_jtj && b.density) {
                const double d = (double)pen.density * b.density[k].r;
                acc += d * d;
            }
        } else if (type == kForeground) {
            const double r = (double)pen.foreground * b.fg[k].r;
            acc += r * r;
        } else if (type == kSmooth) {
            float J[3][6], res[3];
            smooth_rows(b.smooth[k], true, J, res);
            for (int ch = 0; ch < 3; ++ch) acc += (double)pen.smooth * pen.smooth * res[ch] * res[ch];
        } else {
            float J[3][6], res[3];
            p2p_rows(type == kFeature ? b.feat[k] : b.cross[k], J, res);
            const double l = type == kFeature ? pen.feature : pen.cross;
            for (int ch = 0; ch < 3; ++ch) acc += l * l * res[ch] * res[ch];
        }
    }
    acc = wave_sum(acc);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    // one partial per block; the host sums them in block order
    if (threadIdx.x == 0) b.partial[(size_t)kSlotCost * kPartialStride + blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// ------------------------------------------------------------- host launch
void launch_eval_terms(const Buffers& b, const DQ* se3, bool fp64_geometry, hipStream_t stream) {
    if (fp64_geometry)
        hipLaunchKernelGGL(eval_terms_kernel<double>, dim3(1024), dim3(256), 0, stream, b, se3);
    else
        hipLaunchKernelGGL(eval_terms_kernel<float>, dim3(1024), dim3(256), 0, stream, b, se3);
}

void launch_assemble(const Buffers& b, const Penalty& pen, bool fp64_inverse, uint32_t cap_nodes, hipStream_t stream) {
    if (fp64_inverse)
        hipLaunchKernelGGL(assemble_kernel<double>, dim3(cap_nodes), dim3(kAsmThreads), 0, stream, b, pen);
    else
        hipLaunchKernelGGL(assemble_kernel<float>, dim3(cap_nodes), dim3(kAsmThreads), 0, stream, b, pen);
}

// Returns the number of partials written to slot kSlotCost.
uint32_t launch_cost(const Buffers& b, const Penalty& pen, bool density_in_jtj, hipStream_t stream) {
    const uint32_t blocks = 256;
    hipLaunchKernelGGL(cost_kernel, dim3(blocks), dim3(256), 0, stream, b, pen, density_in_jtj ? 1 : 0);
    return blocks;
}

}  // namespace mvdr
