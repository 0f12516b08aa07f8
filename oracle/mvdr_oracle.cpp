// CPU oracle for the non-rigid GN + block-PCG solve.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU and in plain C++,
// the reference algorithm of core/NonRigidSolver so that the HIP path can be
// checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may load it, and only as the checker / the timed CPU
// baseline: the product path never calls it.
//
// It is templated on the arithmetic type: Real = float follows the reference's
// FP32 arithmetic (sequential sums instead of warp_scan / cublasSdot order),
// Real = double is the FP64 oracle the twist-update error is measured against.
//
// Each routine cites the reference file:line it restates.  Where the
// reference's driver or constants are absent from the tree (the streamed
// solve body, Constants::*, the cross-view handler) the behaviour is a
// documented reconstruction; those functions say "parity unpinned".
//
// Pinning (what checks each part, tests/test_oracle_*.py):
//   dual-quaternion / SE3 math ...... closed forms, scipy rotations, round trips
//   term selection .................. brute-force Python restatement on tiny maps
//   term Jacobians .................. central finite differences (FP64); image terms
//                                     on a linear image with bilinear_images set
//   Node2Term / NodePair2Term ....... brute-force closed form (sorted term lists)
//   JtJ / JtR blocks ................ dense J^T J and J^T r from explicit J
//   6x6 inverse ..................... A * inv(A) = I
//   PCG ............................. numpy PCG recurrence and a direct solve
//   warp update ..................... Rodrigues closed form
//   cross-view term ................. parity unpinned (reference implementation absent)
#include "../include/mvdr/crossview.h"
#include "../include/mvdr/render.h"
#include "../include/mvdr/rigid.h"
#include "../include/mvdr/solver.h"
#include "../include/mvdr/warpfield.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct float4_t {
    float x, y, z, w;
};

template <typename R>
struct V3 {
    R x, y, z;
};
template <typename R> inline V3<R> v3(R a, R b, R c) { return {a, b, c}; }
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline V3<R> vcross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// mat33 as three columns (math/MatUtils.h:19-20) + translation (MatUtils.h:401-403)
template <typename R>
struct M34 {
    V3<R> c[3];
    V3<R> t;
    R m(int r, int col) const { const V3<R>& v = c[col]; return r == 0 ? v.x : (r == 1 ? v.y : v.z); }
    void set(int r, int col, R val) { V3<R>& v = c[col]; (r == 0 ? v.x : (r == 1 ? v.y : v.z)) = val; }
    // mat33::operator*(float3) (MatUtils.h:303-309)
    V3<R> rot(V3<R> p) const {
        return {m(0, 0) * p.x + m(0, 1) * p.y + m(0, 2) * p.z, m(1, 0) * p.x + m(1, 1) * p.y + m(1, 2) * p.z,
                m(2, 0) * p.x + m(2, 1) * p.y + m(2, 2) * p.z};
    }
    // mat33::transpose_dot (MatUtils.h:343-349)
    V3<R> rot_t(V3<R> p) const {
        return {m(0, 0) * p.x + m(1, 0) * p.y + m(2, 0) * p.z, m(0, 1) * p.x + m(1, 1) * p.y + m(2, 1) * p.z,
                m(0, 2) * p.x + m(1, 2) * p.y + m(2, 2) * p.z};
    }
    V3<R> apply(V3<R> p) const { return rot(p) + t; }
    // mat34::inverse (MatUtils.h:585-590)
    M34 inverse() const {
        M34 o;
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 3; ++col) o.set(r, col, m(col, r));
        V3<R> rt = o.rot(t);
        o.t = {-rt.x, -rt.y, -rt.z};
        return o;
    }
    // mat34::operator* (MatUtils.h:440-446): this * b
    M34 mul(const M34& b) const {
        M34 o;
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 3; ++col)
                o.set(r, col, m(r, 0) * b.m(0, col) + m(r, 1) * b.m(1, col) + m(r, 2) * b.m(2, col));
        o.t = rot(b.t) + t;
        return o;
    }
    static M34 identity() {
        M34 o;
        o.c[0] = v3<R>(1, 0, 0); o.c[1] = v3<R>(0, 1, 0); o.c[2] = v3<R>(0, 0, 1);
        o.t = v3<R>(0, 0, 0);
        return o;
    }
    static M34 from(const mvdr_mat34& a) {
        M34 o;
        for (int col = 0; col < 3; ++col) o.c[col] = v3<R>(a.rot[3 * col], a.rot[3 * col + 1], a.rot[3 * col + 2]);
        o.t = v3<R>(a.trans[0], a.trans[1], a.trans[2]);
        return o;
    }
};

// Quaternion stored (x, y, z, w) (Quaternion.h:15-16)
template <typename R>
struct Q {
    R x, y, z, w;
};
// operator*(Quaternion, Quaternion) (Quaternion.h:171-181)
template <typename R> inline Q<R> qmul(Q<R> a, Q<R> b) {
    Q<R> q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return q;
}
template <typename R> inline Q<R> conj(Q<R> a) { return {-a.x, -a.y, -a.z, a.w}; }
template <typename R> inline R qdot(Q<R> a, Q<R> b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename R> inline R inv_sqrt(R v) { return R(1) / std::sqrt(v); }

// Quaternion::rotation_matrix(false) (Quaternion.h:104-117)
template <typename R> M34<R> rotation_of(Q<R> q) {
    M34<R> r;
    r.set(0, 0, 1 - 2 * q.y * q.y - 2 * q.z * q.z);
    r.set(0, 1, 2 * q.x * q.y - 2 * q.z * q.w);
    r.set(0, 2, 2 * q.x * q.z + 2 * q.y * q.w);
    r.set(1, 0, 2 * q.x * q.y + 2 * q.z * q.w);
    r.set(1, 1, 1 - 2 * q.x * q.x - 2 * q.z * q.z);
    r.set(1, 2, 2 * q.y * q.z - 2 * q.x * q.w);
    r.set(2, 0, 2 * q.x * q.z - 2 * q.y * q.w);
    r.set(2, 1, 2 * q.y * q.z + 2 * q.x * q.w);
    r.set(2, 2, 1 - 2 * q.x * q.x - 2 * q.y * q.y);
    r.t = v3<R>(0, 0, 0);
    return r;
}

template <typename R>
struct DQ {
    Q<R> q0, q1;
    static DQ from(const mvdr_dq& d) {
        return {{(R)d.q0[0], (R)d.q0[1], (R)d.q0[2], (R)d.q0[3]}, {(R)d.q1[0], (R)d.q1[1], (R)d.q1[2], (R)d.q1[3]}};
    }
    void to(mvdr_dq& d) const {
        d.q0[0] = (float)q0.x; d.q0[1] = (float)q0.y; d.q0[2] = (float)q0.z; d.q0[3] = (float)q0.w;
        d.q1[0] = (float)q1.x; d.q1[1] = (float)q1.y; d.q1[2] = (float)q1.z; d.q1[3] = (float)q1.w;
    }
    // DualQuaternion::normalize (DualQuaternion.h:281-287)
    void normalize() {
        R inv = inv_sqrt(q0.w * q0.w + q0.x * q0.x + q0.y * q0.y + q0.z * q0.z);
        q0 = {inv * q0.x, inv * q0.y, inv * q0.z, inv * q0.w};
        q1 = {inv * q1.x, inv * q1.y, inv * q1.z, inv * q1.w};
        R d = qdot(q0, q1);
        q1 = {q1.x - d * q0.x, q1.y - d * q0.y, q1.z - d * q0.z, q1.w - d * q0.w};
    }
    // DualQuaternion::se3_matrix (DualQuaternion.h:326-335)
    M34<R> se3_matrix() const {
        DQ n = *this;
        n.normalize();
        M34<R> m = rotation_of(n.q0);
        Q<R> two_q1 = {2 * n.q1.x, 2 * n.q1.y, 2 * n.q1.z, 2 * n.q1.w};
        Q<R> tp = qmul(two_q1, conj(n.q0));
        m.t = v3<R>(tp.x, tp.y, tp.z);
        return m;
    }
    // DualQuaternion::operator mat34 (DualQuaternion.h:309-319): normalized(),
    // then Quaternion::matrix() which normalises q0 once more (Quaternion.h:88-103)
    M34<R> to_mat34() const {
        DQ n = *this;
        n.normalize();
        Q<R> qn = n.q0;
        R inv = inv_sqrt(qn.w * qn.w + qn.x * qn.x + qn.y * qn.y + qn.z * qn.z);
        qn = {inv * qn.x, inv * qn.y, inv * qn.z, inv * qn.w};
        M34<R> m = rotation_of(qn);
        Q<R> two_q1 = {2 * n.q1.x, 2 * n.q1.y, 2 * n.q1.z, 2 * n.q1.w};
        Q<R> tp = qmul(two_q1, conj(n.q0));
        m.t = v3<R>(tp.x, tp.y, tp.z);
        return m;
    }
    // DualQuaternion(const mat34&) (DualQuaternion.h:123-130) with Quaternion(mat33)
    // (Quaternion.h:46-77)
    static DQ from_mat34(const M34<R>& T) {
        Q<R> q;
        const R tr = T.m(0, 0) + T.m(1, 1) + T.m(2, 2);
        if (tr > 0) {
            R s = std::sqrt(tr + R(1)) * 2;
            q.w = s * R(0.25);
            q.x = (T.m(2, 1) - T.m(1, 2)) / s;
            q.y = (T.m(0, 2) - T.m(2, 0)) / s;
            q.z = (T.m(1, 0) - T.m(0, 1)) / s;
        } else if ((T.m(0, 0) > T.m(1, 1)) && (T.m(0, 0) > T.m(2, 2))) {
            R s = std::sqrt(R(1) + T.m(0, 0) - T.m(1, 1) - T.m(2, 2)) * 2;
            q.w = (T.m(2, 1) - T.m(1, 2)) / s;
            q.x = R(0.25) * s;
            q.y = (T.m(0, 1) + T.m(1, 0)) / s;
            q.z = (T.m(0, 2) + T.m(2, 0)) / s;
        } else if (T.m(1, 1) > T.m(2, 2)) {
            R s = std::sqrt(R(1) + T.m(1, 1) - T.m(0, 0) - T.m(2, 2)) * 2;
            q.w = (T.m(0, 2) - T.m(2, 0)) / s;
            q.x = (T.m(0, 1) + T.m(1, 0)) / s;
            q.y = R(0.25) * s;
            q.z = (T.m(1, 2) + T.m(2, 1)) / s;
        } else {
            R s = std::sqrt(R(1) + T.m(2, 2) - T.m(0, 0) - T.m(1, 1)) * 2;
            q.w = (T.m(1, 0) - T.m(0, 1)) / s;
            q.x = (T.m(0, 2) + T.m(2, 0)) / s;
            q.y = (T.m(1, 2) + T.m(2, 1)) / s;
            q.z = R(0.25) * s;
        }
        Q<R> tq = {R(0.5) * T.t.x, R(0.5) * T.t.y, R(0.5) * T.t.z, R(0)};
        DQ o;
        o.q0 = q;                 // (1, 0) * (q, 0): real part q
        o.q1 = qmul(tq, q);       // dual part 0.5 t * q
        return o;
    }
};

// averageDualQuaternion (DualQuaternion.h:381-398): plain weighted sum, no
// sign alignment and no weight renormalisation.
template <typename R>
DQ<R> blend(const std::vector<DQ<R>>& field, const uint16_t knn[4], const R w[4]) {
    DQ<R> a{{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int k = 0; k < 4; ++k) {
        const DQ<R>& d = field[knn[k]];
        a.q0 = {a.q0.x + w[k] * d.q0.x, a.q0.y + w[k] * d.q0.y, a.q0.z + w[k] * d.q0.z, a.q0.w + w[k] * d.q0.w};
        a.q1 = {a.q1.x + w[k] * d.q1.x, a.q1.y + w[k] * d.q1.y, a.q1.z + w[k] * d.q1.z, a.q1.w + w[k] * d.q1.w};
    }
    return a;
}

// Rodrigues twist -> SE3, as in applyWarpFieldUpdateKernel
// (SolverIterationData.cu:9-57) and mat34(twist_rot, twist_trans) (MatUtils.cpp:10-40)
template <typename R>
M34<R> twist_to_se3(V3<R> w, V3<R> v) {
    M34<R> T = M34<R>::identity();
    if (std::fabs(w.x) + std::fabs(w.y) + std::fabs(w.z) >= R(1e-4)) {
        const R angle = std::sqrt(dot(w, w));
        const V3<R> a = (R(1) / angle) * w;
        const R c = std::cos(angle), s = std::sin(angle), t = R(1) - c;
        T.set(0, 0, t * a.x * a.x + c);       T.set(0, 1, t * a.x * a.y - a.z * s); T.set(0, 2, t * a.x * a.z + a.y * s);
        T.set(1, 0, t * a.x * a.y + a.z * s); T.set(1, 1, t * a.y * a.y + c);       T.set(1, 2, t * a.y * a.z - a.x * s);
        T.set(2, 0, t * a.x * a.z - a.y * s); T.set(2, 1, t * a.y * a.z + a.x * s); T.set(2, 2, t * a.z * a.z + c);
    }
    T.t = v;
    return T;
}

// is_zero_vertex (math/VectorUtils.h:214-216)
template <typename R> inline bool is_zero_vertex(const float* v) {
    return std::fabs(v[0]) < 1e-3 && std::fabs(v[1]) < 1e-3 && std::fabs(v[2]) < 1e-3;
}

// compute_feature_huber_weight (huber_weight.h:11-27), cutoff 7.5e-2
template <typename R> inline R huber(R residual) {
    const R a = std::fabs(residual);
    return a < R(7.5e-2) ? R(1) : R(7.5e-2) / a;
}

// __float2int_rn of the reference's projection: the expression
// x / (z + 1e-10) * f + c is evaluated in double (1e-10 is a double literal),
// narrowed to float by __float2int_rn's parameter, then rounded to nearest-even.
inline int project_round(double x, double z, double f, double c) {
    const float u = (float)(x / (z + 1e-10) * f + c);
    return (int)std::nearbyint(u);
}

enum TermType { kDense = 0, kSmooth = 1, kForeground = 2, kFeature = 3, kCross = 4 };

// Penalty constants (PenaltyConstants.cpp:9-51)
struct Penalty {
    double smooth, density, foreground, feature, cross;
};

// setElasticPenaltyValue (SolverIterationData.cpp:104-125).  At iteration 0 the
// reference calls it with default flags (SolverIterationData.cpp:49), so the
// first iteration never carries the foreground or cross-view penalty.
Penalty penalty_for(const mvdr_solver_config& cfg, uint32_t iter) {
    Penalty p;
    if (!cfg.use_elastic_penalty) {
        p = {2.5, 0.0, 0.0, 0.7, 0.1};
        return p;
    }
    const bool use_density = iter > 0 && cfg.use_density;
    const bool use_foreground = iter > 0 && cfg.use_foreground;
    const bool use_cross = iter > 0 && cfg.use_cross_view;
    if (iter < cfg.global_iterations) {
        p = {5.0, 0.0, use_foreground ? 2e-3 : 0.0, 1.0, use_cross ? 1.0 : 0.0};
    } else {
        p = {5.0, use_density ? 1e-2 : 0.0, 0.0, 0.0, use_cross ? 1.0 : 0.0};
    }
    return p;
}

// Host copy of one frame (the caller passes mvdr_frame_input with HOST pointers).
struct Frame {
    uint32_t V = 0, W = 0, H = 0;
    mvdr_frame_input in;
    const float* f4(const void* base, uint32_t x, uint32_t y) const {
        return static_cast<const float*>(base) + 4 * ((size_t)y * W + x);
    }
    bool inside(int x, int y) const { return x >= 0 && y >= 0 && x < (int)W && y < (int)H; }
    // texture fetches with cudaAddressModeBorder (base/CommonUtils.cu:47-49): 0 outside
    uint32_t index(uint32_t v, int x, int y) const {
        return inside(x, y) ? static_cast<const uint32_t*>(in.index_map[v])[(size_t)y * W + x] : 0u;
    }
    void fetch4(const void* base, int x, int y, float out[4]) const {
        if (!inside(x, y)) { out[0] = out[1] = out[2] = out[3] = 0.f; return; }
        std::memcpy(out, f4(base, x, y), 16);
    }
    float fetch1(const void* base, int x, int y) const {
        return inside(x, y) ? static_cast<const float*>(base)[(size_t)y * W + x] : 0.f;
    }
    void fetch2(const void* base, int x, int y, float out[2]) const {
        if (!inside(x, y)) { out[0] = out[1] = 0.f; return; }
        const float* p = static_cast<const float*>(base) + 2 * ((size_t)y * W + x);
        out[0] = p[0]; out[1] = p[1];
    }
    uint8_t edge(uint32_t v, int x, int y) const {
        return inside(x, y) ? static_cast<const uint8_t*>(in.edge_mask[v])[(size_t)y * W + x] : 0;
    }
    void surfel_knn(uint32_t s, uint16_t knn[4], float w[4]) const {
        std::memcpy(knn, static_cast<const uint16_t*>(in.surfel_knn) + 4 * (size_t)s, 8);
        std::memcpy(w, static_cast<const float*>(in.surfel_knn_weight) + 4 * (size_t)s, 16);
    }
};

// Term records.  Static parts (knn, weights, pixel, canonical vertex) are
// fixed at index-build time; J / r / warped parts are refreshed per GN iteration.
template <typename R>
struct ScalarTerm {  // dense depth, foreground (ScalarCostTerm2Jacobian, solver_types.h:90-118)
    uint16_t px, py, view;
    uint16_t knn[4];
    R w[4];
    R J[6];
    R r;
};
template <typename R>
struct P2PTerm {  // feature / cross (Point2PointICPTerm2Jacobian, solver_types.h:141-150)
    uint16_t knn[4];
    R w[4];
    V3<R> canonical, target, warped;
};
template <typename R>
struct SmoothTerm {  // NodeGraphSmoothTerm2Jacobian (solver_types.h:128-138)
    uint16_t i, j;
    uint8_t valid;
    V3<R> Ti_xj, Tj_xj;
};

template <typename R>
class Oracle {
public:
    mvdr_solver_config cfg;
    Frame f;
    uint32_t N = 0, E = 0;
    std::vector<DQ<R>> se3;        // current warp field
    std::vector<V3<R>> nodes;      // canonical node positions
    std::vector<uint16_t> graph;   // 2 per edge
    uint32_t iter = 0;
    Penalty pen{};
    // Checker switch (never the reference's behaviour): image terms sample
    // their map bilinearly at the unrounded projection, so that on a linear
    // image the residual is a smooth function of the warp and its Jacobian
    // can be pinned by finite differences.  Off by default.
    bool bilinear_images = false;

    std::vector<ScalarTerm<R>> dense, density, fg;
    std::vector<SmoothTerm<R>> smooth;
    std::vector<P2PTerm<R>> feat, cross;
    std::vector<V3<R>> dense_canon_v, dense_canon_n;  // unused by the reference layout; kept for FD checks

    // indices
    uint32_t off[5] = {0, 0, 0, 0, 0};
    std::vector<uint32_t> n2t_offset, n2t_index;
    std::vector<uint32_t> pair_key, pair_start, pair_end, pair_terms;

    // linear system
    std::vector<R> diag, diag_inv, jtr, offdiag, twist;

    bool indexed = false;

    Oracle(const mvdr_solver_config& c, const mvdr_frame_input& in) : cfg(c) {
        f.in = in;
        f.V = in.num_views; f.W = in.width; f.H = in.height;
        N = in.num_nodes; E = in.num_graph_edges;
        se3.resize(N);
        nodes.resize(N);
        const mvdr_dq* d = static_cast<const mvdr_dq*>(in.node_se3);
        const float* n4 = static_cast<const float*>(in.canonical_nodes);
        for (uint32_t i = 0; i < N; ++i) {
            se3[i] = DQ<R>::from(d[i]);
            nodes[i] = v3<R>(n4[4 * i], n4[4 * i + 1], n4[4 * i + 2]);
        }
        graph.assign(static_cast<const uint16_t*>(in.node_graph), static_cast<const uint16_t*>(in.node_graph) + 2 * (size_t)E);
        pen = penalty_for(cfg, 0);
    }

    M34<R> init(uint32_t v) const { return M34<R>::from(f.in.initial_camera_se3[v]); }
    M34<R> w2c(uint32_t v) const { return M34<R>::from(f.in.world2camera[v]); }

    // ------------------------------------------------------------ index build
    // ImageTermKNNFetcher mark/compact (ImageTermKNNFetcher.cu:19-68) with the
    // per-pixel KNN of queryPixelKNNKernel (NonRigidSolver.cu:14-34): every
    // pixel whose index map is valid becomes a term, in (view, y, x) order.
    void select_dense() {
        dense.clear();
        for (uint32_t v = 0; v < f.V; ++v)
            for (uint32_t y = 0; y < f.H; ++y)
                for (uint32_t x = 0; x < f.W; ++x) {
                    const uint32_t s = f.index(v, x, y);
                    if (s == 0xFFFFFFFFu) continue;
                    ScalarTerm<R> t{};
                    t.px = (uint16_t)x; t.py = (uint16_t)y; t.view = (uint16_t)v;
                    float w[4];
                    f.surfel_knn(s, t.knn, w);
                    for (int k = 0; k < 4; ++k) t.w[k] = w[k];
                    dense.push_back(t);
                }
    }

    // validDepthPixelInWindow (SparseCorrespondenceHandler.cu:67-93).  The
    // `break` leaves only the inner x loop, so the last row holding a valid
    // pixel wins; the centre is preferred when valid.
    void depth_pixel_in_window(uint32_t v, int cx, int cy, int& ox, int& oy) const {
        ox = 0xFFFF; oy = 0xFFFF;
        float vv[4];
        for (int y = cy - 1; y <= cy + 1; ++y)
            for (int x = cx - 1; x <= cx + 1; ++x) {
                f.fetch4(f.in.vertex_confid[v], x, y, vv);
                if (!is_zero_vertex<R>(vv)) { ox = x & 0xFFFF; oy = y & 0xFFFF; break; }
            }
        f.fetch4(f.in.vertex_confid[v], cx, cy, vv);
        if (!is_zero_vertex<R>(vv)) { ox = cx; oy = cy; }
    }

    // validGeometryPixelInWindow, nearest-live-vertex form
    // (SparseCorrespondenceHandler.cu:46-65): window of half size `half`
    // sampled every `step` pixels; out-of-image taps read the border value 0,
    // which counts as a valid index.
    void geometry_pixel_nearest(uint32_t v, const float obs[3], int cx, int cy, int half, int step, int& ox, int& oy) const {
        ox = 0xFFFF; oy = 0xFFFF;
        float best = 100.0f;
        float lv[4];
        for (int y = cy - half; y <= cy + half; y += step)
            for (int x = cx - half; x <= cx + half; x += step) {
                if (f.index(v, x, y) == 0xFFFFFFFFu) continue;
                f.fetch4(f.in.warp_vertex[v], x, y, lv);
                const float d = (obs[0] - lv[0]) * (obs[0] - lv[0]) + (obs[1] - lv[1]) * (obs[1] - lv[1]) +
                                (obs[2] - lv[2]) * (obs[2] - lv[2]);
                if (d < best) { ox = x & 0xFFFF; oy = y & 0xFFFF; best = d; }
            }
    }

    // chooseValidPixelKernel (USE_NEAREST_DISTANCE_PAIRS) + compactQueryValidPairsKernel
    // (SparseCorrespondenceHandler.cu:100-220): candidate order is view-major.
    void select_features() {
        feat.clear();
        for (uint32_t v = 0; v < f.V; ++v) {
            const uint16_t* pairs = static_cast<const uint16_t*>(f.in.feature_pairs[v]);
            const M34<R> I = init(v), c2w = w2c(v).inverse();
            for (uint32_t k = 0; k < f.in.num_feature_pairs[v]; ++k) {
                const uint16_t* c = pairs + 4 * (size_t)k;
                int dx, dy;
                depth_pixel_in_window(v, c[2], c[3], dx, dy);
                if (f.edge(v, dx, dy) == 1) continue;
                if (!(dx < (int)f.W && dy < (int)f.H)) continue;
                float obs[4];
                f.fetch4(f.in.vertex_confid[v], dx, dy, obs);
                int gx, gy;
                geometry_pixel_nearest(v, obs, c[0], c[1], 1, 1, gx, gy);
                if (!(gx < (int)f.W && gy < (int)f.H)) continue;
                float ref[4];
                f.fetch4(f.in.reference_vertex[v], gx, gy, ref);
                const float d2 = (ref[0] - obs[0]) * (ref[0] - obs[0]) + (ref[1] - obs[1]) * (ref[1] - obs[1]) +
                                 (ref[2] - obs[2]) * (ref[2] - obs[2]);
                if (!(d2 < 2.5e-3f)) continue;
                P2PTerm<R> t{};
                t.canonical = I.apply(v3<R>(ref[0], ref[1], ref[2]));
                t.target = c2w.apply(I.apply(v3<R>(obs[0], obs[1], obs[2])));
                float w[4];
                f.surfel_knn(f.index(v, gx, gy), t.knn, w);
                for (int q = 0; q < 4; ++q) t.w[q] = w[q];
                feat.push_back(t);
            }
        }
    }

    // Cross-view pairs.  PARITY UNPINNED: the reference handler exists only as
    // a header (CrossViewCorrespondenceHandler.h:22-158, window half size 2,
    // step 2).  Rebuilt by analogy with the feature term: the observed point
    // is the depth vertex of view 2 at pixel 2, the geometry pixel is searched
    // around pixel 1 in view 1 for the warp vertex nearest to that point
    // expressed in camera 1, and the pair is kept when the reference vertex is
    // within 5 cm of it (the feature term's 2.5e-3 m^2 test).  With a corrMap
    // for view 1 (CrossViewCorrespondenceHandler.h:30-34), pixel 1 is first
    // back-traced to the previous-frame pixel of the feature pair that ends
    // at it (PatchColliderRGBCorrespondence.cu:171 stores (x, y, z, w) at
    // (z, w)).
    void select_cross() {
        cross.clear();
        if (!f.in.num_cross_pairs) return;
        const mvdr_cross_pair* cp = static_cast<const mvdr_cross_pair*>(f.in.cross_pairs);
        for (uint32_t k = 0; k < f.in.num_cross_pairs; ++k) {
            const mvdr_cross_pair& c = cp[k];
            const uint32_t v1 = c.view[0], v2 = c.view[1];
            if (v1 >= f.V || v2 >= f.V) continue;
            int cx = c.pixel[0], cy = c.pixel[1];
            if (f.in.corr_map[v1] && cx < (int)f.W && cy < (int)f.H) {
                const uint16_t* m = static_cast<const uint16_t*>(f.in.corr_map[v1]) + 4 * ((size_t)cy * f.W + cx);
                if (m[2] == cx && m[3] == cy) { cx = m[0]; cy = m[1]; }
            }
            int dx, dy;
            depth_pixel_in_window(v2, c.pixel[2], c.pixel[3], dx, dy);
            if (f.edge(v2, dx, dy) == 1) continue;
            if (!(dx < (int)f.W && dy < (int)f.H)) continue;
            float obs2[4];
            f.fetch4(f.in.vertex_confid[v2], dx, dy, obs2);
            const M34<float> I1 = M34<float>::from(f.in.initial_camera_se3[v1]);
            const M34<float> I2 = M34<float>::from(f.in.initial_camera_se3[v2]);
            const V3<float> p0 = I2.apply(v3<float>(obs2[0], obs2[1], obs2[2]));
            const V3<float> p1 = I1.inverse().apply(p0);
            const float obs1[3] = {p1.x, p1.y, p1.z};
            int gx, gy;
            geometry_pixel_nearest(v1, obs1, cx, cy, 2, 2, gx, gy);
            if (!(gx < (int)f.W && gy < (int)f.H)) continue;
            float ref[4];
            f.fetch4(f.in.reference_vertex[v1], gx, gy, ref);
            const float d2 = (ref[0] - obs1[0]) * (ref[0] - obs1[0]) + (ref[1] - obs1[1]) * (ref[1] - obs1[1]) +
                             (ref[2] - obs1[2]) * (ref[2] - obs1[2]);
            if (!(d2 < 2.5e-3f)) continue;
            P2PTerm<R> t{};
            t.canonical = init(v1).apply(v3<R>(ref[0], ref[1], ref[2]));
            t.target = w2c(v2).inverse().apply(init(v2).apply(v3<R>(obs2[0], obs2[1], obs2[2])));
            float w[4];
            f.surfel_knn(f.index(v1, gx, gy), t.knn, w);
            for (int q = 0; q < 4; ++q) t.w[q] = w[q];
            cross.push_back(t);
        }
    }

    // markValidColorForegroundMapsPixelKernel + compactValidPixelKernel
    // (DensityForegroundMapHandler.cu:34-144): foreground terms are the
    // rendered pixels that, warped by the current field, project inside the
    // image with view-angle cosine > 0.5 onto a filtered-mask value > 1e-5.
    // The selection is evaluated in FP32 whatever R is, as on the device, so
    // that the FP32 and FP64 oracles select the same term set.
    void select_foreground() {
        fg.clear();
        if (!cfg.use_foreground) return;
        std::vector<DQ<float>> field(N);
        for (uint32_t i = 0; i < N; ++i) {
            mvdr_dq d;
            se3[i].to(d);
            field[i] = DQ<float>::from(d);
        }
        for (uint32_t v = 0; v < f.V; ++v) {
            const M34<float> I = M34<float>::from(f.in.initial_camera_se3[v]);
            const M34<float> Iinv = I.inverse(), W2C = M34<float>::from(f.in.world2camera[v]);
            const mvdr_intrinsic K = f.in.intrinsic[v];
            for (uint32_t y = 0; y < f.H; ++y)
                for (uint32_t x = 0; x < f.W; ++x) {
                    const uint32_t s = f.index(v, x, y);
                    if (s == 0xFFFFFFFFu) continue;
                    float cv[4], cn[4];
                    f.fetch4(f.in.reference_vertex[v], x, y, cv);
                    f.fetch4(f.in.reference_normal[v], x, y, cn);
                    const V3<float> v0 = I.apply(v3<float>(cv[0], cv[1], cv[2]));
                    const V3<float> n0 = I.rot(v3<float>(cn[0], cn[1], cn[2]));
                    ScalarTerm<R> t{};
                    t.px = (uint16_t)x; t.py = (uint16_t)y; t.view = (uint16_t)v;
                    float w[4];
                    f.surfel_knn(s, t.knn, w);
                    for (int k = 0; k < 4; ++k) t.w[k] = w[k];
                    const M34<float> T = blend(field, t.knn, w).se3_matrix();
                    V3<float> pc = W2C.apply(T.apply(v0)), nc = W2C.rot(T.rot(n0));
                    pc = Iinv.apply(pc);
                    nc = Iinv.rot(nc);
                    const float inv = 1.0f / std::sqrt(dot(pc, pc));
                    const float cosine = -dot(inv * pc, nc);
                    const int u = project_round(pc.x, pc.z, K.fx, K.cx), q = project_round(pc.y, pc.z, K.fy, K.cy);
                    if (!(u >= 0 && u < (int)f.W && q >= 0 && q < (int)f.H && cosine > 0.5f)) continue;
                    if (!(f.fetch1(f.in.filtered_foreground[v], u, q) > 1e-5f)) continue;
                    fg.push_back(t);
                }
        }
    }

    void build_smooth_static() {
        smooth.assign(E, SmoothTerm<R>{});
        for (uint32_t e = 0; e < E; ++e) { smooth[e].i = graph[2 * e]; smooth[e].j = graph[2 * e + 1]; }
    }

    void set_offsets() {
        off[0] = (uint32_t)dense.size();
        off[1] = off[0] + E;
        off[2] = off[1] + (uint32_t)fg.size();
        off[3] = off[2] + (uint32_t)feat.size();
        off[4] = off[3] + (uint32_t)cross.size();
    }

    const uint16_t* term_knn(uint32_t term) const {
        if (term < off[0]) return dense[term].knn;
        if (term < off[2] && term >= off[1]) return fg[term - off[1]].knn;
        if (term < off[3] && term >= off[2]) return feat[term - off[2]].knn;
        if (term < off[4] && term >= off[3]) return cross[term - off[3]].knn;
        return nullptr;
    }

    // Node2TermsIndex::BuildIndex (Node2TermsIndex.cu:8-187): key/value pairs
    // (4 per knn term, 2 per edge), stable sort by node, CSR offsets.  The
    // reference leaves offsets of term-less nodes unwritten
    // (Node2TermsIndex.cu:119-133); here offset[n] is the first position whose
    // key is >= n, the proper CSR.
    void build_node2term() {
        std::vector<std::pair<uint16_t, uint32_t>> kv;
        kv.reserve((size_t)4 * off[4]);
        for (uint32_t t = 0; t < off[4]; ++t) {
            if (t >= off[0] && t < off[1]) {
                const uint32_t e = t - off[0];
                kv.push_back({graph[2 * e], t});
                kv.push_back({graph[2 * e + 1], t});
            } else {
                const uint16_t* k = term_knn(t);
                for (int q = 0; q < 4; ++q) kv.push_back({k[q], t});
            }
        }
        std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        n2t_index.resize(kv.size());
        for (size_t i = 0; i < kv.size(); ++i) n2t_index[i] = kv[i].second;
        n2t_offset.assign(N + 1, 0);
        size_t p = 0;
        for (uint32_t n = 0; n <= N; ++n) {
            while (p < kv.size() && kv[p].first < n) ++p;
            n2t_offset[n] = (uint32_t)p;
        }
        n2t_offset[N] = (uint32_t)kv.size();
    }

    // NodePair2TermsIndex (NodePair2TermsIndex.cu:11-163): 6 ordered-position
    // pairs (knn[p] < knn[q]) per knn term, 1 per edge, 0xFFFFFFFF filling the
    // unused slots; stable sort by key; unique keys with their [start, end)
    // ranges.  Keys are (i << 16) | j (encode_nodepair, EncodeUtils.h:187-196,
    // widened from i * 4096 + j so that N may reach 65535).
    void build_nodepair() {
        std::vector<std::pair<uint32_t, uint32_t>> kv;
        kv.reserve((size_t)6 * off[4]);
        for (uint32_t t = 0; t < off[4]; ++t) {
            if (t >= off[0] && t < off[1]) {
                const uint32_t e = t - off[0];
                const uint16_t a = graph[2 * e], b = graph[2 * e + 1];
                uint32_t key = 0xFFFFFFFFu;
                if (a < b) key = ((uint32_t)a << 16) | b;
                else if (b < a) key = ((uint32_t)b << 16) | a;
                kv.push_back({key, t});
            } else {
                const uint16_t* k = term_knn(t);
                uint32_t keys[6];
                for (int q = 0; q < 6; ++q) keys[q] = 0xFFFFFFFFu;
                int o = 0;
                for (int p = 0; p < 4; ++p)
                    for (int q = 0; q < 4; ++q)
                        if (k[p] < k[q]) keys[o++] = ((uint32_t)k[p] << 16) | k[q];
                for (int q = 0; q < 6; ++q) kv.push_back({keys[q], t});
            }
        }
        std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        pair_key.clear(); pair_start.clear(); pair_end.clear(); pair_terms.clear();
        for (size_t i = 0; i < kv.size(); ++i) {
            if (kv[i].first == 0xFFFFFFFFu) break;
            if (i == 0 || kv[i].first != kv[i - 1].first) {
                if (!pair_key.empty()) pair_end.push_back((uint32_t)i);
                pair_key.push_back(kv[i].first);
                pair_start.push_back((uint32_t)i);
            }
            pair_terms.push_back(kv[i].second);
        }
        if (!pair_key.empty()) pair_end.push_back((uint32_t)pair_terms.size());
    }

    void build_index() {
        select_dense();
        select_foreground();
        select_features();
        if (cfg.use_cross_view) select_cross(); else cross.clear();
        if (cfg.density_in_jtj) density = dense;
        build_smooth_static();
        set_offsets();
        build_node2term();
        build_nodepair();
        indexed = true;
    }

    // ------------------------------------------------------- term evaluation
    // computeDenseDepthJacobianKernel (DenseDepthJacobianResidual.cu:24-104);
    // the camera of a term is its compacted pixel's view (ushort3.z).
    void eval_dense(ScalarTerm<R>& t) const {
        const uint32_t v = t.view;
        float cv[4], cn[4];
        f.fetch4(f.in.reference_vertex[v], t.px, t.py, cv);
        f.fetch4(f.in.reference_normal[v], t.px, t.py, cn);
        const M34<R> I = init(v), W2C = w2c(v), Iinv = init(v).inverse();
        const V3<R> v0 = I.apply(v3<R>(cv[0], cv[1], cv[2]));
        const V3<R> n0 = I.rot(v3<R>(cn[0], cn[1], cn[2]));
        const M34<R> T = blend(se3, t.knn, t.w).se3_matrix();
        const V3<R> pl = T.apply(v0), nl = T.rot(n0);
        V3<R> pc = W2C.apply(pl), nc = W2C.rot(nl);
        pc = Iinv.apply(pc);
        nc = Iinv.rot(nc);
        for (R& j : t.J) j = 0;
        t.r = 0;
        const mvdr_intrinsic K = f.in.intrinsic[v];
        const int u = project_round(pc.x, pc.z, K.fx, K.cx), q = project_round(pc.y, pc.z, K.fy, K.cy);
        if (!(u >= 0 && u < (int)f.W && q >= 0 && q < (int)f.H)) return;
        float dv[4], dn[4];
        f.fetch4(f.in.vertex_confid[v], u, q, dv);
        f.fetch4(f.in.normal_radius[v], u, q, dn);
        const V3<R> d = v3<R>(dv[0], dv[1], dv[2]), n = v3<R>(dn[0], dn[1], dn[2]);
        bool valid = true;
        if (is_zero_vertex<R>(dv)) valid = false;
        if (dot(n, nc) < R(0.7f)) valid = false;
        const V3<R> diff = d - pc;
        if (dot(diff, diff) > R(0.075f) * R(0.075f)) valid = false;
        if (!valid) return;
        t.r = dot(n, pc - d);
        const V3<R> jt = W2C.rot_t(I.rot(n));
        const V3<R> jr = vcross(pl, jt);
        t.J[0] = jr.x; t.J[1] = jr.y; t.J[2] = jr.z; t.J[3] = jt.x; t.J[4] = jt.y; t.J[5] = jt.z;
    }

    // computeImageDensityJacobainAndResidual (density_map_jacobian.cuh:12-79),
    // called by computeDensityMapJacobian / computeForegroundMaskJacobian
    // (DensityForegroundJacobianResidual.cu:41-142).  The reference warps the
    // camera-v reference vertex without first moving it to camera 0; kept.
    void eval_image(ScalarTerm<R>& t, const void* const* map, const void* const* grad, bool density_term) const {
        const uint32_t v = t.view;
        float cv[4];
        f.fetch4(f.in.reference_vertex[v], t.px, t.py, cv);
        R geometry = 0;
        if (density_term) {
            float rgb[4];
            f.fetch4(f.in.normalized_rgb[v], t.px, t.py, rgb);
            geometry = R(0.299f) * rgb[0] + R(0.587f) * rgb[1] + R(0.114f) * rgb[2];
        }
        const M34<R> I = init(v), W2C = w2c(v), Iinv = init(v).inverse();
        const M34<R> T = blend(se3, t.knn, t.w).se3_matrix();
        const V3<R> pl = T.apply(v3<R>(cv[0], cv[1], cv[2]));
        V3<R> pc = Iinv.apply(W2C.apply(pl));
        for (R& j : t.J) j = 0;
        t.r = 0;
        const mvdr_intrinsic K = f.in.intrinsic[v];
        const int u = project_round(pc.x, pc.z, K.fx, K.cx), q = project_round(pc.y, pc.z, K.fy, K.cy);
        if (!(u >= 0 && u < (int)f.W && q >= 0 && q < (int)f.H)) return;
        R I0 = f.fetch1(map[v], u, q);
        if (bilinear_images) {
            const double U = pc.x / (pc.z + 1e-10) * K.fx + K.cx, Q = pc.y / (pc.z + 1e-10) * K.fy + K.cy;
            const int x0 = (int)std::floor(U), y0 = (int)std::floor(Q);
            if (!(f.inside(x0, y0) && f.inside(x0 + 1, y0 + 1))) return;
            const double ax = U - x0, ay = Q - y0;
            I0 = R((1 - ay) * ((1 - ax) * f.fetch1(map[v], x0, y0) + ax * f.fetch1(map[v], x0 + 1, y0)) +
                   ay * ((1 - ax) * f.fetch1(map[v], x0, y0 + 1) + ax * f.fetch1(map[v], x0 + 1, y0 + 1)));
        }
        float g[2];
        f.fetch2(grad[v], u, q, g);
        V3<R> d;
        d.x = g[0] * R(K.fx) / pc.z;
        d.y = g[1] * R(K.fy) / pc.z;
        d.z = -R(1) / (pc.z * pc.z) * (R(K.fx) * pc.x * g[0] + R(K.fy) * pc.y * g[1]);
        d = I.rot(d);
        d = W2C.rot_t(d);
        t.J[0] = -d.y * pl.z + d.z * pl.y;
        t.J[1] = d.x * pl.z - d.z * pl.x;
        t.J[2] = -d.x * pl.y + d.y * pl.x;
        t.J[3] = d.x; t.J[4] = d.y; t.J[5] = d.z;
        t.r = I0 - geometry;
    }

    // forwardWarpSmootherNodeKernel (NodeGraphSmoothHandler.cu:10-43)
    void eval_smooth(SmoothTerm<R>& s) const {
        const V3<R> xi = nodes[s.i], xj = nodes[s.j];
        const M34<R> Ti = se3[s.i].se3_matrix(), Tj = se3[s.j].se3_matrix();
        s.Ti_xj = Ti.apply(xj);
        s.Tj_xj = Tj.apply(xj);
        const V3<R> d = xi - xj;
        s.valid = dot(d, d) > R(64 * 0.025f * 0.025f) ? 0 : 1;
    }

    // forwardWarpFeatureVertexKernel (SparseCorrespondenceHandler.cu:193-220)
    void eval_p2p(P2PTerm<R>& t) const {
        const M34<R> T = blend(se3, t.knn, t.w).se3_matrix();
        t.warped = T.apply(t.canonical);
    }

    void eval_terms() {
        for (auto& t : dense) eval_dense(t);
        if (!density.empty()) for (auto& t : density) eval_image(t, f.in.gray, f.in.gray_gradient, true);
        for (auto& t : fg) eval_image(t, f.in.filtered_foreground, f.in.foreground_gradient, false);
        for (auto& s : smooth) eval_smooth(s);
        for (auto& t : feat) eval_p2p(t);
        for (auto& t : cross) eval_p2p(t);
    }

    // ------------------------------------------------------ per-term Jacobians
    // Knn slot of a node in a term (PreconditionerBlocked.cu:25-28): the sum
    // 0*(n==x) + 1*(n==y) + 2*(n==z) + 3*(n==w).
    static int slot(const uint16_t knn[4], uint32_t n) {
        return 1 * (n == knn[1]) + 2 * (n == knn[2]) + 3 * (n == knn[3]);
    }

    // computePointToPointICPTermJacobian (geometry_icp_jacobian.cuh:91-117),
    // Huber-weighted (USE_FEATURE_HUBER_WEIGHT)
    static void p2p_jacobian(const P2PTerm<R>& t, R J[3][6], R res[3]) {
        const V3<R> w = t.warped;
        const R r[3] = {w.x - t.target.x, w.y - t.target.y, w.z - t.target.z};
        const R h[3] = {huber(r[0]), huber(r[1]), huber(r[2])};
        const R rot[3][3] = {{0, w.z, -w.y}, {-w.z, 0, w.x}, {w.y, -w.x, 0}};
        for (int c = 0; c < 3; ++c) {
            for (int k = 0; k < 3; ++k) { J[c][k] = h[c] * rot[c][k]; J[c][3 + k] = (k == c) ? h[c] : R(0); }
            res[c] = h[c] * r[c];
        }
    }

    // computeSmoothTermJacobian (geometry_icp_jacobian.cuh:323-343)
    static void smooth_jacobian(const SmoothTerm<R>& s, bool node_i, R J[3][6]) {
        const V3<R> a = node_i ? s.Ti_xj : s.Tj_xj;
        const R sg = node_i ? R(1) : R(-1);
        const R rot[3][3] = {{0, a.z, -a.y}, {-a.z, 0, a.x}, {a.y, -a.x, 0}};
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) { J[c][k] = sg * rot[c][k]; J[c][3 + k] = (k == c) ? sg : R(0); }
    }

    R lambda2(TermType type) const {
        switch (type) {
            case kDense: return 1;
            case kSmooth: return R(pen.smooth) * R(pen.smooth);
            case kForeground: return R(pen.foreground) * R(pen.foreground);
            case kFeature: return R(pen.feature) * R(pen.feature);
            default: return R(pen.cross) * R(pen.cross);
        }
    }
    TermType type_of(uint32_t term, uint32_t& local) const {
        if (term < off[0]) { local = term; return kDense; }
        if (term < off[1]) { local = term - off[0]; return kSmooth; }
        if (term < off[2]) { local = term - off[1]; return kForeground; }
        if (term < off[3]) { local = term - off[2]; return kFeature; }
        local = term - off[3]; return kCross;
    }

    // The (up to 3) Jacobian rows of `term` with respect to node n, already
    // multiplied by the knn weight, plus the matching residuals.  Returns the
    // number of channels; channel rows are added with weight lambda^2.
    int rows_for(uint32_t term, uint32_t n, R J[4][6], R res[4], R lam2[4]) const {
        uint32_t local;
        const TermType type = type_of(term, local);
        switch (type) {
            case kDense: {
                const ScalarTerm<R>& t = dense[local];
                const R w = t.w[slot(t.knn, n)];
                for (int k = 0; k < 6; ++k) J[0][k] = w * t.J[k];
                res[0] = t.r;
                lam2[0] = 1;
                int c = 1;
                if (cfg.density_in_jtj && !density.empty()) {
                    const ScalarTerm<R>& d = density[local];
                    for (int k = 0; k < 6; ++k) J[1][k] = w * d.J[k];
                    res[1] = d.r;
                    lam2[1] = R(pen.density) * R(pen.density);
                    c = 2;
                }
                return c;
            }
            case kForeground: {
                const ScalarTerm<R>& t = fg[local];
                const R w = t.w[slot(t.knn, n)];
                for (int k = 0; k < 6; ++k) J[0][k] = w * t.J[k];
                res[0] = t.r;
                lam2[0] = lambda2(kForeground);
                return 1;
            }
            case kSmooth: {
                const SmoothTerm<R>& s = smooth[local];
                R Jc[3][6];
                smooth_jacobian(s, n == s.i, Jc);
                for (int c = 0; c < 3; ++c) {
                    for (int k = 0; k < 6; ++k) J[c][k] = s.valid ? Jc[c][k] : R(0);
                    res[c] = s.valid ? ((c == 0 ? s.Ti_xj.x - s.Tj_xj.x : (c == 1 ? s.Ti_xj.y - s.Tj_xj.y : s.Ti_xj.z - s.Tj_xj.z))) : R(0);
                    lam2[c] = lambda2(kSmooth);
                }
                return 3;
            }
            default: {
                const P2PTerm<R>& t = type == kFeature ? feat[local] : cross[local];
                R Jc[3][6], r[3];
                p2p_jacobian(t, Jc, r);
                const R w = t.w[slot(t.knn, n)];
                for (int c = 0; c < 3; ++c) {
                    for (int k = 0; k < 6; ++k) J[c][k] = w * Jc[c][k];
                    res[c] = r[c];
                    lam2[c] = lambda2(type);
                }
                return 3;
            }
        }
    }

    // ------------------------------------------------------------- assembly
    // computeBlockDiagonalPreconditionerKernel (PreconditionerBlocked.cu:172-284)
    // and computeJtResidualWithIndexKernel (JtDotResidualIndexed.cu:116-224):
    // D_n = sum lambda^2 J^T J (+1e-2 I), b_n = -sum lambda^2 J^T r over the
    // node's term list, in list order.
    void assemble_node_blocks() {
        diag.assign((size_t)36 * N, 0);
        jtr.assign((size_t)6 * N, 0);
        for (uint32_t n = 0; n < N; ++n) {
            R* D = &diag[36 * (size_t)n];
            R* b = &jtr[6 * (size_t)n];
            for (uint32_t p = n2t_offset[n]; p < n2t_offset[n + 1]; ++p) {
                R J[4][6], res[4], lam2[4];
                const int ch = rows_for(n2t_index[p], n, J, res, lam2);
                for (int c = 0; c < ch; ++c) {
                    for (int a = 0; a < 6; ++a) {
                        for (int k = 0; k < 6; ++k) D[6 * a + k] += lam2[c] * J[c][a] * J[c][k];
                        b[a] -= lam2[c] * J[c][a] * res[c];
                    }
                }
            }
            for (int a = 0; a < 6; ++a) D[7 * a] += R(1e-2f);
        }
    }

    // computeJtJNonDiagonalBlockKernel (JtJMaterializer.cu:270-365) with
    // computeScalarJtJBlockJacobian / computeFeatureJtJBlockJacobian
    // (jtj_block_jacobian.cuh:8-65) and computeSmoothJtJBlock
    // (JtJMaterializer.cu:54-146): B_ij[a][b] = sum lambda^2 J_i[a] J_j[b].
    // Stored once per unique pair, row-major, row node i < column node j.
    void assemble_pair_blocks() {
        const size_t P = pair_key.size();
        offdiag.assign(36 * P, 0);
        for (size_t p = 0; p < P; ++p) {
            const uint32_t i = pair_key[p] >> 16, j = pair_key[p] & 0xFFFF;
            R* B = &offdiag[36 * p];
            for (uint32_t q = pair_start[p]; q < pair_end[p]; ++q) {
                const uint32_t term = pair_terms[q];
                R Ji[4][6], Jj[4][6], ri[4], rj[4], li[4], lj[4];
                const int ch = rows_for(term, i, Ji, ri, li);
                rows_for(term, j, Jj, rj, lj);
                for (int c = 0; c < ch; ++c)
                    for (int a = 0; a < 6; ++a)
                        for (int k = 0; k < 6; ++k) B[6 * a + k] += li[c] * Ji[c][a] * Jj[c][k];
            }
        }
    }

    // DenseGaussian<6>::Inverse (DenseGaussian.h:55-125): Gauss-Jordan with
    // partial pivoting on a column-major matrix, then un-permuting the columns.
    static void gauss_jordan6(const R* in_rowmajor, R* out_rowmajor) {
        const int Nn = 6;
        R m[36];
        auto fl = [](int r, int c) { return r + c * 6; };
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) m[fl(r, c)] = in_rowmajor[6 * r + c];
        int perm[6], reloc[6];
        for (int i = 0; i < Nn; ++i) perm[i] = i;
        for (int i = 0; i < Nn; ++i) {
            int piv = i;
            R pv = std::fabs(m[fl(i, i)]);
            for (int j = i + 1; j < Nn; ++j)
                if (pv < std::fabs(m[fl(j, i)])) { pv = std::fabs(m[fl(j, i)]); piv = j; }
            for (int k = 0; k < Nn; ++k) std::swap(m[fl(i, k)], m[fl(piv, k)]);
            std::swap(perm[i], perm[piv]);
            reloc[i] = perm[i];
            const R inv = R(1) / m[fl(i, i)];
            for (int k = 0; k < Nn; ++k) m[fl(i, k)] *= inv;
            m[fl(i, i)] = inv;
            for (int r = 0; r < Nn; ++r) {
                if (r == i) continue;
                const R a = m[fl(r, i)];
                for (int k = 0; k < Nn; ++k) m[fl(r, k)] -= a * m[fl(i, k)];
                m[fl(r, i)] = -a * m[fl(i, i)];
            }
        }
        // BuildInversed: inversed[6 * reloc[j] + i] = raw[fl(i, j)], read row-major by the PCG
        for (int j = 0; j < Nn; ++j)
            for (int i = 0; i < Nn; ++i) out_rowmajor[6 * reloc[j] + i] = m[fl(i, j)];
    }

    void invert_diag() {
        diag_inv.assign(diag.size(), 0);
        for (uint32_t n = 0; n < N; ++n) gauss_jordan6(&diag[36 * (size_t)n], &diag_inv[36 * (size_t)n]);
    }

    // y = A x with A = blockdiag(D) + sum_pairs (B_ij at (i, j) and B_ij^T at (j, i))
    void spmv(const std::vector<R>& x, std::vector<R>& y) const {
        y.assign((size_t)6 * N, 0);
        for (uint32_t n = 0; n < N; ++n) {
            const R* D = &diag[36 * (size_t)n];
            for (int a = 0; a < 6; ++a) {
                R s = 0;
                for (int k = 0; k < 6; ++k) s += D[6 * a + k] * x[6 * (size_t)n + k];
                y[6 * (size_t)n + a] += s;
            }
        }
        for (size_t p = 0; p < pair_key.size(); ++p) {
            const uint32_t i = pair_key[p] >> 16, j = pair_key[p] & 0xFFFF;
            const R* B = &offdiag[36 * p];
            for (int a = 0; a < 6; ++a) {
                R s = 0, st = 0;
                for (int k = 0; k < 6; ++k) {
                    s += B[6 * a + k] * x[6 * (size_t)j + k];
                    st += B[6 * k + a] * x[6 * (size_t)i + k];
                }
                y[6 * (size_t)i + a] += s;
                y[6 * (size_t)j + a] += st;
            }
        }
    }

    void apply_block(const std::vector<R>& M, const std::vector<R>& x, std::vector<R>& y) const {
        y.assign((size_t)6 * N, 0);
        for (uint32_t n = 0; n < N; ++n)
            for (int a = 0; a < 6; ++a) {
                R s = 0;
                for (int k = 0; k < 6; ++k) s += M[36 * (size_t)n + 6 * a + k] * x[6 * (size_t)n + k];
                y[6 * (size_t)n + a] = s;
            }
    }

    static R vdot(const std::vector<R>& a, const std::vector<R>& b) {
        R s = 0;
        for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
        return s;
    }

    // BlockPCG<6>::SolveNoConvergeCheck (BlockPCG.cuh:65-181,312-458): zero
    // initialisation, fixed iteration count, no convergence test.  A zero
    // denominator (an empty system) gives a zero step instead of the
    // reference's 0/0.
    void pcg(uint32_t iters) {
        const size_t M = (size_t)6 * N;
        std::vector<R> r = jtr, d, q, t(M), s;
        apply_block(diag_inv, r, d);
        twist.assign(M, 0);
        R delta = vdot(r, d);
        for (uint32_t it = 0; it < iters; ++it) {
            spmv(d, q);
            const R dq = vdot(d, q);
            const R alpha = dq != R(0) ? delta / dq : R(0);
            for (size_t k = 0; k < M; ++k) { twist[k] += alpha * d[k]; t[k] = r[k] - alpha * q[k]; }
            apply_block(diag_inv, t, s);
            const R delta_old = delta;
            delta = vdot(t, s);
            const R beta = delta_old != R(0) ? delta / delta_old : R(0);
            for (size_t k = 0; k < M; ++k) d[k] = s[k] + beta * d[k];
            r.swap(t);
        }
        refine();
    }

    // Residual refinement used by the mixed-precision mode: one steepest
    // descent step in the preconditioned norm on the true residual b - A x.
    void refine() {
        for (uint32_t k = 0; k < cfg.refine_steps; ++k) {
            const size_t M = (size_t)6 * N;
            std::vector<R> ax, r(M), z, az;
            spmv(twist, ax);
            for (size_t i = 0; i < M; ++i) r[i] = jtr[i] - ax[i];
            apply_block(diag_inv, r, z);
            spmv(z, az);
            const R num = vdot(r, z), den = vdot(z, az);
            const R a = den > R(0) ? num / den : R(0);
            for (size_t i = 0; i < M; ++i) twist[i] += a * z[i];
        }
    }

    // applyWarpFieldUpdateKernel (SolverIterationData.cu:9-57): T_new = exp(xi) * T_prev
    void apply_update() {
        for (uint32_t n = 0; n < N; ++n) {
            const R* x = &twist[6 * (size_t)n];
            const M34<R> T = twist_to_se3(v3<R>(x[0], x[1], x[2]), v3<R>(x[3], x[4], x[5]));
            const M34<R> prev = se3[n].to_mat34();
            se3[n] = DQ<R>::from_mat34(T.mul(prev));
        }
    }

    // computeTermResidualKernel (ResidualEvaluator.cu:71-142), summed.
    double total_residual() const {
        double total = 0;
        for (const auto& t : dense) total += (double)t.r * t.r;
        if (cfg.density_in_jtj)
            for (const auto& t : density) total += (double)(pen.density * t.r) * (pen.density * t.r);
        for (const auto& s : smooth) {
            if (!s.valid) continue;
            const V3<R> d = s.Ti_xj - s.Tj_xj;
            total += pen.smooth * pen.smooth * ((double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z);
        }
        for (const auto& t : fg) total += (pen.foreground * t.r) * (pen.foreground * t.r);
        for (int which = 0; which < 2; ++which) {
            const auto& vec = which ? cross : feat;
            const double lam = which ? pen.cross : pen.feature;
            for (const auto& t : vec) {
                R J[3][6], r[3];
                p2p_jacobian(t, J, r);
                for (int c = 0; c < 3; ++c) total += lam * lam * (double)r[c] * r[c];
            }
        }
        return total;
    }

    void gn_iteration() {
        pen = penalty_for(cfg, iter);
        eval_terms();
        assemble_node_blocks();
        assemble_pair_blocks();
        invert_diag();
        pcg(cfg.pcg_iterations);
        apply_update();
        ++iter;
    }
};

// Type-erased handle: both instantiations behind one C API.
struct Handle {
    int precision = 1;  // 0 = float, 1 = double
    std::unique_ptr<Oracle<float>> f32;
    std::unique_ptr<Oracle<double>> f64;
    std::vector<uint8_t> frame_copy;  // host maps are referenced, not copied
};

template <typename R, typename Fn>
auto with(Handle* h, Fn fn) {
    if (h->precision == 0) return fn(*h->f32);
    return fn(*h->f64);
}

// ---------------------------------------------------------------- getters
template <typename R>
long long get_array(Oracle<R>& o, const std::string& name, void* dst, size_t bytes) {
    std::vector<double> fd;
    std::vector<uint32_t> iu;
    bool is_float = true;
    auto push_scalar = [&](const ScalarTerm<R>& t) {
        fd.push_back(t.knn[0]); fd.push_back(t.knn[1]); fd.push_back(t.knn[2]); fd.push_back(t.knn[3]);
        for (int k = 0; k < 4; ++k) fd.push_back(t.w[k]);
        for (int k = 0; k < 6; ++k) fd.push_back(t.J[k]);
        fd.push_back(t.r);
        fd.push_back(t.px + 65536.0 * t.py + 65536.0 * 65536.0 * t.view);
    };
    if (name == "dense_terms") for (auto& t : o.dense) push_scalar(t);          // 16 doubles / term
    else if (name == "density_terms") for (auto& t : o.density) push_scalar(t);
    else if (name == "foreground_terms") for (auto& t : o.fg) push_scalar(t);
    else if (name == "jtj_diag") fd.assign(o.diag.begin(), o.diag.end());
    else if (name == "jtj_diag_inv") fd.assign(o.diag_inv.begin(), o.diag_inv.end());
    else if (name == "jtr") fd.assign(o.jtr.begin(), o.jtr.end());
    else if (name == "jtj_offdiag") fd.assign(o.offdiag.begin(), o.offdiag.end());
    else if (name == "twist") fd.assign(o.twist.begin(), o.twist.end());
    else if (name == "node_se3")
        for (auto& d : o.se3) {
            fd.push_back(d.q0.x); fd.push_back(d.q0.y); fd.push_back(d.q0.z); fd.push_back(d.q0.w);
            fd.push_back(d.q1.x); fd.push_back(d.q1.y); fd.push_back(d.q1.z); fd.push_back(d.q1.w);
        }
    else if (name == "penalty") fd = {o.pen.smooth, o.pen.density, o.pen.foreground, o.pen.feature, o.pen.cross};
    else {
        is_float = false;
        if (name == "node2term_offset") iu = o.n2t_offset;
        else if (name == "node2term_index") iu = o.n2t_index;
        else if (name == "nodepair_key") iu = o.pair_key;
        else if (name == "nodepair_range")
            for (size_t p = 0; p < o.pair_key.size(); ++p) { iu.push_back(o.pair_start[p]); iu.push_back(o.pair_end[p]); }
        else if (name == "nodepair_term_index") iu = o.pair_terms;
        else if (name == "term_offsets") iu.assign(o.off, o.off + 5);
        else return -1;
    }
    const size_t size = is_float ? fd.size() * sizeof(double) : iu.size() * sizeof(uint32_t);
    if (dst && bytes >= size && size) std::memcpy(dst, is_float ? (const void*)fd.data() : (const void*)iu.data(), size);
    return (long long)size;
}

// Smooth-term export: 9 doubles per edge (i, j, valid, Ti_xj[3], Tj_xj[3]).
template <typename R>
long long get_smooth(Oracle<R>& o, void* dst, size_t bytes) {
    std::vector<double> fd;
    for (auto& s : o.smooth) {
        fd.push_back(s.i); fd.push_back(s.j); fd.push_back(s.valid);
        fd.push_back(s.Ti_xj.x); fd.push_back(s.Ti_xj.y); fd.push_back(s.Ti_xj.z);
        fd.push_back(s.Tj_xj.x); fd.push_back(s.Tj_xj.y); fd.push_back(s.Tj_xj.z);
    }
    const size_t size = fd.size() * sizeof(double);
    if (dst && bytes >= size && size) std::memcpy(dst, fd.data(), size);
    return (long long)size;
}

// ---------------------------------------------------------------- setters
// Replace the oracle's term records with externally computed ones (the GPU's
// FP32 records), so that assembly and PCG can be checked from identical inputs.
template <typename R>
int set_array(Oracle<R>& o, const std::string& name, const double* src, size_t count) {
    auto load_scalar = [&](std::vector<ScalarTerm<R>>& vec) {
        if (count % 16 || count / 16 != vec.size()) return -1;
        for (size_t i = 0; i < vec.size(); ++i) {
            const double* s = src + 16 * i;
            for (int k = 0; k < 4; ++k) { vec[i].knn[k] = (uint16_t)s[k]; vec[i].w[k] = (R)s[4 + k]; }
            for (int k = 0; k < 6; ++k) vec[i].J[k] = (R)s[8 + k];
            vec[i].r = (R)s[14];
        }
        return 0;
    };
    if (name == "dense_terms") return load_scalar(o.dense);
    if (name == "density_terms") return load_scalar(o.density);
    if (name == "foreground_terms") return load_scalar(o.fg);
    if (name == "smooth_terms") {
        if (count % 9 || count / 9 != o.smooth.size()) return -1;
        for (size_t i = 0; i < o.smooth.size(); ++i) {
            const double* s = src + 9 * i;
            o.smooth[i].valid = (uint8_t)s[2];
            o.smooth[i].Ti_xj = v3<R>((R)s[3], (R)s[4], (R)s[5]);
            o.smooth[i].Tj_xj = v3<R>((R)s[6], (R)s[7], (R)s[8]);
        }
        return 0;
    }
    if (name == "node_se3") {
        if (count != 8 * (size_t)o.N) return -1;
        for (uint32_t i = 0; i < o.N; ++i) {
            const double* s = src + 8 * i;
            o.se3[i].q0 = {(R)s[0], (R)s[1], (R)s[2], (R)s[3]};
            o.se3[i].q1 = {(R)s[4], (R)s[5], (R)s[6], (R)s[7]};
        }
        return 0;
    }
    if (name == "twist") {
        if (count != 6 * (size_t)o.N) return -1;
        o.twist.assign(src, src + count);
        return 0;
    }
    if (name == "bilinear_images") {
        if (count != 1) return -1;
        o.bilinear_images = src[0] != 0.0;
        return 0;
    }
    return -1;
}

// P2P records travel as 20 doubles: knn[4], w[4], canonical[3], target[3], warped[3], pad[3].
template <typename R>
int set_p2p(Oracle<R>& o, bool cross_terms, const double* src, size_t count) {
    auto& vec = cross_terms ? o.cross : o.feat;
    if (count % 20 || count / 20 != vec.size()) return -1;
    for (size_t i = 0; i < vec.size(); ++i) {
        const double* s = src + 20 * i;
        for (int k = 0; k < 4; ++k) { vec[i].knn[k] = (uint16_t)s[k]; vec[i].w[k] = (R)s[4 + k]; }
        vec[i].canonical = v3<R>((R)s[8], (R)s[9], (R)s[10]);
        vec[i].target = v3<R>((R)s[11], (R)s[12], (R)s[13]);
        vec[i].warped = v3<R>((R)s[14], (R)s[15], (R)s[16]);
    }
    return 0;
}

template <typename R>
long long get_p2p(Oracle<R>& o, bool cross_terms, void* dst, size_t bytes) {
    const auto& vec = cross_terms ? o.cross : o.feat;
    std::vector<double> fd;
    for (const auto& t : vec) {
        for (int k = 0; k < 4; ++k) fd.push_back(t.knn[k]);
        for (int k = 0; k < 4; ++k) fd.push_back(t.w[k]);
        fd.push_back(t.canonical.x); fd.push_back(t.canonical.y); fd.push_back(t.canonical.z);
        fd.push_back(t.target.x); fd.push_back(t.target.y); fd.push_back(t.target.z);
        fd.push_back(t.warped.x); fd.push_back(t.warped.y); fd.push_back(t.warped.z);
        fd.push_back(0); fd.push_back(0); fd.push_back(0);
    }
    const size_t size = fd.size() * sizeof(double);
    if (dst && bytes >= size && size) std::memcpy(dst, fd.data(), size);
    return (long long)size;
}

}  // namespace

// ====================================================================== C API
extern "C" {

typedef Handle mvdr_oracle;

const char* mvdr_oracle_last_error(void) { return g_err.c_str(); }

// precision: 0 = FP32 restatement, 1 = FP64 oracle.  `frame` holds HOST
// pointers, which must outlive the handle.
mvdr_oracle* mvdr_oracle_create(const mvdr_solver_config* cfg, const mvdr_frame_input* frame, int precision) {
    if (!cfg || !frame) { g_err = "null argument"; return nullptr; }
    if (frame->num_views < 1 || frame->num_views > MVDR_MAX_VIEWS || frame->num_nodes > 65535) {
        g_err = "frame shape out of range";
        return nullptr;
    }
    auto* h = new Handle();
    h->precision = precision ? 1 : 0;
    if (h->precision) h->f64.reset(new Oracle<double>(*cfg, *frame));
    else h->f32.reset(new Oracle<float>(*cfg, *frame));
    return h;
}

void mvdr_oracle_destroy(mvdr_oracle* h) { delete h; }

int mvdr_oracle_build_index(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) { o.build_index(); return 0; });
}

// Stage entry points; `iter` selects the penalty schedule of that GN iteration.
int mvdr_oracle_set_iteration(mvdr_oracle* h, uint32_t iter) {
    return with<float>(h, [iter](auto& o) { o.iter = iter; o.pen = penalty_for(o.cfg, iter); return 0; });
}
int mvdr_oracle_eval_terms(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) { o.eval_terms(); return 0; });
}
int mvdr_oracle_assemble(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) { o.assemble_node_blocks(); o.assemble_pair_blocks(); o.invert_diag(); return 0; });
}
int mvdr_oracle_pcg(mvdr_oracle* h, uint32_t iters) {
    return with<float>(h, [iters](auto& o) { o.pcg(iters); return 0; });
}
int mvdr_oracle_apply_update(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) { o.apply_update(); return 0; });
}
int mvdr_oracle_gn_iteration(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) { o.gn_iteration(); return 0; });
}
// Full SolveNonRigidAlignment: index build + gn_iterations GN steps.
int mvdr_oracle_solve(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) {
        o.build_index();
        o.iter = 0;
        for (uint32_t k = 0; k < o.cfg.gn_iterations; ++k) o.gn_iteration();
        return 0;
    });
}
double mvdr_oracle_total_residual(mvdr_oracle* h) {
    return with<float>(h, [](auto& o) { return o.total_residual(); });
}

// y = A x with the current assembled matrix (both FP64 arrays of 6N).
int mvdr_oracle_apply_jtj(mvdr_oracle* h, const double* x, double* y) {
    return with<float>(h, [x, y](auto& o) {
        using R = typename std::remove_reference<decltype(o.diag)>::type::value_type;
        std::vector<R> xv(x, x + 6 * (size_t)o.N), yv;
        o.spmv(xv, yv);
        for (size_t i = 0; i < yv.size(); ++i) y[i] = (double)yv[i];
        return 0;
    });
}

// Floating arrays come back as doubles, index arrays as uint32.  Returns the
// byte size (or -1 for an unknown name); copies when dst is large enough.
// Record layouts (doubles): dense / density / foreground terms 16 per term
// (knn[4], w[4], J[6], r, packed pixel); feature / cross terms 20 per term
// (knn[4], w[4], canonical[3], target[3], warped[3], pad[3]); smooth terms 9 per
// edge (i, j, valid, Ti_xj[3], Tj_xj[3]).
long long mvdr_oracle_get(mvdr_oracle* h, const char* name, void* dst, size_t bytes) {
    const std::string n(name ? name : "");
    return with<float>(h, [&](auto& o) -> long long {
        if (n == "smooth_terms") return get_smooth(o, dst, bytes);
        if (n == "feature_terms") return get_p2p(o, false, dst, bytes);
        if (n == "cross_terms") return get_p2p(o, true, dst, bytes);
        return get_array(o, n, dst, bytes);
    });
}

// ---- math hooks, FP64, for pinning the primitives against closed forms ----
static void put_m34(const M34<double>& m, double out[12]) {
    for (int c = 0; c < 3; ++c) { out[3 * c] = m.c[c].x; out[3 * c + 1] = m.c[c].y; out[3 * c + 2] = m.c[c].z; }
    out[9] = m.t.x; out[10] = m.t.y; out[11] = m.t.z;
}
static M34<double> get_m34(const double in[12]) {
    M34<double> m;
    for (int c = 0; c < 3; ++c) m.c[c] = v3<double>(in[3 * c], in[3 * c + 1], in[3 * c + 2]);
    m.t = v3<double>(in[9], in[10], in[11]);
    return m;
}
static DQ<double> get_dq(const double d[8]) { return {{d[0], d[1], d[2], d[3]}, {d[4], d[5], d[6], d[7]}}; }
static void put_dq(const DQ<double>& d, double out[8]) {
    out[0] = d.q0.x; out[1] = d.q0.y; out[2] = d.q0.z; out[3] = d.q0.w;
    out[4] = d.q1.x; out[5] = d.q1.y; out[6] = d.q1.z; out[7] = d.q1.w;
}

// mode 0: DualQuaternion::se3_matrix; mode 1: operator mat34.  Output mat34 as
// 9 column-major rotation entries + 3 translation entries.
void mvdr_oracle_dq_to_mat34(const double dq[8], int mode, double out[12]) {
    const DQ<double> d = get_dq(dq);
    put_m34(mode ? d.to_mat34() : d.se3_matrix(), out);
}
void mvdr_oracle_mat34_to_dq(const double m[12], double dq[8]) { put_dq(DQ<double>::from_mat34(get_m34(m)), dq); }
void mvdr_oracle_twist_to_mat34(const double twist[6], double out[12]) {
    put_m34(twist_to_se3(v3<double>(twist[0], twist[1], twist[2]), v3<double>(twist[3], twist[4], twist[5])), out);
}
// weighted blend of four dual quaternions (4 x 8 doubles) with knn 0..3
void mvdr_oracle_blend(const double dqs[32], const double w[4], double out[8]) {
    std::vector<DQ<double>> f(4);
    for (int k = 0; k < 4; ++k) f[k] = get_dq(dqs + 8 * k);
    const uint16_t knn[4] = {0, 1, 2, 3};
    put_dq(blend(f, knn, w), out);
}
void mvdr_oracle_inverse6(const double A[36], double out[36]) { Oracle<double>::gauss_jordan6(A, out); }
void mvdr_oracle_penalty(const mvdr_solver_config* cfg, uint32_t iter, double out[5]) {
    const Penalty p = penalty_for(*cfg, iter);
    out[0] = p.smooth; out[1] = p.density; out[2] = p.foreground; out[3] = p.feature; out[4] = p.cross;
}

int mvdr_oracle_set(mvdr_oracle* h, const char* name, const double* src, size_t count) {
    const std::string n(name ? name : "");
    return with<float>(h, [&](auto& o) -> int {
        if (n == "feature_terms") return set_p2p(o, false, src, count);
        if (n == "cross_terms") return set_p2p(o, true, src, count);
        return set_array(o, n, src, count);
    });
}

// Cross-view edge pairs (CrossViewEdgeCorrespondence.h:21-224; the
// implementation is absent from the reference, so this follows the rules
// stated in include/mvdr/crossview.h -- PARITY UNPINNED against the
// reference).  Plain loops over every pixel and window tap, FP32 in the
// device's operation order.  Writes up to `cap` pairs (6 u16 each: x1, y1,
// x2, y2, view1, view2) of list 0 (candidates), 1 (unique) or 2
// (back-traced); returns the list's full length.
long long mvdr_oracle_crossview(const mvdr_crossview_config* cfg, const mvdr_crossview_input* in, int list,
                                uint16_t* out, uint32_t cap) {
    if (!cfg || !in || list < 0 || list > 2) return -1;
    const uint32_t V = in->num_views, W = in->width, H = in->height, plane = W * H;
    auto vertex = [&](uint32_t v, int x, int y, float o[4]) {
        std::memcpy(o, static_cast<const float*>(in->vertex[v]) + 4 * ((size_t)y * W + x), 16);
    };
    auto is_edge = [&](uint32_t v, int x, int y, float o[4]) {
        if (x < 0 || y < 0 || x >= (int)W || y >= (int)H) return false;
        if (static_cast<const uint8_t*>(in->edge_mask[v])[(size_t)y * W + x] != 1) return false;
        vertex(v, x, y, o);
        return !is_zero_vertex<float>(o);
    };
    struct Cand { bool ok; uint32_t view2; int x2, y2; float d2; };
    std::vector<Cand> cand((size_t)V * plane, Cand{false, 0, 0, 0, 0.f});
    const int R = (int)cfg->search_radius, S = (int)cfg->search_step, T = (int)cfg->track_radius;
    for (uint32_t v = 0; v < V && V > 1; ++v) {
        const M34<float> Iv = M34<float>::from(in->initial_camera_se3[v]);
        for (int y = 0; y < (int)H; ++y)
            for (int x = 0; x < (int)W; ++x) {
                float src[4];
                if (!is_edge(v, x, y, src)) continue;
                const V3<float> world = Iv.apply(v3<float>(src[0], src[1], src[2]));
                float best = INFINITY;
                Cand c{false, 0, 0, 0, 0.f};
                const uint32_t next = (v + 1) % V, last = (v + V - 1) % V;
                for (int k = 0; k < 2; ++k) {
                    const uint32_t n = k == 0 ? next : last;
                    if (k == 1 && last == next) break;
                    const V3<float> q = M34<float>::from(in->initial_camera_se3[n]).inverse().apply(world);
                    if (!(q.z > 0.f)) continue;
                    const mvdr_intrinsic K = in->intrinsic[n];
                    const int u = project_round(q.x, q.z, K.fx, K.cx), w = project_round(q.y, q.z, K.fy, K.cy);
                    float nb = INFINITY;
                    int bx = -1, by = -1;
                    for (int yy = w - R; yy <= w + R; yy += S)
                        for (int xx = u - R; xx <= u + R; xx += S) {
                            float t[4];
                            if (!is_edge(n, xx, yy, t)) continue;
                            const float dx = t[0] - q.x, dy = t[1] - q.y, dz = t[2] - q.z;
                            const float d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 < nb) { nb = d2; bx = xx; by = yy; }
                        }
                    if (nb < best) { best = nb; c = Cand{true, n, bx, by, nb}; }
                }
                if (c.ok && best < cfg->max_distance2) cand[(size_t)v * plane + (size_t)y * W + x] = c;
            }
    }
    auto tgt = [&](const Cand& c) { return (size_t)c.view2 * plane + (size_t)c.y2 * W + c.x2; };
    // per target: the nearest candidate, ties to the lower source pixel
    std::map<size_t, size_t> winner;
    for (size_t p = 0; p < cand.size(); ++p) {
        if (!cand[p].ok) continue;
        auto it = winner.find(tgt(cand[p]));
        if (it == winner.end() || cand[p].d2 < cand[it->second].d2) winner[tgt(cand[p])] = p;
    }
    auto traceable = [&](uint32_t v, int x, int y) {
        if (!in->corr_map[v]) return false;
        for (int yy = y - T; yy <= y + T; ++yy)
            for (int xx = x - T; xx <= x + T; ++xx) {
                if (xx < 0 || yy < 0 || xx >= (int)W || yy >= (int)H) continue;
                const uint16_t* e = static_cast<const uint16_t*>(in->corr_map[v]) + 4 * ((size_t)yy * W + xx);
                if (e[2] == xx && e[3] == yy) return true;
            }
        return false;
    };
    long long n = 0;
    for (size_t p = 0; p < cand.size(); ++p) {
        const Cand& c = cand[p];
        if (!c.ok) continue;
        const size_t t = tgt(c);
        bool unique = winner[t] == p;
        if (unique && t < p && cand[t].ok && tgt(cand[t]) == p) {
            auto it = winner.find(p);
            if (it != winner.end() && it->second == t) unique = false;
        }
        const uint32_t v = (uint32_t)(p / plane), o = (uint32_t)(p % plane);
        const bool keep = list == 0 || (unique && (list == 1 || (traceable(v, (int)(o % W), (int)(o / W)) &&
                                                                traceable(c.view2, c.x2, c.y2))));
        if (!keep) continue;
        if (out && (uint64_t)n < cap) {
            uint16_t* r = out + 6 * n;
            r[0] = (uint16_t)(o % W); r[1] = (uint16_t)(o / W); r[2] = (uint16_t)c.x2; r[3] = (uint16_t)c.y2;
            r[4] = (uint16_t)v; r[5] = (uint16_t)c.view2;
        }
        ++n;
    }
    return n;
}

// ---------------------------------------------------------- warp field
// BuildNodesGraphKernel (WarpField.cu:15-59): 8 slots, a closer node (squared
// distance in (1e-6, worst)) replaces the worst slot, the new worst is the
// first strict maximum; edges (i, slot) in slot order, unfilled = 0xFFFF.
void mvdr_oracle_node_graph(const float* nodes, uint32_t n, uint16_t* graph) {
    for (uint32_t i = 0; i < n; ++i) {
        float dist[8];
        int idx[8];
        for (int k = 0; k < 8; ++k) { dist[k] = 1e5f; idx[k] = -1; }
        int worst = 0;
        const float* a = nodes + 4 * (size_t)i;
        for (uint32_t k = 0; k < n; ++k) {
            const float* b = nodes + 4 * (size_t)k;
            const float d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
            if (d > 1e-6f && d < dist[worst]) {
                dist[worst] = d;
                idx[worst] = (int)k;
                worst = 0;
                float mx = 0.f;
                for (int j = 0; j < 8; ++j)
                    if (dist[j] > mx) { worst = j; mx = dist[j]; }
            }
        }
        for (int k = 0; k < 8; ++k) {
            graph[2 * (8 * (size_t)i + k)] = (uint16_t)i;
            graph[2 * (8 * (size_t)i + k) + 1] = (uint16_t)idx[k];
        }
    }
}

// skinningVertexAndNodeBruteForceKernel (CanonicalNodesSkinner.cu:356-437) with
// the 4-slot heap of brute_foce_knn.cuh:17-39 over the real nodes; points are
// the vertices, then the nodes.
void mvdr_oracle_skin(const float* points, uint32_t num_points, const float* nodes, uint32_t n, float radius,
                      uint16_t* knn, float* weight) {
    for (uint32_t p = 0; p < num_points; ++p) {
        const float* v = points + 4 * (size_t)p;
        float d[4] = {1e6f, 1e6f, 1e6f, 1e6f};
        uint32_t id[4] = {0, 0, 0, 0};
        for (uint32_t k = 0; k < n; ++k) {
            const float* q = nodes + 4 * (size_t)k;
            const float dx = v[0] - q[0], dy = v[1] - q[1], dz = v[2] - q[2];
            const float dist = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
            if (!(dist < d[0])) continue;
            d[0] = dist; id[0] = k;
            if (d[1] < d[2]) {
                if (d[0] < d[2]) { std::swap(d[0], d[2]); std::swap(id[0], id[2]); }
            } else if (d[0] < d[1]) {
                std::swap(d[0], d[1]); std::swap(id[0], id[1]);
                if (d[1] < d[3]) { std::swap(d[1], d[3]); std::swap(id[1], id[3]); }
            }
        }
        float w[4];
        const float den = (2.0f * radius) * radius;
        for (int s = 0; s < 4; ++s) {
            const float* q = nodes + 4 * (size_t)id[s];
            const float dx = v[0] - q[0], dy = v[1] - q[1], dz = v[2] - q[2];
            w[s] = std::exp(-(dx * dx + dy * dy + dz * dz) / den);
        }
        const float inv = 1.0f / (std::fabs(w[0]) + std::fabs(w[1]) + std::fabs(w[2]) + std::fabs(w[3]));
        for (int s = 0; s < 4; ++s) {
            knn[4 * (size_t)p + s] = (uint16_t)id[s];
            weight[4 * (size_t)p + s] = w[s] * inv;
        }
    }
}

// forwardWarpVertexAndNodeKernel (SurfelNodeDeformer.cu:9-60) for one point
// set: v' = T v (w kept), n' = R n, T = se3_matrix(blend(knn, w)) in FP32.
// normals may be null (nodes).  knn naming a node >= n copies the point.
void mvdr_oracle_warp(const float* field_dq, uint32_t n, const float* points, const float* normals,
                      const uint16_t* knn, const float* weight, uint32_t count, int inverse, float* out_points,
                      float* out_normals) {
    std::vector<DQ<float>> field(n);
    for (uint32_t i = 0; i < n; ++i) {
        mvdr_dq d;
        std::memcpy(&d, field_dq + 8 * (size_t)i, sizeof(d));
        field[i] = DQ<float>::from(d);
    }
    for (uint32_t p = 0; p < count; ++p) {
        const float* v = points + 4 * (size_t)p;
        const uint16_t* k = knn + 4 * (size_t)p;
        float* ov = out_points + 4 * (size_t)p;
        std::memcpy(ov, v, 16);
        if (normals) std::memcpy(out_normals + 4 * (size_t)p, normals + 4 * (size_t)p, 16);
        if (k[0] >= n || k[1] >= n || k[2] >= n || k[3] >= n) continue;
        const M34<float> T = blend(field, k, weight + 4 * (size_t)p).se3_matrix();
        V3<float> a = v3<float>(v[0], v[1], v[2]);
        a = inverse ? T.rot_t(a - T.t) : T.apply(a);
        ov[0] = a.x; ov[1] = a.y; ov[2] = a.z;
        if (normals) {
            const float* nn = normals + 4 * (size_t)p;
            V3<float> b = v3<float>(nn[0], nn[1], nn[2]);
            b = inverse ? T.rot_t(b) : T.rot(b);
            float* on = out_normals + 4 * (size_t)p;
            on[0] = b.x; on[1] = b.y; on[2] = b.z;
        }
    }
}

// CorrectLargeErrorNode, correspondence-guided form (NonRigidSolver.cu:100-138,
// :155-216; rules in include/mvdr/warpfield.h), on host arrays in place, FP32.
namespace {
Q<float> qpow_ref(Q<float> q, float t) {  // Quaternion::pow (Quaternion.h:235-249)
    const float n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
    const float nt = std::atan2(n, q.w) * t;
    if (std::fabs(n) < 1e-6f) return {0.f, 0.f, 0.f, std::cos(nt)};
    const float sc = std::sin(nt) / n;
    return {q.x * sc, q.y * sc, q.z * sc, std::cos(nt)};
}
DQ<float> dq_scale_ref(const DQ<float>& d, float ratio) {  // DualQuaternion::operator*(float) (:165-178)
    float t = ratio;
    if (t > 1.0f) t = 1.0f;
    else if (t < 1e-10f) t = 1e-10f;
    const Q<float> r = qmul(d.q1, conj(d.q0));
    const Q<float> q0s = qpow_ref(d.q0, t);
    return {q0s, qmul(Q<float>{t * r.x, t * r.y, t * r.z, t * r.w}, q0s)};
}
DQ<float> dq_mul_ref(const DQ<float>& a, const DQ<float>& b) {  // DualQuaternion::operator* (:152-157)
    const Q<float> u = qmul(a.q1, b.q0), v = qmul(a.q0, b.q1);
    return {qmul(a.q0, b.q0), {u.x + v.x, u.y + v.y, u.z + v.z, u.w + v.w}};
}
uint32_t pixel_ref(float x, float z, float f, float c) {
    const float p = (float)(((double)x / ((double)z + 1e-10)) * (double)f + (double)c);
    if (!(p > 0.f)) return 0u;
    if (p >= 4294967040.f) return 0xFFFFFFFFu;
    return (uint32_t)std::nearbyint(p);
}
}  // namespace

int mvdr_oracle_correct_nodes(const mvdr_node_correction* in) {
    if (!in || in->num_views < 1 || in->num_views > MVDR_MAX_VIEWS) return -1;
    const uint32_t V = in->num_views, W = in->width, H = in->height, N = in->num_nodes;
    const int hw = 2;                    // CorrSearchWindowHalfSize (NonRigidSolver.h:41)
    const float r2 = 0.025f * 0.025f;    // NODE_RADIUS^2 (GlobalConfigs.h:85-86)
    float4_t* nodes = static_cast<float4_t*>(in->live_nodes);
    const float* err = static_cast<const float*>(in->node_error);
    mvdr_dq* corr = static_cast<mvdr_dq*>(in->correction);
    for (uint32_t n = 0; n < N; ++n) {
        DQ<float> out{{0.f, 0.f, 0.f, 1.f}, {0.f, 0.f, 0.f, 0.f}};
        const float4_t node = nodes[n];
        uint32_t v = 0xFFFFFFFFu;  // decodeNodesCameraView (CommonUtils.h:110-117)
        for (uint32_t i = 0; i < (uint32_t)MVDR_MAX_VIEWS; ++i)
            if ((float)i < node.w && node.w <= (float)(i + 1)) { v = i; break; }
        if (err[n] > 0.2f && v < V) {
            const M34<float> w2c = M34<float>::from(in->world2camera[v]);
            const M34<float> init = M34<float>::from(in->initial_camera_se3[v]);
            const M34<float> w2c_inv = w2c.inverse(), init_inv = init.inverse();
            const V3<float> c = init_inv.apply(w2c.apply(v3<float>(node.x, node.y, node.z)));
            const mvdr_intrinsic K = in->intrinsic[v];
            const uint32_t u = pixel_ref(c.x, c.z, K.fx, K.cx), y = pixel_ref(c.y, c.z, K.fy, K.cy);
            M34<float> sel = M34<float>::identity();
            const bool room = W > 2u * hw && H > 2u * hw;
            if (room && !(u >= W - hw || u < (uint32_t)hw || y >= H - hw || y < (uint32_t)hw)) {
                float best = 1e6f;
                const float4_t* prev = static_cast<const float4_t*>(in->previous_vertex[v]);
                const float* guided = static_cast<const float*>(in->guided_se3[v]);
                const uint8_t* valid = static_cast<const uint8_t*>(in->guided_valid[v]);
                for (uint32_t i = u - hw; i <= u + hw; ++i)
                    for (uint32_t j = y - hw; j <= y + hw; ++j) {
                        const size_t o = (size_t)j * W + i;
                        const float pv[4] = {prev[o].x, prev[o].y, prev[o].z, prev[o].w};
                        if (valid[o] != 1 || is_zero_vertex<float>(pv)) continue;
                        M34<float> G;
                        const float* g = guided + 12 * o;
                        for (int col = 0; col < 3; ++col) G.c[col] = v3<float>(g[3 * col], g[3 * col + 1], g[3 * col + 2]);
                        G.t = v3<float>(g[9], g[10], g[11]);
                        const V3<float> d = c - v3<float>(pv[0], pv[1], pv[2]);
                        const float d2 = d.x * d.x + d.y * d.y + d.z * d.z;
                        const V3<float> m = G.apply(c) - c;
                        if (best > d2 && d2 < 1e-4f && m.x * m.x + m.y * m.y + m.z * m.z < 4 * r2) {
                            best = d2;
                            sel = G;
                        }
                    }
                sel = w2c_inv.mul(init).mul(sel).mul(init_inv).mul(w2c);
            }
            out = dq_scale_ref(DQ<float>::from_mat34(sel), err[n]);
        }
        out.to(corr[n]);
    }
    std::vector<DQ<float>> field(N);
    for (uint32_t n = 0; n < N; ++n) field[n] = DQ<float>::from(corr[n]);
    auto apply = [&](float4_t* pts, float4_t* normals, const uint16_t* knn, const float* w, uint32_t count,
                     mvdr_dq* se3) {
        for (uint32_t p = 0; p < count; ++p) {
            const uint16_t* k = knn + 4 * (size_t)p;
            if (k[0] >= N || k[1] >= N || k[2] >= N || k[3] >= N) continue;
            const DQ<float> avg = blend(field, k, w + 4 * (size_t)p);
            const M34<float> T = avg.se3_matrix();
            const V3<float> a = T.apply(v3<float>(pts[p].x, pts[p].y, pts[p].z));
            pts[p].x = a.x; pts[p].y = a.y; pts[p].z = a.z;
            if (normals) {
                const V3<float> b = T.rot(v3<float>(normals[p].x, normals[p].y, normals[p].z));
                normals[p].x = b.x; normals[p].y = b.y; normals[p].z = b.z;
            }
            if (se3) dq_mul_ref(avg, DQ<float>::from(se3[p])).to(se3[p]);
        }
    };
    apply(static_cast<float4_t*>(in->live_vertex), static_cast<float4_t*>(in->live_normal),
          static_cast<const uint16_t*>(in->surfel_knn), static_cast<const float*>(in->surfel_knn_weight), in->num_surfels,
          nullptr);
    apply(nodes, nullptr, static_cast<const uint16_t*>(in->node_knn), static_cast<const float*>(in->node_knn_weight), N,
          static_cast<mvdr_dq*>(in->node_se3));
    return 0;
}

// ------------------------------------------------------- solver maps
// The point program of render/shaders/SolverMap.vert/.frag as a z-buffer over
// surfels in draw order (rules in include/mvdr/render.h): a surfel passing the
// depth-range and confidence/time tests lands on pixel (floor(u), floor(v)) and
// replaces the pixel's surfel when strictly nearer.  Host arrays; the output
// struct's arrays are written (zbuffer unused).
int mvdr_oracle_render(const mvdr_render_input* in, const mvdr_render_output* out) {
    const uint32_t W = in->width, H = in->height, S = in->num_surfels;
    const float4_t* ct = static_cast<const float4_t*>(in->color_time);
    for (uint32_t v = 0; v < in->num_views; ++v) {
        std::vector<float> depth((size_t)W * H, INFINITY);
        std::vector<uint32_t> idx((size_t)W * H, 0xFFFFFFFFu);
        const M34<float> T = M34<float>::from(in->world2camera[v]);
        const mvdr_intrinsic K = in->intrinsic[v];
        const float4_t* rv = static_cast<const float4_t*>(in->reference_vertex[v]);
        const float4_t* wv = static_cast<const float4_t*>(in->warp_vertex[v]);
        for (uint32_t s = 0; s < S; ++s) {
            const V3<float> p = T.apply(v3<float>(wv[s].x, wv[s].y, wv[s].z));
            if (p.z > in->max_depth + 0.05f || p.z < 0.f) continue;
            if (rv[s].w < in->confidence_threshold && std::fabs(ct[s].z - in->current_time) >= in->time_threshold) continue;
            const float u = (K.fx * p.x) / p.z + K.cx, w = (K.fy * p.y) / p.z + K.cy;
            if (!(u >= 0.f && w >= 0.f && u < (float)W && w < (float)H)) continue;
            const size_t o = (size_t)std::floor(w) * W + (size_t)std::floor(u);
            if (p.z < depth[o]) { depth[o] = p.z; idx[o] = s; }
        }
        const float4_t* rn = static_cast<const float4_t*>(in->reference_normal[v]);
        const float4_t* wn = static_cast<const float4_t*>(in->warp_normal[v]);
        for (size_t o = 0; o < (size_t)W * H; ++o) {
            float4_t a{0, 0, 0, 0}, b{0, 0, 0, 0}, c{0, 0, 0, 0}, d{0, 0, 0, 0}, rgb{0, 0, 0, 0};
            const uint32_t s = idx[o];
            if (s != 0xFFFFFFFFu) {
                a = {rv[s].x, rv[s].y, rv[s].z, 1.f};
                b = {rn[s].x, rn[s].y, rn[s].z, 0.f};
                c = {wv[s].x, wv[s].y, wv[s].z, 1.f};
                d = {wn[s].x, wn[s].y, wn[s].z, 0.f};
                int32_t bits;
                std::memcpy(&bits, &ct[s].x, 4);
                rgb = {(float)((bits & 0x00FF0000) >> 16) / 255.f, (float)((bits & 0x0000FF00) >> 8) / 255.f,
                       (float)(bits & 0x000000FF) / 255.f, ct[s].y};
            }
            static_cast<float4_t*>(out->reference_vertex[v])[o] = a;
            static_cast<float4_t*>(out->reference_normal[v])[o] = b;
            static_cast<float4_t*>(out->warp_vertex[v])[o] = c;
            if (out->warp_normal[v]) static_cast<float4_t*>(out->warp_normal[v])[o] = d;
            static_cast<uint32_t*>(out->index_map[v])[o] = s;
            if (out->normalized_rgb[v]) static_cast<float4_t*>(out->normalized_rgb[v])[o] = rgb;
        }
    }
    return 0;
}

// ------------------------------------------------------------- rigid ICP
// SolveRigidAlignment(solverMaps, observation, LiveField2ObservedField, n)
// (RigidSolver.cpp:258-276, rows RigidSolver.cu:128-226, update
// RigidSolver.cpp:389-446) with the rules of include/mvdr/rigid.h: FP64 sums
// over all pixels (or every 32nd 256-pixel block, times 32, with
// reference_reduction), Cholesky, twist -> Rodrigues transform on the left.
// Host maps; `poses` holds the start poses on entry and the result on exit.
// The optional `system` receives the last iteration's 27 sums of view 0.
int mvdr_oracle_rigid(const mvdr_rigid_input* in, int reference_reduction, uint32_t iterations, mvdr_mat34* poses,
                      double* system) {
    const uint32_t W = in->width, H = in->height;
    for (uint32_t v = 0; v < in->num_views; ++v) {
        M34<float> T = M34<float>::from(poses[v]);
        const float4_t* cv = static_cast<const float4_t*>(in->conversion_vertex[v]);
        const float4_t* cn = static_cast<const float4_t*>(in->conversion_normal[v]);
        const float4_t* rv = static_cast<const float4_t*>(in->reference_vertex[v]);
        const float4_t* rn = static_cast<const float4_t*>(in->reference_normal[v]);
        const mvdr_intrinsic K = in->intrinsic[v];
        for (uint32_t it = 0; it < iterations; ++it) {
            double S[27] = {0};
            for (uint32_t pix = 0; pix < W * H; ++pix) {
                if (reference_reduction && ((pix / 256) % 32) != 0) continue;
                const V3<float> p = T.apply(v3<float>(cv[pix].x, cv[pix].y, cv[pix].z));
                const V3<float> pn = T.rot(v3<float>(cn[pix].x, cn[pix].y, cn[pix].z));
                const float u = (float)((double)p.x / ((double)p.z + 1e-10) * (double)K.fx + (double)K.cx);
                const float w = (float)((double)p.y / ((double)p.z + 1e-10) * (double)K.fy + (double)K.cy);
                auto px16 = [](float f) -> uint32_t {
                    uint32_t r;
                    if (!(f > 0.f)) r = 0u;
                    else if (f >= 4294967040.f) r = 0xFFFFFFFFu;
                    else r = (uint32_t)std::nearbyint(f);
                    return r & 0xFFFFu;
                };
                const uint32_t x = px16(u), y = px16(w);
                if (!(x < W && y < H)) continue;
                const float4_t q = rv[y * W + x], qn = rn[y * W + x];
                const float qa[4] = {q.x, q.y, q.z, q.w};
                const float ndot = pn.x * qn.x + pn.y * qn.y + pn.z * qn.z;
                const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (is_zero_vertex<float>(qa) || ndot < 0.8f || d2 > 0.03f * 0.03f) continue;
                const V3<float> n = v3<float>(qn.x, qn.y, qn.z);
                const V3<float> pxn = vcross(p, n);
                const double b = -(double)(n.x * dx + n.y * dy + n.z * dz);
                const double a[6] = {pxn.x, pxn.y, pxn.z, n.x, n.y, n.z};
                int k = 0;
                for (int i = 0; i < 6; ++i)
                    for (int j = i; j < 6; ++j) S[k++] += a[i] * a[j];
                for (int i = 0; i < 6; ++i) S[k++] += a[i] * b;
            }
            if (reference_reduction)
                for (double& x : S) x *= 32.0;
            if (system && v == 0) std::memcpy(system, S, sizeof(S));
            double A[6][6], y[6];
            int k = 0;
            for (int i = 0; i < 6; ++i)
                for (int j = i; j < 6; ++j) { A[i][j] = A[j][i] = S[k]; ++k; }
            for (int i = 0; i < 6; ++i) y[i] = S[21 + i];
            bool ok = true;
            for (int j = 0; j < 6 && ok; ++j) {
                double d = A[j][j];
                for (int q2 = 0; q2 < j; ++q2) d -= A[j][q2] * A[j][q2];
                if (!(d > 0.0)) { ok = false; break; }
                A[j][j] = std::sqrt(d);
                for (int i = j + 1; i < 6; ++i) {
                    double t = A[i][j];
                    for (int q2 = 0; q2 < j; ++q2) t -= A[i][q2] * A[j][q2];
                    A[i][j] = t / A[j][j];
                }
            }
            if (!ok) continue;
            for (int i = 0; i < 6; ++i) {
                double t = y[i];
                for (int q2 = 0; q2 < i; ++q2) t -= A[i][q2] * y[q2];
                y[i] = t / A[i][i];
            }
            for (int i = 5; i >= 0; --i) {
                double t = y[i];
                for (int q2 = i + 1; q2 < 6; ++q2) t -= A[q2][i] * y[q2];
                y[i] = t / A[i][i];
            }
            const M34<float> U = twist_to_se3(v3<float>((float)y[0], (float)y[1], (float)y[2]),
                                              v3<float>((float)y[3], (float)y[4], (float)y[5]));
            T = U.mul(T);
        }
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) poses[v].rot[3 * c + r] = T.m(r, c);
        poses[v].trans[0] = T.t.x; poses[v].trans[1] = T.t.y; poses[v].trans[2] = T.t.z;
    }
    return 0;
}

}  // extern "C"
