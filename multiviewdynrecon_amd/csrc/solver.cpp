// Host side of the C ABI (include/mvdr/solver.h): buffer ownership, input
// validation, the GN schedule and stage launches.  Mirrors
// SparseSurfelFusion::NonRigidSolver (core/NonRigidSolver/NonRigidSolver.h:78-541).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "tuning.h"

#include "common.h"

namespace mvdr {
void launch_index_build(const Batch& B, bool use_foreground, bool use_cross, int mode, hipStream_t stream);
void launch_pair_term_lists(const Batch& B, hipStream_t stream);
uint32_t index_chunks(uint64_t max_terms);
void launch_eval_terms(const Batch& B, uint32_t which, bool fp64_geometry, hipStream_t stream);
void launch_pair_blocks(const Batch& B, const Penalty& pen, hipStream_t stream);
void launch_diag_blocks(const Batch& B, hipStream_t stream);
void launch_assemble(const Batch& B, const Penalty& pen, bool gram, bool diag_only, hipStream_t stream);
void launch_pcg_matrix_free(const Batch& B, uint32_t iters, uint32_t refine, bool fp64, const Penalty& pen,
                            bool recompute, uint32_t which, hipStream_t stream);
void launch_mf_apply_once(const Batch& B, bool fp64, const Penalty& pen, bool recompute, uint32_t which,
                          hipStream_t stream);
void launch_debug_jacobian(const Batch& B, bool fp64, bool transpose, uint32_t which, hipStream_t stream);
void launch_mf_entry_weights(const Batch& B, hipStream_t stream);
void launch_spmv_once(const Batch& B, bool fp64, hipStream_t stream);
void launch_pcg_update_once(const Batch& B, bool fp64, hipStream_t stream);
uint32_t launch_cost(const Batch& B, const Penalty& pen, bool density_in_jtj, hipStream_t stream);
void launch_pcg(const Batch& B, uint32_t iters, uint32_t refine, bool fp64, hipStream_t stream);
void launch_apply_update(const Batch& B, uint32_t k, hipStream_t stream);
void launch_cg_solve(const Batch& B, uint32_t iters, uint32_t refine, uint32_t k, const Penalty& pen,
                     hipStream_t stream);
void launch_cg_iter_once(const Batch& B, hipStream_t stream);
void launch_cg_transform_once(const Batch& B, const Penalty& pen, hipStream_t stream);
uint32_t cg_workgroups_for(uint32_t nodes, uint32_t frames);
uint32_t transform_workgroups_for(uint32_t nodes, uint32_t frames);
uint32_t gram_rows_for(uint32_t nodes, uint32_t frames);
void launch_gram_rows_once(const Batch& B, const Penalty& pen, bool diag_only, hipStream_t stream);
uint32_t eval_workgroups_for(uint32_t nodes, uint32_t frames);
void launch_copy_se3(const Batch& B, hipStream_t stream);
}  // namespace mvdr

using namespace mvdr;

struct mvdr_solver {
    mvdr_solver_config cfg{};
    Buffers b{};
    FrameDev* frame_dev = nullptr;      // device copy of the pointer table
    FrameDev frame_host{};
    std::vector<void*> allocations;
    uint32_t cap_tiles = 0;
    uint32_t cap_chunks = 0;            // Node2Term counting-sort chunks
    Buffers* b_dev = nullptr;           // device copy of b (a batch of one)
    Batch self{};
    bool has_input = false;
    bool indexed = false;
    bool pair_terms = false;            // NodePair2Term lists of the current index present
    // offdiag holds B_ij, diag / diag_inv the damped blocks and their inverses
    // (mixed mode: S_ij and the Gram sums until formed on demand)
    bool pair_blocks = true;
    uint32_t graph_nodes = 0;           // Batch::nodes the captured graph was sized for
    uint64_t graph_table = 0;           // launch-table version the graph was captured with
    uint32_t iter = 0;                  // GN iterations done on the current frame
    hipEvent_t done = nullptr;          // end of the last enqueued work
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t graph_stream = nullptr;
    std::string error;

    ~mvdr_solver() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (done) (void)hipEventDestroy(done);
        for (void* p : allocations) (void)hipFree(p);
    }
};

// One solve lane of a batch: a contiguous run of members whose launches form
// their own chain (and captured graph) on their own stream.  Lanes run
// concurrently, so one lane's latency-bound CG launches overlap another's
// evaluation and assembly.  Every kernel indexes its frame, so a member's
// result does not depend on the lane it runs in.
struct BatchLane {
    Batch batch{};
    uint32_t first = 0, count = 0;  // members [first, first + count)
    uint32_t graph_nodes = 0;
    uint64_t graph_table = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t stream = nullptr;  // lanes 1..: the lane's own stream (lane 0 uses the caller's)
    hipEvent_t joined = nullptr;   // lanes 1..: recorded on the lane's stream after its solve
};

struct mvdr_batch {
    std::vector<mvdr_solver*> members;
    mvdr_solver_config cfg{};
    Buffers* dev = nullptr;             // the members' Buffers, one per frame
    Batch batch{};                      // all members (stage timing)
    std::vector<BatchLane> lanes;
    hipEvent_t fork = nullptr;          // recorded on the caller's stream before the lanes start
    hipEvent_t done = nullptr;
    std::string error;

    void drop_graphs() {
        for (BatchLane& l : lanes) {
            if (l.graph_exec) (void)hipGraphExecDestroy(l.graph_exec);
            if (l.graph) (void)hipGraphDestroy(l.graph);
            l.graph_exec = nullptr;
            l.graph = nullptr;
        }
    }
    void drop_lane_streams() {
        for (BatchLane& l : lanes) {
            if (l.stream) (void)hipStreamSynchronize(l.stream);
            if (l.joined) (void)hipEventDestroy(l.joined);
            if (l.stream) (void)hipStreamDestroy(l.stream);
            l.joined = nullptr;
            l.stream = nullptr;
        }
    }
    ~mvdr_batch() {
        if (done) (void)hipEventSynchronize(done);
        drop_graphs();
        drop_lane_streams();
        if (fork) (void)hipEventDestroy(fork);
        if (done) (void)hipEventDestroy(done);
        if (dev) (void)hipFree(dev);
    }
};

namespace {

mvdr_status fail(mvdr_solver* s, mvdr_status st, const std::string& msg) {
    if (s) s->error = msg;
    return st;
}

#define HIP_TRY(s, call)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return fail((s), MVDR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// setElasticPenaltyValue (SolverIterationData.cpp:104-125) with the constants
// of PenaltyConstants.cpp:9-51.  Iteration 0 uses the reference's default flags
// (SolverIterationData.cpp:49): no foreground, density or cross-view penalty.
Penalty penalty_for(const mvdr_solver_config& c, uint32_t iter) {
    if (!c.use_elastic_penalty) return Penalty{2.5f, 0.0f, 0.0f, 0.7f, 0.1f};
    const bool density = iter > 0 && c.use_density;
    const bool foreground = iter > 0 && c.use_foreground;
    const bool cross = iter > 0 && c.use_cross_view;
    if (iter < c.global_iterations) return Penalty{5.0f, 0.0f, foreground ? 2e-3f : 0.0f, 1.0f, cross ? 1.0f : 0.0f};
    return Penalty{5.0f, density ? 1e-2f : 0.0f, 0.0f, 0.0f, cross ? 1.0f : 0.0f};
}

// The density channel enters J^T J only under USE_DENSE_IMAGE_DENSITY_TERM
// (PreconditionerBlocked.cu:216-219), which the reference build leaves off.
Penalty kernel_penalty(const mvdr_solver_config& c, uint32_t iter) {
    Penalty p = penalty_for(c, iter);
    if (!c.density_in_jtj) p.density = 0.f;
    return p;
}

const char* error_text(uint32_t bits) {
    if (bits & kErrDense) return "dense term count exceeds max_views*max_width*max_height";
    if (bits & kErrFg) return "foreground term count exceeds capacity";
    if (bits & kErrFeat) return "feature term count exceeds max_feature_pairs";
    if (bits & kErrCross) return "cross-view term count exceeds max_cross_pairs";
    if (bits & kErrKv1) return "node-to-term entries exceed capacity";
    if (bits & kErrKv2) return "node-pair term entries exceed capacity";
    if (bits & kErrPairs) return "unique node pairs exceed 64 * max_nodes";
    if (bits & kErrRow) return "a node shares terms with more than 512 other nodes";
    if (bits & kErrNodes) return "an input node / surfel index is out of range";
    return "capacity exceeded";
}

// J^T J applied matrix-free (stored or recomputed Jacobians).
bool matrix_free(const mvdr_solver_config& c) { return c.jtj_mode != MVDR_JTJ_MATERIALIZED; }

// Index build and/or GN iterations [first_iter, first_iter + iters) of every
// frame of the batch B, all sharing the schedule of config c.
void enqueue_batch(const Batch& B, const mvdr_solver_config& c, hipStream_t stream, bool index, uint32_t first_iter,
                   uint32_t iters) {
    const bool fp64 = c.precision == MVDR_PRECISION_MIXED;
    const bool mf = matrix_free(c);
    if (index) {
        launch_copy_se3(B, stream);
        launch_index_build(B, c.use_foreground != 0, c.use_cross_view != 0,
                           mf ? kIndexNodes : (fp64 ? kIndexGram : kIndexLists), stream);
        if (c.jtj_mode == MVDR_JTJ_RECOMPUTE) launch_mf_entry_weights(B, stream);
    }
    for (uint32_t k = first_iter; k < first_iter + iters; ++k) {
        const Penalty pen = kernel_penalty(c, k);
        launch_eval_terms(B, k & 1, fp64, stream);
        if (mf) {
            // diagonal blocks and J^T r from the row Gram assembly; J^T J applied matrix-free
            launch_assemble(B, pen, true, true, stream);
            launch_pcg_matrix_free(B, c.pcg_iterations, c.refine_steps, fp64, pen, c.jtj_mode == MVDR_JTJ_RECOMPUTE,
                                   k & 1, stream);
            launch_apply_update(B, k, stream);
            continue;
        }
        launch_assemble(B, pen, fp64, false, stream);
        if (fp64) {
            launch_cg_solve(B, c.pcg_iterations, c.refine_steps, k, pen, stream);
        } else {
            launch_pcg(B, c.pcg_iterations, 0, false, stream);
            launch_apply_update(B, k, stream);
        }
    }
}

// Grid sizing node count: the frame's node count rounded up to 64.
uint32_t live_nodes(uint32_t n) { return std::max<uint32_t>(64, (n + 63) / 64 * 64); }

// After an assembly: the mixed mode's fused CG leaves S_ij in offdiag and the
// Gram sums in diag (no FP32 inverses).
bool blocks_after_assembly(const mvdr_solver_config& c) {
    return !(c.precision == MVDR_PRECISION_MIXED && !matrix_free(c));
}

void enqueue_solve(mvdr_solver* s, hipStream_t stream, bool index, uint32_t first_iter, uint32_t iters) {
    s->self.nodes = live_nodes(s->frame_host.N);
    if (index) s->pair_terms = s->cfg.precision != MVDR_PRECISION_MIXED && !matrix_free(s->cfg);
    enqueue_batch(s->self, s->cfg, stream, index, first_iter, iters);
}

// Node-centred <-> reference coordinates (see to_node_frame in kernels_solve.hip):
// x_ref = T x_node with T = [[I, 0], [[g]x, I]].  Row-major 6x6 in double.
void node_transform(const float4& g, bool inverse, double T[36]) {
    const double s = inverse ? -1.0 : 1.0;
    for (int q = 0; q < 36; ++q) T[q] = (q % 7 == 0) ? 1.0 : 0.0;
    const double G[3][3] = {{0, -g.z, g.y}, {g.z, 0, -g.x}, {-g.y, g.x, 0}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) T[6 * (3 + r) + c] = s * G[r][c];
}

// out = X^T M Y for row-major 6x6 matrices, M / out in float.
void xt_m_y(const double X[36], const float* M, const double Y[36], float* out) {
    double tmp[36];
    for (int p = 0; p < 6; ++p)
        for (int k = 0; k < 6; ++k) {
            double acc = 0;
            for (int q = 0; q < 6; ++q) acc += (double)M[6 * p + q] * Y[6 * q + k];
            tmp[6 * p + k] = acc;
        }
    for (int a = 0; a < 6; ++a)
        for (int k = 0; k < 6; ++k) {
            double acc = 0;
            for (int p = 0; p < 6; ++p) acc += X[6 * p + a] * tmp[6 * p + k];
            out[6 * a + k] = (float)acc;
        }
}

void transpose6(const double A[36], double At[36]) {
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) At[6 * c + r] = A[6 * r + c];
}

// Map the node-centred system buffers copied from the device back to the
// reference coordinates of JtJDiagonalBlockValue / JtResidualValue.
mvdr_status to_reference_frame(mvdr_solver* s, int tap, float* data, uint32_t pairs) {
    const uint32_t N = s->frame_host.N;
    std::vector<float4> g(N);
    HIP_TRY(s, hipMemcpy(g.data(), s->frame_host.nodes, sizeof(float4) * N, hipMemcpyDeviceToHost));
    double Ti[36], Tinv[36], Tt[36];
    if (tap == MVDR_TAP_SPMV_INPUT || tap == MVDR_TAP_SPMV_OUTPUT) {
        // d_ref = T d', q_ref = T^-T q'
        for (uint32_t n = 0; n < N; ++n) {
            node_transform(g[n], tap == MVDR_TAP_SPMV_OUTPUT, Ti);
            float* v = data + 6 * (size_t)n;
            double out[6];
            for (int a = 0; a < 6; ++a) {
                double acc = 0;
                for (int p = 0; p < 6; ++p)
                    acc += (tap == MVDR_TAP_SPMV_INPUT ? Ti[6 * a + p] : Ti[6 * p + a]) * v[p];
                out[a] = acc;
            }
            for (int a = 0; a < 6; ++a) v[a] = (float)out[a];
        }
        return MVDR_OK;
    }
    if (tap == MVDR_TAP_JTJ_DIAG || tap == MVDR_TAP_JTJ_DIAG_INV || tap == MVDR_TAP_JTR) {
        for (uint32_t n = 0; n < N; ++n) {
            node_transform(g[n], true, Tinv);
            if (tap == MVDR_TAP_JTJ_DIAG) {
                xt_m_y(Tinv, data + 36 * (size_t)n, Tinv, data + 36 * (size_t)n);  // T^-T D' T^-1
            } else if (tap == MVDR_TAP_JTJ_DIAG_INV) {
                node_transform(g[n], false, Ti);
                transpose6(Ti, Tt);
                xt_m_y(Tt, data + 36 * (size_t)n, Tt, data + 36 * (size_t)n);      // T inv(D') T^T
            } else {
                float* bn = data + 6 * (size_t)n;
                double out[6];
                for (int a = 0; a < 6; ++a) {
                    double acc = 0;
                    for (int p = 0; p < 6; ++p) acc += Tinv[6 * p + a] * bn[p];   // T^-T b'
                    out[a] = acc;
                }
                for (int a = 0; a < 6; ++a) bn[a] = (float)out[a];
            }
        }
        return MVDR_OK;
    }
    // off-diagonal: B = T_i^-T B' T_j^-1
    std::vector<uint32_t> key(pairs);
    if (pairs) HIP_TRY(s, hipMemcpy(key.data(), s->b.pair_key, 4ull * pairs, hipMemcpyDeviceToHost));
    double Tj[36];
    for (uint32_t p = 0; p < pairs; ++p) {
        node_transform(g[key[p] >> 16], true, Tinv);
        node_transform(g[key[p] & 0xFFFF], true, Tj);
        xt_m_y(Tinv, data + 36 * (size_t)p, Tj, data + 36 * (size_t)p);
    }
    return MVDR_OK;
}

mvdr_status check_counts(mvdr_solver* s, Counts& out) {
    HIP_TRY(s, hipMemcpy(&out, s->b.counts, sizeof(Counts), hipMemcpyDeviceToHost));
    if (out.error) return fail(s, MVDR_E_CAPACITY, error_text(out.error));
    return MVDR_OK;
}

}  // namespace

extern "C" {

const char* mvdr_status_string(mvdr_status st) {
    switch (st) {
        case MVDR_OK: return "MVDR_OK";
        case MVDR_E_INVALID: return "MVDR_E_INVALID";
        case MVDR_E_SHAPE: return "MVDR_E_SHAPE";
        case MVDR_E_CAPACITY: return "MVDR_E_CAPACITY";
        case MVDR_E_HIP: return "MVDR_E_HIP";
        case MVDR_E_STATE: return "MVDR_E_STATE";
    }
    return "MVDR_E_UNKNOWN";
}

void mvdr_solver_default_config(mvdr_solver_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_views = 3;
    c->max_width = 640;
    c->max_height = 480;
    c->max_nodes = 8192;
    c->max_graph_edges = 8 * 8192;
    c->max_feature_pairs = 3 * 1024;
    c->max_cross_pairs = 4096;
    c->gn_iterations = 5;
    c->pcg_iterations = 10;
    c->global_iterations = 1;
    c->use_elastic_penalty = 1;
    c->use_density = 1;
    c->use_foreground = 0;
    c->use_cross_view = 1;
    c->density_in_jtj = 0;
    c->precision = MVDR_PRECISION_FP32;
    c->jtj_mode = MVDR_JTJ_MATERIALIZED;
    c->refine_steps = 0;
    c->use_graph = 0;
}

const char* mvdr_solver_last_error(const mvdr_solver* s) { return s ? s->error.c_str() : "null solver"; }

// Shape checks of a config that precede any device work.
static bool config_shape_ok(const mvdr_solver_config& c) {
    if (c.max_views < 1 || c.max_views > MVDR_MAX_VIEWS || c.max_width < 1 || c.max_height < 1 || c.max_width > 65535 ||
        c.max_height > 65535 || c.max_nodes < 1 || c.max_nodes > 65535 || c.precision > 1 || c.jtj_mode > 2 ||
        c.pcg_iterations > MVDR_MAX_PCG_ITERATIONS)
        return false;
    const uint64_t pixels = (uint64_t)c.max_views * c.max_width * c.max_height;
    if (pixels > 0x7FFFFFFFull) return false;
    const uint64_t fg = c.use_foreground ? pixels : 0;
    const uint64_t knn_terms = pixels + fg + std::max<uint32_t>(c.max_feature_pairs, 1) + std::max<uint32_t>(c.max_cross_pairs, 1);
    return 6 * knn_terms + std::max<uint32_t>(c.max_graph_edges, 1) <= 0xFFFFFFFFull;
}

// Every per-frame device buffer of config c as 256-byte aligned pieces of one
// block at `base` (base == nullptr: pointers stay null and only the size is
// computed).  Also places the device copies of the frame table and of b.
// Returns the block's size in bytes.
static size_t layout_buffers(const mvdr_solver_config& c, Buffers& b, FrameDev*& frame_dev, Buffers*& b_dev,
                             uint32_t& cap_tiles, uint32_t& cap_chunks, char* base) {
    std::memset(&b, 0, sizeof(b));
    const uint64_t pixels = (uint64_t)c.max_views * c.max_width * c.max_height;
    b.cap_pixels = (uint32_t)pixels;
    b.cap_feat = std::max<uint32_t>(c.max_feature_pairs, 1);
    b.cap_cross = std::max<uint32_t>(c.max_cross_pairs, 1);
    b.cap_edges = std::max<uint32_t>(c.max_graph_edges, 1);
    b.cap_nodes = c.max_nodes;
    const uint64_t fg = c.use_foreground ? pixels : 0;
    const uint64_t knn_terms = pixels + fg + b.cap_feat + b.cap_cross;
    const uint64_t kv1 = 4 * knn_terms + 2ull * b.cap_edges;
    const uint64_t kv2 = 6 * knn_terms + b.cap_edges;
    b.cap_kv1 = (uint32_t)kv1;
    b.cap_kv2 = (uint32_t)kv2;
    b.cap_pairs = (uint32_t)std::min<uint64_t>(kv2, 64ull * c.max_nodes);
    cap_tiles = (uint32_t)((pixels + 4095) / 4096);
    cap_chunks = index_chunks(knn_terms + b.cap_edges);
    const uint32_t N1 = c.max_nodes + 1;

    size_t off = 0;
    auto A = [&](auto*& p, size_t n) {
        using T = std::remove_reference_t<decltype(*p)>;
        off = (off + 255) & ~(size_t)255;
        p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += std::max<size_t>(n, 1) * sizeof(T);
    };
    A(frame_dev, 1);
    A(b_dev, 1);
    A(b.counts, 1);
    A(b.dense_px, b.cap_pixels);
    A(b.dense_knn, b.cap_pixels);
    A(b.dense, b.cap_pixels);
    if (c.density_in_jtj) A(b.density, b.cap_pixels);
    if (c.use_foreground) { A(b.fg_px, b.cap_pixels); A(b.fg, b.cap_pixels); }
    A(b.smooth, b.cap_edges);
    A(b.feat, b.cap_feat);
    A(b.feat_canon, b.cap_feat);
    A(b.cross, b.cap_cross);
    A(b.cross_canon, b.cap_cross);
    A(b.sel_ok, (size_t)b.cap_feat + b.cap_cross);
    A(b.sel_rec, (size_t)b.cap_feat + b.cap_cross);
    A(b.sel_canon, (size_t)b.cap_feat + b.cap_cross);
    A(b.tile_count, cap_tiles);
    A(b.tile_offset, cap_tiles);
    A(b.feat_flag, 1);
    A(b.n2t_hist, (size_t)cap_chunks * c.max_nodes);
    A(b.n2t_count, N1);
    A(b.n2t_offset, N1);
    A(b.n2t_index, b.cap_kv1);
    A(b.n2t_node, b.cap_kv1);
    A(b.row_up, N1);
    A(b.row_low, N1);
    A(b.row_kv2, N1);
    A(b.pair_offset, N1);
    A(b.kv2_offset, N1);
    A(b.pair_key, b.cap_pairs);
    A(b.pair_range, b.cap_pairs);
    A(b.pair_terms, b.cap_kv2);
    A(b.kv2_pair, b.cap_kv2);
    const bool mf = matrix_free(c);
    if (mf) {
        A(b.mf_u, knn_terms + b.cap_edges);
        if (c.jtj_mode == MVDR_JTJ_RECOMPUTE) {
            A(b.mf_g, 2 * (knn_terms + b.cap_edges));
            A(b.n2t_w, b.cap_kv1);
        }
    } else if (c.precision == MVDR_PRECISION_MIXED) {
        A(b.n2t_col, b.cap_kv1);
        A(b.row_part, (size_t)kRowPartners * c.max_nodes);
        A(b.pair_edge_n, b.cap_pairs);
        A(b.pair_edge, (size_t)kPairEdges * b.cap_pairs);
    } else {
        A(b.part_diag, ((size_t)b.cap_kv1 / 64 + 1) * 2 * 27);
        A(b.part_pair, ((size_t)b.cap_kv2 / 64 + 1) * 2 * 36);
    }
    A(b.low_offset, N1);
    A(b.low_ent, b.cap_pairs);
    A(b.diag, 36 * (size_t)c.max_nodes);
    A(b.diag_inv, 36 * (size_t)c.max_nodes);
    // FP64 factors: the mixed mode, and the row Gram assembly of the matrix-free solve
    const bool factors64 = c.precision == MVDR_PRECISION_MIXED || mf;
    if (factors64) A(b.diag_inv64, 36 * (size_t)c.max_nodes);
    A(b.jtr, 6 * (size_t)c.max_nodes);
    A(b.offdiag, 36 * (size_t)b.cap_pairs);
    for (float** v : {&b.x, &b.r, &b.s, &b.q, &b.d0, &b.d1}) A(*v, 6 * (size_t)c.max_nodes);
    if (factors64) A(b.linv64, kTri * (size_t)c.max_nodes);
    if (c.precision == MVDR_PRECISION_MIXED && !mf) {
        A(b.offhat, 36 * (size_t)b.cap_pairs);
        A(b.row_ent, 2 * (size_t)b.cap_pairs);
        A(b.row_ptr, (size_t)c.max_nodes + 1);
        A(b.low_pos, b.cap_pairs);
        A(b.cg_rws[0], kRws * (size_t)c.max_nodes);
        A(b.cg_rws[1], kRws * (size_t)c.max_nodes);
        for (float** v : {&b.cg_p, &b.cg_x, &b.cg_b}) A(*v, 6 * (size_t)c.max_nodes);
        A(b.cg_scal, 2 * 72);
    }
    // per-block sums, or one per 32-row group in the fused CG
    b.partial_stride = std::max<uint32_t>(kPartialStride, (c.max_nodes + 31) / 32);
    A(b.partial, (size_t)kPartialSlots * b.partial_stride);
    A(b.se3[0], c.max_nodes);
    A(b.se3[1], c.max_nodes);
    b.frame = frame_dev;
    return (off + 255) & ~(size_t)255;
}

mvdr_status mvdr_solver_create(const mvdr_solver_config* config, mvdr_solver** out) {
    if (!config || !out) return MVDR_E_INVALID;
    *out = nullptr;
    const mvdr_solver_config& c = *config;
    if (!config_shape_ok(c)) return MVDR_E_SHAPE;
    auto* s = new mvdr_solver();
    s->cfg = c;
    Buffers& b = s->b;
    const size_t bytes = layout_buffers(c, b, s->frame_dev, s->b_dev, s->cap_tiles, s->cap_chunks, nullptr);
    void* block = nullptr;
    hipError_t e = hipMalloc(&block, bytes);
    if (e == hipSuccess) {
        s->allocations.push_back(block);
        layout_buffers(c, b, s->frame_dev, s->b_dev, s->cap_tiles, s->cap_chunks, static_cast<char*>(block));
        e = hipMemset(b.counts, 0, sizeof(Counts));
    }
    if (e == hipSuccess) e = hipMemset(b.partial, 0, (size_t)kPartialSlots * b.partial_stride * sizeof(double));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemcpy(s->b_dev, &b, sizeof(Buffers), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        std::fprintf(stderr, "mvdr_solver_create: %s\n", hipGetErrorString(e));
        delete s;
        return MVDR_E_HIP;
    }
    s->self = Batch{s->b_dev, 1, s->cap_tiles, c.max_nodes, s->cap_chunks};
    *out = s;
    return MVDR_OK;
}

void mvdr_solver_destroy(mvdr_solver* s) {
    if (!s) return;
    if (s->done) (void)hipEventSynchronize(s->done);
    delete s;
}

// Validates a frame against config c (SetSolverInput's checks) and fills the
// device pointer table f.  On failure `err` holds the message.
static mvdr_status frame_from_input(const mvdr_solver_config& c, const mvdr_frame_input* in, FrameDev& f,
                                    const char*& err) {
    auto fail = [&](mvdr_status st, const char* msg) { err = msg; return st; };
    if (in->num_views < 1 || in->num_views > c.max_views || in->width < 1 || in->height < 1 || in->width > c.max_width ||
        in->height > c.max_height)
        return fail(MVDR_E_SHAPE, "views / image size outside the configured maximum");
    if (in->num_nodes < 1 || in->num_nodes > c.max_nodes)
        return fail(MVDR_E_SHAPE, "num_nodes outside [1, max_nodes]");
    if (in->num_graph_edges > c.max_graph_edges) return fail(MVDR_E_SHAPE, "num_graph_edges > max_graph_edges");
    uint64_t feat_total = 0;
    for (uint32_t v = 0; v < in->num_views; ++v) feat_total += in->num_feature_pairs[v];
    if (feat_total > c.max_feature_pairs) return fail(MVDR_E_SHAPE, "feature pairs > max_feature_pairs");
    if (in->num_cross_pairs > c.max_cross_pairs) return fail(MVDR_E_SHAPE, "cross pairs > max_cross_pairs");
    if (!in->surfel_knn || !in->surfel_knn_weight || !in->node_se3 || !in->canonical_nodes || in->num_surfels == 0)
        return fail(MVDR_E_INVALID, "surfel knn / node arrays must be set");
    if (in->num_graph_edges && !in->node_graph) return fail(MVDR_E_INVALID, "node_graph is null");
    if (in->num_cross_pairs && c.use_cross_view && !in->cross_pairs) return fail(MVDR_E_INVALID, "cross_pairs is null");
    for (uint32_t v = 0; v < in->num_views; ++v) {
        if (!in->vertex_confid[v] || !in->normal_radius[v] || !in->reference_vertex[v] || !in->reference_normal[v] ||
            !in->warp_vertex[v] || !in->index_map[v])
            return fail(MVDR_E_INVALID, "a required per-view map is null");
        if ((in->num_feature_pairs[v] || (in->num_cross_pairs && c.use_cross_view)) && !in->edge_mask[v])
            return fail(MVDR_E_INVALID, "edge_mask is required with feature or cross-view pairs");
        if (in->num_feature_pairs[v] && !in->feature_pairs[v]) return fail(MVDR_E_INVALID, "feature_pairs is null");
        if (c.use_foreground && (!in->filtered_foreground[v] || !in->foreground_gradient[v]))
            return fail(MVDR_E_INVALID, "foreground maps are required when use_foreground is set");
        if (c.density_in_jtj && (!in->gray[v] || !in->gray_gradient[v] || !in->normalized_rgb[v]))
            return fail(MVDR_E_INVALID, "gray / rgb maps are required when density_in_jtj is set");
    }
    std::memset(&f, 0, sizeof(f));
    f.V = in->num_views;
    f.W = in->width;
    f.H = in->height;
    f.N = in->num_nodes;
    f.E = in->num_graph_edges;
    f.num_surfels = in->num_surfels;
    f.feat_base[0] = 0;
    for (uint32_t v = 0; v < MVDR_MAX_VIEWS; ++v) {
        f.num_feat[v] = v < in->num_views ? in->num_feature_pairs[v] : 0;
        f.feat_base[v + 1] = f.feat_base[v] + f.num_feat[v];
    }
    f.num_cross = c.use_cross_view ? in->num_cross_pairs : 0;
    for (uint32_t v = 0; v < in->num_views; ++v) {
        f.fx[v] = in->intrinsic[v].fx;
        f.fy[v] = in->intrinsic[v].fy;
        f.cx[v] = in->intrinsic[v].cx;
        f.cy[v] = in->intrinsic[v].cy;
        f.init[v] = to_mat34(in->initial_camera_se3[v]);
        f.init_inv[v] = inverse(f.init[v]);
        f.w2c[v] = to_mat34(in->world2camera[v]);
        f.c2w[v] = inverse(f.w2c[v]);
        f.vertex[v] = static_cast<const float4*>(in->vertex_confid[v]);
        f.normal[v] = static_cast<const float4*>(in->normal_radius[v]);
        f.ffg[v] = static_cast<const float*>(in->filtered_foreground[v]);
        f.fg_grad[v] = static_cast<const float2*>(in->foreground_gradient[v]);
        f.gray[v] = static_cast<const float*>(in->gray[v]);
        f.gray_grad[v] = static_cast<const float2*>(in->gray_gradient[v]);
        f.edge[v] = static_cast<const uint8_t*>(in->edge_mask[v]);
        f.ref_v[v] = static_cast<const float4*>(in->reference_vertex[v]);
        f.ref_n[v] = static_cast<const float4*>(in->reference_normal[v]);
        f.warp_v[v] = static_cast<const float4*>(in->warp_vertex[v]);
        f.index[v] = static_cast<const uint32_t*>(in->index_map[v]);
        f.rgb[v] = static_cast<const float4*>(in->normalized_rgb[v]);
        f.feat[v] = static_cast<const ushort4*>(in->feature_pairs[v]);
        f.corr[v] = static_cast<const ushort4*>(in->corr_map[v]);
    }
    f.surfel_knn = static_cast<const ushort4*>(in->surfel_knn);
    f.surfel_w = static_cast<const float4*>(in->surfel_knn_weight);
    f.node_se3_in = static_cast<const DQ*>(in->node_se3);
    f.nodes = static_cast<const float4*>(in->canonical_nodes);
    f.graph = static_cast<const ushort2*>(in->node_graph);
    f.cross = static_cast<const mvdr_cross_pair*>(in->cross_pairs);
    return MVDR_OK;
}


mvdr_status mvdr_solver_set_input(mvdr_solver* s, const mvdr_frame_input* in) {
    if (!s || !in) return MVDR_E_INVALID;
    FrameDev f;
    const char* err = nullptr;
    const mvdr_status rc = frame_from_input(s->cfg, in, f, err);
    if (rc != MVDR_OK) return fail(s, rc, err);
    s->frame_host = f;
    // the previous solve may still read the pointer table
    HIP_TRY(s, hipEventSynchronize(s->done));
    HIP_TRY(s, hipMemcpy(s->frame_dev, &f, sizeof(FrameDev), hipMemcpyHostToDevice));
    s->has_input = true;
    s->indexed = false;
    s->iter = 0;
    return MVDR_OK;
}

mvdr_status mvdr_solver_build_index(mvdr_solver* s, void* stream) {
    if (!s) return MVDR_E_INVALID;
    if (!s->has_input) return fail(s, MVDR_E_STATE, "build_index before set_input");
    hipStream_t st = static_cast<hipStream_t>(stream);
    enqueue_solve(s, st, true, 0, 0);
    HIP_TRY(s, hipGetLastError());
    HIP_TRY(s, hipEventRecord(s->done, st));
    s->indexed = true;
    s->iter = 0;
    return MVDR_OK;
}

mvdr_status mvdr_solver_gn_iteration(mvdr_solver* s, void* stream) {
    if (!s) return MVDR_E_INVALID;
    if (!s->indexed) return fail(s, MVDR_E_STATE, "gn_iteration before build_index");
    hipStream_t st = static_cast<hipStream_t>(stream);
    enqueue_solve(s, st, false, s->iter, 1);
    HIP_TRY(s, hipGetLastError());
    HIP_TRY(s, hipEventRecord(s->done, st));
    ++s->iter;
    s->pair_blocks = blocks_after_assembly(s->cfg);
    return MVDR_OK;
}

mvdr_status mvdr_solver_solve(mvdr_solver* s, void* stream) {
    if (!s) return MVDR_E_INVALID;
    if (!s->has_input) return fail(s, MVDR_E_STATE, "solve before set_input");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t gn = s->cfg.gn_iterations;
    if (s->cfg.use_graph) {
        // The pointer table and every buffer are device-resident, so one
        // captured graph serves every frame of this solver whose node count
        // rounds to the same grid size; another is captured when it, or the
        // launch table, changes.
        if (s->graph_exec &&
            (s->graph_nodes != live_nodes(s->frame_host.N) || s->graph_table != launch_table_version())) {
            (void)hipGraphExecDestroy(s->graph_exec);
            (void)hipGraphDestroy(s->graph);
            s->graph_exec = nullptr;
            s->graph = nullptr;
        }
        if (!s->graph_exec) {
            s->graph_nodes = live_nodes(s->frame_host.N);
            s->graph_table = launch_table_version();
            hipStream_t cap = nullptr;
            HIP_TRY(s, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                enqueue_solve(s, cap, true, 0, gn);
                e = hipStreamEndCapture(cap, &s->graph);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0);
            (void)hipStreamDestroy(cap);
            if (e != hipSuccess) return fail(s, MVDR_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        }
        HIP_TRY(s, hipGraphLaunch(s->graph_exec, st));
        s->pair_terms = s->cfg.precision != MVDR_PRECISION_MIXED && !matrix_free(s->cfg);
    } else {
        enqueue_solve(s, st, true, 0, gn);
        HIP_TRY(s, hipGetLastError());
    }
    HIP_TRY(s, hipEventRecord(s->done, st));
    s->indexed = true;
    s->iter = gn;
    s->pair_blocks = gn == 0 || blocks_after_assembly(s->cfg);
    return MVDR_OK;
}

mvdr_status mvdr_solver_get_node_se3(mvdr_solver* s, const void** dq, uint32_t* n) {
    if (!s || !dq || !n) return MVDR_E_INVALID;
    if (!s->has_input) return fail(s, MVDR_E_STATE, "no input");
    *dq = s->indexed ? (const void*)s->b.se3[s->iter & 1] : (const void*)s->frame_host.node_se3_in;
    *n = s->frame_host.N;
    return MVDR_OK;
}

mvdr_status mvdr_solver_total_residual(mvdr_solver* s, void* stream, double* cost) {
    if (!s || !cost) return MVDR_E_INVALID;
    if (!s->indexed) return fail(s, MVDR_E_STATE, "total_residual before build_index");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const Penalty pen = penalty_for(s->cfg, s->iter);
    if (!s->pair_blocks && s->iter > 0) {
        // The mixed mode forms B_ij on demand from S_ij and the graph-edge rows
        // of the smooth records, which the evaluation below overwrites at the
        // new warp: form the blocks of the last iteration first.
        s->self.nodes = live_nodes(s->frame_host.N);
        launch_pair_blocks(s->self, kernel_penalty(s->cfg, s->iter - 1), st);
        launch_diag_blocks(s->self, st);
        s->pair_blocks = true;
    }
    launch_eval_terms(s->self, s->iter & 1, s->cfg.precision == MVDR_PRECISION_MIXED, st);
    Penalty kp = pen;
    if (!s->cfg.density_in_jtj) kp.density = 0.f;
    const uint32_t parts = launch_cost(s->self, kp, s->cfg.density_in_jtj != 0, st);
    HIP_TRY(s, hipGetLastError());
    HIP_TRY(s, hipEventRecord(s->done, st));
    HIP_TRY(s, hipStreamSynchronize(st));
    Counts cnt;
    mvdr_status rc = check_counts(s, cnt);
    if (rc != MVDR_OK) return rc;
    std::vector<double> partials(parts);
    HIP_TRY(s, hipMemcpy(partials.data(), s->b.partial + (size_t)kSlotCost * s->b.partial_stride, parts * sizeof(double),
                         hipMemcpyDeviceToHost));
    double total = 0.0;
    for (double p : partials) total += p;
    *cost = total;
    return MVDR_OK;
}

mvdr_status mvdr_solver_get_sizes(mvdr_solver* s, mvdr_solver_sizes* out) {
    if (!s || !out) return MVDR_E_INVALID;
    HIP_TRY(s, hipEventSynchronize(s->done));
    Counts c;
    HIP_TRY(s, hipMemcpy(&c, s->b.counts, sizeof(Counts), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->num_dense_terms = c.dense;
    out->num_smooth_terms = c.smooth;
    out->num_foreground_terms = c.fg;
    out->num_feature_terms = c.feat;
    out->num_cross_terms = c.cross;
    out->num_node2term = c.kv1;
    out->num_nodepair_terms = c.kv2;
    out->num_node_pairs = c.pairs;
    out->num_nodes = s->frame_host.N;
    out->gn_iterations_done = s->iter;
    if (c.error) return fail(s, MVDR_E_CAPACITY, error_text(c.error));
    return MVDR_OK;
}

mvdr_status mvdr_solver_debug_get(mvdr_solver* s, int tap, void* dst, size_t bytes, size_t* needed) {
    if (!s) return MVDR_E_INVALID;
    if (!s->has_input) return fail(s, MVDR_E_STATE, "no input");
    HIP_TRY(s, hipEventSynchronize(s->done));
    if ((tap == MVDR_TAP_NODEPAIR_RANGE || tap == MVDR_TAP_NODEPAIR_TERM_INDEX) && s->indexed && !s->pair_terms &&
        !matrix_free(s->cfg)) {
        // the row Gram assembly needs no NodePair2Term lists: build them for inspection
        s->self.nodes = live_nodes(s->frame_host.N);
        launch_pair_term_lists(s->self, nullptr);
        HIP_TRY(s, hipGetLastError());
        HIP_TRY(s, hipStreamSynchronize(nullptr));
        s->pair_terms = true;
    }
    const bool block_tap = tap == MVDR_TAP_JTJ_OFFDIAG || tap == MVDR_TAP_JTJ_DIAG || tap == MVDR_TAP_JTJ_DIAG_INV;
    if (block_tap && !s->pair_blocks && s->iter > 0) {
        s->self.nodes = live_nodes(s->frame_host.N);
        launch_pair_blocks(s->self, kernel_penalty(s->cfg, s->iter - 1), nullptr);
        launch_diag_blocks(s->self, nullptr);
        HIP_TRY(s, hipGetLastError());
        HIP_TRY(s, hipStreamSynchronize(nullptr));
        s->pair_blocks = true;
    }
    Counts c;
    HIP_TRY(s, hipMemcpy(&c, s->b.counts, sizeof(Counts), hipMemcpyDeviceToHost));
    const Buffers& b = s->b;
    const uint32_t N = s->frame_host.N;
    const void* src = nullptr;
    size_t size = 0;
    float pen[5];
    switch (tap) {
        case MVDR_TAP_NODE2TERM_OFFSET: src = b.n2t_offset; size = 4ull * (N + 1); break;
        case MVDR_TAP_NODE2TERM_INDEX: src = b.n2t_index; size = 4ull * c.kv1; break;
        case MVDR_TAP_NODEPAIR_KEY: src = b.pair_key; size = 4ull * c.pairs; break;
        case MVDR_TAP_NODEPAIR_RANGE: src = b.pair_range; size = 8ull * c.pairs; break;
        case MVDR_TAP_NODEPAIR_TERM_INDEX: src = b.pair_terms; size = 4ull * c.kv2; break;
        case MVDR_TAP_DENSE_TERMS: src = b.dense; size = sizeof(ScalarRec) * (size_t)c.dense; break;
        case MVDR_TAP_SMOOTH_TERMS: src = b.smooth; size = sizeof(SmoothRec) * (size_t)c.smooth; break;
        case MVDR_TAP_FEATURE_TERMS: src = b.feat; size = sizeof(P2PRec) * (size_t)c.feat; break;
        case MVDR_TAP_CROSS_TERMS: src = b.cross; size = sizeof(P2PRec) * (size_t)c.cross; break;
        case MVDR_TAP_FOREGROUND_TERMS: src = b.fg; size = b.fg ? sizeof(ScalarRec) * (size_t)c.fg : 0; break;
        case MVDR_TAP_DENSE_PIXELS: src = b.dense_px; size = 8ull * c.dense; break;
        case MVDR_TAP_JTJ_DIAG: src = b.diag; size = 144ull * N; break;
        case MVDR_TAP_JTJ_DIAG_INV: src = b.diag_inv; size = 144ull * N; break;
        case MVDR_TAP_JTR: src = b.jtr; size = 24ull * N; break;
        case MVDR_TAP_JTJ_OFFDIAG: src = b.offdiag; size = 144ull * c.pairs; break;
        case MVDR_TAP_TWIST: src = b.x; size = 24ull * N; break;
        case MVDR_TAP_NODE_SE3: {
            const void* p;
            uint32_t n;
            mvdr_solver_get_node_se3(s, &p, &n);
            src = p;
            size = 32ull * n;
            break;
        }
        case MVDR_TAP_PENALTY: {
            // constants of the most recent iteration (iteration 0 before any)
            const Penalty p = penalty_for(s->cfg, s->iter > 0 ? s->iter - 1 : 0);
            pen[0] = p.smooth; pen[1] = p.density; pen[2] = p.foreground; pen[3] = p.feature; pen[4] = p.cross;
            if (needed) *needed = sizeof(pen);
            if (!dst || bytes < sizeof(pen)) return fail(s, MVDR_E_CAPACITY, "buffer too small");
            std::memcpy(dst, pen, sizeof(pen));
            return MVDR_OK;
        }
        case MVDR_TAP_DENSITY_TERMS: src = b.density; size = b.density ? sizeof(ScalarRec) * (size_t)c.dense : 0; break;
        case MVDR_TAP_SPMV_INPUT: src = b.d1; size = 24ull * N; break;
        case MVDR_TAP_SPMV_OUTPUT: src = b.q; size = 24ull * N; break;
        case MVDR_TAP_FOREGROUND_PIXELS: src = b.fg_px; size = b.fg_px ? 8ull * c.fg : 0; break;
        default: return fail(s, MVDR_E_INVALID, "unknown tap");
    }
    if (needed) *needed = size;
    if (size && (!dst || bytes < size)) return fail(s, MVDR_E_CAPACITY, "buffer too small");
    if (size) HIP_TRY(s, hipMemcpy(dst, src, size, hipMemcpyDeviceToHost));
    if (size && (tap == MVDR_TAP_JTJ_DIAG || tap == MVDR_TAP_JTJ_DIAG_INV || tap == MVDR_TAP_JTR ||
                 tap == MVDR_TAP_JTJ_OFFDIAG || tap == MVDR_TAP_SPMV_INPUT || tap == MVDR_TAP_SPMV_OUTPUT)) {
        const mvdr_status rc = to_reference_frame(s, tap, static_cast<float*>(dst), c.pairs);
        if (rc != MVDR_OK) return rc;
    }
    if (c.error) return fail(s, MVDR_E_CAPACITY, error_text(c.error));
    return MVDR_OK;
}

mvdr_status mvdr_solver_debug_jacobian(mvdr_solver* s, int transpose, const void* in, void* out, void* stream) {
    if (!s || !in || !out || transpose < 0 || transpose > 1) return MVDR_E_INVALID;
    if (s->cfg.jtj_mode != MVDR_JTJ_RECOMPUTE) return fail(s, MVDR_E_INVALID, "debug_jacobian needs jtj_mode RECOMPUTE");
    if (!s->indexed) return fail(s, MVDR_E_STATE, "debug_jacobian needs a built index");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(s, hipEventSynchronize(s->done));
    Counts c;
    HIP_TRY(s, hipMemcpy(&c, s->b.counts, sizeof(Counts), hipMemcpyDeviceToHost));
    if (c.error) return fail(s, MVDR_E_CAPACITY, error_text(c.error));
    const size_t twists = 24ull * s->frame_host.N, rows = 16ull * c.off[4];
    s->self.nodes = live_nodes(s->frame_host.N);
    HIP_TRY(s, hipMemcpyAsync(transpose ? (void*)s->b.mf_u : (void*)s->b.d1, in, transpose ? rows : twists,
                              hipMemcpyDeviceToDevice, st));
    launch_debug_jacobian(s->self, s->cfg.precision == MVDR_PRECISION_MIXED, transpose != 0, s->iter & 1, st);
    HIP_TRY(s, hipGetLastError());
    HIP_TRY(s, hipMemcpyAsync(out, transpose ? (const void*)s->b.q : (const void*)s->b.mf_u, transpose ? twists : rows,
                              hipMemcpyDeviceToDevice, st));
    HIP_TRY(s, hipStreamSynchronize(st));
    return MVDR_OK;
}

// Stage micro-benchmark over a batch (a solver is a batch of one): `reps`
// launches of one stage on the data of GN iteration k, timed with hipEvents
// on `st`.
static mvdr_status bench_stage_impl(const Batch& B, const mvdr_solver_config& c, uint32_t k, int stage, uint32_t reps,
                                    hipStream_t st, float* avg_ms, std::string& err, bool& pair_blocks) {
    const bool fp64 = c.precision == MVDR_PRECISION_MIXED;
    const bool mf = matrix_free(c);
    if (stage < MVDR_STAGE_EVAL_TERMS || stage > MVDR_STAGE_CG_ITERATION) {
        err = "unknown stage";
        return MVDR_E_INVALID;
    }
    if (stage == MVDR_STAGE_CG_ITERATION && (!fp64 || mf)) {
        err = "the fused CG iteration runs in the mixed precision mode with materialized J^T J";
        return MVDR_E_INVALID;
    }
    const Penalty pen = kernel_penalty(c, k);
    if (stage == MVDR_STAGE_SPMV && !mf && !pair_blocks) {
        launch_pair_blocks(B, pen, st);  // the materialized SpMV reads B_ij and the damped D_i
        launch_diag_blocks(B, st);
        pair_blocks = true;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
        switch (stage) {
            case MVDR_STAGE_EVAL_TERMS: launch_eval_terms(B, k & 1, fp64, st); break;
            case MVDR_STAGE_ASSEMBLE: launch_assemble(B, pen, fp64 || mf, mf, st); break;
            case MVDR_STAGE_PCG_UPDATE: launch_pcg_update_once(B, fp64, st); break;
            case MVDR_STAGE_SPMV:
                if (mf) launch_mf_apply_once(B, fp64, pen, c.jtj_mode == MVDR_JTJ_RECOMPUTE, k & 1, st);
                else launch_spmv_once(B, fp64, st);
                break;
            case MVDR_STAGE_INDEX: enqueue_batch(B, c, st, true, 0, 0); break;
            default: launch_cg_iter_once(B, st); break;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        err = std::string("HIP: ") + hipGetErrorString(e);
        return MVDR_E_HIP;
    }
    if (stage == MVDR_STAGE_ASSEMBLE) pair_blocks = blocks_after_assembly(c);
    *avg_ms = ms / (float)reps;
    return MVDR_OK;
}

mvdr_status mvdr_solver_bench_stage(mvdr_solver* s, int stage, uint32_t reps, void* stream, float* avg_ms) {
    if (!s || !avg_ms || reps == 0) return MVDR_E_INVALID;
    if (!s->indexed || s->iter == 0) return fail(s, MVDR_E_STATE, "bench_stage needs a completed solve");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const mvdr_status rc = bench_stage_impl(s->self, s->cfg, s->iter - 1, stage, reps, st, avg_ms, s->error, s->pair_blocks);
    if (rc == MVDR_OK) HIP_TRY(s, hipEventRecord(s->done, st));
    return rc;
}

// ------------------------------------------------------------------ batches
// Members must agree on everything that shapes the launch sequence; their
// capacities may differ (grids take the largest, kernels their own sizes).
static bool same_schedule(const mvdr_solver_config& a, const mvdr_solver_config& b) {
    return a.gn_iterations == b.gn_iterations && a.pcg_iterations == b.pcg_iterations &&
           a.global_iterations == b.global_iterations && a.use_elastic_penalty == b.use_elastic_penalty &&
           a.use_density == b.use_density && a.use_foreground == b.use_foreground &&
           a.use_cross_view == b.use_cross_view && a.density_in_jtj == b.density_in_jtj &&
           a.precision == b.precision && a.jtj_mode == b.jtj_mode && a.refine_steps == b.refine_steps;
}

}  // extern "C"

namespace {

// Lanes of a new batch: two from 8 members on (each lane then still fills the
// device in its evaluation and assembly launches), else one.
uint32_t default_lanes(uint32_t members) { return members >= 8 ? 2u : 1u; }

// Splits the members into `n` contiguous lanes of near-equal size, each with
// its own stream and join event (lane 0 runs on the caller's stream).
hipError_t make_lanes(mvdr_batch* bt, uint32_t n) {
    bt->drop_graphs();
    bt->drop_lane_streams();
    const uint32_t count = (uint32_t)bt->members.size();
    bt->lanes.assign(n, BatchLane{});
    hipError_t e = hipSuccess;
    for (uint32_t l = 0; l < n; ++l) {
        BatchLane& lane = bt->lanes[l];
        lane.first = (uint32_t)((uint64_t)count * l / n);
        lane.count = (uint32_t)((uint64_t)count * (l + 1) / n) - lane.first;
        Batch B{bt->dev + lane.first, lane.count, 0, 64, 0};
        for (uint32_t k = lane.first; k < lane.first + lane.count; ++k) {
            B.cap_tiles = std::max(B.cap_tiles, bt->members[k]->self.cap_tiles);
            B.cap_chunks = std::max(B.cap_chunks, bt->members[k]->self.cap_chunks);
        }
        lane.batch = B;
        if (l > 0 && e == hipSuccess) {
            e = hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&lane.joined, hipEventDisableTiming);
        }
    }
    return e;
}

}  // namespace

extern "C" {

mvdr_status mvdr_batch_create(mvdr_solver* const* solvers, uint32_t count, mvdr_batch** out) {
    if (!out) return MVDR_E_INVALID;
    *out = nullptr;
    if (!solvers || count == 0 || count > 65535) return MVDR_E_INVALID;
    for (uint32_t k = 0; k < count; ++k) {
        if (!solvers[k]) return MVDR_E_INVALID;
        if (!same_schedule(solvers[k]->cfg, solvers[0]->cfg)) return MVDR_E_SHAPE;
        for (uint32_t q = 0; q < k; ++q)
            if (solvers[q] == solvers[k]) return MVDR_E_INVALID;
    }
    auto* bt = new mvdr_batch();
    bt->members.assign(solvers, solvers + count);
    bt->cfg = solvers[0]->cfg;
    std::vector<Buffers> host(count);
    Batch B{nullptr, count, 0, 64, 0};
    for (uint32_t k = 0; k < count; ++k) {
        const mvdr_solver* m = solvers[k];
        host[k] = m->b;
        B.cap_tiles = std::max(B.cap_tiles, m->self.cap_tiles);
        B.cap_chunks = std::max(B.cap_chunks, m->self.cap_chunks);
    }
    hipError_t e = hipMalloc(&bt->dev, sizeof(Buffers) * count);
    if (e == hipSuccess) e = hipMemcpy(bt->dev, host.data(), sizeof(Buffers) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->fork, hipEventDisableTiming);
    if (e == hipSuccess) {
        B.dev = bt->dev;
        bt->batch = B;
        e = make_lanes(bt, default_lanes(count));
    }
    if (e != hipSuccess) {
        std::fprintf(stderr, "mvdr_batch_create: %s\n", hipGetErrorString(e));
        delete bt;
        return MVDR_E_HIP;
    }
    *out = bt;
    return MVDR_OK;
}

mvdr_status mvdr_batch_set_lanes(mvdr_batch* bt, uint32_t lanes) {
    if (!bt || lanes == 0 || lanes > MVDR_MAX_BATCH_LANES || lanes > bt->members.size()) return MVDR_E_INVALID;
    if (lanes == bt->lanes.size()) return MVDR_OK;
    (void)hipEventSynchronize(bt->done);
    const hipError_t e = make_lanes(bt, lanes);
    if (e != hipSuccess) {
        bt->error = std::string("HIP: ") + hipGetErrorString(e);
        return MVDR_E_HIP;
    }
    return MVDR_OK;
}

uint32_t mvdr_batch_lanes(const mvdr_batch* bt) { return bt ? (uint32_t)bt->lanes.size() : 0u; }

void mvdr_batch_destroy(mvdr_batch* bt) { delete bt; }

mvdr_status mvdr_batch_bench_stage(mvdr_batch* bt, int stage, uint32_t reps, void* stream, float* avg_ms) {
    if (!bt || !avg_ms || reps == 0) return MVDR_E_INVALID;
    const mvdr_solver* m0 = bt->members[0];
    for (const mvdr_solver* m : bt->members)
        if (!m->indexed || m->iter != m0->iter || m->iter == 0) {
            bt->error = "bench_stage needs a completed batch solve";
            return MVDR_E_STATE;
        }
    hipStream_t st = static_cast<hipStream_t>(stream);
    bool pair_blocks = m0->pair_blocks;
    const mvdr_status rc = bench_stage_impl(bt->batch, bt->cfg, m0->iter - 1, stage, reps, st, avg_ms, bt->error, pair_blocks);
    for (mvdr_solver* m : bt->members) m->pair_blocks = pair_blocks;
    if (rc == MVDR_OK)
        for (mvdr_solver* m : bt->members) (void)hipEventRecord(m->done, st);
    return rc;
}

// Launch-shape tuning (tuning.h): for the batch's shape each tunable launch is
// timed over its candidates on the data of GN iteration k (the stages of a
// completed solve, as bench_stage) and the fastest is kept in the process-wide
// launch table.  The parameters spread work over workgroups and change no
// result; the timed launches rewrite only what they would write in a solve.
static mvdr_status autotune_impl(const Batch& B, const mvdr_solver_config& c, uint32_t k, uint32_t reps, hipStream_t st,
                                 mvdr_launch_params* chosen, std::string& err) {
    const bool fp64 = c.precision == MVDR_PRECISION_MIXED;
    const bool fused = fp64 && !matrix_free(c);
    const Penalty pen = kernel_penalty(c, k);
    LaunchParams p = launch_params(B.nodes, B.frames);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto timed = [&](auto&& launch) {
        float ms = 0.f;
        if (e != hipSuccess) return ms;
        launch();  // warm-up
        e = hipEventRecord(e0, st);
        for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) launch();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess) e = hipGetLastError();
        return ms;
    };
    // term evaluation: workgroups per frame of its grid-stride loop
    {
        float best = 0.f;
        uint32_t arg = 0;
        for (uint32_t cand : {256u, 512u, 1024u, 2048u}) {
            p.eval_workgroups = cand;
            set_launch_params(B.nodes, B.frames, p);
            const float t = timed([&] { launch_eval_terms(B, k & 1, fp64, st); });
            if (arg == 0 || t < best) { best = t; arg = cand; }
        }
        p.eval_workgroups = arg;
    }
    if (fused) {
        // fused CG row kernels: one pass of rows per workgroup, or at most
        // 128 / 64 / 32 workgroups per frame (the resolved counts, each once)
        const uint32_t g = (B.nodes + 31) / 32;
        float best = 0.f;
        uint32_t arg = 0, seen[4] = {0, 0, 0, 0}, n_seen = 0;
        for (uint32_t cand : {g, 128u, 64u, 32u}) {
            p.cg_workgroups = cand;
            set_launch_params(B.nodes, B.frames, p);
            const uint32_t resolved = cg_workgroups_for(B.nodes, B.frames);
            bool dup = false;
            for (uint32_t q = 0; q < n_seen; ++q) dup = dup || seen[q] == resolved;
            if (dup) continue;
            seen[n_seen++] = resolved;
            const float t = timed([&] { launch_cg_iter_once(B, st); });
            if (arg == 0 || t < best) { best = t; arg = cand; }
        }
        p.cg_workgroups = arg;
        // CG transform: workgroups per frame of its grid-stride loop
        best = 0.f;
        arg = 0;
        for (uint32_t cand : {256u, 512u, 1024u, 2048u}) {
            p.transform_workgroups = cand;
            set_launch_params(B.nodes, B.frames, p);
            const float t = timed([&] { launch_cg_transform_once(B, pen, st); });
            if (arg == 0 || t < best) { best = t; arg = cand; }
        }
        p.transform_workgroups = arg;
    }
    if (fp64) {
        // row Gram assembly: rows per row wave (the diagonal-only form for the
        // matrix-free solve)
        float best = 0.f;
        uint32_t arg = 0;
        for (uint32_t cand : {1u, 2u, 4u}) {
            p.gram_rows = cand;
            set_launch_params(B.nodes, B.frames, p);
            const float t = timed([&] { launch_gram_rows_once(B, pen, !fused, st); });
            if (arg == 0 || t < best) { best = t; arg = cand; }
        }
        p.gram_rows = arg;
    }
    set_launch_params(B.nodes, B.frames, p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        err = std::string("HIP: ") + hipGetErrorString(e);
        return MVDR_E_HIP;
    }
    if (chosen) {
        chosen->cg_workgroups = fused ? cg_workgroups_for(B.nodes, B.frames) : 0;
        chosen->eval_workgroups = eval_workgroups_for(B.nodes, B.frames);
        chosen->transform_workgroups = fused ? transform_workgroups_for(B.nodes, B.frames) : 0;
        chosen->gram_rows = fp64 ? gram_rows_for(B.nodes, B.frames) : 0;
    }
    return MVDR_OK;
}

mvdr_status mvdr_solver_autotune(mvdr_solver* s, uint32_t reps, void* stream, mvdr_launch_params* chosen) {
    if (!s || reps == 0) return MVDR_E_INVALID;
    if (!s->indexed || s->iter == 0) return fail(s, MVDR_E_STATE, "autotune needs a completed solve");
    hipStream_t st = static_cast<hipStream_t>(stream);
    s->self.nodes = live_nodes(s->frame_host.N);
    const mvdr_status rc = autotune_impl(s->self, s->cfg, s->iter - 1, reps, st, chosen, s->error);
    if (rc == MVDR_OK) HIP_TRY(s, hipEventRecord(s->done, st));
    return rc;
}

mvdr_status mvdr_batch_autotune(mvdr_batch* bt, uint32_t reps, void* stream, mvdr_launch_params* chosen) {
    if (!bt || reps == 0) return MVDR_E_INVALID;
    const mvdr_solver* m0 = bt->members[0];
    for (const mvdr_solver* m : bt->members)
        if (!m->indexed || m->iter != m0->iter || m->iter == 0) {
            bt->error = "autotune needs a completed batch solve";
            return MVDR_E_STATE;
        }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t k = m0->iter - 1;
    // launch shapes of the current lanes: one per lane (frames per lane, the
    // lane's node count)
    auto tune_shapes = [&](mvdr_launch_params* out) {
        mvdr_status rc = MVDR_OK;
        for (size_t l = 0; l < bt->lanes.size() && rc == MVDR_OK; ++l) {
            Batch B = bt->lanes[l].batch;
            uint32_t nodes = 64;
            for (uint32_t q = bt->lanes[l].first; q < bt->lanes[l].first + bt->lanes[l].count; ++q)
                nodes = std::max(nodes, live_nodes(bt->members[q]->frame_host.N));
            B.nodes = nodes;
            mvdr_launch_params got{};
            rc = autotune_impl(B, bt->cfg, k, reps, st, l == 0 && out ? out : &got, bt->error);
        }
        return rc;
    };
    // then the lane count: whole solves (same inputs, same results) timed for
    // each candidate with its tuned shapes
    const uint32_t members = (uint32_t)bt->members.size();
    uint32_t best = (uint32_t)bt->lanes.size();
    if (members > 1) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        float best_ms = 0.f;
        for (uint32_t L = 1; L <= std::min<uint32_t>(members, MVDR_MAX_BATCH_LANES) && e == hipSuccess; ++L) {
            (void)hipEventSynchronize(bt->done);
            e = make_lanes(bt, L);
            if (e != hipSuccess) break;
            mvdr_status rc = tune_shapes(nullptr);
            if (rc == MVDR_OK) rc = mvdr_batch_solve(bt, st);  // captures the lanes' graphs
            if (rc != MVDR_OK) {
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                return rc;
            }
            e = hipEventRecord(e0, st);
            for (uint32_t r = 0; r < reps && e == hipSuccess; ++r)
                if (mvdr_batch_solve(bt, st) != MVDR_OK) e = hipErrorUnknown;
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && (L == 1 || ms < best_ms)) {
                best_ms = ms;
                best = L;
            }
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e == hipSuccess) e = make_lanes(bt, best);
        if (e != hipSuccess) {
            bt->error = std::string("HIP: ") + hipGetErrorString(e);
            return MVDR_E_HIP;
        }
    }
    const mvdr_status rc = tune_shapes(chosen);
    if (rc == MVDR_OK)
        for (mvdr_solver* m : bt->members) (void)hipEventRecord(m->done, st);
    return rc;
}

mvdr_status mvdr_launch_params_get(uint32_t nodes, uint32_t frames, mvdr_launch_params* out) {
    if (!out || nodes == 0 || nodes > 65535 || frames == 0) return MVDR_E_INVALID;
    const uint32_t n = live_nodes(nodes);
    out->cg_workgroups = cg_workgroups_for(n, frames);
    out->eval_workgroups = eval_workgroups_for(n, frames);
    out->transform_workgroups = transform_workgroups_for(n, frames);
    out->gram_rows = gram_rows_for(n, frames);
    return MVDR_OK;
}

mvdr_status mvdr_launch_params_set(uint32_t nodes, uint32_t frames, const mvdr_launch_params* p) {
    if (!p || nodes == 0 || nodes > 65535 || frames == 0) return MVDR_E_INVALID;
    set_launch_params(live_nodes(nodes), frames,
                      LaunchParams{p->cg_workgroups, p->eval_workgroups, p->transform_workgroups, p->gram_rows});
    return MVDR_OK;
}

mvdr_status mvdr_launch_table_clear(void) {
    clear_launch_params();
    return MVDR_OK;
}

mvdr_status mvdr_launch_table_save(const char* path) {
    if (!path) return MVDR_E_INVALID;
    FILE* f = std::fopen(path, "w");
    if (!f) return MVDR_E_INVALID;
    const std::string text = launch_table_text();
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    return (std::fclose(f) == 0 && ok) ? MVDR_OK : MVDR_E_INVALID;
}

mvdr_status mvdr_launch_table_load(const char* path) {
    if (!path) return MVDR_E_INVALID;
    FILE* f = std::fopen(path, "r");
    if (!f) return MVDR_E_INVALID;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    std::string err;
    if (!load_launch_table_text(text, err)) {
        std::fprintf(stderr, "mvdr_launch_table_load: %s\n", err.c_str());
        return MVDR_E_INVALID;
    }
    return MVDR_OK;
}

const char* mvdr_batch_last_error(const mvdr_batch* bt) { return bt ? bt->error.c_str() : "null batch"; }

mvdr_status mvdr_batch_solve(mvdr_batch* bt, void* stream) {
    if (!bt) return MVDR_E_INVALID;
    for (mvdr_solver* m : bt->members)
        if (!m->has_input) {
            bt->error = "a member solver has no input";
            return MVDR_E_STATE;
        }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t gn = bt->cfg.gn_iterations;
    uint32_t all_nodes = 64;
    for (const mvdr_solver* m : bt->members) all_nodes = std::max(all_nodes, live_nodes(m->frame_host.N));
    bt->batch.nodes = all_nodes;
    const bool forked = bt->lanes.size() > 1;
    hipError_t e = forked ? hipEventRecord(bt->fork, st) : hipSuccess;
    for (BatchLane& lane : bt->lanes) {
        if (e != hipSuccess) break;
        hipStream_t ls = lane.stream ? lane.stream : st;
        if (lane.stream) e = hipStreamWaitEvent(ls, bt->fork, 0);
        if (e != hipSuccess) break;
        uint32_t nodes = 64;
        for (uint32_t k = lane.first; k < lane.first + lane.count; ++k)
            nodes = std::max(nodes, live_nodes(bt->members[k]->frame_host.N));
        lane.batch.nodes = nodes;
        if (bt->cfg.use_graph) {
            if (lane.graph_exec && (lane.graph_nodes != nodes || lane.graph_table != launch_table_version())) {
                (void)hipGraphExecDestroy(lane.graph_exec);
                (void)hipGraphDestroy(lane.graph);
                lane.graph_exec = nullptr;
                lane.graph = nullptr;
            }
            if (!lane.graph_exec) {
                lane.graph_nodes = nodes;
                lane.graph_table = launch_table_version();
                hipStream_t cap = nullptr;
                e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
                if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    enqueue_batch(lane.batch, bt->cfg, cap, true, 0, gn);
                    e = hipStreamEndCapture(cap, &lane.graph);
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&lane.graph_exec, lane.graph, nullptr, nullptr, 0);
                if (cap) (void)hipStreamDestroy(cap);
            }
            if (e == hipSuccess) e = hipGraphLaunch(lane.graph_exec, ls);
        } else {
            enqueue_batch(lane.batch, bt->cfg, ls, true, 0, gn);
            e = hipGetLastError();
        }
        if (e == hipSuccess && lane.stream) e = hipEventRecord(lane.joined, ls);
        if (e == hipSuccess && lane.stream) e = hipStreamWaitEvent(st, lane.joined, 0);
    }
    if (e == hipSuccess) e = hipEventRecord(bt->done, st);
    for (mvdr_solver* m : bt->members) {
        if (e == hipSuccess) e = hipEventRecord(m->done, st);
        m->indexed = true;
        m->pair_terms = bt->cfg.precision != MVDR_PRECISION_MIXED && !matrix_free(bt->cfg);
        m->iter = gn;
        m->pair_blocks = gn == 0 || blocks_after_assembly(bt->cfg);
    }
    if (e != hipSuccess) {
        bt->error = std::string("HIP: ") + hipGetErrorString(e);
        return MVDR_E_HIP;
    }
    return MVDR_OK;
}

}  // extern "C"

// ------------------------------------------------------ allocation-free solve
// mvdr_solve_noalloc: SolveNonRigidAlignment for callers that cannot allocate
// device memory.  Every per-frame buffer of the config is carved from one
// file-scope device array of MVDR_NOALLOC_SCRATCH_BYTES, by the same layout the
// handle-based solver allocates; the kernel sequence is the handle's.  Each
// device holds its own copy of the array (until the process ends); its address
// there is taken once per device with hipGetSymbolAddress and used for every
// launch.  Calls that overlap on one device share the array and race.
namespace {

constexpr size_t kScratchBytes = (size_t)MVDR_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};  // device address of `scratch` per device

// The scratch array's device address on the current device (taken on the
// first call on that device, with the device current).
hipError_t scratch_address(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// Device copies of the frame table and the buffer table (passed by value, so
// the launch carries them), and the zeroed counters and partial sums the
// handle-based solver clears at create.
__global__ void __launch_bounds__(256) noalloc_setup_kernel(FrameDev f, Buffers b, FrameDev* fd, Buffers* bd) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *fd = f;
        *bd = b;
        *b.counts = Counts{};
    }
    const size_t n = (size_t)kPartialSlots * b.partial_stride;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b.partial[i] = 0.0;
}

// The solved warp field into the caller's array; a frame whose solve hit a
// capacity error keeps its input warp (its kernels skipped their work).
__global__ void __launch_bounds__(256) noalloc_finish_kernel(const Buffers* __restrict__ bs, uint32_t slot,
                                                             DQ* __restrict__ out, uint32_t* __restrict__ error) {
    const Buffers b = *bs;
    const FrameDev& f = *b.frame;
    const uint32_t err = b.counts->error;
    const DQ* src = err ? f.node_se3_in : b.se3[slot];
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += gridDim.x * blockDim.x) out[n] = src[n];
    if (error && blockIdx.x == 0 && threadIdx.x == 0) *error = err;
}

}  // namespace

extern "C" mvdr_status mvdr_solve_noalloc(const mvdr_solver_config* config, const mvdr_frame_input* input,
                                          void* node_se3_out, uint32_t* device_error, void* stream) {
    if (!config || !input || !node_se3_out) return MVDR_E_INVALID;
    const mvdr_solver_config& c = *config;
    if (!config_shape_ok(c)) return MVDR_E_SHAPE;
    Buffers b;
    FrameDev* fd = nullptr;
    Buffers* bd = nullptr;
    uint32_t tiles = 0, chunks = 0;
    if (layout_buffers(c, b, fd, bd, tiles, chunks, nullptr) > kScratchBytes) return MVDR_E_CAPACITY;
    FrameDev f;
    const char* err = nullptr;
    const mvdr_status rc = frame_from_input(c, input, f, err);
    if (rc != MVDR_OK) return rc;
    void* base = nullptr;
    if (scratch_address(&base) != hipSuccess) return MVDR_E_HIP;
    layout_buffers(c, b, fd, bd, tiles, chunks, static_cast<char*>(base));
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(64), dim3(256), 0, st, f, b, fd, bd);
    const Batch B{bd, 1, tiles, live_nodes(f.N), chunks};
    enqueue_batch(B, c, st, true, 0, c.gn_iterations);
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(64), dim3(256), 0, st, bd, c.gn_iterations & 1,
                       static_cast<DQ*>(node_se3_out), device_error);
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}

