/*
 * C ABI of the MI355X non-rigid warp-field solver.
 *
 * Drop-in boundary for SparseSurfelFusion::NonRigidSolver
 * (core/NonRigidSolver/NonRigidSolver.h:78-541).  Entry points map onto the
 * reference class as follows:
 *
 *   mvdr_solver_create          NonRigidSolver::NonRigidSolver(Intrinsic*)
 *                               + AllocateBuffer()      (NonRigidSolver.cpp:10-48)
 *   mvdr_solver_destroy         ReleaseBuffer()         (NonRigidSolver.cpp:50-74)
 *   mvdr_solver_set_input       SetSolverInput(...)     (NonRigidSolver.h:133-140,
 *                                                        NonRigidSolver.cpp:77-97)
 *   mvdr_solver_solve           SolveNonRigidAlignment() (NonRigidSolver.h:254)
 *   mvdr_solver_build_index     buildSolverIndexStreamed()          (NonRigidSolver.h:256)
 *   mvdr_solver_gn_iteration    solverIteration{Global,Local}IterationStreamed()
 *                                                       (NonRigidSolver.h:257-258)
 *   mvdr_solver_get_node_se3    SolvedNodeSE3()         (NonRigidSolver.h:214)
 *   mvdr_solver_total_residual  ComputeTotalResidualSynced() (NonRigidSolver.h:306)
 *   mvdr_solver_debug_get       JtResidualValue(), JtJDiagonalBlockValue()
 *                               (NonRigidSolver.h:329-330) and the other
 *                               per-stage buffers, for parity checks.
 *
 * Conventions
 *   - All map / array pointers in mvdr_frame_input are DEVICE pointers owned
 *     by the caller; they must stay valid until the solve that uses them has
 *     completed on its stream (same non-owning rule as the reference's
 *     DeviceArrayView / texture inputs).
 *   - Per-view maps are dense row-major images of width*height elements
 *     (no pitch), element types as documented per field.
 *   - Streams are hipStream_t passed as void* (NULL = the null stream).
 *   - No exception crosses the ABI; every call returns an mvdr_status, and
 *     mvdr_solver_last_error() gives the message of the last failure.
 *   - A handle is driven by one host thread at a time (the reference class is
 *     not thread-safe either).
 */
#ifndef MVDR_SOLVER_H
#define MVDR_SOLVER_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mvdr_solver mvdr_solver;

/* Largest pcg_iterations a solver accepts (one FP64 partial-sum slot per
 * iteration of each inner product). */
#define MVDR_MAX_PCG_ITERATIONS 64

typedef enum {
    MVDR_OK = 0,
    MVDR_E_INVALID = 1,   /* null pointer / malformed argument              */
    MVDR_E_SHAPE = 2,     /* views / image / node counts outside the config */
    MVDR_E_CAPACITY = 3,  /* a data-dependent size exceeded its capacity    */
    MVDR_E_HIP = 4,       /* HIP runtime failure                            */
    MVDR_E_STATE = 5      /* call order, e.g. solve before set_input        */
} mvdr_status;

typedef enum {
    MVDR_PRECISION_FP32 = 0,   /* everything FP32 (reference precision)          */
    MVDR_PRECISION_MIXED = 1   /* FP64 term geometry (J, r stored FP32), FP64
                                  dots / alpha / beta / 6x6 factors / SpMV sums;
                                  block-Jacobi PCG run as one-pass CG on
                                  L^-1 A L^-T (same iterates, one launch each) */
} mvdr_precision;

typedef enum {
    MVDR_JTJ_MATERIALIZED = 0, /* USE_MATERIALIZED_JTJ (solver_constants.h:12)   */
    MVDR_JTJ_MATRIX_FREE = 1,  /* J^T(J x) with per-term adjoints                */
    MVDR_JTJ_RECOMPUTE = 2     /* matrix-free with J re-evaluated from the maps and
                                  the current warp inside every PCG apply (the
                                  stored J feeds only the block-Jacobi
                                  preconditioner and J^T r)                        */
} mvdr_jtj_mode;

typedef struct {
    /* capacities (buffers are allocated once, at create) */
    uint32_t max_views;            /* <= MVDR_MAX_VIEWS                               */
    uint32_t max_width;
    uint32_t max_height;
    uint32_t max_nodes;            /* <= 65535 (u16 node ids)                          */
    uint32_t max_graph_edges;      /* node graph pairs, normally 8 * max_nodes         */
    uint32_t max_feature_pairs;    /* sum over views of candidate feature pairs        */
    uint32_t max_cross_pairs;      /* candidate cross-view pairs                       */
    /* schedule (values of the absent Constants::*; base/Constants.h:30-62) */
    uint32_t gn_iterations;        /* kNumGaussNewtonIterations                        */
    uint32_t pcg_iterations;       /* SolvePCGMaterialized(pcg_iterations), at most
                                      MVDR_MAX_PCG_ITERATIONS (MVDR_E_SHAPE above)     */
    uint32_t global_iterations;    /* kNumGlobalSolverItarations                       */
    uint32_t use_elastic_penalty;  /* kUseElasticPenalty                               */
    /* term switches (ConfigParser::Use*Term, base/ConfigParser.h:43-53) */
    uint32_t use_density;          /* density penalty in local iterations              */
    uint32_t use_foreground;       /* foreground-mask term                             */
    uint32_t use_cross_view;       /* cross-view correspondence term                   */
    uint32_t density_in_jtj;       /* USE_DENSE_IMAGE_DENSITY_TERM                     */
    /* numerics / execution */
    uint32_t precision;            /* mvdr_precision                                   */
    uint32_t jtj_mode;             /* mvdr_jtj_mode                                    */
    uint32_t refine_steps;         /* residual refinement steps after each PCG (mixed) */
    uint32_t use_graph;            /* capture the GN iterations into one hipGraph      */
} mvdr_solver_config;

typedef struct {
    uint32_t num_views;
    uint32_t width;
    uint32_t height;
    mvdr_intrinsic intrinsic[MVDR_MAX_VIEWS];          /* ClipColorIntrinsic[v]       */
    mvdr_mat34 world2camera[MVDR_MAX_VIEWS];           /* canonical2Live[v]           */
    mvdr_mat34 initial_camera_se3[MVDR_MAX_VIEWS];     /* camera v -> camera 0        */

    /* CameraObservation (base/CameraObservation.h:16-63) */
    const void* vertex_confid[MVDR_MAX_VIEWS];         /* float4  vertexConfidenceMap */
    const void* normal_radius[MVDR_MAX_VIEWS];         /* float4  normalRadiusMap     */
    const void* filtered_foreground[MVDR_MAX_VIEWS];   /* float   filteredForegroundMask */
    const void* foreground_gradient[MVDR_MAX_VIEWS];   /* float2  foregroundMaskGradientMap */
    const void* gray[MVDR_MAX_VIEWS];                  /* float   grayScaleMap        */
    const void* gray_gradient[MVDR_MAX_VIEWS];         /* float2  grayScaleGradientMap */
    const void* edge_mask[MVDR_MAX_VIEWS];             /* uint8   edgeMaskMap         */

    /* Renderer::SolverMaps (render/Renderer.h:381-388) */
    const void* reference_vertex[MVDR_MAX_VIEWS];      /* float4 */
    const void* reference_normal[MVDR_MAX_VIEWS];      /* float4 */
    const void* warp_vertex[MVDR_MAX_VIEWS];           /* float4 */
    const void* index_map[MVDR_MAX_VIEWS];             /* uint32, 0xFFFFFFFF = empty  */
    const void* normalized_rgb[MVDR_MAX_VIEWS];        /* float4 */

    /* SurfelGeometry::NonRigidSolverInput (core/Geometry/SurfelGeometry.h:179-182) */
    const void* surfel_knn;                            /* ushort4 */
    const void* surfel_knn_weight;                     /* float4  */
    uint32_t num_surfels;

    /* WarpField::NonRigidSolverInput (core/NonRigidSolver/WarpField.h:157-161) */
    const void* node_se3;                              /* mvdr_dq, the initial warp   */
    const void* canonical_nodes;                       /* float4  */
    const void* node_graph;                            /* ushort2 */
    uint32_t num_nodes;
    uint32_t num_graph_edges;

    /* correspondences */
    const void* feature_pairs[MVDR_MAX_VIEWS];         /* ushort4 correspondencePixelPairs[v] */
    uint32_t num_feature_pairs[MVDR_MAX_VIEWS];
    const void* cross_pairs;                           /* mvdr_cross_pair crossCorrPairs */
    uint32_t num_cross_pairs;

    /* CameraObservation::corrMap[v] (base/CameraObservation.h:55), passed to
     * the cross-view handler (NonRigidSolver.cpp:422,
     * CrossViewCorrespondenceHandler.h:30-34): per view a W*H ushort4 map whose
     * entry at current-frame pixel (z, w) holds the feature pair (x, y, z, w)
     * that ends there (PatchColliderRGBCorrespondence.cu:171); other entries
     * are zero.  Nullable per view.  When set for view v1 of a cross-view pair,
     * pixel 1 is back-traced through it: an entry whose (z, w) is pixel 1 moves
     * the geometry search to its previous-frame pixel (x, y); without an entry
     * (or a map) the search stays at pixel 1. */
    const void* corr_map[MVDR_MAX_VIEWS];
} mvdr_frame_input;

/* Sizes the index build found (valid after build_index / solve). */
typedef struct {
    uint32_t num_dense_terms;       /* T_d  */
    uint32_t num_smooth_terms;      /* E    */
    uint32_t num_foreground_terms;  /* T_fg */
    uint32_t num_feature_terms;     /* T_f  */
    uint32_t num_cross_terms;       /* T_c  */
    uint32_t num_node2term;         /* KV1  */
    uint32_t num_nodepair_terms;    /* valid KV2 entries */
    uint32_t num_node_pairs;        /* P (unique i<j pairs) */
    uint32_t num_nodes;
    uint32_t gn_iterations_done;
} mvdr_solver_sizes;

typedef enum {
    MVDR_TAP_NODE2TERM_OFFSET = 0,   /* u32[N+1]                                        */
    MVDR_TAP_NODE2TERM_INDEX = 1,    /* u32[KV1]                                        */
    MVDR_TAP_NODEPAIR_KEY = 2,       /* u32[P]  (i << 16) | j, i < j, ascending          */
    MVDR_TAP_NODEPAIR_RANGE = 3,     /* u32[2P] [start, end) into NODEPAIR_TERM_INDEX    */
    MVDR_TAP_NODEPAIR_TERM_INDEX = 4,/* u32[valid KV2]                                   */
    MVDR_TAP_DENSE_TERMS = 5,        /* 16 x u32/f32 per dense term: [0:2] knn u16x4,
                                        [2] r, [4:8] knn weights, [8:14] J (w, v)     */
    MVDR_TAP_SMOOTH_TERMS = 6,       /* 8 x u32/f32 per edge: [0] i | j << 16, [1] valid,
                                        [2:5] T_i x_j, [5:8] r = T_i x_j - T_j x_j    */
    MVDR_TAP_FEATURE_TERMS = 7,      /* 16 x u32/f32 per feature term: [0:2] knn u16x4,
                                        [4:8] weights, [8:11] target, [12:15] warped  */
    MVDR_TAP_CROSS_TERMS = 8,        /* 16 x u32/f32 per cross term (feature layout)     */
    MVDR_TAP_FOREGROUND_TERMS = 9,   /* 16 x u32/f32 per foreground term (dense layout)  */
    MVDR_TAP_DENSE_PIXELS = 10,      /* u16[4] (x, y, view, 0) per dense term            */
    MVDR_TAP_JTJ_DIAG = 11,          /* f32[36N] row-major, damping included             */
    MVDR_TAP_JTJ_DIAG_INV = 12,      /* f32[36N] row-major                               */
    MVDR_TAP_JTR = 13,               /* f32[6N]  b = -sum lambda^2 w J^T r               */
    MVDR_TAP_JTJ_OFFDIAG = 14,       /* f32[36P] row-major block (i, j), i < j           */
    MVDR_TAP_TWIST = 15,             /* f32[6N]  last PCG solution                       */
    MVDR_TAP_NODE_SE3 = 16,          /* mvdr_dq[N] current warp field                    */
    MVDR_TAP_PENALTY = 17,           /* f32[5] lambda: smooth, density, foreground, feature, cross */
    MVDR_TAP_DENSITY_TERMS = 18,     /* 16 x u32/f32 per dense term (density channel)     */
    MVDR_TAP_SPMV_INPUT = 19,        /* f32[6N] vector d of the last J^T J apply (reference twist coordinates) */
    MVDR_TAP_SPMV_OUTPUT = 20,       /* f32[6N] its product J^T J d (reference coordinates); a STAGE_SPMV bench
                                        run leaves d = M^-1 b of the last GN iteration here */
    MVDR_TAP_FOREGROUND_PIXELS = 21  /* u16[4] (x, y, view, 0) per foreground term (use_foreground) */
} mvdr_tap;

const char* mvdr_status_string(mvdr_status status);

/* Fill a config with the defaults used by the reference build. */
void mvdr_solver_default_config(mvdr_solver_config* config);

mvdr_status mvdr_solver_create(const mvdr_solver_config* config, mvdr_solver** solver);
void mvdr_solver_destroy(mvdr_solver* solver);
const char* mvdr_solver_last_error(const mvdr_solver* solver);

/* SetSolverInput: validates shapes against the config, stores the views,
 * resets the warp field to input->node_se3 and the GN counter to zero. */
mvdr_status mvdr_solver_set_input(mvdr_solver* solver, const mvdr_frame_input* input);

/* SolveNonRigidAlignment: index build + gn_iterations GN steps on `stream`. */
mvdr_status mvdr_solver_solve(mvdr_solver* solver, void* stream);

/* The two halves of the solve, for stage-wise parity checks. */
mvdr_status mvdr_solver_build_index(mvdr_solver* solver, void* stream);
mvdr_status mvdr_solver_gn_iteration(mvdr_solver* solver, void* stream);

/* SolvedNodeSE3: device view of the current warp field (solver-owned; valid
 * until the next set_input / solve / gn_iteration on this handle). */
mvdr_status mvdr_solver_get_node_se3(mvdr_solver* solver, const void** dq_device, uint32_t* num_nodes);

/* ComputeTotalResidualSynced: sum of squared weighted residuals at the
 * current warp field, with the current penalty constants.  Synchronises. */
mvdr_status mvdr_solver_total_residual(mvdr_solver* solver, void* stream, double* cost);

mvdr_status mvdr_solver_get_sizes(mvdr_solver* solver, mvdr_solver_sizes* sizes);

/* Copy a stage buffer to host memory (synchronises).  If dst is NULL or
 * bytes is too small, *needed receives the required byte count and the call
 * returns MVDR_E_CAPACITY without copying (needed may be NULL). */
mvdr_status mvdr_solver_debug_get(mvdr_solver* solver, int tap, void* dst, size_t bytes, size_t* needed);

/* Debug, MVDR_JTJ_RECOMPUTE only: the term Jacobian at the current warp (the
 * field after the GN iterations run so far), re-evaluated from the maps by the
 * kernels of the recomputing matrix-free apply, with unit term weights (no
 * penalty).  transpose 0: out = J in, with in = device float[6 * num_nodes]
 * (reference twists: rotation xyz, translation xyz per node) and out = device
 * float4[num_terms] (term t's rows: dense depth then density when
 * density_in_jtj; the three components of an edge / feature / cross-view
 * term; foreground one row; unused rows 0).  transpose 1: out = J^T in with
 * the two shapes swapped -- the exact adjoint of transpose 0.  Terms in the
 * global order of the term taps (dense, smooth, foreground, feature, cross).
 * Overwrites the solve's PCG work vectors (SPMV taps); synchronises `stream`.
 * No counterpart in the reference (its matrix-free path has no such hook);
 * it exposes the J x and J^T y of ApplyJtJMatrixFreeIndexed.cu:29-349 for
 * adjoint and finite-difference checks. */
mvdr_status mvdr_solver_debug_jacobian(mvdr_solver* solver, int transpose, const void* in, void* out, void* stream);

typedef enum {
    MVDR_STAGE_EVAL_TERMS = 0,     /* term residuals / Jacobians                      */
    MVDR_STAGE_ASSEMBLE = 1,       /* J^T J diagonal + off-diagonal blocks, J^T r,
                                      block inverse (one fused pass per node row)   */
    MVDR_STAGE_PCG_UPDATE = 2,     /* one PCG vector update (x, r, M^-1 r, dot)       */
    MVDR_STAGE_SPMV = 3,           /* one block-sparse ApplyJtJ (q = A d) + dot(d, q) */
    MVDR_STAGE_INDEX = 4,          /* the whole index build                           */
    MVDR_STAGE_CG_ITERATION = 5    /* one fused CG iteration of the mixed mode:
                                      L^-1 A L^-T apply, vector updates, both dots  */
} mvdr_stage;

/* Kernel micro-benchmark (the reference keeps a commented 1000x loop around
 * its JtJ kernel, JtJMaterializer.cpp:52-60): re-launch one stage `reps` times
 * on `stream` on the data of the last solve and return the mean time per
 * launch measured with hipEvents on that stream.  Needs a completed solve;
 * stage buffers hold re-computed values afterwards, which the next solve
 * overwrites.  Synchronises. */
mvdr_status mvdr_solver_bench_stage(mvdr_solver* solver, int stage, uint32_t reps, void* stream, float* avg_ms);

/* ------------------------------------------------------------------ batches
 * No counterpart in the reference, which solves one frame at a time: a batch
 * solves the current frames of several solver instances in one launch per
 * stage (kernels index the frame by blockIdx.y), for many-scene throughput.
 * Members must share gn/pcg iterations, penalties, term switches, precision
 * and refinement (MVDR_E_SHAPE otherwise); capacities and frame sizes may
 * differ.  The members must outlive the batch.  After mvdr_batch_solve every
 * member's warp field is solved exactly as mvdr_solver_solve would, and its
 * SolvedNodeSE3 / taps / sizes read as after a single solve. */
typedef struct mvdr_batch mvdr_batch;

mvdr_status mvdr_batch_create(mvdr_solver* const* solvers, uint32_t count, mvdr_batch** batch);
void mvdr_batch_destroy(mvdr_batch* batch);
const char* mvdr_batch_last_error(const mvdr_batch* batch);
/* SolveNonRigidAlignment of every member (use_graph of the first member's
 * config: the whole batch solve is captured once and replayed). */
mvdr_status mvdr_batch_solve(mvdr_batch* batch, void* stream);
/* mvdr_solver_bench_stage over all members at once (one launch per rep). */
mvdr_status mvdr_batch_bench_stage(mvdr_batch* batch, int stage, uint32_t reps, void* stream, float* avg_ms);
/* Solve lanes (no counterpart in the reference): the members are split into
 * `lanes` contiguous runs of near-equal size, each solved by its own chain of
 * launches (and captured graph) on its own stream, forked from and joined
 * back to the stream given to mvdr_batch_solve, so that one lane's
 * latency-bound CG launches overlap another's evaluation and assembly.  Every
 * member's result is the same for any lane count.  A new batch has two lanes
 * from 8 members on, else one.  1 <= lanes <= min(members,
 * MVDR_MAX_BATCH_LANES); waits for the batch's last solve. */
#define MVDR_MAX_BATCH_LANES 4
mvdr_status mvdr_batch_set_lanes(mvdr_batch* batch, uint32_t lanes);
uint32_t mvdr_batch_lanes(const mvdr_batch* batch);

/* ------------------------------------------------------- launch-shape tuning
 * No counterpart in the reference (its launches use fixed block sizes).  The
 * workgroup counts below spread work over the device and change no result
 * (the fused CG's inner products are summed per 32-row group, the term
 * evaluation and the CG transform are grid-stride loops over independent
 * items).  A process-wide table holds them per (node count in 1024-node
 * buckets, frames per launch); launches without an entry use the built-in
 * rule.  Captured graphs are re-captured after the table changes.  The table
 * is read from the file named by MVDR_LAUNCH_TABLE, if set, at the first
 * launch. */
typedef struct {
    uint32_t cg_workgroups;         /* fused CG row kernels, per frame (0 in a set: the rule) */
    uint32_t eval_workgroups;       /* term evaluation, per frame                             */
    uint32_t transform_workgroups;  /* CG transform, per frame                                */
    uint32_t gram_rows;             /* rows per row wave of the Gram assembly: 1, 2 or 4      */
} mvdr_launch_params;

/* Times each tunable launch over its candidates (reps launches each, on the
 * data of the last GN iteration of a completed solve, as bench_stage) and
 * keeps the fastest in the table for this solver's / batch's shape; `chosen`
 * (nullable) receives the resolved counts (0 for launches the config does not
 * use).  A batch also times reps whole solves (its members' inputs solved
 * again, with the same results) for each lane count 1 .. min(members,
 * MVDR_MAX_BATCH_LANES), with the shapes tuned for its lanes, and keeps the
 * fastest (mvdr_batch_lanes reads it); `chosen` then describes lane 0.
 * Synchronises `stream`. */
mvdr_status mvdr_solver_autotune(mvdr_solver* solver, uint32_t reps, void* stream, mvdr_launch_params* chosen);
mvdr_status mvdr_batch_autotune(mvdr_batch* batch, uint32_t reps, void* stream, mvdr_launch_params* chosen);
/* The counts launches for `nodes` nodes and `frames` frames use. */
mvdr_status mvdr_launch_params_get(uint32_t nodes, uint32_t frames, mvdr_launch_params* out);
mvdr_status mvdr_launch_params_set(uint32_t nodes, uint32_t frames, const mvdr_launch_params* params);
mvdr_status mvdr_launch_table_clear(void);
/* Text file, one line per entry: node_bucket frames cg eval transform. */
mvdr_status mvdr_launch_table_save(const char* path);
mvdr_status mvdr_launch_table_load(const char* path);

/* ------------------------------------------------------ allocation-free solve
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): SolveNonRigidAlignment of one frame without a handle --
 * index build and config->gn_iterations GN steps enqueued on `stream`, the
 * solved warp field written to node_se3_out (device, mvdr_dq[num_nodes]).
 * Every buffer of the solve lives in one library-owned device array of
 * MVDR_NOALLOC_SCRATCH_BYTES per device (held until the process ends); a
 * config whose buffers do not fit returns MVDR_E_CAPACITY before any GPU work
 * (max_nodes about 2,000 at one 320x240 view in the mixed precision, 4,000 in
 * FP32).  Nothing is allocated; the call does not synchronise; use_graph is
 * ignored.  device_error (nullable, device u32) receives the solve's
 * capacity error bits (0 = none; on an error the output is the input warp).
 * Calls that overlap on one device share that array and race: serialise them
 * (one stream, or events).  The results equal mvdr_solver_solve's for the
 * same config and frame. */
#define MVDR_NOALLOC_SCRATCH_BYTES (64u << 20)

mvdr_status mvdr_solve_noalloc(const mvdr_solver_config* config, const mvdr_frame_input* input, void* node_se3_out,
                               uint32_t* device_error, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MVDR_SOLVER_H */
