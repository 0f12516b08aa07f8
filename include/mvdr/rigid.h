/*
 * Rigid multi-view point-to-plane ICP: the per-frame rigid solver that runs
 * before the non-rigid solve with the same 6x6 normal-equation pattern.
 * Drop-in boundary for the pose-refinement path of
 * SparseSurfelFusion::RigidSolver (core/RigidSolver/RigidSolver.h:467-477):
 *
 *   mvdr_rigid_create / _destroy       RigidSolver(...) + buffer allocation / ~RigidSolver
 *   mvdr_rigid_solve                   SolveRigidAlignment(SolverMaps*, CameraObservation&,
 *                                      const mat34* LiveField2ObservedField, MaxInteration)
 *                                      (RigidSolver.cpp:258-276)
 *   mvdr_rigid_get_pose                GetWorld2Camera(cameraID), GetWorld2Camera()
 *                                      (RigidSolver.h:483-489)
 *
 * One iteration for view v (RigidSolver.cu:128-226, RigidSolver.cpp:389-446):
 * every pixel's conversion vertex / normal (the live solver map) is mapped by
 * the current world2camera[v] and projected (round to nearest, the
 * reference's float-to-u16 conversion: negative -> 0, low 16 bits kept); at a
 * reference pixel with a non-zero vertex, n_ref . n_conv >= 0.8 and squared
 * distance <= 0.03^2 it adds the row A = [p x n, n], b = -n . (p - q) to
 * A^T A and A^T b; the 6x6 system is solved by Cholesky, the twist
 * (rot, trans) becomes a transform (Rodrigues; identity rotation when
 * |rot|_1 < 1e-4) composed on the left of world2camera[v].  The combined
 * canonical-to-live transform is the average of view 0's pose and
 * relative_pose[v] * pose[v] * relative_pose[v]^-1 of the others, rotation
 * matrices averaged entrywise (RigidSolver.cpp:279-308).
 *
 * Differences from the reference: sums are FP64 per 256-pixel block and over
 * blocks (the reference sums in FP32 warp scans); the solve runs on the device
 * with no host round trip per iteration.  The reference's block reduction
 * (columnReduceKernel, RigidSolver.cu:368-383) reads only every 32nd block's
 * partial (each lane re-reads column i, not i + lane) and scales it by 32;
 * `reference_reduction` = 1 reproduces that sampling, 0 (default) sums every
 * block.
 */
#ifndef MVDR_RIGID_H
#define MVDR_RIGID_H

#include "solver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mvdr_rigid mvdr_rigid;

typedef struct {
    uint32_t max_views;
    uint32_t max_width;
    uint32_t max_height;
    uint32_t reference_reduction;   /* 1: every 32nd 256-pixel block, as the reference */
} mvdr_rigid_config;

typedef struct {
    uint32_t num_views;
    uint32_t width;
    uint32_t height;
    mvdr_intrinsic intrinsic[MVDR_MAX_VIEWS];
    const void* reference_vertex[MVDR_MAX_VIEWS];    /* float4 observation vertexConfidenceMap */
    const void* reference_normal[MVDR_MAX_VIEWS];    /* float4 observation normalRadiusMap     */
    const void* conversion_vertex[MVDR_MAX_VIEWS];   /* float4 solver map warp_vertex_map      */
    const void* conversion_normal[MVDR_MAX_VIEWS];   /* float4 solver map warp_normal_map      */
    mvdr_mat34 world2camera[MVDR_MAX_VIEWS];         /* LiveField2ObservedField[v], the start  */
    mvdr_mat34 relative_pose[MVDR_MAX_VIEWS];        /* CamerasRelativePose[v] (view 0 unused) */
} mvdr_rigid_input;

mvdr_status mvdr_rigid_create(const mvdr_rigid_config* config, mvdr_rigid** handle);
void mvdr_rigid_destroy(mvdr_rigid* handle);
/* `iterations` ICP iterations of every view, enqueued on `stream` (asynchronous). */
mvdr_status mvdr_rigid_solve(mvdr_rigid* handle, const mvdr_rigid_input* input, uint32_t iterations, void* stream);
/* The solved world2camera of each view and their average (synchronises). */
mvdr_status mvdr_rigid_get_pose(mvdr_rigid* handle, mvdr_mat34* world2camera, mvdr_mat34* canonical2live);

#ifdef __cplusplus
}
#endif

#endif /* MVDR_RIGID_H */
