/*
 * Warp-field maintenance around the solve: the node graph, surfel / node
 * skinning, and the forward / inverse warp of surfels and nodes by a solved
 * warp field.  Each entry point replaces one reference kernel launch:
 *
 *   mvdr_warp_build_node_graph  WarpField::BuildNodesGraph
 *                               (core/NonRigidSolver/WarpField.cu:15-59, :62-71)
 *   mvdr_warp_skin              CanonicalNodesSkinner, skinningVertexAndNodeBruteForceKernel
 *                               (core/NonRigidSolver/CanonicalNodesSkinner.cu:356-437,
 *                               search core/Geometry/brute_foce_knn.cuh:8-99)
 *   mvdr_warp_forward           SurfelNodeDeformer::ForwardWarpSurfelsAndNodes,
 *                               forwardWarpVertexAndNodeKernel
 *                               (core/Geometry/SurfelNodeDeformer.cu:9-60)
 *   mvdr_warp_inverse           inverseWarpVertexNormalKernel without the correction
 *                               transform (core/Geometry/SurfelNodeDeformer.cu:96-124)
 *   mvdr_warp_correct_nodes     NonRigidSolver::CorrectLargeErrorNode, correspondence-guided
 *                               form (core/NonRigidSolver/NonRigidSolver.cu:100-233, :365-426)
 *
 * All arrays are caller-owned DEVICE memory; the calls enqueue on `stream`
 * (void*, hipStream_t) and allocate nothing.  Types: float4 positions (w
 * carried through), ushort4 knn, float4 knn weights, ushort2 graph edges,
 * mvdr_dq warp field.  Node ids are u16: num_nodes <= 65535.
 *
 * Rules (the reference's, restated; deviations marked):
 *   node graph: for node i, the nodes k with squared distance in (1e-6,
 *     current worst) replace the worst of 8 slots (slot scan for the new
 *     worst: first strict maximum); edges (i, slot_k) in slot order, an
 *     unfilled slot names node 0xFFFF.
 *   skinning: the 4 nearest nodes of each vertex and each node (itself
 *     included) by the reference's 4-slot heap update, distance
 *     fma(dz, dz, fma(dy, dy, dx * dx)); weights exp(-d / (2 * r * r)) with
 *     r = node_radius (evaluated (2 * r) * r), divided by their sum.
 *     Deviation: the reference scans the node array padded to a multiple of
 *     4 (reading past the last node); here only the real nodes are scanned
 *     (the same result whenever there are >= 4 nodes and the padding lies
 *     farther than the 4th neighbour).
 *   forward warp: v' = T v, n' = R n with T = se3_matrix(blend(knn, w)).
 *   inverse warp: v = R^T (v' - t), n = R^T n'.
 *   num_nodes is the warp field's size: a point whose knn names a node at or
 *   beyond it is copied unwarped (the reference reads out of bounds).
 */
#ifndef MVDR_WARPFIELD_H
#define MVDR_WARPFIELD_H

#include "solver.h"

/*
 * Large-error node correction (mvdr_warp_correct_nodes), after a solve:
 *   1. per node n with node_error[n] > 0.2: view v from the node's w
 *      (v < w <= v + 1); its live position c = initial_camera_se3[v]^-1 *
 *      world2camera[v] * node, projected to pixel (u, y) = round-to-nearest of
 *      x / (z + 1e-10) * f + principal point (evaluated in double, as the
 *      reference's double literal makes it, then rounded to float; negative
 *      values give pixel 0).  Inside the border (2 <= u < W - 2, likewise y)
 *      the 5 x 5 window (u outer, y inner) is scanned for the pixel with
 *      guided_valid == 1, a non-zero previous vertex p, |c - p|^2 < 1e-4 and
 *      |G c - c|^2 < 4 * 0.025^2 (G = guided_se3 there) whose |c - p|^2 is
 *      strictly smallest; S = world2camera^-1 * initial * G * initial^-1 *
 *      world2camera (identity G when none qualifies or outside the border).
 *      correction[n] = DualQuaternion(S) scaled by node_error[n] (clamped to
 *      [1e-10, 1]): rotation q0^t, dual part ((q1 conj(q0)) t) q0^t.  Nodes at
 *      or below 0.2 get the identity.  Deviation: a node whose w names no
 *      view below num_views gets the identity (the reference reads out of
 *      bounds).
 *   2. every surfel and every node: the weighted sum of its knn's corrections
 *      (no normalisation), as a transform (normalised), moves live_vertex /
 *      live_normal (w kept) and live_nodes; each node's node_se3 becomes that
 *      sum times node_se3 (the product of the unnormalised sum, as in the
 *      reference).  A point whose knn names a node >= num_nodes is left as is.
 * The optical-flow-guided form (computeOpticalFlowGuidedNodesSE3,
 * NonRigidSolver.cu:142-153) sets every correction to the identity after
 * computing it, so it leaves the field unchanged; it is not reproduced.
 * live_vertex / live_normal are one copy: the reference writes the same
 * values into every view's surfel arrays.
 */
typedef struct {
    uint32_t num_views;
    uint32_t width;
    uint32_t height;
    mvdr_intrinsic intrinsic[MVDR_MAX_VIEWS];
    mvdr_mat34 initial_camera_se3[MVDR_MAX_VIEWS];
    mvdr_mat34 world2camera[MVDR_MAX_VIEWS];
    const void* previous_vertex[MVDR_MAX_VIEWS];  /* float4 W*H, observation.PreviousVertexConfidenceMap */
    const void* guided_se3[MVDR_MAX_VIEWS];       /* mvdr_mat34 W*H, correspondenceSe3Map                */
    const void* guided_valid[MVDR_MAX_VIEWS];     /* uint8 W*H, markValidSe3Map                          */
    uint32_t num_nodes;
    const void* node_error;                       /* float [num_nodes], nodeUnitedError                   */
    void* live_nodes;                             /* float4 [num_nodes], in place                         */
    void* node_se3;                               /* mvdr_dq [num_nodes], in place                        */
    const void* node_knn;                         /* ushort4 [num_nodes]                                  */
    const void* node_knn_weight;                  /* float4 [num_nodes]                                   */
    uint32_t num_surfels;
    void* live_vertex;                            /* float4 [num_surfels], in place                       */
    void* live_normal;                            /* float4 [num_surfels], in place                       */
    const void* surfel_knn;                       /* ushort4 [num_surfels]                                */
    const void* surfel_knn_weight;                /* float4 [num_surfels]                                 */
    void* correction;                             /* mvdr_dq [num_nodes], out (dqCorrectNodes)            */
} mvdr_node_correction;

#ifdef __cplusplus
extern "C" {
#endif

mvdr_status mvdr_warp_build_node_graph(const void* canonical_nodes, uint32_t num_nodes, void* node_graph, void* stream);

mvdr_status mvdr_warp_skin(const void* vertices, uint32_t num_vertices, const void* canonical_nodes, uint32_t num_nodes,
                           float node_radius, void* vertex_knn, void* vertex_knn_weight, void* node_knn,
                           void* node_knn_weight, void* stream);

mvdr_status mvdr_warp_forward(const void* node_se3, const void* canonical_vertex, const void* canonical_normal,
                              const void* vertex_knn, const void* vertex_knn_weight, uint32_t num_vertices,
                              const void* canonical_nodes, const void* node_knn, const void* node_knn_weight,
                              uint32_t num_nodes, void* live_vertex, void* live_normal, void* live_nodes,
                              void* stream);

mvdr_status mvdr_warp_inverse(const void* node_se3, uint32_t num_nodes, const void* live_vertex,
                              const void* live_normal, const void* vertex_knn, const void* vertex_knn_weight,
                              uint32_t num_vertices, void* reference_vertex, void* reference_normal, void* stream);

mvdr_status mvdr_warp_correct_nodes(const mvdr_node_correction* input, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MVDR_WARPFIELD_H */
