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

#ifdef __cplusplus
}
#endif

#endif /* MVDR_WARPFIELD_H */
