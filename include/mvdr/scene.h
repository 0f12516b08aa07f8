/*
 * Seeded synthetic scene generator (host only) producing the inputs of one
 * non-rigid solve: surfels with 4-NN skinning, a voxel-subsampled node graph,
 * a ground-truth warp, per-view observation maps (depth vertex/normal,
 * foreground, gray, edge mask), per-view solver maps (index / reference /
 * warp vertex / rgb), sparse feature pairs and cross-view pairs.
 *
 * The reference renders its solver maps with OpenGL (render/Renderer.h:381-395)
 * and reads observations from cameras (ImageProcess/); neither is on the
 * solver path, so benchmarks and tests synthesise both here.  Node graph and
 * skinning follow WarpField.cu:15-59 (8-NN, self excluded) and
 * CanonicalNodesSkinner.cu:154-166 (Gaussian weights, radius 0.025 m).
 */
#ifndef MVDR_SCENE_H
#define MVDR_SCENE_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mvdr_scene mvdr_scene;

typedef struct {
    uint64_t seed;
    uint32_t shape;               /* 0 = sphere, 1 = cylinder-plus-limbs body          */
    uint32_t num_views;           /* 1..3, cameras 120 degrees apart around the object */
    uint32_t width, height;
    float fx, fy;                 /* principal point is the image centre               */
    uint32_t target_surfels;
    uint32_t target_nodes;
    float cam_distance_min;       /* camera distances drawn uniformly in [min, max]    */
    float cam_distance_max;
    uint32_t features_per_view;
    uint32_t cross_pairs;
    float depth_noise_sigma;      /* metres                                            */
    float outlier_fraction;       /* fraction of valid depth pixels replaced           */
    float max_rot_deg;            /* ground-truth per-node rotation bound              */
    float max_trans;              /* ground-truth per-node translation bound (metres)  */
    float feature_jitter_px;
} mvdr_scene_params;

typedef struct {
    uint32_t num_surfels;
    uint32_t num_nodes;
    uint32_t num_graph_edges;
    uint32_t num_views;
    uint32_t width, height;
    uint32_t num_feature_pairs[MVDR_MAX_VIEWS];
    uint32_t num_cross_pairs;
    float voxel_size;
} mvdr_scene_info;

/* Defaults for the named workloads: 0 = single-view sphere (~1k nodes,
 * 20k surfels, 320x240), 1 = three-view body (~4k nodes, 150k surfels,
 * 640x480), 3 = scaled body (~8k nodes, 300k surfels).  Returns 0 on success. */
int mvdr_scene_default_params(int workload, mvdr_scene_params* params);

int mvdr_scene_create(const mvdr_scene_params* params, mvdr_scene** scene);
void mvdr_scene_destroy(mvdr_scene* scene);
int mvdr_scene_get_info(const mvdr_scene* scene, mvdr_scene_info* info);
const char* mvdr_scene_last_error(void);

/* Copy the named array into dst.  Per-view arrays take `view`; global ones
 * ignore it.  Returns the byte size of the array, or -1 for an unknown name;
 * copies only when dst is non-NULL and bytes >= size.  Names:
 *   global : surfel_pos f4, surfel_normal f4, surfel_knn u16x4,
 *            surfel_knn_weight f4, nodes f4, node_graph u16x2,
 *            node_se3_init dq, node_se3_gt dq, cross_pairs mvdr_cross_pair
 *   view   : intrinsic, world2camera, initial_camera_se3 (one struct each),
 *            vertex_confid f4, normal_radius f4, foreground_mask u8,
 *            filtered_foreground f32, foreground_gradient f2, gray f32,
 *            gray_gradient f2, edge_mask u8, reference_vertex f4,
 *            reference_normal f4, warp_vertex f4, index_map u32,
 *            normalized_rgb f4, feature_pairs u16x4, corr_map u16x4 (per
 *            pixel: the feature pair observed there, else zero)            */
long long mvdr_scene_get(const mvdr_scene* scene, const char* name, uint32_t view, void* dst, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* MVDR_SCENE_H */
