/*
 * Solver-map generation: the per-view index / reference / warp maps the solve
 * reads, rasterised from the surfels on the GPU instead of by OpenGL.
 * Replaces Renderer::DrawSolverMapsWithRecentObservation /
 * DrawSolverMapsConfidentObservation + MapSolverMapsToCuda
 * (render/Renderer.h:381-395) with the point program of
 * render/shaders/SolverMap.vert and SolverMap.frag:
 *
 *   surfel s of view v is drawn when its warp vertex, mapped by world2camera[v],
 *   has 0 <= z <= max_depth + 0.05 and it is not an old unconfident surfel
 *   (reference_vertex.w < confidence_threshold with
 *   |color_time.z - current_time| >= time_threshold);
 *   it lands on pixel (floor(fx x / z + cx), floor(fy y / z + cy)) -- a unit
 *   point sprite covers the pixel whose centre lies within half a pixel;
 *   the nearest z wins the pixel (depth test LESS in draw order: equal depths
 *   keep the lower surfel index);
 *   the winner writes reference_vertex (xyz, 1), reference_normal (xyz, 0),
 *   warp_vertex (xyz, 1), warp_normal (xyz, 0), index_map = s and
 *   normalized_rgb = (r, g, b decoded from the bits of color_time.x, / 255,
 *   color_time.y); pixels without a surfel read 0 and index 0xFFFFFFFF.
 *
 * The sized-sprite program (SolverMapSized.vert) is not reproduced: every
 * surfel covers one pixel.  All arrays are caller-owned device memory; the
 * call enqueues three kernels on `stream` and allocates nothing (the depth
 * buffer `zbuffer[v]`, u64 per pixel, is caller scratch).  Per-view surfel
 * attribute pointers may all name the same arrays.
 */
#ifndef MVDR_RENDER_H
#define MVDR_RENDER_H

#include "solver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t num_views;
    uint32_t width;
    uint32_t height;
    uint32_t num_surfels;
    mvdr_intrinsic intrinsic[MVDR_MAX_VIEWS];
    mvdr_mat34 world2camera[MVDR_MAX_VIEWS];        /* the shader's world2camera uniform  */
    const void* reference_vertex[MVDR_MAX_VIEWS];   /* float4 reference_vertex_confid     */
    const void* reference_normal[MVDR_MAX_VIEWS];   /* float4 reference_normal_radius     */
    const void* warp_vertex[MVDR_MAX_VIEWS];        /* float4                             */
    const void* warp_normal[MVDR_MAX_VIEWS];        /* float4                             */
    const void* color_time;                         /* float4 (rgb bits, y, time, -)      */
    float max_depth;                                /* width_height_maxdepth_currtime.z   */
    float current_time;                             /* width_height_maxdepth_currtime.w   */
    float confidence_threshold;                     /* confid_time_threshold.x            */
    float time_threshold;                           /* confid_time_threshold.y            */
} mvdr_render_input;

typedef struct {
    void* reference_vertex[MVDR_MAX_VIEWS];   /* float4 W*H */
    void* reference_normal[MVDR_MAX_VIEWS];   /* float4 W*H */
    void* warp_vertex[MVDR_MAX_VIEWS];        /* float4 W*H */
    void* warp_normal[MVDR_MAX_VIEWS];        /* float4 W*H, nullable */
    void* index_map[MVDR_MAX_VIEWS];          /* uint32 W*H */
    void* normalized_rgb[MVDR_MAX_VIEWS];     /* float4 W*H, nullable */
    void* zbuffer[MVDR_MAX_VIEWS];            /* uint64 W*H scratch */
} mvdr_render_output;

mvdr_status mvdr_render_solver_maps(const mvdr_render_input* input, const mvdr_render_output* output, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MVDR_RENDER_H */
