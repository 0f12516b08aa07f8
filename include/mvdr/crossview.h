/*
 * Cross-view edge pair production: the step upstream of the solver's
 * cross-view term.  Drop-in boundary for
 * SparseSurfelFusion::CrossViewEdgeCorrespondence
 * (core/CrossViewEdgeCorrespondence/CrossViewEdgeCorrespondence.h:94-223):
 *
 *   mvdr_crossview_create      CrossViewEdgeCorrespondence(Intrinsic*, mat34*)
 *                              + AllocateBuffer()            (:101-103, :151)
 *   mvdr_crossview_destroy     ~CrossViewEdgeCorrespondence / ReleaseBuffer (:103, :157)
 *   mvdr_crossview_find        SetCrossViewMatchingInput + FindCrossViewEdgeMatchedPairs
 *                                                            (:197-203)
 *   mvdr_crossview_get_pairs   GetCrossViewMatchingCorrPairs / ...UniqueCorrPairs /
 *                              ...BackTracingCorrPairs       (:210-222)
 *
 * The reference ships only this header (its .cu is absent from the tree), so
 * the rules below are this library's reconstruction from the declared
 * kernels, their comments and constants (CrossViewSearchSize 20,
 * CrossViewSearchStep 2, OpticalSearchSize 3, OpticalSearchStep 1, :24-29):
 *
 *   1. candidates (MarkValidPairPixels, GetClosetCrossViewPairs): every edge
 *      pixel (edge mask 1, non-zero vertex) of view v is moved to camera 0
 *      (initial_camera_se3[v]) and into each neighbouring view n (v + 1 and
 *      v - 1 mod V), projected there (rounded to nearest), and the window of
 *      half size search_radius sampled every search_step pixels around the
 *      projection is scanned (rows, then columns) for the edge pixel of view n
 *      whose vertex is nearest to the moved point (strict <; view v + 1 is
 *      scanned first).  A nearest squared distance below max_distance2 gives
 *      the candidate (x, y) in v -> (z, w) in n.  At most one candidate per
 *      source pixel, in (view, y, x) order.
 *   2. unique pairs (FilterSeveralForOneMatchingPairs,
 *      SelectUniqueCrossViewMatchingPairs): of the candidates naming the same
 *      target pixel the nearest stays (ties: the lower source pixel); a pair
 *      found from both ends (a -> b and b -> a both kept) stays once, from the
 *      lower source pixel (views, then rows, then columns).
 *   3. back-traced pairs (SelectCrossViewBackTracingPairs,
 *      GetTrackedCrossViewPairs): the unique pairs whose two pixels each have
 *      an optical-flow correspondence within the window of half size
 *      track_radius (step 1) in their view's corr_map (an entry (x, y, z, w)
 *      stored at pixel (z, w) of the current frame; other entries are not
 *      correspondences).  Without corr maps this list is empty.
 *
 * Outputs are mvdr_cross_pair records (pixel 1 in view[0], pixel 2 in
 * view[1]), each list in source-pixel order, solver-owned device memory valid
 * until the next find.  Inputs are caller-owned device pointers, dense
 * row-major images.  Errors: status codes; more candidates than max_pairs set
 * MVDR_E_CAPACITY at get_pairs (the lists are then truncated).
 */
#ifndef MVDR_CROSSVIEW_H
#define MVDR_CROSSVIEW_H

#include "solver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mvdr_crossview mvdr_crossview;

typedef struct {
    uint32_t max_views;      /* <= MVDR_MAX_VIEWS                                 */
    uint32_t max_width;
    uint32_t max_height;
    uint32_t max_pairs;      /* capacity of each output list                      */
    uint32_t search_radius;  /* CrossViewSearchSize (20)                          */
    uint32_t search_step;    /* CrossViewSearchStep (2)                           */
    uint32_t track_radius;   /* OpticalSearchSize 3 -> half size 1                */
    float max_distance2;     /* pairsSquDisThreshold, squared metres              */
} mvdr_crossview_config;

typedef struct {
    uint32_t num_views;
    uint32_t width;
    uint32_t height;
    mvdr_intrinsic intrinsic[MVDR_MAX_VIEWS];
    mvdr_mat34 initial_camera_se3[MVDR_MAX_VIEWS];   /* camera v -> camera 0           */
    const void* vertex[MVDR_MAX_VIEWS];             /* float4 vertex map (camera v)   */
    const void* edge_mask[MVDR_MAX_VIEWS];          /* uint8, 1 = edge pixel          */
    const void* corr_map[MVDR_MAX_VIEWS];           /* ushort4, nullable (stage 3)    */
} mvdr_crossview_input;

typedef enum {
    MVDR_CROSSVIEW_CANDIDATES = 0,   /* GetCrossViewMatchingCorrPairs        */
    MVDR_CROSSVIEW_UNIQUE = 1,       /* GetCrossViewMatchingUniqueCorrPairs  */
    MVDR_CROSSVIEW_BACKTRACED = 2    /* GetCrossViewBackTracingCorrPairs     */
} mvdr_crossview_list;

void mvdr_crossview_default_config(mvdr_crossview_config* config);
mvdr_status mvdr_crossview_create(const mvdr_crossview_config* config, mvdr_crossview** handle);
void mvdr_crossview_destroy(mvdr_crossview* handle);
/* Enqueues the three stages on `stream` (asynchronous). */
mvdr_status mvdr_crossview_find(mvdr_crossview* handle, const mvdr_crossview_input* input, void* stream);
/* Device view of one list and its length (synchronises with the last find). */
mvdr_status mvdr_crossview_get_pairs(mvdr_crossview* handle, int list, const void** pairs_device, uint32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* MVDR_CROSSVIEW_H */
