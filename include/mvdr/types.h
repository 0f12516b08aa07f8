/*
 * Plain-C value types shared by the solver boundary (solver.h) and the
 * synthetic scene generator (scene.h).  Every layout here is byte-compatible
 * with the reference's own device types so that a caller holding reference
 * buffers can pass them through unchanged:
 *
 *   mvdr_dq          == SparseSurfelFusion::DualQuaternion
 *                       (math/DualQuaternion/DualQuaternion.h:99-102; each
 *                        Quaternion is a float4 x,y,z,w with w the real part,
 *                        Quaternion.h:15-16)
 *   mvdr_mat34       == SparseSurfelFusion::mat34 (math/MatUtils.h:19-20,401-403):
 *                       rotation as three float3 COLUMNS, then float3 trans
 *   mvdr_intrinsic   == base/Camera.h Intrinsic {focal_x, focal_y,
 *                       principal_x, principal_y}
 *   mvdr_cross_pair  == CrossViewCorrPairs (base/CommonTypes.h:179-182):
 *                       ushort4 PixelPairs + ushort2 PixelViews
 */
#ifndef MVDR_TYPES_H
#define MVDR_TYPES_H

#include <stddef.h>
#include <stdint.h>

#define MVDR_MAX_VIEWS 3

typedef struct {
    float fx, fy, cx, cy;
} mvdr_intrinsic;

typedef struct {
    float rot[9];   /* column-major: rot[3*c + r] is row r of column c */
    float trans[3];
} mvdr_mat34;

typedef struct {
    float q0[4];    /* real part, (x, y, z, w) */
    float q1[4];    /* dual part, (x, y, z, w) */
} mvdr_dq;

typedef struct {
    uint16_t pixel[4];  /* (x1, y1, x2, y2) */
    uint16_t view[2];   /* (view of pixel 1, view of pixel 2) */
} mvdr_cross_pair;

#endif /* MVDR_TYPES_H */
