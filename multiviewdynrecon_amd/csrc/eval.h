// Term residuals and Jacobians at a warp field, evaluated from the solver
// maps, the observation maps and the node transforms: the evaluation pass of
// every GN iteration and the matrix-free apply that recomputes J inside each
// PCG iteration share these functions.
#pragma once
#include "common.h"

namespace mvdr {

// ---------------------------------------------------------- term evaluation
// R = float is the reference's arithmetic; R = double (mixed precision mode)
// evaluates the warp, camera chain and projection in FP64 and stores J and r in
// FP32.  In FP32, p_live carries ~1e-7 m of rounding: enough to move ~0.1% of
// the projections to a neighbouring depth pixel and to cost ~1e-3 in the twist.
template <typename R> __device__ inline M34<R> init_inverse(const FrameDev& f, uint32_t v) {
    if constexpr (sizeof(R) == sizeof(float)) return f.init_inv[v];
    else return inverse(widen<R>(f.init[v]));
}

// computeDenseDepthJacobianKernel (DenseDepthJacobianResidual.cu:24-104).  The
// camera of a term is the view stored with its compacted pixel.
template <typename R>
__device__ inline void eval_dense(const FrameDev& f, const DQ* se3, ushort4 px, ScalarRec& rec) {
    const uint32_t v = px.z;
    const uint32_t o = px.y * f.W + px.x;
    const V3<R> cv = xyz_as<R>(f.ref_v[v][o]), cn = xyz_as<R>(f.ref_n[v][o]);
    const M34<R> I = widen<R>(f.init[v]), W2C = widen<R>(f.w2c[v]), Iinv = init_inverse<R>(f, v);
    const V3<R> v0 = I.apply(cv), n0 = I.rot(cn);
    const M34<R> T = warp_of<R>(se3, rec.knn, rec.w);
    const V3<R> pl = T.apply(v0), nl = T.rot(n0);
    V3<R> pc = W2C.apply(pl), nc = W2C.rot(nl);
    pc = Iinv.apply(pc);
    nc = Iinv.rot(nc);
    float r = 0.f, J[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int u = project_round<R>(pc.x, pc.z, f.fx[v], f.cx[v]);
    const int w = project_round<R>(pc.y, pc.z, f.fy[v], f.cy[v]);
    if (u >= 0 && u < (int)f.W && w >= 0 && w < (int)f.H) {
        const float4 dv = f.vertex[v][w * f.W + u];
        const float4 dn = f.normal[v][w * f.W + u];
        const V3<R> d = xyz_as<R>(dv), n = xyz_as<R>(dn);
        bool valid = !is_zero_vertex(dv);
        if (dot3(n, nc) < R(0.7f)) valid = false;
        const V3<R> diff = d - pc;
        if (dot3(diff, diff) > R(0.075f) * R(0.075f)) valid = false;
        if (valid) {
            r = (float)dot3(n, pc - d);
            const V3<R> jt = W2C.rot_t(I.rot(n));
            const V3<R> jr = cross3(pl, jt);
            J[0] = (float)jr.x; J[1] = (float)jr.y; J[2] = (float)jr.z;
            J[3] = (float)jt.x; J[4] = (float)jt.y; J[5] = (float)jt.z;
        }
    }
    rec.r = r;
#pragma unroll
    for (int k = 0; k < 6; ++k) rec.J[k] = J[k];
}

// computeImageDensityJacobainAndResidual (density_map_jacobian.cuh:12-79) as
// called for density and foreground terms (DensityForegroundJacobianResidual.cu:41-142).
// The camera-v reference vertex is warped without the move to camera 0, as there.
template <typename R>
__device__ inline void eval_image(const FrameDev& f, const DQ* se3, ushort4 px, bool density, ScalarRec& rec) {
    const uint32_t v = px.z;
    const uint32_t o = px.y * f.W + px.x;
    const V3<R> cv = xyz_as<R>(f.ref_v[v][o]);
    R geometry = 0;
    if (density) {
        const float4 rgb = f.rgb[v][o];
        geometry = R(0.299f) * rgb.x + R(0.587f) * rgb.y + R(0.114f) * rgb.z;
    }
    const M34<R> T = warp_of<R>(se3, rec.knn, rec.w);
    const V3<R> pl = T.apply(cv);
    const V3<R> pc = init_inverse<R>(f, v).apply(widen<R>(f.w2c[v]).apply(pl));
    float r = 0.f, J[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const R fx = f.fx[v], fy = f.fy[v];
    const int u = project_round<R>(pc.x, pc.z, f.fx[v], f.cx[v]);
    const int w = project_round<R>(pc.y, pc.z, f.fy[v], f.cy[v]);
    if (u >= 0 && u < (int)f.W && w >= 0 && w < (int)f.H) {
        const uint32_t q = w * f.W + u;
        const R img = density ? f.gray[v][q] : f.ffg[v][q];
        const float2 g = density ? f.gray_grad[v][q] : f.fg_grad[v][q];
        V3<R> d;
        d.x = R(g.x) * fx / pc.z;
        d.y = R(g.y) * fy / pc.z;
        d.z = -R(1) / (pc.z * pc.z) * (fx * pc.x * R(g.x) + fy * pc.y * R(g.y));
        d = widen<R>(f.init[v]).rot(d);
        d = widen<R>(f.w2c[v]).rot_t(d);
        J[0] = (float)(-d.y * pl.z + d.z * pl.y);
        J[1] = (float)(d.x * pl.z - d.z * pl.x);
        J[2] = (float)(-d.x * pl.y + d.y * pl.x);
        J[3] = (float)d.x; J[4] = (float)d.y; J[5] = (float)d.z;
        r = (float)(img - geometry);
    }
    rec.r = r;
#pragma unroll
    for (int k = 0; k < 6; ++k) rec.J[k] = J[k];
}

// A graph edge (i, j) at the warp se3 (NodeGraphSmoothHandler.cu:10-43):
// T_i x_j, the residual T_i x_j - T_j x_j and the edge's validity.
template <typename R>
__device__ inline SmoothRec eval_smooth(const FrameDev& f, const DQ* se3, uint32_t e) {
    const ushort2 ij = f.graph[e];
    const float4 xi = f.nodes[ij.x], xj = f.nodes[ij.y];
    const M34<R> Ti = node_warp<R>(se3[ij.x]), Tj = node_warp<R>(se3[ij.y]);
    const V3<R> a = Ti.apply(xyz_as<R>(xj)), s = Tj.apply(xyz_as<R>(xj));
    const V3<R> r = a - s;
    const f3 d = xyz(xi) - xyz(xj);
    SmoothRec rec;
    rec.ij = (uint32_t)ij.x | ((uint32_t)ij.y << 16);
    rec.valid = dot3(d, d) > 64 * 0.025f * 0.025f ? 0u : 1u;
    rec.Ti_xj[0] = (float)a.x; rec.Ti_xj[1] = (float)a.y; rec.Ti_xj[2] = (float)a.z;
    rec.r[0] = (float)r.x; rec.r[1] = (float)r.y; rec.r[2] = (float)r.z;
    return rec;
}

// A feature / cross-view term's canonical vertex warped by the blend of its
// knn (SparseCorrespondenceHandler.cu:193-220).
template <typename R>
__device__ inline float4 eval_warped(const DQ* se3, ushort4 knn, float4 w, float4 can) {
    const M34<R> T = warp_of<R>(se3, knn, w);
    const V3<R> p = T.apply(xyz_as<R>(can));
    return make_float4((float)p.x, (float)p.y, (float)p.z, 1.0f);
}

}  // namespace mvdr
