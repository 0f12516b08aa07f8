// Solver-map rasterisation (include/mvdr/render.h): a z-buffer of packed
// 64-bit keys (depth bits << 32 | surfel index) per pixel, filled with
// atomicMin by one thread per (view, surfel) and resolved by one thread per
// (view, pixel).  The minimum of the packed key is order-independent, so the
// maps are deterministic whatever the scheduling: the nearest depth wins and
// equal depths keep the lower surfel index, as the reference's depth test
// does in draw order.  Both passes stream their arrays once (HBM bound:
// 64 B of surfel attributes read per surfel, 8 B atomic per drawn surfel,
// ~88 B written per pixel).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/mvdr/render.h"
#include "common.h"

namespace mvdr {
namespace {

struct RenderViews {
    uint32_t V, W, H, S;
    float fx[kMaxViews], fy[kMaxViews], cx[kMaxViews], cy[kMaxViews];
    Mat34 w2c[kMaxViews];
    const float4* ref_v[kMaxViews];
    const float4* ref_n[kMaxViews];
    const float4* warp_v[kMaxViews];
    const float4* warp_n[kMaxViews];
    const float4* color_time;
    float max_depth, now, conf_thr, time_thr;
};

struct RenderMaps {
    float4* ref_v[kMaxViews];
    float4* ref_n[kMaxViews];
    float4* warp_v[kMaxViews];
    float4* warp_n[kMaxViews];
    uint32_t* index[kMaxViews];
    float4* rgb[kMaxViews];
    unsigned long long* z[kMaxViews];
};

__global__ void __launch_bounds__(256) zclear_kernel(RenderMaps m, uint32_t V, uint32_t pixels) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < V * pixels; p += gridDim.x * blockDim.x) {
        const uint32_t v = p / pixels;
        m.z[v][p - v * pixels] = ~0ull;
    }
}

__global__ void __launch_bounds__(256) splat_kernel(RenderViews r, RenderMaps m) {
    const uint32_t total = r.V * r.S;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const uint32_t v = g / r.S, s = g - v * r.S;
        const f3 p = r.w2c[v].apply(xyz(r.warp_v[v][s]));
        if (p.z > r.max_depth + 0.05f || p.z < 0.f) continue;
        if (r.ref_v[v][s].w < r.conf_thr && fabsf(r.color_time[s].z - r.now) >= r.time_thr) continue;
        const float u = (r.fx[v] * p.x) / p.z + r.cx[v];
        const float w = (r.fy[v] * p.y) / p.z + r.cy[v];
        if (!(u >= 0.f && w >= 0.f && u < (float)r.W && w < (float)r.H)) continue;
        const uint32_t px = (uint32_t)floorf(u), py = (uint32_t)floorf(w);
        atomicMin(&m.z[v][py * r.W + px], ((unsigned long long)__float_as_uint(p.z) << 32) | s);
    }
}

__device__ inline float4 decode_rgb(float4 ct) {
    const int c = __float_as_int(ct.x);
    return make_float4((float)((c & 0x00FF0000) >> 16) / 255.f, (float)((c & 0x0000FF00) >> 8) / 255.f,
                       (float)(c & 0x000000FF) / 255.f, ct.y);
}

__global__ void __launch_bounds__(256) resolve_kernel(RenderViews r, RenderMaps m) {
    const uint32_t pixels = r.W * r.H;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < r.V * pixels; g += gridDim.x * blockDim.x) {
        const uint32_t v = g / pixels, o = g - v * pixels;
        const unsigned long long key = m.z[v][o];
        const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 rv = zero, rn = zero, wv = zero, wn = zero, rgb = zero;
        uint32_t idx = kInvalidIndex;
        if (key != ~0ull) {
            idx = (uint32_t)(key & 0xFFFFFFFFull);
            const float4 a = r.ref_v[v][idx], b = r.ref_n[v][idx], c = r.warp_v[v][idx], d = r.warp_n[v][idx];
            rv = make_float4(a.x, a.y, a.z, 1.f);
            rn = make_float4(b.x, b.y, b.z, 0.f);
            wv = make_float4(c.x, c.y, c.z, 1.f);
            wn = make_float4(d.x, d.y, d.z, 0.f);
            rgb = decode_rgb(r.color_time[idx]);
        }
        m.ref_v[v][o] = rv;
        m.ref_n[v][o] = rn;
        m.warp_v[v][o] = wv;
        if (m.warp_n[v]) m.warp_n[v][o] = wn;
        m.index[v][o] = idx;
        if (m.rgb[v]) m.rgb[v][o] = rgb;
    }
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)); }

}  // namespace
}  // namespace mvdr

using namespace mvdr;

extern "C" mvdr_status mvdr_render_solver_maps(const mvdr_render_input* in, const mvdr_render_output* out,
                                               void* stream) {
    if (!in || !out) return MVDR_E_INVALID;
    if (in->num_views < 1 || in->num_views > MVDR_MAX_VIEWS || in->width < 1 || in->height < 1 ||
        in->width > 65535 || in->height > 65535)
        return MVDR_E_SHAPE;
    if ((uint64_t)in->num_views * in->width * in->height > 0x7FFFFFFFull ||
        (uint64_t)in->num_views * in->num_surfels > 0x7FFFFFFFull)
        return MVDR_E_SHAPE;
    RenderViews r;
    RenderMaps m;
    std::memset(&r, 0, sizeof(r));
    std::memset(&m, 0, sizeof(m));
    r.V = in->num_views;
    r.W = in->width;
    r.H = in->height;
    r.S = in->num_surfels;
    r.color_time = static_cast<const float4*>(in->color_time);
    if (r.S && !r.color_time) return MVDR_E_INVALID;
    r.max_depth = in->max_depth;
    r.now = in->current_time;
    r.conf_thr = in->confidence_threshold;
    r.time_thr = in->time_threshold;
    for (uint32_t v = 0; v < r.V; ++v) {
        r.fx[v] = in->intrinsic[v].fx;
        r.fy[v] = in->intrinsic[v].fy;
        r.cx[v] = in->intrinsic[v].cx;
        r.cy[v] = in->intrinsic[v].cy;
        r.w2c[v] = to_mat34(in->world2camera[v]);
        r.ref_v[v] = static_cast<const float4*>(in->reference_vertex[v]);
        r.ref_n[v] = static_cast<const float4*>(in->reference_normal[v]);
        r.warp_v[v] = static_cast<const float4*>(in->warp_vertex[v]);
        r.warp_n[v] = static_cast<const float4*>(in->warp_normal[v]);
        if (r.S && (!r.ref_v[v] || !r.ref_n[v] || !r.warp_v[v] || !r.warp_n[v])) return MVDR_E_INVALID;
        m.ref_v[v] = static_cast<float4*>(out->reference_vertex[v]);
        m.ref_n[v] = static_cast<float4*>(out->reference_normal[v]);
        m.warp_v[v] = static_cast<float4*>(out->warp_vertex[v]);
        m.warp_n[v] = static_cast<float4*>(out->warp_normal[v]);
        m.index[v] = static_cast<uint32_t*>(out->index_map[v]);
        m.rgb[v] = static_cast<float4*>(out->normalized_rgb[v]);
        m.z[v] = static_cast<unsigned long long*>(out->zbuffer[v]);
        if (!m.ref_v[v] || !m.ref_n[v] || !m.warp_v[v] || !m.index[v] || !m.z[v]) return MVDR_E_INVALID;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t pixels = r.W * r.H;
    hipLaunchKernelGGL(zclear_kernel, dim3(grid_for((uint64_t)r.V * pixels)), dim3(256), 0, st, m, r.V, pixels);
    if (r.S) hipLaunchKernelGGL(splat_kernel, dim3(grid_for((uint64_t)r.V * r.S)), dim3(256), 0, st, r, m);
    hipLaunchKernelGGL(resolve_kernel, dim3(grid_for((uint64_t)r.V * pixels)), dim3(256), 0, st, r, m);
    return hipGetLastError() == hipSuccess ? MVDR_OK : MVDR_E_HIP;
}
