// Rigid multi-view point-to-plane ICP (include/mvdr/rigid.h).
//
// Per iteration, two launches for all views at once:
//   rigid_rows_kernel  one 256-thread workgroup per 256-pixel block and view:
//                      each pixel's A row and b, the 27 products summed in FP64
//                      through the wave (shuffles) and the workgroup (LDS) into
//                      one partial per block -- the pixel maps are read once
//                      (HBM bound: 32 B conversion + 32 B gathered reference
//                      per pixel);
//   rigid_solve_kernel one workgroup per view: the block partials summed in a
//                      fixed order, the 6x6 system solved by an FP64 Cholesky
//                      in registers, the pose updated in device memory -- the
//                      next iteration reads it there, so the host never waits.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/mvdr/rigid.h"
#include "common.h"
#include "device_util.h"

namespace mvdr {
namespace {

constexpr uint32_t kRigidBlock = 256;   // pixels per block (RigidSolverDevice::blockSize)
constexpr int kSums = 27;               // 21 of A^T A (upper, row-major) + 6 of A^T b

struct RigidViews {
    uint32_t V, W, H, blocks;           // blocks per view
    int reference_reduction;
    float fx[kMaxViews], fy[kMaxViews], cx[kMaxViews], cy[kMaxViews];
    const float4* ref_v[kMaxViews];
    const float4* ref_n[kMaxViews];
    const float4* conv_v[kMaxViews];
    const float4* conv_n[kMaxViews];
    Mat34* pose;                        // world2camera per view (device, updated in place)
    double* partial;                    // V x blocks x kSums
};

// __float2uint_rn into the reference's ushort2 pixel: negative -> 0, NaN -> 0,
// beyond u32 -> 0xFFFFFFFF, then the low 16 bits.
__device__ inline uint32_t pixel_u16(float u) {
    uint32_t r;
    if (!(u > 0.f)) r = 0u;
    else if (u >= 4294967040.f) r = 0xFFFFFFFFu;
    else r = (uint32_t)rintf(u);
    return r & 0xFFFFu;
}

__device__ inline bool zero_vertex4(float4 v) { return fabsf(v.x) < 1e-3f && fabsf(v.y) < 1e-3f && fabsf(v.z) < 1e-3f; }

__global__ void __launch_bounds__(kRigidBlock) rigid_rows_kernel(RigidViews r) {
    const uint32_t v = blockIdx.y, blk = blockIdx.x;
    const uint32_t pixel = blk * kRigidBlock + threadIdx.x;
    double row[6] = {0, 0, 0, 0, 0, 0}, b = 0.0;
    const bool take = !r.reference_reduction || (blk % 32u) == 0u;
    if (take && pixel < r.W * r.H) {
        const Mat34 T = r.pose[v];
        const float4 cv = r.conv_v[v][pixel], cn = r.conv_n[v][pixel];
        const f3 p = T.apply(xyz(cv)), pn = T.rot(xyz(cn));
        const float u = (float)((double)p.x / ((double)p.z + 1e-10) * (double)r.fx[v] + (double)r.cx[v]);
        const float w = (float)((double)p.y / ((double)p.z + 1e-10) * (double)r.fy[v] + (double)r.cy[v]);
        const uint32_t x = pixel_u16(u), y = pixel_u16(w);
        if (x < r.W && y < r.H) {
            const float4 q = r.ref_v[v][y * r.W + x], qn = r.ref_n[v][y * r.W + x];
            const float ndot = pn.x * qn.x + pn.y * qn.y + pn.z * qn.z;
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (!(zero_vertex4(q) || ndot < 0.8f || d2 > 0.03f * 0.03f)) {
                const f3 n = mk(qn.x, qn.y, qn.z);
                const f3 pxn = cross3(p, n);
                b = -(double)(n.x * dx + n.y * dy + n.z * dz);
                row[0] = pxn.x; row[1] = pxn.y; row[2] = pxn.z;
                row[3] = n.x; row[4] = n.y; row[5] = n.z;
            }
        }
    }
    // 27 block sums: wave butterflies, then the four waves through LDS
    __shared__ double part[kSums][kRigidBlock / 64];
    int s = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            const double t = wave_sum(row[i] * row[j]);
            if ((threadIdx.x & 63) == 0) part[s][threadIdx.x >> 6] = t;
            ++s;
        }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double t = wave_sum(row[i] * b);
        if ((threadIdx.x & 63) == 0) part[s][threadIdx.x >> 6] = t;
        ++s;
    }
    __syncthreads();
    if (threadIdx.x < kSums) {
        double t = 0.0;
        for (uint32_t w = 0; w < kRigidBlock / 64; ++w) t += part[threadIdx.x][w];
        r.partial[((size_t)v * r.blocks + blk) * kSums + threadIdx.x] = t;
    }
}

// mat34(const float3& twist_rot, const float3& twist_trans) (MatUtils.cpp:11-40)
__device__ inline Mat34 twist_transform(f3 w, f3 t) {
    Mat34 R;
    R.c0 = mk(1, 0, 0); R.c1 = mk(0, 1, 0); R.c2 = mk(0, 0, 1);
    if (!(fabsf(w.x) + fabsf(w.y) + fabsf(w.z) < 1e-4f)) {
        const float angle = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
        const f3 a = (1.0f / angle) * w;
        const float c = cosf(angle), s = sinf(angle), k = 1.0f - c;
        R.c0 = mk(k * a.x * a.x + c, k * a.x * a.y + a.z * s, k * a.x * a.z - a.y * s);
        R.c1 = mk(k * a.x * a.y - a.z * s, k * a.y * a.y + c, k * a.y * a.z + a.x * s);
        R.c2 = mk(k * a.x * a.z + a.y * s, k * a.y * a.z - a.x * s, k * a.z * a.z + c);
    }
    R.t = t;
    return R;
}

__device__ inline Mat34 compose(const Mat34& A, const Mat34& B) {
    Mat34 o;
    o.c0 = A.rot(B.c0);
    o.c1 = A.rot(B.c1);
    o.c2 = A.rot(B.c2);
    o.t = A.rot(B.t) + A.t;
    return o;
}

__global__ void __launch_bounds__(64) rigid_solve_kernel(RigidViews r) {
    const uint32_t v = blockIdx.x;
    __shared__ double sums[kSums];
    if (threadIdx.x < kSums) {
        double t = 0.0;
        for (uint32_t blk = 0; blk < r.blocks; ++blk) t += r.partial[((size_t)v * r.blocks + blk) * kSums + threadIdx.x];
        // the reference's sampled reduction scales the sampled partials by the lane count
        sums[threadIdx.x] = r.reference_reduction ? 32.0 * t : t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double A[6][6], y[6];
    int s = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) { A[i][j] = A[j][i] = sums[s]; ++s; }
    for (int i = 0; i < 6; ++i) y[i] = sums[s + i];
    // Cholesky A = L L^T (lower triangle in place); a non-positive pivot
    // (no pixel pairs) leaves the pose unchanged
    bool ok = true;
    for (int j = 0; j < 6 && ok; ++j) {
        double d = A[j][j];
        for (int k = 0; k < j; ++k) d -= A[j][k] * A[j][k];
        if (!(d > 0.0)) { ok = false; break; }
        const double l = sqrt(d);
        A[j][j] = l;
        for (int i = j + 1; i < 6; ++i) {
            double t = A[i][j];
            for (int k = 0; k < j; ++k) t -= A[i][k] * A[j][k];
            A[i][j] = t / l;
        }
    }
    if (!ok) return;
    for (int i = 0; i < 6; ++i) {
        double t = y[i];
        for (int k = 0; k < i; ++k) t -= A[i][k] * y[k];
        y[i] = t / A[i][i];
    }
    for (int i = 5; i >= 0; --i) {
        double t = y[i];
        for (int k = i + 1; k < 6; ++k) t -= A[k][i] * y[k];
        y[i] = t / A[i][i];
    }
    const Mat34 U = twist_transform(mk((float)y[0], (float)y[1], (float)y[2]), mk((float)y[3], (float)y[4], (float)y[5]));
    r.pose[v] = compose(U, r.pose[v]);
}

struct PoseSet {
    Mat34 m[kMaxViews];
};

__global__ void rigid_start_kernel(PoseSet start, Mat34* pose, uint32_t V) {
    if (threadIdx.x < V) pose[threadIdx.x] = start.m[threadIdx.x];
}

}  // namespace
}  // namespace mvdr

using namespace mvdr;

struct mvdr_rigid {
    mvdr_rigid_config cfg{};
    void* block = nullptr;
    Mat34* pose = nullptr;
    double* partial = nullptr;
    uint32_t views = 0;
    mvdr_mat34 relative[MVDR_MAX_VIEWS];
    hipEvent_t done = nullptr;
    bool solved = false;
    ~mvdr_rigid() {
        if (done) (void)hipEventDestroy(done);
        if (block) (void)hipFree(block);
    }
};

extern "C" {

mvdr_status mvdr_rigid_create(const mvdr_rigid_config* c, mvdr_rigid** out) {
    if (!c || !out) return MVDR_E_INVALID;
    *out = nullptr;
    if (c->max_views < 1 || c->max_views > MVDR_MAX_VIEWS || c->max_width < 1 || c->max_height < 1 ||
        c->max_width > 65535 || c->max_height > 65535 || c->reference_reduction > 1)
        return MVDR_E_SHAPE;
    auto* h = new mvdr_rigid();
    h->cfg = *c;
    const size_t blocks = ((size_t)c->max_width * c->max_height + kRigidBlock - 1) / kRigidBlock;
    const size_t pose_bytes = 256;
    const size_t bytes = pose_bytes + (size_t)c->max_views * blocks * kSums * sizeof(double);
    hipError_t e = hipMalloc(&h->block, bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete h;
        return MVDR_E_HIP;
    }
    h->pose = static_cast<Mat34*>(h->block);
    h->partial = reinterpret_cast<double*>(static_cast<char*>(h->block) + pose_bytes);
    *out = h;
    return MVDR_OK;
}

void mvdr_rigid_destroy(mvdr_rigid* h) {
    if (!h) return;
    if (h->done) (void)hipEventSynchronize(h->done);
    delete h;
}

mvdr_status mvdr_rigid_solve(mvdr_rigid* h, const mvdr_rigid_input* in, uint32_t iterations, void* stream) {
    if (!h || !in) return MVDR_E_INVALID;
    if (in->num_views < 1 || in->num_views > h->cfg.max_views || in->width < 1 || in->height < 1 ||
        in->width > h->cfg.max_width || in->height > h->cfg.max_height)
        return MVDR_E_SHAPE;
    RigidViews r;
    std::memset(&r, 0, sizeof(r));
    r.V = in->num_views;
    r.W = in->width;
    r.H = in->height;
    r.blocks = (r.W * r.H + kRigidBlock - 1) / kRigidBlock;
    r.reference_reduction = (int)h->cfg.reference_reduction;
    r.pose = h->pose;
    r.partial = h->partial;
    PoseSet start;
    for (uint32_t v = 0; v < r.V; ++v) {
        if (!in->reference_vertex[v] || !in->reference_normal[v] || !in->conversion_vertex[v] ||
            !in->conversion_normal[v])
            return MVDR_E_INVALID;
        r.fx[v] = in->intrinsic[v].fx;
        r.fy[v] = in->intrinsic[v].fy;
        r.cx[v] = in->intrinsic[v].cx;
        r.cy[v] = in->intrinsic[v].cy;
        r.ref_v[v] = static_cast<const float4*>(in->reference_vertex[v]);
        r.ref_n[v] = static_cast<const float4*>(in->reference_normal[v]);
        r.conv_v[v] = static_cast<const float4*>(in->conversion_vertex[v]);
        r.conv_n[v] = static_cast<const float4*>(in->conversion_normal[v]);
        start.m[v] = to_mat34(in->world2camera[v]);
        h->relative[v] = in->relative_pose[v];
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(rigid_start_kernel, dim3(1), dim3(64), 0, st, start, h->pose, r.V);
    for (uint32_t it = 0; it < iterations; ++it) {
        hipLaunchKernelGGL(rigid_rows_kernel, dim3(r.blocks, r.V), dim3(kRigidBlock), 0, st, r);
        hipLaunchKernelGGL(rigid_solve_kernel, dim3(r.V), dim3(64), 0, st, r);
    }
    if (hipGetLastError() != hipSuccess) return MVDR_E_HIP;
    if (hipEventRecord(h->done, st) != hipSuccess) return MVDR_E_HIP;
    h->views = r.V;
    h->solved = true;
    return MVDR_OK;
}

mvdr_status mvdr_rigid_get_pose(mvdr_rigid* h, mvdr_mat34* w2c, mvdr_mat34* c2l) {
    if (!h || !w2c) return MVDR_E_INVALID;
    if (!h->solved) return MVDR_E_STATE;
    if (hipEventSynchronize(h->done) != hipSuccess) return MVDR_E_HIP;
    Mat34 pose[MVDR_MAX_VIEWS];
    if (hipMemcpy(pose, h->pose, h->views * sizeof(Mat34), hipMemcpyDeviceToHost) != hipSuccess) return MVDR_E_HIP;
    auto put = [](const Mat34& m, mvdr_mat34& o) {
        const f3 cols[3] = {m.c0, m.c1, m.c2};
        for (int c = 0; c < 3; ++c) {
            o.rot[3 * c] = cols[c].x;
            o.rot[3 * c + 1] = cols[c].y;
            o.rot[3 * c + 2] = cols[c].z;
        }
        o.trans[0] = m.t.x; o.trans[1] = m.t.y; o.trans[2] = m.t.z;
    };
    for (uint32_t v = 0; v < h->views; ++v) put(pose[v], w2c[v]);
    if (c2l) {
        // AverageCanonicalFieldRigidSE3 (RigidSolver.cpp:279-308): view 0's pose
        // plus relative * pose * relative^-1 of the others, averaged entrywise
        float rot[9], trans[3];
        for (uint32_t v = 0; v < h->views; ++v) {
            Mat34 m = pose[v];
            if (v > 0) {
                const Mat34 rel = to_mat34(h->relative[v]);
                auto mul = [](const Mat34& A, const Mat34& B) {
                    Mat34 o;
                    o.c0 = A.rot(B.c0); o.c1 = A.rot(B.c1); o.c2 = A.rot(B.c2);
                    o.t = A.rot(B.t) + A.t;
                    return o;
                };
                m = mul(rel, mul(pose[v], inverse(rel)));
            }
            const float cur[9] = {m.c0.x, m.c0.y, m.c0.z, m.c1.x, m.c1.y, m.c1.z, m.c2.x, m.c2.y, m.c2.z};
            for (int k = 0; k < 9; ++k) rot[k] = v == 0 ? cur[k] : rot[k] + cur[k];
            const float ct[3] = {m.t.x, m.t.y, m.t.z};
            for (int k = 0; k < 3; ++k) trans[k] = v == 0 ? ct[k] : trans[k] + ct[k];
        }
        for (int k = 0; k < 9; ++k) c2l->rot[k] = rot[k] / (float)h->views;
        for (int k = 0; k < 3; ++k) c2l->trans[k] = trans[k] / (float)h->views;
    }
    return MVDR_OK;
}

}  // extern "C"
