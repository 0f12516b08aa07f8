// Seeded synthetic scene generator for the non-rigid solve (host only).
//
// Produces everything mvdr_frame_input needs.  Geometry conventions follow the
// reference solver (core/NonRigidSolver/DenseDepthJacobianResidual.cu:24-104):
//   * canonical surfels and nodes live in the camera-0 frame;
//   * initial_camera_se3[v] maps camera-v coordinates into camera 0;
//   * world2camera[v] is the rigid canonical->live alignment (same for all views);
//   * reference / warp vertex maps hold camera-v coordinates;
//   * the observed vertex / normal maps hold camera-v coordinates.
// Node graph: 8-NN with self excluded, WarpField.cu:15-59.
// Skinning: 4-NN, w = exp(-d^2 / (2 * 0.025^2)) (CanonicalNodesSkinner.cu:154-166),
// normalised to sum 1 so the blended dual quaternion stays well scaled.
#include "../../../include/mvdr/scene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_last_error;

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
};
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(const V3& a) { double n = norm(a); return n > 0 ? a * (1.0 / n) : a; }

// Row-major 3x3 rotation + translation.
struct Rigid {
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    V3 t;
    V3 apply(const V3& p) const { return rot(p) + t; }
    V3 rot(const V3& p) const {
        return {R[0] * p.x + R[1] * p.y + R[2] * p.z, R[3] * p.x + R[4] * p.y + R[5] * p.z,
                R[6] * p.x + R[7] * p.y + R[8] * p.z};
    }
    V3 rot_t(const V3& p) const {
        return {R[0] * p.x + R[3] * p.y + R[6] * p.z, R[1] * p.x + R[4] * p.y + R[7] * p.z,
                R[2] * p.x + R[5] * p.y + R[8] * p.z};
    }
    Rigid inverse() const {
        Rigid o;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) o.R[3 * r + c] = R[3 * c + r];
        o.t = o.rot(t) * -1.0;
        return o;
    }
    Rigid compose(const Rigid& b) const {  // this * b
        Rigid o;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += R[3 * r + k] * b.R[3 * k + c];
                o.R[3 * r + c] = s;
            }
        o.t = rot(b.t) + t;
        return o;
    }
};

Rigid axis_angle(const V3& w) {
    Rigid o;
    double a = norm(w);
    if (a < 1e-12) return o;
    V3 k = w * (1.0 / a);
    double c = std::cos(a), s = std::sin(a), t = 1 - c;
    o.R[0] = t * k.x * k.x + c;       o.R[1] = t * k.x * k.y - k.z * s; o.R[2] = t * k.x * k.z + k.y * s;
    o.R[3] = t * k.x * k.y + k.z * s; o.R[4] = t * k.y * k.y + c;       o.R[5] = t * k.y * k.z - k.x * s;
    o.R[6] = t * k.x * k.z - k.y * s; o.R[7] = t * k.y * k.z + k.x * s; o.R[8] = t * k.z * k.z + c;
    return o;
}

mvdr_mat34 to_mat34(const Rigid& r) {
    mvdr_mat34 m;
    for (int c = 0; c < 3; ++c)
        for (int row = 0; row < 3; ++row) m.rot[3 * c + row] = (float)r.R[3 * row + c];
    m.trans[0] = (float)r.t.x; m.trans[1] = (float)r.t.y; m.trans[2] = (float)r.t.z;
    return m;
}

// Rotation matrix -> unit quaternion (x, y, z, w), same branch structure as
// Quaternion(mat33) (math/DualQuaternion/Quaternion.h:46-77), in double.
void quat_from_rot(const Rigid& r, double q[4]) {
    auto m = [&](int i, int j) { return r.R[3 * i + j]; };
    double tr = m(0, 0) + m(1, 1) + m(2, 2);
    if (tr > 0) {
        double s = std::sqrt(tr + 1.0) * 2;
        q[3] = 0.25 * s; q[0] = (m(2, 1) - m(1, 2)) / s; q[1] = (m(0, 2) - m(2, 0)) / s; q[2] = (m(1, 0) - m(0, 1)) / s;
    } else if (m(0, 0) > m(1, 1) && m(0, 0) > m(2, 2)) {
        double s = std::sqrt(1.0 + m(0, 0) - m(1, 1) - m(2, 2)) * 2;
        q[3] = (m(2, 1) - m(1, 2)) / s; q[0] = 0.25 * s; q[1] = (m(0, 1) + m(1, 0)) / s; q[2] = (m(0, 2) + m(2, 0)) / s;
    } else if (m(1, 1) > m(2, 2)) {
        double s = std::sqrt(1.0 + m(1, 1) - m(0, 0) - m(2, 2)) * 2;
        q[3] = (m(0, 2) - m(2, 0)) / s; q[0] = (m(0, 1) + m(1, 0)) / s; q[1] = 0.25 * s; q[2] = (m(1, 2) + m(2, 1)) / s;
    } else {
        double s = std::sqrt(1.0 + m(2, 2) - m(0, 0) - m(1, 1)) * 2;
        q[3] = (m(1, 0) - m(0, 1)) / s; q[0] = (m(0, 2) + m(2, 0)) / s; q[1] = (m(1, 2) + m(2, 1)) / s; q[2] = 0.25 * s;
    }
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] /= n;
}

// Quaternion product a*b with (x, y, z, w) storage.
void qmul(const double a[4], const double b[4], double o[4]) {
    o[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
    o[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
    o[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
    o[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
}

mvdr_dq dq_from_rigid(const Rigid& r) {
    double q0[4];
    quat_from_rot(r, q0);
    const double tq[4] = {0.5 * r.t.x, 0.5 * r.t.y, 0.5 * r.t.z, 0.0};
    double q1[4];
    qmul(tq, q0, q1);  // dual part = 0.5 * t * q0
    mvdr_dq dq;
    for (int i = 0; i < 4; ++i) { dq.q0[i] = (float)q0[i]; dq.q1[i] = (float)q1[i]; }
    return dq;
}

struct Rng {
    std::mt19937_64 gen;
    explicit Rng(uint64_t seed) : gen(seed) {}
    double uniform() { return (double)(gen() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {  // Box-Muller, one value per call
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    V3 unit_vector() {
        double z = uniform(-1, 1), phi = uniform(0, 2 * M_PI), s = std::sqrt(std::max(0.0, 1 - z * z));
        return {s * std::cos(phi), s * std::sin(phi), z};
    }
    V3 in_ball() {
        V3 d = unit_vector();
        return d * std::cbrt(uniform());
    }
    uint64_t below(uint64_t n) { return n ? gen() % n : 0; }
};

struct Surfel {
    V3 p, n;
    double radius;
    double color[3];
};

// ---------------------------------------------------------------- geometry
void sample_sphere(std::vector<Surfel>& out, const V3& c, double r, uint32_t count, Rng& rng) {
    const double golden = M_PI * (3.0 - std::sqrt(5.0));
    const double spacing = std::sqrt(4 * M_PI * r * r / count);
    for (uint32_t i = 0; i < count; ++i) {
        double y = 1 - 2 * (i + 0.5) / count;
        double rad = std::sqrt(std::max(0.0, 1 - y * y));
        double th = golden * i;
        V3 d(std::cos(th) * rad, y, std::sin(th) * rad);
        Surfel s;
        s.n = normalized(d + rng.unit_vector() * 0.02);
        s.p = c + d * r;
        s.radius = 0.75 * spacing;
        out.push_back(s);
    }
}

// Open cylinder around `axis` from `base`, plus flat caps when requested.
void sample_cylinder(std::vector<Surfel>& out, const V3& base, const V3& axis_dir, double r, double len, double spacing,
                     bool caps, Rng& rng) {
    V3 a = normalized(axis_dir);
    V3 tmp = std::fabs(a.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
    V3 u = normalized(cross(a, tmp)), v = cross(a, u);
    int nt = std::max(8, (int)std::ceil(2 * M_PI * r / spacing));
    int nh = std::max(2, (int)std::ceil(len / spacing));
    for (int ih = 0; ih < nh; ++ih)
        for (int it = 0; it < nt; ++it) {
            double th = 2 * M_PI * (it + 0.5 * (ih & 1)) / nt + rng.uniform(-0.1, 0.1) / nt;
            double h = (ih + rng.uniform(0.2, 0.8)) * len / nh;
            V3 radial = u * std::cos(th) + v * std::sin(th);
            Surfel s;
            s.p = base + a * h + radial * r;
            s.n = normalized(radial + rng.unit_vector() * 0.02);
            s.radius = 0.75 * spacing;
            out.push_back(s);
        }
    if (caps) {
        for (int side = 0; side < 2; ++side) {
            V3 centre = base + a * (side ? len : 0.0);
            V3 nrm = a * (side ? 1.0 : -1.0);
            int nr = std::max(1, (int)std::ceil(r / spacing));
            for (int ir = 0; ir < nr; ++ir) {
                double rr = (ir + 0.5) * r / nr;
                int nth = std::max(4, (int)std::ceil(2 * M_PI * rr / spacing));
                for (int it = 0; it < nth; ++it) {
                    double th = 2 * M_PI * (it + rng.uniform(0, 0.5)) / nth;
                    Surfel s;
                    s.p = centre + (u * std::cos(th) + v * std::sin(th)) * rr;
                    s.n = nrm;
                    s.radius = 0.75 * spacing;
                    out.push_back(s);
                }
            }
        }
    }
}

void sample_sphere_spacing(std::vector<Surfel>& out, const V3& c, double r, double spacing, Rng& rng) {
    uint32_t count = (uint32_t)std::max(16.0, 4 * M_PI * r * r / (spacing * spacing));
    sample_sphere(out, c, r, count, rng);
    for (size_t i = out.size() - count; i < out.size(); ++i) out[i].radius = 0.75 * spacing;
}

// Body centred at `c` (y axis points down, the camera convention).
void sample_body(std::vector<Surfel>& out, const V3& c, uint32_t target, Rng& rng) {
    const double tr = 0.13, th = 0.50;  // torso radius / height
    const double hr = 0.09;             // head radius
    const double ar = 0.04, al = 0.42;  // arms
    const double lr = 0.055, ll = 0.42; // legs
    double area = 2 * M_PI * tr * th + 2 * M_PI * tr * tr + 4 * M_PI * hr * hr + 2 * (2 * M_PI * ar * al) +
                  2 * (2 * M_PI * lr * ll);
    double spacing = std::sqrt(area / target);
    V3 up(0, -1, 0);
    V3 torso_base = c + V3(0, th / 2, 0);
    sample_cylinder(out, torso_base, up, tr, th, spacing, true, rng);
    sample_sphere_spacing(out, c + V3(0, -th / 2 - hr - 0.01, 0), hr, spacing, rng);
    for (int side = -1; side <= 1; side += 2) {
        V3 shoulder = c + V3(side * (tr + ar + 0.005), -th / 2 + 0.04, 0);
        sample_cylinder(out, shoulder, V3(side * 0.35, 1.0, 0.15), ar, al, spacing, true, rng);
        V3 hip = c + V3(side * 0.065, th / 2 + 0.01, 0);
        sample_cylinder(out, hip, V3(side * 0.12, 1.0, -0.05), lr, ll, spacing, true, rng);
    }
}

// --------------------------------------------------------------- the scene
struct View {
    mvdr_intrinsic intr;
    Rigid init;        // camera v -> camera 0
    Rigid world2cam;   // canonical -> live (rigid)
    std::vector<float> vertex, normal, filtered_fg, fg_grad, gray, gray_grad, ref_vertex, ref_normal, warp_vertex, rgb;
    std::vector<uint8_t> fg_mask, edge;
    std::vector<uint32_t> index;
    std::vector<uint16_t> features;  // ushort4 per pair
    std::vector<uint16_t> corr;      // ushort4 per pixel: the feature pair ending there, else 0
};

}  // namespace

struct mvdr_scene {
    mvdr_scene_params params;
    uint32_t W = 0, H = 0;
    std::vector<Surfel> surfels;
    std::vector<float> surfel_pos, surfel_normal, surfel_w, nodes;
    std::vector<uint16_t> surfel_knn, graph;
    std::vector<mvdr_dq> se3_init, se3_gt;
    std::vector<mvdr_cross_pair> cross;
    std::vector<View> views;
    double voxel = 0;
};

namespace {

// Voxel subsampling: one node per occupied voxel, the surfel closest to the
// voxel's centroid; nodes ordered by voxel key for determinism.
std::vector<uint32_t> voxel_subsample(const std::vector<Surfel>& s, double voxel) {
    struct Acc { V3 sum; uint32_t count = 0; };
    std::map<std::tuple<long, long, long>, Acc> cells;
    auto key = [&](const V3& p) {
        return std::make_tuple((long)std::floor(p.x / voxel), (long)std::floor(p.y / voxel), (long)std::floor(p.z / voxel));
    };
    for (const auto& x : s) { auto& a = cells[key(x.p)]; a.sum += x.p; a.count++; }
    std::map<std::tuple<long, long, long>, std::pair<double, uint32_t>> best;
    for (uint32_t i = 0; i < s.size(); ++i) {
        auto k = key(s[i].p);
        const auto& a = cells[k];
        V3 cen = a.sum * (1.0 / a.count);
        double d = dot(s[i].p - cen, s[i].p - cen);
        auto it = best.find(k);
        if (it == best.end() || d < it->second.first) best[k] = {d, i};
    }
    std::vector<uint32_t> out;
    out.reserve(best.size());
    for (auto& kv : best) out.push_back(kv.second.second);
    return out;
}

// Uniform grid over node positions for k-NN queries.
struct Grid {
    double cell;
    V3 lo;
    int nx, ny, nz;
    std::vector<std::vector<uint32_t>> bins;
    const std::vector<V3>* pts;
    Grid(const std::vector<V3>& p, double c) : cell(c), pts(&p) {
        V3 mn(1e30, 1e30, 1e30), mx(-1e30, -1e30, -1e30);
        for (auto& q : p) {
            mn = {std::min(mn.x, q.x), std::min(mn.y, q.y), std::min(mn.z, q.z)};
            mx = {std::max(mx.x, q.x), std::max(mx.y, q.y), std::max(mx.z, q.z)};
        }
        lo = mn;
        nx = (int)((mx.x - mn.x) / cell) + 1; ny = (int)((mx.y - mn.y) / cell) + 1; nz = (int)((mx.z - mn.z) / cell) + 1;
        bins.assign((size_t)nx * ny * nz, {});
        for (uint32_t i = 0; i < p.size(); ++i) {
            int ix, iy, iz; locate(p[i], ix, iy, iz);
            bins[((size_t)iz * ny + iy) * nx + ix].push_back(i);
        }
    }
    void locate(const V3& q, int& ix, int& iy, int& iz) const {
        ix = std::min(nx - 1, std::max(0, (int)((q.x - lo.x) / cell)));
        iy = std::min(ny - 1, std::max(0, (int)((q.y - lo.y) / cell)));
        iz = std::min(nz - 1, std::max(0, (int)((q.z - lo.z) / cell)));
    }
    // k nearest points (ascending distance, ties by index); exclude index `skip`.
    void knn(const V3& q, int k, uint32_t skip, std::vector<std::pair<double, uint32_t>>& out) const {
        int ix, iy, iz; locate(q, ix, iy, iz);
        for (int ring = 1;; ++ring) {
            out.clear();
            for (int z = iz - ring; z <= iz + ring; ++z)
                for (int y = iy - ring; y <= iy + ring; ++y)
                    for (int x = ix - ring; x <= ix + ring; ++x) {
                        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
                        for (uint32_t j : bins[((size_t)z * ny + y) * nx + x]) {
                            if (j == skip) continue;
                            V3 d = (*pts)[j] - q;
                            out.push_back({dot(d, d), j});
                        }
                    }
            // the k-th neighbour is trustworthy once it lies within the searched radius
            if ((int)out.size() >= k) {
                std::partial_sort(out.begin(), out.begin() + k, out.end());
                double reach = ring * cell;
                bool covers = ix - ring <= 0 && iy - ring <= 0 && iz - ring <= 0 && ix + ring >= nx - 1 &&
                              iy + ring >= ny - 1 && iz + ring >= nz - 1;
                if (out[k - 1].first <= reach * reach || covers) { out.resize(k); return; }
            } else {
                bool covers = ix - ring <= 0 && iy - ring <= 0 && iz - ring <= 0 && ix + ring >= nx - 1 &&
                              iy + ring >= ny - 1 && iz + ring >= nz - 1;
                if (covers) { std::sort(out.begin(), out.end()); return; }
            }
        }
    }
};

struct RenderTarget {
    std::vector<double> depth;
    std::vector<int32_t> id;
    std::vector<V3> vtx;
};

// Splat oriented surfels into a W x H z-buffer: depth from the ray / tangent
// plane intersection, square footprint of the projected radius.
void splat(const std::vector<V3>& pos, const std::vector<V3>& nrm, const std::vector<double>& radius,
           const mvdr_intrinsic& K, uint32_t W, uint32_t H, RenderTarget& rt) {
    rt.depth.assign((size_t)W * H, 1e30);
    rt.id.assign((size_t)W * H, -1);
    rt.vtx.assign((size_t)W * H, V3());
    for (uint32_t s = 0; s < pos.size(); ++s) {
        const V3& p = pos[s];
        const V3& n = nrm[s];
        if (p.z < 0.1) continue;
        if (dot(n, p) >= 0) continue;  // back facing
        double u = K.fx * p.x / p.z + K.cx, v = K.fy * p.y / p.z + K.cy;
        double r = std::max(0.5, radius[s] * K.fx / p.z);
        int x0 = (int)std::floor(u - r + 0.5), x1 = (int)std::floor(u + r + 0.5);
        int y0 = (int)std::floor(v - r + 0.5), y1 = (int)std::floor(v + r + 0.5);
        for (int y = std::max(0, y0); y <= std::min((int)H - 1, y1); ++y)
            for (int x = std::max(0, x0); x <= std::min((int)W - 1, x1); ++x) {
                V3 d((x - K.cx) / K.fx, (y - K.cy) / K.fy, 1.0);
                double den = dot(n, d);
                if (std::fabs(den) < 0.15) continue;
                double t = dot(n, p) / den;
                if (t <= 0.1 || std::fabs(t - p.z) > 3 * radius[s] + 0.01) continue;
                // disc footprint: the hit must lie within the surfel's radius
                const V3 hit = d * t - p;
                if (dot(hit, hit) > radius[s] * radius[s]) continue;
                size_t o = (size_t)y * W + x;
                if (t < rt.depth[o]) { rt.depth[o] = t; rt.id[o] = (int32_t)s; rt.vtx[o] = d * t; }
            }
    }
}

}  // namespace

extern "C" {

const char* mvdr_scene_last_error(void) { return g_last_error.c_str(); }

int mvdr_scene_default_params(int workload, mvdr_scene_params* p) {
    if (!p) return -1;
    std::memset(p, 0, sizeof(*p));
    p->depth_noise_sigma = 0.001f;
    p->outlier_fraction = 0.05f;
    p->max_rot_deg = 5.0f;
    p->max_trans = 0.01f;
    p->feature_jitter_px = 1.0f;
    p->seed = 0x4D564452ull + (uint64_t)workload;
    switch (workload) {
        case 0:
            p->shape = 0; p->num_views = 1; p->width = 320; p->height = 240; p->fx = p->fy = 285.f;
            p->target_surfels = 20000; p->target_nodes = 1000;
            p->cam_distance_min = p->cam_distance_max = 1.2f;
            p->features_per_view = 0; p->cross_pairs = 0;
            return 0;
        case 1:
        case 2:
            p->shape = 1; p->num_views = 3; p->width = 640; p->height = 480; p->fx = p->fy = 570.f;
            p->target_surfels = 150000; p->target_nodes = 4000;
            p->cam_distance_min = 1.0f; p->cam_distance_max = 1.5f;
            p->features_per_view = 300; p->cross_pairs = 2000;
            return 0;
        case 3:
            p->shape = 1; p->num_views = 3; p->width = 640; p->height = 480; p->fx = p->fy = 570.f;
            p->target_surfels = 300000; p->target_nodes = 8000;
            p->cam_distance_min = 1.0f; p->cam_distance_max = 1.5f;
            p->features_per_view = 300; p->cross_pairs = 2000;
            return 0;
        case 4:  // one scene of the uneven batch; the caller sets seed, nodes, surfels and views per scene
            p->seed = 0x4D564452ull + 100;
            p->shape = 1; p->num_views = 3; p->width = 640; p->height = 480; p->fx = p->fy = 570.f;
            p->target_surfels = 150000; p->target_nodes = 4000;
            p->cam_distance_min = 1.0f; p->cam_distance_max = 1.5f;
            p->features_per_view = 300; p->cross_pairs = 2000;
            return 0;
        default:
            return -1;
    }
}

int mvdr_scene_create(const mvdr_scene_params* params, mvdr_scene** out) {
    if (!params || !out) { g_last_error = "null argument"; return -1; }
    const mvdr_scene_params P = *params;
    if (P.num_views < 1 || P.num_views > MVDR_MAX_VIEWS || P.width < 8 || P.height < 8 || P.target_surfels < 64 ||
        P.target_nodes < 16 || P.width > 4096 || P.height > 4096) {
        g_last_error = "scene parameters out of range";
        return -1;
    }
    auto* sc = new mvdr_scene();
    sc->params = P;
    sc->W = P.width; sc->H = P.height;
    Rng rng(P.seed);

    // cameras: object centre on camera 0's optical axis
    std::vector<double> dist(P.num_views);
    for (auto& d : dist) d = rng.uniform(P.cam_distance_min, P.cam_distance_max);
    const V3 centre(0, 0, dist[0]);

    if (P.shape == 0) sample_sphere(sc->surfels, centre, 0.3, P.target_surfels, rng);
    else sample_body(sc->surfels, centre, P.target_surfels, rng);
    for (auto& s : sc->surfels) {
        s.color[0] = 0.5 + 0.5 * std::sin(23.0 * s.p.x + 3.0 * s.p.y);
        s.color[1] = 0.5 + 0.5 * std::sin(19.0 * s.p.y - 5.0 * s.p.z);
        s.color[2] = 0.5 + 0.5 * std::sin(29.0 * s.p.z + 7.0 * s.p.x);
    }
    const uint32_t S = (uint32_t)sc->surfels.size();

    // nodes: bisect the voxel size onto the requested node count
    double lo = 0.004, hi = 0.2;
    std::vector<uint32_t> node_src;
    for (int it = 0; it < 40; ++it) {
        double mid = std::sqrt(lo * hi);
        node_src = voxel_subsample(sc->surfels, mid);
        if (node_src.size() > P.target_nodes) lo = mid; else hi = mid;
        if (std::fabs((double)node_src.size() - P.target_nodes) < 0.01 * P.target_nodes) break;
    }
    sc->voxel = std::sqrt(lo * hi);
    node_src = voxel_subsample(sc->surfels, sc->voxel);
    const uint32_t N = (uint32_t)node_src.size();
    if (N > 65535 || N < 9) { delete sc; g_last_error = "node count out of range"; return -1; }
    std::vector<V3> node_pos(N);
    for (uint32_t i = 0; i < N; ++i) node_pos[i] = sc->surfels[node_src[i]].p;

    // 8-NN node graph, WarpField.cu:15-59: keep the 8 smallest squared
    // distances > 1e-6, replacing the current maximum on strict improvement.
    sc->graph.resize((size_t)N * 16);
    for (uint32_t i = 0; i < N; ++i) {
        float dv[8]; int iv[8];
        for (int k = 0; k < 8; ++k) { dv[k] = 1e5f; iv[k] = -1; }
        int maxi = 0;
        const float px = (float)node_pos[i].x, py = (float)node_pos[i].y, pz = (float)node_pos[i].z;
        for (uint32_t k = 0; k < N; ++k) {
            const float dx = px - (float)node_pos[k].x, dy = py - (float)node_pos[k].y, dz = pz - (float)node_pos[k].z;
            const float d = dx * dx + dy * dy + dz * dz;
            if (d > 1e-6f && d < dv[maxi]) {
                dv[maxi] = d; iv[maxi] = (int)k;
                maxi = 0; float m = 0;
                for (int j = 0; j < 8; ++j) if (dv[j] > m) { maxi = j; m = dv[j]; }
            }
        }
        for (int k = 0; k < 8; ++k) {
            sc->graph[(size_t)(8 * i + k) * 2 + 0] = (uint16_t)i;
            sc->graph[(size_t)(8 * i + k) * 2 + 1] = (uint16_t)iv[k];
        }
    }

    // 4-NN skinning
    const double node_r2 = 0.025 * 0.025;
    Grid grid(node_pos, std::max(0.02, 2 * sc->voxel));
    sc->surfel_knn.resize((size_t)S * 4);
    sc->surfel_w.resize((size_t)S * 4);
    std::vector<std::pair<double, uint32_t>> nn;
    for (uint32_t s = 0; s < S; ++s) {
        grid.knn(sc->surfels[s].p, 4, 0xFFFFFFFFu, nn);
        double w[4], sum = 0;
        for (int k = 0; k < 4; ++k) { w[k] = std::exp(-nn[k].first / (2 * node_r2)); sum += w[k]; }
        if (sum < 1e-30) { sum = 4; for (double& x : w) x = 1; }
        for (int k = 0; k < 4; ++k) {
            sc->surfel_knn[4 * s + k] = (uint16_t)nn[k].second;
            sc->surfel_w[4 * s + k] = (float)(w[k] / sum);
        }
    }

    // ground-truth warp: per node local rotation about the node + translation,
    // three Laplacian sweeps over the graph.
    std::vector<V3> rv(N), tv(N);
    const double max_rot = P.max_rot_deg * M_PI / 180.0;
    for (uint32_t i = 0; i < N; ++i) {
        rv[i] = rng.unit_vector() * rng.uniform(0, max_rot);
        tv[i] = rng.in_ball() * P.max_trans;
    }
    for (int sweep = 0; sweep < 3; ++sweep) {
        std::vector<V3> r2(N), t2(N);
        for (uint32_t i = 0; i < N; ++i) {
            V3 ar = rv[i], at = tv[i];
            for (int k = 0; k < 8; ++k) { uint16_t j = sc->graph[(size_t)(8 * i + k) * 2 + 1]; ar += rv[j]; at += tv[j]; }
            r2[i] = ar * (1.0 / 9); t2[i] = at * (1.0 / 9);
        }
        rv.swap(r2); tv.swap(t2);
    }
    // averaging shrinks the field; rescale so the largest node motion meets the bounds
    double rmax = 1e-12, tmax = 1e-12;
    for (uint32_t i = 0; i < N; ++i) { rmax = std::max(rmax, norm(rv[i])); tmax = std::max(tmax, norm(tv[i])); }
    for (uint32_t i = 0; i < N; ++i) { rv[i] = rv[i] * (max_rot / rmax); tv[i] = tv[i] * (P.max_trans / tmax); }
    std::vector<Rigid> gt(N);
    sc->se3_gt.resize(N);
    sc->se3_init.resize(N);
    for (uint32_t i = 0; i < N; ++i) {
        Rigid r = axis_angle(rv[i]);
        r.t = node_pos[i] - r.rot(node_pos[i]) + tv[i];
        gt[i] = r;
        sc->se3_gt[i] = dq_from_rigid(r);
        mvdr_dq id{};
        id.q0[3] = 1.0f;
        sc->se3_init[i] = id;
    }

    // rigid canonical -> live alignment shared by all views
    Rigid W2C = axis_angle(normalized(V3(0.2, 1.0, -0.1)) * (0.8 * M_PI / 180.0));
    W2C.t = V3(0.003, -0.002, 0.004);

    // live (ground truth) and initial-warp positions
    std::vector<V3> live_p(S), live_n(S), init_p(S), init_n(S);
    std::vector<double> radius(S);
    // ground-truth node dual quaternions in double for the blend below
    std::vector<std::array<double, 8>> gt_dq(N);
    for (uint32_t i = 0; i < N; ++i) {
        double q0[4];
        quat_from_rot(gt[i], q0);
        const double tq[4] = {0.5 * gt[i].t.x, 0.5 * gt[i].t.y, 0.5 * gt[i].t.z, 0.0};
        double q1[4];
        qmul(tq, q0, q1);
        for (int k = 0; k < 4; ++k) { gt_dq[i][k] = q0[k]; gt_dq[i][4 + k] = q1[k]; }
    }
    for (uint32_t s = 0; s < S; ++s) {
        // dual-quaternion blend, the warp model of the solver (averageDualQuaternion
        // then normalise and convert, DualQuaternion.h:281-335)
        double d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) {
            const auto& g = gt_dq[sc->surfel_knn[4 * s + k]];
            const double w = sc->surfel_w[4 * s + k];
            for (int j = 0; j < 8; ++j) d[j] += w * g[j];
        }
        const double inv = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3]);
        for (double& x : d) x *= inv;
        const double dd = d[0] * d[4] + d[1] * d[5] + d[2] * d[6] + d[3] * d[7];
        for (int j = 0; j < 4; ++j) d[4 + j] -= dd * d[j];
        Rigid b;
        {
            const double x = d[0], y = d[1], z = d[2], w = d[3];
            b.R[0] = 1 - 2 * y * y - 2 * z * z; b.R[1] = 2 * x * y - 2 * z * w;     b.R[2] = 2 * x * z + 2 * y * w;
            b.R[3] = 2 * x * y + 2 * z * w;     b.R[4] = 1 - 2 * x * x - 2 * z * z; b.R[5] = 2 * y * z - 2 * x * w;
            b.R[6] = 2 * x * z - 2 * y * w;     b.R[7] = 2 * y * z + 2 * x * w;     b.R[8] = 1 - 2 * x * x - 2 * y * y;
            const double a[4] = {2 * d[4], 2 * d[5], 2 * d[6], 2 * d[7]};
            const double c[4] = {-x, -y, -z, w};
            double t[4];
            qmul(a, c, t);
            b.t = V3(t[0], t[1], t[2]);
        }
        const Surfel& sf = sc->surfels[s];
        live_p[s] = W2C.apply(b.apply(sf.p));
        // normal of the deformed surface: finite-difference tangents of the warp
        // with the skinning weights re-evaluated along the surface (the rotated
        // canonical normal misses the warp's stretch and shear)
        {
            const uint16_t* kn = &sc->surfel_knn[4 * (size_t)s];
            auto warp_at = [&](const V3& x) {
                double wk[4], ws = 0, dsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (int k = 0; k < 4; ++k) {
                    const V3 dv = x - node_pos[kn[k]];
                    wk[k] = std::exp(-dot(dv, dv) / (2 * node_r2));
                    ws += wk[k];
                }
                for (int k = 0; k < 4; ++k)
                    for (int j = 0; j < 8; ++j) dsum[j] += (wk[k] / ws) * gt_dq[kn[k]][j];
                const double iv = 1.0 / std::sqrt(dsum[0] * dsum[0] + dsum[1] * dsum[1] + dsum[2] * dsum[2] + dsum[3] * dsum[3]);
                for (double& v : dsum) v *= iv;
                const double pr = dsum[0] * dsum[4] + dsum[1] * dsum[5] + dsum[2] * dsum[6] + dsum[3] * dsum[7];
                for (int j = 0; j < 4; ++j) dsum[4 + j] -= pr * dsum[j];
                Rigid m;
                const double x0 = dsum[0], y0 = dsum[1], z0 = dsum[2], w0 = dsum[3];
                m.R[0] = 1 - 2 * y0 * y0 - 2 * z0 * z0; m.R[1] = 2 * x0 * y0 - 2 * z0 * w0; m.R[2] = 2 * x0 * z0 + 2 * y0 * w0;
                m.R[3] = 2 * x0 * y0 + 2 * z0 * w0; m.R[4] = 1 - 2 * x0 * x0 - 2 * z0 * z0; m.R[5] = 2 * y0 * z0 - 2 * x0 * w0;
                m.R[6] = 2 * x0 * z0 - 2 * y0 * w0; m.R[7] = 2 * y0 * z0 + 2 * x0 * w0; m.R[8] = 1 - 2 * x0 * x0 - 2 * y0 * y0;
                const double a2[4] = {2 * dsum[4], 2 * dsum[5], 2 * dsum[6], 2 * dsum[7]};
                const double c2[4] = {-x0, -y0, -z0, w0};
                double tt[4];
                qmul(a2, c2, tt);
                m.t = V3(tt[0], tt[1], tt[2]);
                return m.apply(x);
            };
            const V3 helper = std::fabs(sf.n.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
            const V3 t1 = normalized(cross(sf.n, helper)), t2 = cross(sf.n, t1);
            const double hs = 1e-3;
            const V3 d1 = warp_at(sf.p + t1 * hs) - warp_at(sf.p - t1 * hs);
            const V3 d2 = warp_at(sf.p + t2 * hs) - warp_at(sf.p - t2 * hs);
            V3 nl = normalized(cross(d1, d2));
            if (dot(nl, b.rot(sf.n)) < 0) nl = nl * -1.0;
            live_n[s] = normalized(W2C.rot(nl));
        }
        init_p[s] = W2C.apply(sf.p);
        init_n[s] = W2C.rot(sf.n);
        radius[s] = sf.radius;
    }

    sc->views.resize(P.num_views);
    const size_t WH = (size_t)P.width * P.height;
    std::vector<std::vector<int32_t>> observed_id(P.num_views);
    std::vector<std::vector<double>> observed_depth(P.num_views);
    for (uint32_t v = 0; v < P.num_views; ++v) {
        View& view = sc->views[v];
        view.intr = {P.fx, P.fy, 0.5f * (P.width - 1), 0.5f * (P.height - 1)};
        const double ang = 2.0 * M_PI * v / 3.0;
        V3 cpos = centre + V3(std::sin(ang), 0, -std::cos(ang)) * dist[v];
        V3 z = normalized(centre - cpos), y(0, 1, 0), x = normalized(cross(y, z));
        y = cross(z, x);
        Rigid init;
        init.R[0] = x.x; init.R[1] = y.x; init.R[2] = z.x;
        init.R[3] = x.y; init.R[4] = y.y; init.R[5] = z.y;
        init.R[6] = x.z; init.R[7] = y.z; init.R[8] = z.z;
        init.t = cpos;
        if (v == 0) init = Rigid();
        view.init = init;
        view.world2cam = W2C;
        const Rigid inv = init.inverse();

        // ---- observation: ground-truth live surfels seen from camera v
        std::vector<V3> cp(S), cn(S);
        for (uint32_t s = 0; s < S; ++s) { cp[s] = inv.apply(live_p[s]); cn[s] = inv.rot(live_n[s]); }
        RenderTarget obs;
        splat(cp, cn, radius, view.intr, P.width, P.height, obs);
        view.vertex.assign(WH * 4, 0.f);
        view.normal.assign(WH * 4, 0.f);
        view.fg_mask.assign(WH, 0);
        std::vector<double> depth(WH, 0.0);
        for (size_t o = 0; o < WH; ++o) {
            if (obs.id[o] < 0) continue;
            view.fg_mask[o] = 1;
            double zt = obs.depth[o] + P.depth_noise_sigma * rng.normal();
            if (rng.uniform() < P.outlier_fraction) zt = rng.uniform(0.5, 2.5);
            depth[o] = zt;
        }
        observed_id[v] = obs.id;
        observed_depth[v] = depth;
        auto V = [&](int x, int yy) {
            size_t o = (size_t)yy * P.width + x;
            return V3((x - view.intr.cx) / view.intr.fx * depth[o], (yy - view.intr.cy) / view.intr.fy * depth[o], depth[o]);
        };
        for (uint32_t yy = 0; yy < P.height; ++yy)
            for (uint32_t xx = 0; xx < P.width; ++xx) {
                size_t o = (size_t)yy * P.width + xx;
                if (depth[o] <= 0) continue;
                V3 p = V(xx, yy);
                view.vertex[4 * o + 0] = (float)p.x; view.vertex[4 * o + 1] = (float)p.y;
                view.vertex[4 * o + 2] = (float)p.z; view.vertex[4 * o + 3] = 1.0f;
                if (xx == 0 || yy == 0 || xx + 1 >= P.width || yy + 1 >= P.height) continue;
                size_t ol = o - 1, orr = o + 1, ou = o - P.width, od = o + P.width;
                if (depth[ol] <= 0 || depth[orr] <= 0 || depth[ou] <= 0 || depth[od] <= 0) continue;
                if (std::fabs(depth[ol] - depth[orr]) > 0.05 || std::fabs(depth[ou] - depth[od]) > 0.05) continue;
                V3 n = normalized(cross(V(xx + 1, yy) - V(xx - 1, yy), V(xx, yy + 1) - V(xx, yy - 1)));
                if (dot(n, p) > 0) n = n * -1.0;
                view.normal[4 * o + 0] = (float)n.x; view.normal[4 * o + 1] = (float)n.y;
                view.normal[4 * o + 2] = (float)n.z; view.normal[4 * o + 3] = (float)(p.z / view.intr.fx);
            }
        // edge mask: valid pixel with an invalid 4-neighbour or a depth jump
        view.edge.assign(WH, 0);
        for (uint32_t yy = 0; yy < P.height; ++yy)
            for (uint32_t xx = 0; xx < P.width; ++xx) {
                size_t o = (size_t)yy * P.width + xx;
                if (depth[o] <= 0) continue;
                bool e = false;
                const int dx[4] = {-1, 1, 0, 0}, dy[4] = {0, 0, -1, 1};
                for (int k = 0; k < 4 && !e; ++k) {
                    int nx = (int)xx + dx[k], ny = (int)yy + dy[k];
                    if (nx < 0 || ny < 0 || nx >= (int)P.width || ny >= (int)P.height) { e = true; break; }
                    double dn = depth[(size_t)ny * P.width + nx];
                    if (dn <= 0 || std::fabs(dn - depth[o]) > 0.03) e = true;
                }
                view.edge[o] = e ? 1 : 0;
            }
        // filtered foreground: separable box blur (radius 3) of (1 - mask):
        // 0 deep inside the silhouette, 1 far outside
        {
            std::vector<float> a(WH), b(WH);
            for (size_t o = 0; o < WH; ++o) a[o] = view.fg_mask[o] ? 0.f : 1.f;
            const int R = 3;
            for (uint32_t yy = 0; yy < P.height; ++yy)
                for (uint32_t xx = 0; xx < P.width; ++xx) {
                    double s = 0; int c = 0;
                    for (int k = -R; k <= R; ++k) {
                        int x2 = std::min((int)P.width - 1, std::max(0, (int)xx + k));
                        s += a[(size_t)yy * P.width + x2]; ++c;
                    }
                    b[(size_t)yy * P.width + xx] = (float)(s / c);
                }
            for (uint32_t yy = 0; yy < P.height; ++yy)
                for (uint32_t xx = 0; xx < P.width; ++xx) {
                    double s = 0; int c = 0;
                    for (int k = -R; k <= R; ++k) {
                        int y2 = std::min((int)P.height - 1, std::max(0, (int)yy + k));
                        s += b[(size_t)y2 * P.width + xx]; ++c;
                    }
                    a[(size_t)yy * P.width + xx] = (float)(s / c);
                }
            view.filtered_fg = a;
        }
        // gray image from the surfel colours
        view.gray.assign(WH, 0.f);
        for (size_t o = 0; o < WH; ++o)
            if (obs.id[o] >= 0) {
                const double* c = sc->surfels[obs.id[o]].color;
                view.gray[o] = (float)(0.299 * c[0] + 0.587 * c[1] + 0.114 * c[2]);
            }
        auto gradient = [&](const std::vector<float>& img, std::vector<float>& g) {
            g.assign(WH * 2, 0.f);
            for (uint32_t yy = 1; yy + 1 < P.height; ++yy)
                for (uint32_t xx = 1; xx + 1 < P.width; ++xx) {
                    size_t o = (size_t)yy * P.width + xx;
                    g[2 * o + 0] = 0.5f * (img[o + 1] - img[o - 1]);
                    g[2 * o + 1] = 0.5f * (img[o + P.width] - img[o - P.width]);
                }
        };
        gradient(view.filtered_fg, view.fg_grad);
        gradient(view.gray, view.gray_grad);

        // ---- solver maps: canonical surfels at the initial warp, seen from camera v
        for (uint32_t s = 0; s < S; ++s) { cp[s] = inv.apply(init_p[s]); cn[s] = inv.rot(init_n[s]); }
        RenderTarget geo;
        splat(cp, cn, radius, view.intr, P.width, P.height, geo);
        view.index.assign(WH, 0xFFFFFFFFu);
        view.ref_vertex.assign(WH * 4, 0.f);
        view.ref_normal.assign(WH * 4, 0.f);
        view.warp_vertex.assign(WH * 4, 0.f);
        view.rgb.assign(WH * 4, 0.f);
        for (size_t o = 0; o < WH; ++o) {
            int32_t s = geo.id[o];
            if (s < 0) continue;
            view.index[o] = (uint32_t)s;
            const Surfel& sf = sc->surfels[s];
            V3 rv3 = inv.apply(sf.p), rn3 = inv.rot(sf.n), wv = cp[s];
            float* r4 = &view.ref_vertex[4 * o];
            r4[0] = (float)rv3.x; r4[1] = (float)rv3.y; r4[2] = (float)rv3.z; r4[3] = (float)sf.radius;
            float* n4 = &view.ref_normal[4 * o];
            n4[0] = (float)rn3.x; n4[1] = (float)rn3.y; n4[2] = (float)rn3.z; n4[3] = 0.f;
            float* w4 = &view.warp_vertex[4 * o];
            w4[0] = (float)wv.x; w4[1] = (float)wv.y; w4[2] = (float)wv.z; w4[3] = 1.f;
            float* c4 = &view.rgb[4 * o];
            c4[0] = (float)sf.color[0]; c4[1] = (float)sf.color[1]; c4[2] = (float)sf.color[2]; c4[3] = 1.f;
        }
        // ---- sparse feature pairs: rendered pixel -> observed pixel of the same surfel
        view.corr.assign(WH * 4, 0);
        std::vector<uint32_t> valid_px;
        for (size_t o = 0; o < WH; ++o) if (view.index[o] != 0xFFFFFFFFu) valid_px.push_back((uint32_t)o);
        for (uint32_t f = 0, tries = 0; f < P.features_per_view && tries < 100 * P.features_per_view + 100 && !valid_px.empty();
             ++tries) {
            uint32_t o = valid_px[rng.below(valid_px.size())];
            uint32_t s = view.index[o];
            V3 q = inv.apply(live_p[s]);
            if (q.z < 0.1) continue;
            double u = view.intr.fx * q.x / q.z + view.intr.cx + P.feature_jitter_px * rng.normal();
            double w = view.intr.fy * q.y / q.z + view.intr.cy + P.feature_jitter_px * rng.normal();
            long ui = std::lround(u), wi = std::lround(w);
            if (ui < 0 || wi < 0 || ui >= (long)P.width || wi >= (long)P.height) continue;
            view.features.push_back((uint16_t)(o % P.width));
            view.features.push_back((uint16_t)(o / P.width));
            view.features.push_back((uint16_t)ui);
            view.features.push_back((uint16_t)wi);
            // correspondence map (PatchColliderRGBCorrespondence.cu:171): the pair at its observed pixel
            uint16_t* cm = &view.corr[4 * ((size_t)wi * P.width + ui)];
            cm[0] = (uint16_t)(o % P.width); cm[1] = (uint16_t)(o / P.width); cm[2] = (uint16_t)ui; cm[3] = (uint16_t)wi;
            ++f;
        }
    }

    // ---- cross-view pairs: surfel rendered in view v1, observed (visible) in view v2
    if (P.num_views >= 2 && P.cross_pairs > 0) {
        std::vector<std::vector<uint32_t>> valid(P.num_views);
        for (uint32_t v = 0; v < P.num_views; ++v)
            for (size_t o = 0; o < WH; ++o) if (sc->views[v].index[o] != 0xFFFFFFFFu) valid[v].push_back((uint32_t)o);
        uint32_t made = 0;
        for (uint64_t tries = 0; made < P.cross_pairs && tries < 400ull * P.cross_pairs; ++tries) {
            uint32_t v1 = (uint32_t)rng.below(P.num_views);
            uint32_t v2 = (v1 + 1 + (uint32_t)rng.below(P.num_views - 1)) % P.num_views;
            if (valid[v1].empty()) continue;
            uint32_t o1 = valid[v1][rng.below(valid[v1].size())];
            uint32_t s = sc->views[v1].index[o1];
            const View& view2 = sc->views[v2];
            V3 q = view2.init.inverse().apply(live_p[s]);
            if (q.z < 0.1 || dot(view2.init.inverse().rot(live_n[s]), q) >= 0) continue;
            long u = std::lround(view2.intr.fx * q.x / q.z + view2.intr.cx);
            long w = std::lround(view2.intr.fy * q.y / q.z + view2.intr.cy);
            if (u < 0 || w < 0 || u >= (long)P.width || w >= (long)P.height) continue;
            size_t o2 = (size_t)w * P.width + u;
            if (observed_id[v2][o2] < 0) continue;
            const double zt = observed_depth[v2][o2];
            if (std::fabs(zt - q.z) > 0.01) continue;  // occluded, or an outlier pixel
            long uj = std::lround(u + P.feature_jitter_px * rng.normal());
            long wj = std::lround(w + P.feature_jitter_px * rng.normal());
            uj = std::min<long>(P.width - 1, std::max<long>(0, uj));
            wj = std::min<long>(P.height - 1, std::max<long>(0, wj));
            mvdr_cross_pair cpair;
            cpair.pixel[0] = (uint16_t)(o1 % P.width); cpair.pixel[1] = (uint16_t)(o1 / P.width);
            cpair.pixel[2] = (uint16_t)uj; cpair.pixel[3] = (uint16_t)wj;
            cpair.view[0] = (uint16_t)v1; cpair.view[1] = (uint16_t)v2;
            sc->cross.push_back(cpair);
            ++made;
        }
    }

    // flat arrays
    sc->surfel_pos.resize((size_t)S * 4);
    sc->surfel_normal.resize((size_t)S * 4);
    for (uint32_t s = 0; s < S; ++s) {
        const Surfel& sf = sc->surfels[s];
        float* p = &sc->surfel_pos[4 * s];
        p[0] = (float)sf.p.x; p[1] = (float)sf.p.y; p[2] = (float)sf.p.z; p[3] = 1.f;
        float* n = &sc->surfel_normal[4 * s];
        n[0] = (float)sf.n.x; n[1] = (float)sf.n.y; n[2] = (float)sf.n.z; n[3] = (float)sf.radius;
    }
    sc->nodes.resize((size_t)N * 4);
    for (uint32_t i = 0; i < N; ++i) {
        sc->nodes[4 * i + 0] = (float)node_pos[i].x; sc->nodes[4 * i + 1] = (float)node_pos[i].y;
        sc->nodes[4 * i + 2] = (float)node_pos[i].z; sc->nodes[4 * i + 3] = 1.f;
    }
    *out = sc;
    return 0;
}

void mvdr_scene_destroy(mvdr_scene* scene) { delete scene; }

int mvdr_scene_get_info(const mvdr_scene* sc, mvdr_scene_info* info) {
    if (!sc || !info) return -1;
    std::memset(info, 0, sizeof(*info));
    info->num_surfels = (uint32_t)sc->surfels.size();
    info->num_nodes = (uint32_t)(sc->nodes.size() / 4);
    info->num_graph_edges = (uint32_t)(sc->graph.size() / 2);
    info->num_views = sc->params.num_views;
    info->width = sc->W; info->height = sc->H;
    for (uint32_t v = 0; v < sc->params.num_views; ++v) info->num_feature_pairs[v] = (uint32_t)(sc->views[v].features.size() / 4);
    info->num_cross_pairs = (uint32_t)sc->cross.size();
    info->voxel_size = (float)sc->voxel;
    return 0;
}

long long mvdr_scene_get(const mvdr_scene* sc, const char* name, uint32_t view, void* dst, size_t bytes) {
    if (!sc || !name) return -1;
    const std::string n(name);
    const void* src = nullptr;
    size_t size = 0;
    mvdr_intrinsic intr;
    mvdr_mat34 m34;
    auto pick = [&](const auto& vec) { src = vec.data(); size = vec.size() * sizeof(vec[0]); };
    if (n == "surfel_pos") pick(sc->surfel_pos);
    else if (n == "surfel_normal") pick(sc->surfel_normal);
    else if (n == "surfel_knn") pick(sc->surfel_knn);
    else if (n == "surfel_knn_weight") pick(sc->surfel_w);
    else if (n == "nodes") pick(sc->nodes);
    else if (n == "node_graph") pick(sc->graph);
    else if (n == "node_se3_init") pick(sc->se3_init);
    else if (n == "node_se3_gt") pick(sc->se3_gt);
    else if (n == "cross_pairs") pick(sc->cross);
    else {
        if (view >= sc->views.size()) return -1;
        const View& v = sc->views[view];
        if (n == "intrinsic") { intr = v.intr; src = &intr; size = sizeof(intr); }
        else if (n == "world2camera") { m34 = to_mat34(v.world2cam); src = &m34; size = sizeof(m34); }
        else if (n == "initial_camera_se3") { m34 = to_mat34(v.init); src = &m34; size = sizeof(m34); }
        else if (n == "vertex_confid") pick(v.vertex);
        else if (n == "normal_radius") pick(v.normal);
        else if (n == "foreground_mask") pick(v.fg_mask);
        else if (n == "filtered_foreground") pick(v.filtered_fg);
        else if (n == "foreground_gradient") pick(v.fg_grad);
        else if (n == "gray") pick(v.gray);
        else if (n == "gray_gradient") pick(v.gray_grad);
        else if (n == "edge_mask") pick(v.edge);
        else if (n == "reference_vertex") pick(v.ref_vertex);
        else if (n == "reference_normal") pick(v.ref_normal);
        else if (n == "warp_vertex") pick(v.warp_vertex);
        else if (n == "index_map") pick(v.index);
        else if (n == "normalized_rgb") pick(v.rgb);
        else if (n == "feature_pairs") pick(v.features);
        else if (n == "corr_map") pick(v.corr);
        else return -1;
    }
    if (dst && bytes >= size && size) std::memcpy(dst, src, size);
    return (long long)size;
}

}  // extern "C"
