"""Pin the oracle's term residuals and Jacobians.

* Jacobians: central finite differences (FP64, h = 1e-4) of the residuals under
  the reference's own update rule T <- [Rodrigues(w) | v] T (SolverIterationData.cu:9-57).
  A global twist applied to every node moves each blended transform by exactly
  that twist, so d r / d xi = sum_k w_k J = J for normalised skinning weights.
* Residuals: with a noise-free observation the dense residuals at the ground
  truth warp are far smaller than at the initial warp (forward-model check).
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from multiviewdynrecon_amd.config import workload_config
from multiviewdynrecon_amd.frame import HostFrame
from oracle import Oracle


def dq_rows(R, t, n):
    q = Rotation.from_matrix(R).as_quat()
    x1, y1, z1, w1 = t[0] / 2, t[1] / 2, t[2] / 2, 0.0
    x2, y2, z2, w2 = q
    dual = [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2]
    return np.tile(np.concatenate([q, dual]), (n, 1))


def twist_se3(xi):
    return Rotation.from_rotvec(xi[:3]).as_matrix(), np.asarray(xi[3:], dtype=np.float64)


@pytest.fixture(scope="module")
def o0(scene0):
    cfg = workload_config(0)
    o = Oracle(cfg, HostFrame(scene0), "fp64")
    o.build_index()
    o.set_iteration(0)
    return o


def eval_at(o, se3):
    o.set("node_se3", se3)
    o.eval_terms()
    return o.get("dense_terms"), o.get("smooth_terms")


def test_dense_jacobian_matches_finite_differences(o0, scene0):
    N = scene0.num_nodes
    base, _ = eval_at(o0, dq_rows(np.eye(3), np.zeros(3), N))
    valid = np.abs(base[:, 8:14]).sum(1) > 0
    assert valid.sum() > 5000
    # the residual reads the observation at the rounded projection, so a step
    # that moves a term across a pixel boundary jumps; h = 1e-6 keeps the
    # moves at ~3e-4 px (h = 1e-4 already re-projects ~5% of the terms)
    h = 1e-6
    ok_frac = []
    for k in range(6):
        xi = np.zeros(6)
        xi[k] = h
        rp, _ = eval_at(o0, dq_rows(*twist_se3(xi), N))
        rm, _ = eval_at(o0, dq_rows(*twist_se3(-xi), N))
        both = valid & (np.abs(rp[:, 8:14]).sum(1) > 0) & (np.abs(rm[:, 8:14]).sum(1) > 0)
        fd = (rp[both, 14] - rm[both, 14]) / (2 * h)
        J = base[both, 8 + k]
        close = np.abs(fd - J) <= 1e-4 + 1e-3 * np.abs(J)
        ok_frac.append(close.mean())
    # a handful of terms re-project onto a neighbouring pixel inside +-h
    assert min(ok_frac) > 0.995, ok_frac


def test_smooth_jacobian_matches_finite_differences(o0, scene0):
    N = scene0.num_nodes
    ident = dq_rows(np.eye(3), np.zeros(3), N)
    _, base = eval_at(o0, ident)
    rng = np.random.default_rng(3)
    edges = rng.choice(len(base), 40, replace=False)
    h = 1e-5
    for e in edges:
        i, j = int(base[e, 0]), int(base[e, 1])
        if i == j:
            continue
        a = base[e, 3:6]
        s = base[e, 6:9]
        J_i = np.array([[0, a[2], -a[1], 1, 0, 0], [-a[2], 0, a[0], 0, 1, 0], [a[1], -a[0], 0, 0, 0, 1]])
        J_j = -np.array([[0, s[2], -s[1], 1, 0, 0], [-s[2], 0, s[0], 0, 1, 0], [s[1], -s[0], 0, 0, 0, 1]])
        for node, J in ((i, J_i), (j, J_j)):
            for k in range(6):
                xi = np.zeros(6)
                xi[k] = h
                rows = []
                for sign in (1, -1):
                    se3 = ident.copy()
                    se3[node] = dq_rows(*twist_se3(sign * xi), 1)[0]
                    _, sm = eval_at(o0, se3)
                    rows.append(sm[e, 3:6] - sm[e, 6:9])
                fd = (rows[0] - rows[1]) / (2 * h)
                np.testing.assert_allclose(fd, J[:, k], atol=1e-6)


@pytest.mark.parametrize("name", ["feature_terms", "cross_terms"])
def test_feature_warp_jacobian_matches_finite_differences(scene1, name):
    """Feature and cross-view terms share the point-to-point warp
    (forwardWarpFeatureVertexKernel, SparseCorrespondenceHandler.cu:193-220)."""
    cfg = workload_config(1)
    assert cfg.use_cross_view
    o = Oracle(cfg, HostFrame(scene1), "fp64")
    o.build_index()
    N = scene1.num_nodes
    o.set("node_se3", dq_rows(np.eye(3), np.zeros(3), N))
    o.eval_terms()
    base = o.get(name)
    assert len(base) > 200, len(base)
    h = 1e-5
    w = base[:, 14:17]
    for k in range(6):
        xi = np.zeros(6)
        xi[k] = h
        out = []
        for sign in (1, -1):
            o.set("node_se3", dq_rows(*twist_se3(sign * xi), N))
            o.eval_terms()
            out.append(o.get(name)[:, 14:17])
        fd = (out[0] - out[1]) / (2 * h)
        # d warped / d xi = [-[w]x | I]: column k
        expect = np.zeros_like(w)
        if k < 3:
            e = np.zeros(3)
            e[k] = 1
            expect = np.cross(e, w)
        else:
            expect[:, k - 3] = 1
        np.testing.assert_allclose(fd, expect, atol=1e-6)


def test_dense_residuals_vanish_at_ground_truth_without_noise(built_libs):
    from multiviewdynrecon_amd.scene import make_scene

    scene = make_scene(0, depth_noise_sigma=0.0, outlier_fraction=0.0)
    o = Oracle(workload_config(0), HostFrame(scene), "fp64")
    o.build_index()
    o.set_iteration(0)
    res = {}
    for name, se3 in (("init", scene.arrays["node_se3_init"]), ("gt", scene.arrays["node_se3_gt"])):
        d, _ = eval_at(o, se3.astype(np.float64))
        v = np.abs(d[:, 8:14]).sum(1) > 0
        res[name] = (np.median(np.abs(d[v, 14])), v.sum())
    assert res["gt"][1] > 0.9 * res["init"][1]
    # ~1.9 mm at the initial warp, tens of micrometres (splat discretisation) at the truth
    assert res["gt"][0] < 0.05 * res["init"][0], res


def test_dense_residual_matches_independent_restatement(scene0):
    """r = n . (p_cam - d) at the projected pixel, recomputed with numpy from the
    scene arrays (DenseDepthJacobianResidual.cu:24-104, view 0: identity camera pose)."""
    o = Oracle(workload_config(0), HostFrame(scene0), "fp64")
    o.build_index()
    o.set_iteration(0)
    se3 = scene0.arrays["node_se3_gt"].astype(np.float64)
    d, _ = eval_at(o, se3)
    v = scene0.views[0]
    K = v.intrinsic.astype(np.float64)
    Rw = v.world2camera[:9].reshape(3, 3).T.astype(np.float64)
    tw = v.world2camera[9:].astype(np.float64)
    im = v.maps["index_map"]
    ys, xs = np.nonzero(im != 0xFFFFFFFF)
    s = im[ys, xs]
    knn = scene0.arrays["surfel_knn"][s].astype(np.int64)
    w = scene0.arrays["surfel_knn_weight"][s].astype(np.float64)
    b = (se3[knn] * w[..., None]).sum(1)
    q0 = b[:, :4] / np.linalg.norm(b[:, :4], axis=1, keepdims=True)
    q1 = b[:, 4:] / np.linalg.norm(b[:, :4], axis=1, keepdims=True)
    q1 = q1 - (q0 * q1).sum(1, keepdims=True) * q0
    R = Rotation.from_quat(q0).as_matrix()
    conj = q0 * np.array([-1, -1, -1, 1])
    # t = 2 q1 conj(q0) (vector part), written out
    a, c = q1, conj
    t = 2 * np.stack([a[:, 3] * c[:, 0] + a[:, 0] * c[:, 3] + a[:, 1] * c[:, 2] - a[:, 2] * c[:, 1],
                      a[:, 3] * c[:, 1] - a[:, 0] * c[:, 2] + a[:, 1] * c[:, 3] + a[:, 2] * c[:, 0],
                      a[:, 3] * c[:, 2] + a[:, 0] * c[:, 1] - a[:, 1] * c[:, 0] + a[:, 2] * c[:, 3]], 1)
    ref = v.maps["reference_vertex"][ys, xs, :3].astype(np.float64)
    P = np.einsum("nij,nj->ni", R, ref) + t
    P = P @ Rw.T + tw
    u = np.rint((P[:, 0] / (P[:, 2] + 1e-10) * K[0] + K[2]).astype(np.float32)).astype(np.int64)
    q = np.rint((P[:, 1] / (P[:, 2] + 1e-10) * K[1] + K[3]).astype(np.float32)).astype(np.int64)
    inside = (u >= 0) & (u < scene0.width) & (q >= 0) & (q < scene0.height)
    r = np.zeros(len(s))
    dv = v.maps["vertex_confid"][q[inside], u[inside], :3].astype(np.float64)
    dn = v.maps["normal_radius"][q[inside], u[inside], :3].astype(np.float64)
    nref = v.maps["reference_normal"][ys, xs, :3].astype(np.float64)[inside]
    nc = np.einsum("nij,nj->ni", R[inside], nref) @ Rw.T
    ok = (np.abs(dv).max(1) >= 1e-3) & ((dn * nc).sum(1) >= 0.7) & (((dv - P[inside]) ** 2).sum(1) <= 0.075 ** 2)
    ri = np.where(ok, (dn * (P[inside] - dv)).sum(1), 0.0)
    r[inside] = ri
    np.testing.assert_allclose(d[:, 14], r, atol=1e-9)


# Image terms (density_map_jacobian.cuh:12-79): the residual reads the map at
# the rounded projection, a step function of the warp, so finite differences
# see zero.  On a linear image I(u, q) = A u + B q + C with a constant gradient
# map (A, B), the oracle's checker switch `bilinear_images` samples the map at
# the unrounded projection instead: the residual is then exactly
# A U(xi) + B Q(xi) + C - g, whose derivative is the reference's chain (focal
# scaling, 1/z, the initial-camera and world-to-camera rotations, the twist's
# cross product).  A, B, C are powers of two, so every sample of the image is
# exact in float32 and bilinear sampling reproduces the plane.  fx != fy, so a
# swapped focal length fails too.
LIN_A, LIN_B, LIN_C = 2.0 ** -7, -(2.0 ** -8), 8.0


@pytest.fixture(scope="module")
def linear_image_oracle(scene1):
    import copy

    sc = copy.deepcopy(scene1)
    q, u = np.mgrid[0:sc.height, 0:sc.width]
    img = (LIN_A * u + LIN_B * q + LIN_C).astype(np.float32)
    assert np.array_equal(img.astype(np.float64), LIN_A * u + LIN_B * q + LIN_C)
    grad = np.empty((sc.height, sc.width, 2), np.float32)
    grad[..., 0], grad[..., 1] = LIN_A, LIN_B
    for v in sc.views:
        v.intrinsic = v.intrinsic.copy()
        v.intrinsic[1] = 540.0
        for name in ("gray", "filtered_foreground"):
            v.maps[name] = img.copy()
        for name in ("gray_gradient", "foreground_gradient"):
            v.maps[name] = grad.copy()
    cfg = workload_config(1, use_foreground=1, density_in_jtj=1)
    o = Oracle(cfg, HostFrame(sc), "fp64")
    o.scene = sc  # the frame points into these maps
    o.build_index()
    o.set_iteration(0)
    return o


def image_terms_at(o, name, xi, bilinear):
    o.set("bilinear_images", [1.0 if bilinear else 0.0])
    try:
        o.set("node_se3", dq_rows(*twist_se3(xi), o.scene.num_nodes))
        o.eval_terms()
        return o.get(name)
    finally:
        o.set("bilinear_images", [0.0])


@pytest.mark.parametrize("name", ["density_terms", "foreground_terms"])
def test_image_jacobian_matches_finite_differences(linear_image_oracle, name):
    o = linear_image_oracle
    base = image_terms_at(o, name, np.zeros(6), True)
    valid = np.abs(base[:, 8:14]).sum(1) > 0
    assert valid.sum() > 1000, valid.sum()
    h = 1e-5
    for k in range(6):
        xi = np.zeros(6)
        xi[k] = h
        rp = image_terms_at(o, name, xi, True)
        rm = image_terms_at(o, name, -xi, True)
        both = valid & (np.abs(rp[:, 8:14]).sum(1) > 0) & (np.abs(rm[:, 8:14]).sum(1) > 0)
        assert both.sum() > 0.99 * valid.sum()
        fd = (rp[both, 14] - rm[both, 14]) / (2 * h)
        np.testing.assert_allclose(fd, base[both, 8 + k], rtol=1e-6, atol=1e-7, err_msg=f"{name} twist {k}")


@pytest.mark.parametrize("name", ["density_terms", "foreground_terms"])
def test_image_residual_reads_the_rounded_projection(linear_image_oracle, name):
    """The reference lookup (nearest pixel) against the smooth one: within half a
    pixel step of the plane, on the plane's 2^-8 lattice once the rendered
    density g = 0.299 R + 0.587 G + 0.114 B (ColorTypeTransfer.h:51-56) is
    added back, and with the same Jacobian (the gradient map is constant)."""
    o = linear_image_oracle
    xi = np.array([0.01, -0.02, 0.015, 0.01, 0.005, -0.01])
    near = image_terms_at(o, name, xi, False)
    smooth = image_terms_at(o, name, xi, True)
    both = (np.abs(near[:, 8:14]).sum(1) > 0) & (np.abs(smooth[:, 8:14]).sum(1) > 0)
    assert both.sum() > 1000
    np.testing.assert_array_equal(near[both, 8:14], smooth[both, 8:14])
    diff = np.abs(near[both, 14] - smooth[both, 14])
    assert diff.max() <= (abs(LIN_A) + abs(LIN_B)) / 2 + 1e-9
    assert np.median(diff) > 1e-4  # the lookup does round
    g = np.zeros(len(near))
    if name == "density_terms":
        code = near[:, 15].astype(np.int64)
        px, py, view = code & 0xFFFF, (code >> 16) & 0xFFFF, code >> 32
        for v in range(len(o.scene.views)):
            sel = view == v
            rgb = o.scene.views[v].maps["normalized_rgb"][py[sel], px[sel], :3].astype(np.float64)
            coef = np.array([0.299, 0.587, 0.114], np.float32).astype(np.float64)
            g[sel] = rgb @ coef
    lattice = (near[both, 14] + g[both]) * 256.0
    np.testing.assert_allclose(lattice, np.rint(lattice), atol=1e-6)
