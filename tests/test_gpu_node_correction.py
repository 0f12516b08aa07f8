"""Large-error node correction on the GPU (mvdr_warp_correct_nodes, through the
C ABI) against the oracle: the planted frame of the oracle's pins, and the body
scene (~4k nodes with errors U[0, 1] spread over its three views, its
observation vertex maps as the previous frame, random guided transforms of up
to 1 degree and 5 mm on half the pixels).  The window selection is exact (the
same nodes corrected, from the same pixels); corrections, points and warps
agree within 1e-5 (the device's atan2 / sin / cos against the host's)."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from multiviewdynrecon_amd.warpfield import CorrectLargeErrorNode
from test_oracle_node_correction import H as H_PLANTED
from test_oracle_node_correction import W as W_PLANTED
from test_oracle_node_correction import planted_inputs, run_oracle
from test_oracle_warpfield import oracle_skin

pytestmark = pytest.mark.gpu


def dev(a):
    a = np.ascontiguousarray(a).copy()
    if a.dtype == np.uint16:  # knn: moved as int16 bits
        a = a.view(np.int16)
    return torch.from_numpy(a).cuda()


def run_gpu(views, W, H, nodes, surfels):
    dv = [{k: (dev(a) if isinstance(a, np.ndarray) and a.ndim >= 2 else a) for k, a in d.items()} for d in views]
    dn = {k: dev(v) for k, v in nodes.items()}
    ds = {k: dev(v) for k, v in surfels.items()}
    corr = CorrectLargeErrorNode(dv, W, H, dn, ds)
    torch.cuda.synchronize()
    return corr.cpu().numpy(), {k: v.cpu().numpy() for k, v in dn.items()}, {k: v.cpu().numpy() for k, v in ds.items()}


def identity_mask(corr):
    """Nodes left (up to rounding) at the identity: no candidate was selected.
    A selected guided transform moves the entries by >= ~1e-4; the conjugated
    identity of a node without one differs from it by rounding only."""
    return np.all(np.abs(corr - np.array([0, 0, 0, 1, 0, 0, 0, 0], np.float32)) < 1e-5, axis=1)


def compare(views, W, H, nodes, surfels, min_corrected):
    c_ref, n_ref, s_ref = run_oracle(views, W, H, nodes, surfels)
    c_gpu, n_gpu, s_gpu = run_gpu(views, W, H, nodes, surfels)
    np.testing.assert_array_equal(identity_mask(c_gpu), identity_mask(c_ref))
    assert (~identity_mask(c_ref)).sum() >= min_corrected
    np.testing.assert_allclose(c_gpu, c_ref, atol=1e-5)
    np.testing.assert_allclose(n_gpu["live"], n_ref["live"], atol=1e-5)
    np.testing.assert_allclose(n_gpu["se3"], n_ref["se3"], atol=1e-5)
    np.testing.assert_allclose(s_gpu["vertex"], s_ref["vertex"], atol=1e-5)
    np.testing.assert_allclose(s_gpu["normal"], s_ref["normal"], atol=1e-5)
    return c_ref


def test_planted_frame_gpu():
    views, nodes, surfels, _ = planted_inputs()
    compare(views, W_PLANTED, H_PLANTED, nodes, surfels, min_corrected=4)


def test_body_scene_gpu(scene1):
    rng = np.random.default_rng(31)
    A = scene1.arrays
    W, H, V = scene1.width, scene1.height, scene1.num_views
    nodes_xyz = A["nodes"].astype(np.float32)
    n = len(nodes_xyz)
    live = nodes_xyz.copy()
    live[:, 3] = (np.arange(n) % V + 1).astype(np.float32)  # view code v + 1
    nk, nw = oracle_skin(nodes_xyz, nodes_xyz)
    views = []
    for v, view in enumerate(scene1.views):
        rot = Rotation.from_rotvec(rng.normal(size=(H * W, 3)) * np.radians(0.6)).as_matrix()
        tr = rng.uniform(-0.005, 0.005, size=(H * W, 3))
        guided = np.concatenate([rot.transpose(0, 2, 1).reshape(-1, 9), tr], axis=1).astype(np.float32)
        views.append(dict(intrinsic=view.intrinsic, initial_camera_se3=view.initial_camera_se3,
                          world2camera=view.world2camera, previous_vertex=view.maps["vertex_confid"],
                          guided_se3=guided.reshape(H, W, 12), guided_valid=(rng.random((H, W)) < 0.5).astype(np.uint8)))
    nodes = dict(error=rng.random(n).astype(np.float32), live=live, se3=A["node_se3_gt"].astype(np.float32),
                 knn=nk.astype(np.uint16), weight=nw.astype(np.float32))
    surfels = dict(vertex=A["surfel_pos"].astype(np.float32), normal=A["surfel_normal"].astype(np.float32),
                   knn=A["surfel_knn"].astype(np.uint16), weight=A["surfel_knn_weight"].astype(np.float32))
    c = compare(views, W, H, nodes, surfels, min_corrected=50)
    print(f"body scene: {(~identity_mask(c)).sum()} of {n} nodes corrected")
