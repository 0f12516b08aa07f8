"""Warp-field routines on the GPU (mvdr_warp_*, through the C ABI) against the
oracle on the body scene (~150k surfels, ~4k nodes): the node graph and the
skinning knn bit for bit, the skinning weights within 1e-6 (the device's expf
against the host's), the forward warp under the scene's ground-truth field
within 2e-6 m (rsqrt against 1/sqrt in the normalisation), and the inverse
warp returning the canonical surfels."""
import numpy as np
import pytest

from multiviewdynrecon_amd import warpfield as wf
from test_oracle_warpfield import oracle_graph, oracle_skin, oracle_warp

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_node_graph_and_skinning_body(scene1):
    A = scene1.arrays
    nodes, pts = A["nodes"], A["surfel_pos"]
    g = host(wf.BuildNodesGraph(dev(nodes))).view(np.uint16)
    np.testing.assert_array_equal(g, oracle_graph(nodes))
    vk, vw, nk, nw = (host(t) for t in wf.PerformSkinning(dev(pts), dev(nodes)))
    k_ref, w_ref = oracle_skin(np.concatenate([pts, nodes]), nodes)
    np.testing.assert_array_equal(np.concatenate([vk.view(np.uint16), nk.view(np.uint16)]), k_ref)
    np.testing.assert_allclose(np.concatenate([vw, nw]), w_ref, rtol=1e-6, atol=1e-7)
    # the scene generator's skinning is the same rule: its knn sets agree
    same = (np.sort(vk.view(np.uint16), 1) == np.sort(A["surfel_knn"], 1)).all(1).mean()
    assert same > 0.99


def test_forward_and_inverse_warp_body(scene1):
    A = scene1.arrays
    field = A["node_se3_gt"]
    pts, normals, knn, w, nodes = A["surfel_pos"], A["surfel_normal"], A["surfel_knn"], A["surfel_knn_weight"], A["nodes"]
    nk, nw = oracle_skin(nodes, nodes)
    lv, ln, lnodes = (host(t) for t in wf.ForwardWarpSurfelsAndNodes(
        dev(field), dev(pts), dev(normals), dev(knn.view(np.int16)), dev(w), dev(nodes), dev(nk.view(np.int16)),
        dev(nw)))
    rv_ref, rn_ref = oracle_warp(field, pts, normals, knn, w)
    np.testing.assert_allclose(lv, rv_ref, rtol=0, atol=2e-6)
    np.testing.assert_allclose(ln, rn_ref, rtol=0, atol=2e-6)
    np.testing.assert_allclose(lnodes, oracle_warp(field, nodes, None, nk, nw)[0], rtol=0, atol=2e-6)
    rv, rn = (host(t) for t in wf.InverseWarpSurfels(dev(field), dev(lv), dev(ln), dev(knn.view(np.int16)), dev(w)))
    np.testing.assert_allclose(rv, pts, rtol=0, atol=4e-6)
    np.testing.assert_allclose(rn, normals, rtol=0, atol=4e-6)
