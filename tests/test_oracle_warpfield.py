"""Pin the oracle's warp-field routines (include/mvdr/warpfield.h):

* node graph (WarpField.cu:15-59): on random nodes every row holds exactly the
  8 nearest other nodes (numpy argsort); on a planted row the slots follow the
  replace-the-worst order traced by hand, a coincident node is excluded and a
  short graph fills its free slots with 0xFFFF;
* skinning (CanonicalNodesSkinner.cu:356-437, brute_foce_knn.cuh:17-39): the 4
  knn are the 4 nearest nodes, the weights normalised exp(-d / (2 r^2))
  computed here in FP64, and on a planted sequence the slot order is the one
  traced by hand through the reference's heap update;
* warps (SurfelNodeDeformer.cu:9-124): with one rigid transform on every node
  the forward warp is that transform (scipy), and the inverse warp undoes the
  forward warp of any field.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import lib as oracle_lib
from test_oracle_math import dq_of

R = 0.025


def L():
    lib = oracle_lib()
    lib.mvdr_oracle_node_graph.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    lib.mvdr_oracle_skin.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p, C.c_void_p]
    lib.mvdr_oracle_warp.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_uint32, C.c_int, C.c_void_p, C.c_void_p]
    return lib


def oracle_graph(nodes):
    nodes = np.ascontiguousarray(nodes, np.float32)
    g = np.zeros((len(nodes) * 8, 2), np.uint16)
    L().mvdr_oracle_node_graph(nodes.ctypes.data, len(nodes), g.ctypes.data)
    return g


def oracle_skin(points, nodes, radius=R):
    points = np.ascontiguousarray(points, np.float32)
    nodes = np.ascontiguousarray(nodes, np.float32)
    k = np.zeros((len(points), 4), np.uint16)
    w = np.zeros((len(points), 4), np.float32)
    L().mvdr_oracle_skin(points.ctypes.data, len(points), nodes.ctypes.data, len(nodes), radius, k.ctypes.data,
                         w.ctypes.data)
    return k, w


def oracle_warp(field, points, normals, knn, w, inverse=False):
    field = np.ascontiguousarray(field, np.float32)
    points = np.ascontiguousarray(points, np.float32)
    knn = np.ascontiguousarray(knn, np.uint16)
    w = np.ascontiguousarray(w, np.float32)
    op = np.zeros_like(points)
    on = None
    if normals is not None:
        normals = np.ascontiguousarray(normals, np.float32)
        on = np.zeros_like(normals)
    L().mvdr_oracle_warp(field.ctypes.data, len(field), points.ctypes.data,
                         normals.ctypes.data if normals is not None else None, knn.ctypes.data, w.ctypes.data,
                         len(points), int(inverse), op.ctypes.data, on.ctypes.data if on is not None else None)
    return op, on


def cloud(rng, n, spread=0.2):
    return np.concatenate([rng.uniform(-spread, spread, (n, 3)) + [0, 0, 1.0], np.ones((n, 1))], 1).astype(np.float32)


def test_node_graph_rows_are_the_8_nearest(built_libs):
    rng = np.random.default_rng(3)
    nodes = cloud(rng, 300)
    g = oracle_graph(nodes).reshape(300, 8, 2)
    d2 = ((nodes[:, None, :3].astype(np.float64) - nodes[None, :, :3]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    want = np.sort(np.argsort(d2, 1)[:, :8], 1)
    assert (g[:, :, 0] == np.arange(300)[:, None]).all()
    np.testing.assert_array_equal(np.sort(g[:, :, 1].astype(np.int64), 1), want)


def test_node_graph_planted_slot_order(built_libs):
    # node 0 at the origin, then nodes along x at squared distances
    # .25 .01 .16 .04 .81 .09 .64 .36 fill slots 0..7; the worst (.81, slot 4)
    # is replaced by node 9 (.0025); node 10 coincides with node 0: excluded
    xs = [0.5, 0.1, 0.4, 0.2, 0.9, 0.3, 0.8, 0.6, 0.05, 0.0]
    nodes = np.array([[0, 0, 0, 1]] + [[x, 0, 0, 1] for x in xs], np.float32)
    row = oracle_graph(nodes).reshape(-1, 8, 2)[0]
    assert row[:, 1].tolist() == [1, 2, 3, 4, 9, 6, 7, 8]
    small = oracle_graph(nodes[:4]).reshape(-1, 8, 2)[0]
    assert small[:, 1].tolist() == [1, 2, 3] + [0xFFFF] * 5


def test_skinning_nearest_four_and_weights(built_libs):
    rng = np.random.default_rng(4)
    nodes = cloud(rng, 120, 0.1)
    pts = cloud(rng, 500, 0.1)
    k, w = oracle_skin(np.concatenate([pts, nodes]), nodes)
    allp = np.concatenate([pts, nodes])[:, :3].astype(np.float64)
    d2 = ((allp[:, None, :] - nodes[None, :, :3]) ** 2).sum(-1)
    np.testing.assert_array_equal(np.sort(k.astype(np.int64), 1), np.sort(np.argsort(d2, 1)[:, :4], 1))
    dk = np.take_along_axis(d2, k.astype(np.int64), 1)
    ew = np.exp(-dk / (2 * R * R))
    np.testing.assert_allclose(w, ew / ew.sum(1, keepdims=True), rtol=2e-5, atol=1e-7)
    assert (k[len(pts):] == np.arange(120)[:, None]).any(1).all()  # a node is its own nearest


def test_skinning_planted_heap_order(built_libs):
    # squared distances 4, 1, 3, 2, 0.5, 5 in scan order: the heap ends with
    # slots (node 2, node 3, node 4, node 1) -- traced through the update rule
    d2 = [4.0, 1.0, 3.0, 2.0, 0.5, 5.0]
    nodes = np.array([[np.sqrt(d), 0, 0, 1] for d in d2], np.float32)
    k, _ = oracle_skin(np.array([[0, 0, 0, 1]], np.float32), nodes)
    assert k[0].tolist() == [2, 3, 4, 1]


def test_forward_warp_is_the_rigid_transform_and_inverse_undoes_it(built_libs):
    rng = np.random.default_rng(5)
    nodes = cloud(rng, 50, 0.1)
    pts = cloud(rng, 400, 0.1)
    normals = np.concatenate([rng.normal(size=(400, 3)), np.zeros((400, 1))], 1).astype(np.float32)
    normals[:, :3] /= np.linalg.norm(normals[:, :3], axis=1, keepdims=True)
    k, w = oracle_skin(pts, nodes)
    Rm = Rotation.from_rotvec([0.2, -0.1, 0.3]).as_matrix()
    t = np.array([0.05, -0.02, 0.01])
    field = np.tile(dq_of(Rm, t), (50, 1)).astype(np.float32)
    lv, ln = oracle_warp(field, pts, normals, k, w)
    np.testing.assert_allclose(lv[:, :3], pts[:, :3] @ Rm.T + t, atol=2e-6)
    np.testing.assert_allclose(ln[:, :3], normals[:, :3] @ Rm.T, atol=2e-6)
    np.testing.assert_array_equal(lv[:, 3], pts[:, 3])
    # a non-rigid field: the inverse warp returns the canonical points
    field = np.stack([dq_of(Rotation.from_rotvec(rng.normal(scale=0.1, size=3)).as_matrix(),
                            rng.normal(scale=0.02, size=3)) for _ in range(50)]).astype(np.float32)
    lv, ln = oracle_warp(field, pts, normals, k, w)
    rv, rn = oracle_warp(field, lv, ln, k, w, inverse=True)
    np.testing.assert_allclose(rv, pts, atol=3e-6)
    np.testing.assert_allclose(rn, normals, atol=3e-6)
