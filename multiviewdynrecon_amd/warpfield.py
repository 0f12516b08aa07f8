"""Warp-field maintenance on the GPU (include/mvdr/warpfield.h), named after the
reference's entry points: WarpField::BuildNodesGraph (WarpField.cu:62-71),
CanonicalNodesSkinner's brute-force skinning (CanonicalNodesSkinner.cu:356-437),
SurfelNodeDeformer::ForwardWarpSurfelsAndNodes and the inverse warp
(SurfelNodeDeformer.cu:9-124), NonRigidSolver::CorrectLargeErrorNode
(NonRigidSolver.cu:365-426).  Arguments are torch tensors on the GPU; the
calls enqueue on the current stream.  Loads libmvdr_solver.so (no fallback)."""
import ctypes as C

from . import _abi
from .render import mat34
from .solver import SolverError, _stream_handle, lib as _solver_lib

NODE_RADIUS = 0.025  # GlobalConfigs.h:85
_BOUND = False


def lib():
    global _BOUND
    L = _solver_lib()
    if not _BOUND:
        vp, u32 = C.c_void_p, C.c_uint32
        L.mvdr_warp_build_node_graph.argtypes = [vp, u32, vp, vp]
        L.mvdr_warp_skin.argtypes = [vp, u32, vp, u32, C.c_float, vp, vp, vp, vp, vp]
        L.mvdr_warp_forward.argtypes = [vp, vp, vp, vp, vp, u32, vp, vp, vp, u32, vp, vp, vp, vp]
        L.mvdr_warp_inverse.argtypes = [vp, u32, vp, vp, vp, vp, u32, vp, vp, vp]
        L.mvdr_warp_correct_nodes.argtypes = [C.POINTER(_abi.NodeCorrection), vp]
        _BOUND = True
    return L


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _check(rc, what):
    if rc != 0:
        raise SolverError(rc, f"{what} failed")


def BuildNodesGraph(nodes, stream=None):
    """ushort2[8 N] node graph (as int16 pairs viewed from uint16) of float4 nodes[N]."""
    import torch

    n = nodes.shape[0]
    graph = torch.empty((n * 8, 2), dtype=torch.int16, device=nodes.device)
    _check(lib().mvdr_warp_build_node_graph(_p(nodes), n, _p(graph), _stream_handle(stream)), "node graph")
    return graph


def PerformSkinning(vertices, nodes, radius=NODE_RADIUS, stream=None):
    """(vertex knn, vertex weights, node knn, node weights): 4-NN skinning."""
    import torch

    dev = nodes.device
    S, N = vertices.shape[0], nodes.shape[0]
    vk = torch.empty((S, 4), dtype=torch.int16, device=dev)
    vw = torch.empty((S, 4), dtype=torch.float32, device=dev)
    nk = torch.empty((N, 4), dtype=torch.int16, device=dev)
    nw = torch.empty((N, 4), dtype=torch.float32, device=dev)
    _check(lib().mvdr_warp_skin(_p(vertices), S, _p(nodes), N, radius, _p(vk), _p(vw), _p(nk), _p(nw),
                                _stream_handle(stream)), "skinning")
    return vk, vw, nk, nw


def ForwardWarpSurfelsAndNodes(node_se3, vertex, normal, vknn, vw, nodes, nknn, nw, stream=None):
    """(live vertex, live normal, live nodes) under the warp field node_se3[N]."""
    import torch

    lv, ln, lnodes = torch.empty_like(vertex), torch.empty_like(normal), torch.empty_like(nodes)
    _check(lib().mvdr_warp_forward(_p(node_se3), _p(vertex), _p(normal), _p(vknn), _p(vw), vertex.shape[0],
                                   _p(nodes), _p(nknn), _p(nw), nodes.shape[0], _p(lv), _p(ln), _p(lnodes),
                                   _stream_handle(stream)), "forward warp")
    return lv, ln, lnodes


def InverseWarpSurfels(node_se3, live_vertex, live_normal, vknn, vw, stream=None):
    """(reference vertex, reference normal): the inverse of the forward warp."""
    import torch

    rv, rn = torch.empty_like(live_vertex), torch.empty_like(live_normal)
    _check(lib().mvdr_warp_inverse(_p(node_se3), node_se3.shape[0], _p(live_vertex), _p(live_normal), _p(vknn),
                                   _p(vw), live_vertex.shape[0], _p(rv), _p(rn), _stream_handle(stream)),
           "inverse warp")
    return rv, rn


def correction_input(views, width, height, nodes, surfels, correction, ptr):
    """mvdr_node_correction from views (dicts: intrinsic (4,), initial_camera_se3 /
    world2camera (12,), previous_vertex (H, W, 4), guided_se3 (H, W, 12),
    guided_valid (H, W) uint8), nodes (dict: error, live, se3, knn, weight) and
    surfels (dict: vertex, normal, knn, weight); ptr maps an array to its
    address (host or device)."""
    ci = _abi.NodeCorrection()
    ci.num_views, ci.width, ci.height = len(views), width, height
    for v, d in enumerate(views):
        ci.intrinsic[v] = _abi.Intrinsic(*[float(x) for x in d["intrinsic"]])
        ci.initial_camera_se3[v] = mat34(d["initial_camera_se3"])
        ci.world2camera[v] = mat34(d["world2camera"])
        ci.previous_vertex[v] = ptr(d["previous_vertex"])
        ci.guided_se3[v] = ptr(d["guided_se3"])
        ci.guided_valid[v] = ptr(d["guided_valid"])
    ci.num_nodes = len(nodes["error"])
    ci.node_error, ci.live_nodes, ci.node_se3 = ptr(nodes["error"]), ptr(nodes["live"]), ptr(nodes["se3"])
    ci.node_knn, ci.node_knn_weight = ptr(nodes["knn"]), ptr(nodes["weight"])
    ci.num_surfels = len(surfels["vertex"]) if surfels else 0
    if surfels:
        ci.live_vertex, ci.live_normal = ptr(surfels["vertex"]), ptr(surfels["normal"])
        ci.surfel_knn, ci.surfel_knn_weight = ptr(surfels["knn"]), ptr(surfels["weight"])
    ci.correction = ptr(correction)
    return ci


def CorrectLargeErrorNode(views, width, height, nodes, surfels, stream=None):
    """Correspondence-guided correction of the nodes whose error exceeds 0.2, in
    place on nodes["live"], nodes["se3"], surfels["vertex"] / ["normal"] (torch
    tensors on the GPU); returns the per-node correction dual quaternions."""
    import torch

    corr = torch.empty((len(nodes["error"]), 8), dtype=torch.float32, device=nodes["live"].device)
    ci = correction_input(views, width, height, nodes, surfels, corr, _p)
    _check(lib().mvdr_warp_correct_nodes(C.byref(ci), _stream_handle(stream)), "node correction")
    return corr
