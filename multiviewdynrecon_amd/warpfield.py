"""Warp-field maintenance on the GPU (include/mvdr/warpfield.h), named after the
reference's entry points: WarpField::BuildNodesGraph (WarpField.cu:62-71),
CanonicalNodesSkinner's brute-force skinning (CanonicalNodesSkinner.cu:356-437),
SurfelNodeDeformer::ForwardWarpSurfelsAndNodes and the inverse warp
(SurfelNodeDeformer.cu:9-124).  Arguments are torch tensors on the GPU; the
calls enqueue on the current stream.  Loads libmvdr_solver.so (no fallback)."""
import ctypes as C

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
