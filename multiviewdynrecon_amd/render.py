"""Solver-map rasterisation on the GPU (include/mvdr/render.h), in place of the
reference's OpenGL solver-map pass (Renderer::DrawSolverMaps*,
render/Renderer.h:381-395; render/shaders/SolverMap.vert/.frag).  Loads
libmvdr_solver.so (no fallback)."""
import ctypes as C

import numpy as np

from . import _abi
from .solver import SolverError, _stream_handle, lib as _solver_lib

MAPS = ("reference_vertex", "reference_normal", "warp_vertex", "warp_normal", "index_map", "normalized_rgb")
_BOUND = False


def lib():
    global _BOUND
    L = _solver_lib()
    if not _BOUND:
        L.mvdr_render_solver_maps.argtypes = [C.POINTER(_abi.RenderInput), C.POINTER(_abi.RenderOutput), C.c_void_p]
        _BOUND = True
    return L


def mat34(m):
    out = _abi.Mat34()
    out.rot[:] = [float(x) for x in m[:9]]
    out.trans[:] = [float(x) for x in m[9:12]]
    return out


def render_input(views, width, height, surfels, max_depth=3.0, current_time=0.0, confidence_threshold=0.0,
                 time_threshold=30.0, ptr=None):
    """mvdr_render_input: views = [(intrinsic (fx, fy, cx, cy), world2camera mat34 (12,))],
    surfels = dict of per-view or shared float4 arrays (reference_vertex, reference_normal,
    warp_vertex, warp_normal) and color_time; `ptr` maps an array to its pointer."""
    ri = _abi.RenderInput()
    ri.num_views, ri.width, ri.height = len(views), width, height
    ri.num_surfels = int(surfels["color_time"].shape[0])
    for v, (K, w2c) in enumerate(views):
        ri.intrinsic[v] = _abi.Intrinsic(*[float(x) for x in K])
        ri.world2camera[v] = mat34(w2c)
        for name in ("reference_vertex", "reference_normal", "warp_vertex", "warp_normal"):
            a = surfels[name]
            getattr(ri, name)[v] = ptr(a[v] if isinstance(a, (list, tuple)) else a)
    ri.color_time = ptr(surfels["color_time"])
    ri.max_depth, ri.current_time = max_depth, current_time
    ri.confidence_threshold, ri.time_threshold = confidence_threshold, time_threshold
    return ri


def RenderSolverMaps(views, width, height, surfels, stream=None, **kw):
    """Device maps (torch tensors) of every view: dict map name -> list per view."""
    import torch

    keep = []

    def ptr(a):
        t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a)).cuda()
        keep.append(t)
        return t.data_ptr()

    ri = render_input(views, width, height, surfels, ptr=ptr, **kw)
    ro = _abi.RenderOutput()
    maps = {name: [] for name in MAPS}
    for v in range(len(views)):
        for name in MAPS:
            t = torch.empty((height, width) if name == "index_map" else (height, width, 4),
                            dtype=torch.int32 if name == "index_map" else torch.float32, device="cuda")
            maps[name].append(t)
            getattr(ro, name)[v] = t.data_ptr()
        z = torch.empty((height, width), dtype=torch.int64, device="cuda")
        keep.append(z)
        ro.zbuffer[v] = z.data_ptr()
    rc = lib().mvdr_render_solver_maps(C.byref(ri), C.byref(ro), _stream_handle(stream))
    if rc != 0:
        raise SolverError(rc, "mvdr_render_solver_maps failed")
    torch.cuda.synchronize()
    return maps
