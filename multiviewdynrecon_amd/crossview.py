"""Cross-view edge pair production (include/mvdr/crossview.h), mirroring
SparseSurfelFusion::CrossViewEdgeCorrespondence
(core/CrossViewEdgeCorrespondence/CrossViewEdgeCorrespondence.h:94-223):
SetCrossViewMatchingInput, FindCrossViewEdgeMatchedPairs and the three
getters.  Loads libmvdr_solver.so; there is no fallback."""
import ctypes as C

import numpy as np

from . import _abi
from .solver import SolverError, _stream_handle, lib as _solver_lib

_BOUND = False


def lib():
    global _BOUND
    L = _solver_lib()
    if not _BOUND:
        L.mvdr_crossview_default_config.argtypes = [C.POINTER(_abi.CrossViewConfig)]
        L.mvdr_crossview_create.argtypes = [C.POINTER(_abi.CrossViewConfig), C.POINTER(C.c_void_p)]
        L.mvdr_crossview_destroy.argtypes = [C.c_void_p]
        L.mvdr_crossview_find.argtypes = [C.c_void_p, C.POINTER(_abi.CrossViewInput), C.c_void_p]
        L.mvdr_crossview_get_pairs.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]
        _BOUND = True
    return L


def default_config(**overrides):
    c = _abi.CrossViewConfig()
    lib().mvdr_crossview_default_config(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def crossview_input(scene, device=None, corr_map=True, keep=None):
    """mvdr_crossview_input over a scene's per-view vertex / edge / correspondence
    maps: host pointers (device=None) or device copies (torch, device='cuda')."""
    keep = [] if keep is None else keep
    ci = _abi.CrossViewInput()
    ci.num_views, ci.width, ci.height = scene.num_views, scene.width, scene.height

    def ptr(a):
        a = np.ascontiguousarray(a)
        if device is None:
            keep.append(a)
            return a.ctypes.data
        import torch

        t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy()).to(device)
        keep.append(t)
        return t.data_ptr()

    for v, view in enumerate(scene.views):
        ci.intrinsic[v] = _abi.Intrinsic(*[float(x) for x in view.intrinsic])
        m = _abi.Mat34()
        m.rot[:] = [float(x) for x in view.initial_camera_se3[:9]]
        m.trans[:] = [float(x) for x in view.initial_camera_se3[9:12]]
        ci.initial_camera_se3[v] = m
        ci.vertex[v] = ptr(view.maps["vertex_confid"])
        ci.edge_mask[v] = ptr(view.maps["edge_mask"])
        if corr_map and "corr_map" in view.maps:
            ci.corr_map[v] = ptr(view.maps["corr_map"])
    return ci, keep


class CrossViewEdgeCorrespondence:
    def __init__(self, config=None, **overrides):
        cfg = config if config is not None else default_config(**overrides)
        self.config = cfg
        self.h = C.c_void_p()
        rc = lib().mvdr_crossview_create(C.byref(cfg), C.byref(self.h))
        if rc != 0:
            raise SolverError(rc, "mvdr_crossview_create failed")
        self._input = None

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mvdr_crossview_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def SetCrossViewMatchingInput(self, crossview_input):
        self._input = crossview_input

    def FindCrossViewEdgeMatchedPairs(self, stream=None):
        rc = lib().mvdr_crossview_find(self.h, C.byref(self._input), _stream_handle(stream))
        if rc != 0:
            raise SolverError(rc, "mvdr_crossview_find failed")

    def _get(self, which):
        import torch

        p, n = C.c_void_p(), C.c_uint32()
        rc = lib().mvdr_crossview_get_pairs(self.h, _abi.CROSSVIEW_LISTS[which], C.byref(p), C.byref(n))
        if rc not in (0, 3):
            raise SolverError(rc, "mvdr_crossview_get_pairs failed")
        out = np.empty((n.value, 6), np.uint16)
        if n.value:
            torch.cuda.synchronize()
            buf = (C.c_uint16 * (6 * n.value))()
            from .solver import _memcpy_d2h

            _memcpy_d2h(C.addressof(buf), p.value, 12 * n.value)
            out[:] = np.frombuffer(buf, np.uint16).reshape(-1, 6)
        if rc == 3:
            raise SolverError(rc, f"{which}: more pairs than max_pairs ({self.config.max_pairs})")
        return out

    def GetCrossViewMatchingCorrPairs(self):
        return self._get("candidates")

    def GetCrossViewMatchingUniqueCorrPairs(self):
        return self._get("unique")

    def GetCrossViewBackTracingCorrPairs(self):
        return self._get("backtraced")
