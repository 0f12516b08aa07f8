"""Rigid multi-view point-to-plane ICP on the GPU (include/mvdr/rigid.h),
mirroring SparseSurfelFusion::RigidSolver's pose refinement
(SolveRigidAlignment with solver maps, GetWorld2Camera; RigidSolver.h:467-489).
Loads libmvdr_solver.so (no fallback)."""
import ctypes as C

import numpy as np

from . import _abi
from .render import mat34
from .solver import SolverError, _stream_handle, lib as _solver_lib

_BOUND = False


def lib():
    global _BOUND
    L = _solver_lib()
    if not _BOUND:
        L.mvdr_rigid_create.argtypes = [C.POINTER(_abi.RigidConfig), C.POINTER(C.c_void_p)]
        L.mvdr_rigid_destroy.argtypes = [C.c_void_p]
        L.mvdr_rigid_solve.argtypes = [C.c_void_p, C.POINTER(_abi.RigidInput), C.c_uint32, C.c_void_p]
        L.mvdr_rigid_get_pose.argtypes = [C.c_void_p, C.POINTER(_abi.Mat34), C.POINTER(_abi.Mat34)]
        _BOUND = True
    return L


def rigid_input(views, width, height, ptr):
    """views: list of dicts with intrinsic (4,), reference_vertex / reference_normal /
    conversion_vertex / conversion_normal (H, W, 4) arrays, world2camera (12,) and
    relative_pose (12,)."""
    ri = _abi.RigidInput()
    ri.num_views, ri.width, ri.height = len(views), width, height
    for v, d in enumerate(views):
        ri.intrinsic[v] = _abi.Intrinsic(*[float(x) for x in d["intrinsic"]])
        for name in ("reference_vertex", "reference_normal", "conversion_vertex", "conversion_normal"):
            getattr(ri, name)[v] = ptr(d[name])
        ri.world2camera[v] = mat34(d["world2camera"])
        ri.relative_pose[v] = mat34(d.get("relative_pose", np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)))
    return ri


def mat34_array(m):
    return np.array(list(m.rot) + list(m.trans), np.float32)


class RigidSolver:
    def __init__(self, max_views=3, max_width=640, max_height=480, reference_reduction=0):
        cfg = _abi.RigidConfig(max_views, max_width, max_height, reference_reduction)
        self.h = C.c_void_p()
        rc = lib().mvdr_rigid_create(C.byref(cfg), C.byref(self.h))
        if rc != 0:
            raise SolverError(rc, "mvdr_rigid_create failed")
        self._keep = []

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mvdr_rigid_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def SolveRigidAlignment(self, views, width, height, iterations=3, stream=None):
        import torch

        keep = []

        def ptr(a):
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a)).cuda()
            keep.append(t)
            return t.data_ptr()

        ri = rigid_input(views, width, height, ptr)
        self._keep = keep
        self._views = len(views)
        rc = lib().mvdr_rigid_solve(self.h, C.byref(ri), iterations, _stream_handle(stream))
        if rc != 0:
            raise SolverError(rc, "mvdr_rigid_solve failed")

    def GetWorld2Camera(self):
        """(per-view world2camera (V, 12), canonical-to-live (12,)) as mvdr_mat34 floats."""
        poses = (_abi.Mat34 * 3)()
        avg = _abi.Mat34()
        rc = lib().mvdr_rigid_get_pose(self.h, poses, C.byref(avg))
        if rc != 0:
            raise SolverError(rc, "mvdr_rigid_get_pose failed")
        return np.stack([mat34_array(poses[v]) for v in range(self._views)]), mat34_array(avg)
